"""GPU: the Krylov iterations replayed from a captured hipGraph give the
same results, bit for bit, as launching them kernel by kernel: several design
iterations (each with its own hierarchy refresh and re-capture, and the
program's own synchronising host work between the replays), every iteration
compared, for the FP64 and mixed single-design solves and a design batch; and
the device timers recorded inside the graph still count the fine applies."""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

PRM = dict(Emax=1.0, volfrac=0.3, rmin=0.3, rtol_el=1e-10, rtol_f=1e-10)


@pytest.mark.parametrize("precision", [0, 2])
def test_graph_replays_match_eager(precision):
    from uft.capi import cantilever
    print(f"[case] graph vs eager precision={precision}", flush=True)
    a = cantilever(8, 12, 6, 6, jitter=0.15, seed=3, graph=1, precision=precision, **PRM)
    b = cantilever(8, 12, 6, 6, jitter=0.15, seed=3, graph=0, precision=precision, **PRM)
    a.init_design()
    b.init_design()
    for it in range(1, 5):
        ra, rb = a.design_iteration(it), b.design_iteration(it)
        assert ra == rb, it
        sa, sb = a.stats(), b.stats()
        assert sa["it_el"] == sb["it_el"] and sa["it_el"] > 2
        assert np.array_equal(a.get("U"), b.get("U"))
    assert np.array_equal(a.get("X"), b.get("X"))


def test_graph_batch_matches_eager():
    from uft.capi import batch_cantilever
    print("[case] graph vs eager, design batch", flush=True)
    ds = [dict(nx=10, ny=5, nz=5, volfrac=0.3, seed=1), dict(nx=14, ny=7, nz=6, volfrac=0.25, seed=2)]
    prm = {k: v for k, v in PRM.items() if k not in ("volfrac", "rmin")}
    a = batch_cantilever(ds, graph=1, precision=1, **prm)
    b = batch_cantilever(ds, graph=0, precision=1, **prm)
    a.init_design()
    b.init_design()
    for it in range(1, 4):
        assert a.design_iteration(it) == b.design_iteration(it)
        ra, rb = a.batch_results(), b.batch_results()
        for k in ("fx", "gx", "ch", "it_el"):
            assert np.array_equal(ra[k], rb[k]), (it, k)
    assert np.array_equal(a.get("X"), b.get("X"))


def test_timers_switch_to_eager_launches():
    """With the device timers on, a graph context launches kernel by kernel
    (timers inside captured graphs are not readable): the timers count the
    fine applies and the results equal the graph replays'."""
    from uft.capi import cantilever
    print("[case] device timers with graph=1", flush=True)
    a = cantilever(8, 12, 6, 6, jitter=0.15, seed=3, graph=1, **PRM)
    b = cantilever(8, 12, 6, 6, jitter=0.15, seed=3, graph=1, **PRM)
    a.init_design()
    b.init_design()
    L = a.L
    L.uft_profile_enable(a.h, 1)
    ra, rb = a.design_iteration(1), b.design_iteration(1)
    assert ra == rb
    it = a.stats()["it_el"]
    cnt, ms = ctypes.c_int64(), ctypes.c_double()
    assert L.uft_profile_read(a.h, 0, ctypes.byref(cnt), ctypes.byref(ms)) == 0
    L.uft_profile_enable(a.h, 0)
    # one FP64 apply per Krylov iteration plus the cycle's fine applies and
    # the true residual, all timed
    assert cnt.value > it and ms.value > 0.0


def test_autotune_smoother():
    """The smoother-schedule autotuner keeps the fastest converged schedule,
    and the design iteration it leaves gives the same compliance (to the
    solver tolerance) as the default schedule."""
    from uft.capi import cantilever, batch_cantilever
    print("[case] smoother autotuner", flush=True)
    a = cantilever(8, 12, 6, 6, jitter=0.15, seed=3, precision=2, **PRM)
    b = cantilever(8, 12, 6, 6, jitter=0.15, seed=3, precision=2, **PRM)
    a.init_design()
    b.init_design()
    d, ct, ms = a.autotune_smoother()
    print("picked degree", d, "cycle", ct, "ms", ms[:, :3])
    assert 1 <= d <= 3 and 0 <= ct <= 2
    assert ms[d - 1, ct] == ms[:, :3].min() and ms[d - 1, ct] < 1e300
    p = a.get_params()
    assert (p.cheb_degree, p.cycle_type) == (d, ct)
    fa, fb = a.design_iteration(1)[0], b.design_iteration(1)[0]
    assert abs(fa - fb) <= 1e-8 * abs(fb)
    ds = [dict(nx=10, ny=5, nz=5, volfrac=0.3, seed=1), dict(nx=14, ny=7, nz=6, volfrac=0.25, seed=2)]
    prm = {k: v for k, v in PRM.items() if k not in ("volfrac", "rmin")}
    c = batch_cantilever(ds, precision=1, **prm)
    c.init_design()
    d, ct, ms = c.autotune_smoother()
    assert 0 <= ct <= 2 and ms[d - 1, ct] == ms[:, :3].min()
