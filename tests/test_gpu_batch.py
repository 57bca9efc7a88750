"""GPU: independent designs of uneven size in one context (uft_batch_set).
Every design of a batch is checked against the oracle running that design
alone (compliance, volume constraint and design change per iteration, the
density field after the loop), the batch against single-design contexts at
a larger size, and the block-diagonal coarsest inverse against the coarsest
operator."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

TIGHT = dict(rtol_el=1e-12, rtol_f=1e-12, Emax=1.0)
SMALL = [dict(nx=8, ny=4, nz=4, volfrac=0.30, seed=5),
         dict(nx=10, ny=5, nz=4, volfrac=0.25, seed=6),
         dict(nx=6, ny=3, nz=3, volfrac=0.35, seed=7)]


def _oracle(O, D, **prm):
    return O.cantilever(8, D["nx"], D["ny"], D["nz"], jitter=0.15, seed=D["seed"], volfrac=D["volfrac"],
                        rmin=1.5 * 2.0 / D["nx"], **prm)


@pytest.mark.parametrize("precision", [0, 2])
def test_batch_vs_oracle(O, precision):
    from uft.capi import batch_cantilever, batch_mesh
    print(f"[case] batch of {len(SMALL)} vs oracle, precision={precision}", flush=True)
    ctx = batch_cantilever(SMALL, jitter=0.15, precision=precision, **TIGHT)
    refs = [_oracle(O, D, **TIGHT) for D in SMALL]
    eoff = batch_mesh(SMALL)[5]
    ctx.init_design()
    for r in refs:
        r.init_design()
    # FP64: the drop-in bar; mixed forms: 1e-6 relative compliance
    tol = 1e-10 if precision == 0 else 1e-6
    for it in range(1, 7):
        fx, gx, ch = ctx.design_iteration(it)
        res = ctx.batch_results()
        want = [r.design_iteration(it) for r in refs]
        for d, (rfx, rgx, rch, _) in enumerate(want):
            assert abs(res["fx"][d] - rfx) <= tol * abs(rfx), (it, d)
            assert abs(res["gx"][d] - rgx[0]) <= max(tol, 1e-10), (it, d)
            assert abs(res["ch"][d] - rch) <= 1e-6, (it, d)
            assert res["relres_el"][d] <= 1e-12
        assert abs(fx - sum(w[0] for w in want)) <= tol * abs(fx)
        assert ch == pytest.approx(max(res["ch"]), abs=0)
    X = ctx.get("X")
    for d, r in enumerate(refs):
        err = np.abs(X[eoff[d]:eoff[d + 1]] - r.vec("X")).max()
        print("design", d, "density difference", err)
        assert err <= (1e-8 if precision == 0 else 1e-6)


def test_batch_vs_single_contexts():
    """A batch of three uneven designs against three contexts holding one
    design each (both GPU, rtol 1e-10): the same compliances per iteration."""
    from uft.capi import batch_cantilever, cantilever
    designs = [dict(nx=16, ny=8, nz=8, volfrac=0.30, seed=21),
               dict(nx=22, ny=11, nz=10, volfrac=0.22, seed=22),
               dict(nx=12, ny=6, nz=6, volfrac=0.38, seed=23)]
    prm = dict(rtol_el=1e-10, rtol_f=1e-10, Emax=1.0, precision=2)
    print("[case] batch vs single contexts", flush=True)
    ctx = batch_cantilever(designs, jitter=0.15, **prm)
    singles = [cantilever(8, D["nx"], D["ny"], D["nz"], jitter=0.15, seed=D["seed"], volfrac=D["volfrac"],
                          rmin=1.5 * 2.0 / D["nx"], **prm) for D in designs]
    ctx.init_design()
    for c in singles:
        c.init_design()
    for it in range(1, 4):
        ctx.design_iteration(it)
        res = ctx.batch_results()
        for d, c in enumerate(singles):
            fx, gx, ch = c.design_iteration(it)
            assert abs(res["fx"][d] - fx) <= 1e-8 * abs(fx), (it, d)
            assert abs(res["gx"][d] - gx[0]) <= 1e-9, (it, d)
        print("iteration", it, "Krylov iterations per design", list(res["it_el"]), "filter", list(res["it_f"]))


def test_coarsest_inverse_blocks():
    """The batched Gauss-Jordan inverse of every design's coarsest block
    times the coarsest operator is the identity (blocks off the diagonal are
    zero: designs do not couple)."""
    from uft.capi import batch_cantilever
    print("[case] batch coarsest inverse", flush=True)
    ctx = batch_cantilever(SMALL, jitter=0.15, **TIGHT)
    ctx.fem_obj_sens()
    n, _ = ctx.solver_info()
    lc = len(n) - 1
    assert lc >= 2
    rp, ci, val = ctx.amg_operator(lc)
    nc = n[lc]
    A = np.zeros((nc, nc))
    for I in range(nc // 6):
        for k in range(rp[I], rp[I + 1]):
            A[6 * I:6 * I + 6, 6 * ci[k]:6 * ci[k] + 6] = val[k]
    inv = ctx.amg_coarse_inverse()
    assert np.allclose(inv, inv.T, atol=0)
    err = np.abs(A @ inv - np.eye(nc)).max()
    print("||A Ainv - I||_max", err)
    assert err <= 1e-9


def test_batch_tet4_with_projection(O):
    """Tet4 designs of uneven size with the Heaviside projection on: every
    design against the oracle running it alone."""
    from uft.capi import batch_cantilever, batch_mesh
    print("[case] batch tet4 + projection", flush=True)
    ds = [dict(nx=5, ny=3, nz=3, volfrac=0.30, seed=31), dict(nx=4, ny=2, nz=2, volfrac=0.35, seed=32)]
    prm = dict(TIGHT, proj=1, beta0=2.0)
    ctx = batch_cantilever(ds, npe=4, jitter=0.15, **prm)
    refs = [O.cantilever(4, D["nx"], D["ny"], D["nz"], jitter=0.15, seed=D["seed"], volfrac=D["volfrac"],
                         rmin=1.5 * 2.0 / D["nx"], **prm) for D in ds]
    eoff = batch_mesh(ds, npe=4)[5]
    ctx.init_design()
    for r in refs:
        r.init_design()
    for it in range(1, 5):
        ctx.design_iteration(it)
        res = ctx.batch_results()
        for d, r in enumerate(refs):
            rfx, rgx, rch, _ = r.design_iteration(it)
            assert abs(res["fx"][d] - rfx) <= 1e-10 * abs(rfx), (it, d)
            assert abs(res["gx"][d] - rgx[0]) <= 1e-10, (it, d)
    X = ctx.get("X")
    for d, r in enumerate(refs):
        assert np.abs(X[eoff[d]:eoff[d + 1]] - r.vec("X")).max() <= 1e-8
