"""GPU: the mixed-precision forms (FP32 or FP16 element blocks and an FP32
multigrid cycle inside an FP64 flexible outer loop) against the FP64 oracle
and the FP64 GPU path.
relative compliance error <= 1e-6 against the FP64 oracle."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("precision", [1, 2])
def test_mixed_loop_vs_oracle(O, precision):
    """precision 1: FP32 preconditioner blocks; 2: FP16 smoother blocks."""
    from uft.capi import cantilever
    args = dict(nx=10, ny=5, nz=5, jitter=0.15, seed=13)
    prm = dict(Emax=1.0, volfrac=0.3, rmin=0.3, rtol_el=1e-12, rtol_f=1e-12)
    print(f"[case] mixed precision={precision} hex8 10x5x5 vs oracle", flush=True)
    ctx = cantilever(8, **args, precision=precision, **prm)
    ref = O.cantilever(8, **args, **prm)
    ctx.init_design()
    ref.init_design()
    worst = 0.0
    for it in range(1, 11):
        fx, gx, ch = ctx.design_iteration(it)
        rfx, rgx, rch, _ = ref.design_iteration(it)
        worst = max(worst, abs(fx - rfx) / abs(rfx))
        assert abs(fx - rfx) <= 1e-6 * abs(rfx)
    assert np.max(np.abs(ctx.get("X") - ref.vec("X"))) <= 1e-6
    assert ctx.stats()["relres_el"] <= 1e-12
    print(f"worst relative compliance error {worst:.3e}")


@pytest.mark.parametrize("cycle", [0, 1, 2])
def test_cycle_types_same_solution(O, cycle):
    """V, W and K cycles are preconditioners only: the FP64 solution agrees."""
    from uft.capi import cantilever
    args = dict(nx=12, ny=6, nz=6, jitter=0.1, seed=3)
    prm = dict(Emax=1.0, volfrac=0.3, rmin=0.25, rtol_el=1e-12, rtol_f=1e-12)
    print(f"[case] cycle {cycle}", flush=True)
    ctx = cantilever(8, **args, cycle_type=cycle, **prm)
    ref = O.cantilever(8, **args, **prm)
    rng = np.random.default_rng(1)
    x = rng.uniform(0.01, 1.0, ref.ne)
    ctx.set("XPhys", x)
    fx, gx = ctx.fem_obj_sens()
    rfx = ref.fem_obj_sens(x)[0]
    assert abs(fx - rfx) <= 1e-10 * abs(rfx)
    assert ctx.stats()["relres_el"] <= 1e-12
