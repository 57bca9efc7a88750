"""GPU: the von Mises stress kernel and the MMA options (robust asymptotes,
constraint modification, user penalisation, restart state, KKT residual)
through the C-ABI, against the oracle."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

PRM = dict(Emax=1.0, volfrac=0.3, rmin=0.3, rtol_el=1e-12, rtol_f=1e-12)


@pytest.mark.parametrize("npe,dims", [(8, (8, 4, 4)), (4, (5, 3, 3))])
def test_stress_vs_oracle(O, npe, dims):
    from uft.capi import cantilever
    print(f"[case] stress npe={npe}", flush=True)
    ctx = cantilever(npe, *dims, jitter=0.15, seed=6, Emax=2.0)
    ref = O.cantilever(npe, *dims, jitter=0.15, seed=6, Emax=2.0)
    rng = np.random.default_rng(1)
    x = rng.uniform(0.05, 1.0, ref.ne)
    ctx.set("XPhys", x)
    ctx.fem_obj_sens()
    u = ctx.get("U")
    got = ctx.stress()
    want = ref.stress(x, u)
    assert np.allclose(got, want, rtol=1e-12, atol=1e-14 * np.abs(want).max())


@pytest.mark.parametrize("robust,conmod", [(0, 1), (1, 0), (1, 1)])
def test_mma_options_loop_vs_oracle(O, robust, conmod):
    from uft.capi import cantilever
    print(f"[case] mma options robust={robust} conmod={conmod}", flush=True)
    args = dict(nx=8, ny=4, nz=4, jitter=0.1, seed=17)
    ctx = cantilever(8, **args, **PRM)
    ref = O.cantilever(8, **args, **PRM)
    ctx.mma_settings(0.4, 0.65, 1.15, robust, conmod)
    ref.mma_settings(0.4, 0.65, 1.15, robust, conmod)
    ctx.mma_set_penalization([0.0], [500.0], [0.0])
    ref.mma_set_acd([0.0], [500.0], [0.0])
    ctx.init_design()
    ref.init_design()
    for it in range(1, 7):
        fx, gx, ch = ctx.design_iteration(it)
        rfx, rgx, rch, _ = ref.design_iteration(it)
        assert abs(fx - rfx) <= 1e-10 * abs(rfx) and abs(gx[0] - rgx[0]) <= 1e-10
    assert np.max(np.abs(ctx.get("X") - ref.vec("X"))) <= 1e-8
    n2, ni = ctx.mma_kkt(gx)
    rn2, rni = ref.mma_kkt(ref.vec("X"), ref.vec("dfdx"), rgx, ref.vec("dgdx"), ref.vec("xmin"), ref.vec("xmax"))
    assert abs(n2 - rn2) <= 1e-8 * rn2 and abs(ni - rni) <= 1e-8 * rni


@pytest.mark.parametrize("carry", [0, 1])
def test_restart_in_fresh_context(carry):
    """Export the state after 3 iterations; a new context loaded with the
    design vectors, the filter warm start and the MMA state reproduces
    iterations 4-6 of the uninterrupted run.  Without solver state carried
    from one hierarchy refresh to the next (carry = 0: tentative prolongator,
    no lagged coarsest inverse) bit for bit; with the defaults (carry = 1:
    the prolongator's smoothing factor comes from the previous refresh's
    eigenvalue estimate, which a fresh context does not have) to the solver
    tolerance."""
    from uft.capi import cantilever
    print(f"[case] restart carry={carry}", flush=True)
    args = dict(nx=8, ny=4, nz=4, jitter=0.1, seed=19)
    prm = dict(PRM) if carry else dict(PRM, prolong_smooth=0, coarse_lag=0)
    a = cantilever(8, **args, **prm)
    a.init_design()
    for it in (1, 2, 3):
        a.design_iteration(it)
    st = a.mma_state()
    saved = {k: a.get(k) for k in ("X", "XFiled", "XPhys", "Xold", "RX")}
    first = [a.design_iteration(it) for it in (4, 5, 6)]
    b = cantilever(8, **args, **prm)
    for k, v in saved.items():
        b.set(k, v)
    b.mma_set_state(*st)
    again = [b.design_iteration(it) for it in (4, 5, 6)]
    if not carry:
        assert first == again
        assert np.array_equal(a.get("X"), b.get("X"))
        return
    for (fa, ga, ca), (fb, gb, cb) in zip(first, again):
        assert abs(fa - fb) <= 1e-10 * abs(fa) and abs(ga[0] - gb[0]) <= 1e-10 and abs(ca - cb) <= 1e-8
    assert np.abs(a.get("X") - b.get("X")).max() <= 1e-8


def test_restart_warns_below_k3():
    from uft.capi import cantilever
    print("[case] restart k<3", flush=True)
    a = cantilever(8, 4, 2, 2, **PRM)
    k, xo1, xo2, U, L = a.mma_state()
    a.mma_set_state(1, xo1, xo2, U, L)
    assert b"legal restart point" in a.L.uft_last_error(a.h)
    assert a.L.uft_mma_set_robust_asymptotes(a.h, 5) == 1
