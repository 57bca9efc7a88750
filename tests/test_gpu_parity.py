"""GPU parity: the HIP path (through the C-ABI) against the CPU oracle on the
same seeded meshes.  Tolerances follow the FP64 drop-in bar: element
matrices and operator applies to ~1e-13 relative, compliance and volume to
1e-10 relative per design iteration, densities to 1e-8 after 10 MMA
iterations (both sides solved at rtol 1e-12).
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

TIGHT = dict(rtol_el=1e-12, rtol_f=1e-12)


def _pair(O, npe, nx, ny, nz, jitter=0.15, seed=11, load_kind=0, **prm):
    from uft.capi import cantilever
    print(f"[case] npe={npe} mesh={nx}x{ny}x{nz} jitter={jitter} seed={seed} prm={prm}", flush=True)
    ctx = cantilever(npe, nx, ny, nz, jitter=jitter, seed=seed, load_kind=load_kind, **prm)
    ref = O.cantilever(npe, nx, ny, nz, jitter=jitter, seed=seed, load_kind=load_kind, **prm)
    return ctx, ref


def _rel(a, b):
    a, b = np.asarray(a, float), np.asarray(b, float)
    return np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300)


@pytest.mark.parametrize("npe", [8, 4])
def test_element_matrices(O, npe):
    ctx, ref = _pair(O, npe, 6, 3, 3, rmin=0.5)
    xyz, elem = ref.xyz, ref.elem
    for e in [0, 1, len(elem) // 2, len(elem) - 1]:
        X = xyz[elem[e]]
        KE = ctx.element_ke(e)
        KEo = (O.hex8_ke if npe == 8 else O.tet4_ke)(X)[0]
        assert _rel(KE, KEo) < 1e-13, e
        KF, T = ctx.element_kf(e)
        KFo, To = (O.hex8_kehel if npe == 8 else O.tet4_kehel)(X, 0.5)
        assert _rel(KF, KFo) < 1e-13 and _rel(T, To) < 1e-13, e
    vol = ctx.get("Vol")
    assert _rel(vol, ref.vec("Vol")) < 1e-14
    assert abs(ctx.total_volume - ref.total_volume) < 1e-13 * ref.total_volume


@pytest.mark.parametrize("npe", [8, 4])
def test_apply_K(O, npe):
    ctx, ref = _pair(O, npe, 7, 4, 3, Emax=1.0)
    rng = np.random.default_rng(3)
    x = rng.uniform(0.01, 1.0, ref.ne)
    u = rng.standard_normal(3 * ref.nn)
    ctx.set("XPhys", x)
    y = ctx.apply_K(u)
    yo = ref.apply_K(x, u)
    assert _rel(y, yo) < 1e-13
    # fixed DOFs: unit diagonal, eliminated columns
    fixed = O.box_bcs(7, 4, 3)[0]
    assert np.array_equal(y[fixed], u[fixed])


@pytest.mark.parametrize("npe", [8, 4])
def test_obj_sens(O, npe):
    ctx, ref = _pair(O, npe, 8, 4, 4, Emax=1.0, volfrac=0.3, **TIGHT)
    rng = np.random.default_rng(5)
    x = rng.uniform(0.05, 1.0, ref.ne)
    ctx.set("XPhys", x)
    fx, gx = ctx.fem_obj_sens()
    rfx, rgx, rdf, rdg, _ = ref.fem_obj_sens(x)
    st = ctx.stats()
    assert st["relres_el"] <= 1e-12
    assert abs(fx - rfx) <= 1e-10 * abs(rfx)
    assert abs(gx[0] - rgx[0]) <= 1e-12
    assert _rel(ctx.get("dfdx"), rdf) < 1e-9
    assert _rel(ctx.get("dgdx0"), rdg[0]) < 1e-14
    assert _rel(ctx.get("U"), ref.vec("U")) < 1e-9


@pytest.mark.parametrize("npe,assembled", [(8, 1), (4, 1), (8, 0), (4, 0)])
def test_filter(O, npe, assembled):
    """The Helmholtz filter, its operator assembled on the node pattern
    (fine_assembled = 1) or applied element by element, against the oracle."""
    from uft.capi import cantilever
    ctx, ref = _pair(O, npe, 8, 4, 4, rmin=0.375, **TIGHT)
    ctx = cantilever(npe, 8, 4, 4, jitter=0.15, seed=11, load_kind=0, rmin=0.375, fine_assembled=assembled, **TIGHT)
    rng = np.random.default_rng(9)
    x = rng.uniform(0.0, 1.0, ref.ne)
    ctx.set("X", x)
    ctx.filter("X", "XFiled")
    xf, _ = ref.filter(x)
    assert _rel(ctx.get("XFiled"), xf) < 1e-10
    # the same Jacobi-CG: the same number of iterations (the GPU tests its
    # convergence on the device and queues iterations in rounds)
    print("filter CG iterations", ctx.stats()["it_f"], ref.diag()["it_f"])
    assert ctx.stats()["it_f"] == ref.diag()["it_f"] > 0
    df = rng.standard_normal(ref.ne)
    ctx.set("dfdx", df)
    ctx.set("dgdx0", np.abs(df))
    ctx.filter_sens()
    rdf, rdg, _ = ref.filter_sens(df, np.abs(df)[None, :])
    assert _rel(ctx.get("dfdx"), rdf) < 1e-10
    assert _rel(ctx.get("dgdx0"), rdg[0]) < 1e-10
    print("filter CG iterations, last sensitivity", ctx.stats()["it_f"], ref.diag()["it_f"])
    assert ctx.stats()["it_f"] == ref.diag()["it_f"] > 0


def test_filter_iteration_limit(O):
    """At its iteration limit the filter CG reports UFT_ERR_NOCONV after
    exactly maxit_f iterations (the device-side tests stop the queued ones),
    and its iterate is the oracle's after as many."""
    from uft.capi import UFTError
    ctx, ref = _pair(O, 8, 8, 4, 4, rmin=0.375, rtol_el=1e-12, rtol_f=1e-14, maxit_f=3)
    x = np.random.default_rng(9).uniform(0.0, 1.0, ref.ne)
    ctx.set("X", x)
    with pytest.raises(UFTError) as e:
        ctx.filter("X", "XFiled")
    assert e.value.code == 3  # UFT_ERR_NOCONV
    assert ctx.stats()["it_f"] == 3
    xf, _ = ref.filter(x)
    assert ref.diag()["it_f"] == 3
    assert _rel(ctx.get("XFiled"), xf) < 1e-10


def _loop(ctx, ref, iters):
    ctx.init_design()
    ref.init_design()
    rows = []
    for it in range(1, iters + 1):
        fx, gx, ch = ctx.design_iteration(it)
        rfx, rgx, rch, _ = ref.design_iteration(it)
        rows.append((fx, rfx, gx[0], rgx[0], ch, rch))
    return rows


@pytest.mark.parametrize("npe,mesh", [(8, (10, 5, 5)), (4, (6, 3, 3))])
def test_design_loop_10(O, npe, mesh):
    nx, ny, nz = mesh
    prm = dict(Emax=1.0, volfrac=0.3 if npe == 8 else 0.25, rmin=1.5 * 2.0 / nx, **TIGHT)
    ctx, ref = _pair(O, npe, nx, ny, nz, load_kind=0 if npe == 8 else 1, **prm)
    rows = _loop(ctx, ref, 10)
    for it, (fx, rfx, g, rg, ch, rch) in enumerate(rows, 1):
        assert abs(fx - rfx) <= 1e-10 * abs(rfx), (it, fx, rfx)
        assert abs(g - rg) <= 1e-10 * max(1.0, abs(rg)), (it, g, rg)
        assert abs(ch - rch) <= 1e-8, (it, ch, rch)
    assert np.max(np.abs(ctx.get("X") - ref.vec("X"))) <= 1e-8
    assert np.max(np.abs(ctx.get("XPhys") - ref.vec("XPhys"))) <= 1e-8


def test_projection_path(O):
    prm = dict(Emax=1.0, volfrac=0.3, rmin=0.375, proj=1, beta0=2.0, **TIGHT)
    ctx, ref = _pair(O, 8, 8, 4, 4, **prm)
    rows = _loop(ctx, ref, 4)
    for fx, rfx, g, rg, ch, rch in rows:
        assert abs(fx - rfx) <= 1e-10 * abs(rfx)
        assert abs(ch - rch) <= 1e-8
    assert np.max(np.abs(ctx.get("XPhys") - ref.vec("XPhys"))) <= 1e-8


def _beta_table(itr, beta0, betamax):
    """PDEFilter.cpp:259-269 for a call at every itr: beta0 below 30, then +1
    on every itr divisible by 10, capped at betamax."""
    return beta0 if itr < 30 else min(betamax, beta0 + (itr // 10) - 2)


@pytest.mark.parametrize("betamax", [8.0, 4.5])
def test_heaviside_schedule(O, betamax):
    """uft_heaviside at every itr 0..100 against the oracle and the hand
    schedule (x~ = 0.75 projects to 1/2 + tanh(b/4) / (2 tanh(b/2)) for
    x0 = 0.5), and uft_projected_sens with the beta of the last call."""
    prm = dict(Emax=1.0, volfrac=0.3, rmin=0.375, proj=1, beta0=1.0, betamax=betamax, **TIGHT)
    ctx, ref = _pair(O, 8, 8, 4, 4, **prm)
    rng = np.random.default_rng(21)
    xf = rng.uniform(0.0, 1.0, ref.ne)
    xf[0] = 0.75
    ctx.set("XFiled", xf)
    for itr in range(101):
        ctx.heaviside(itr)
        xp = ctx.get("XPhys")
        rxp, rbeta = ref.heaviside(xf, itr)
        b = _beta_table(itr, 1.0, betamax)
        assert rbeta == b, (itr, rbeta, b)
        assert np.max(np.abs(xp - rxp)) <= 2e-15, itr
        assert abs(xp[0] - (0.5 + np.tanh(b / 4) / (2 * np.tanh(b / 2)))) <= 2e-15, itr
        if itr in (0, 29, 30, 31, 39, 40, 59, 60, 90, 100):
            df = rng.standard_normal(ref.ne)
            dg = rng.uniform(0.5, 1.5, ref.ne)
            ctx.set("dfdx", df)
            ctx.set("dgdx0", dg)
            ctx.projected_sens()
            rdf, rdg = ref.projected_sens(xf, df, dg[None, :])
            assert _rel(ctx.get("dfdx"), rdf) <= 1e-14, itr
            assert _rel(ctx.get("dgdx0"), rdg[0]) <= 1e-14, itr


def test_projection_continuation_loop(O):
    """42 design iterations with the projection on, through the beta
    continuation at itr 30 and 40 (PDEFilter.cpp:259-269) on both sides;
    the loop is driven past the ch <= 0.01 stop on purpose."""
    prm = dict(Emax=1.0, volfrac=0.3, rmin=0.375, proj=1, beta0=1.0, **TIGHT)
    ctx, ref = _pair(O, 8, 8, 4, 4, **prm)
    rows = _loop(ctx, ref, 42)
    for it, (fx, rfx, g, rg, ch, rch) in enumerate(rows, 1):
        assert abs(fx - rfx) <= 1e-10 * abs(rfx), (it, fx, rfx)
        assert abs(g - rg) <= 1e-10 * max(1.0, abs(rg)), (it, g, rg)
        assert abs(ch - rch) <= 1e-8, (it, ch, rch)
    assert np.max(np.abs(ctx.get("X") - ref.vec("X"))) <= 1e-8
    assert np.max(np.abs(ctx.get("XPhys") - ref.vec("XPhys"))) <= 1e-8
