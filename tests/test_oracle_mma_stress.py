"""Oracle: the von Mises stress routine (stress/ComputeStress.cpp) and the
MMA options around the update -- robust asymptotes, constraint modification,
restart state and the KKT residual (MMA.cpp:5-176, 302-480, 506-637).

Pins: the isoparametric patch property (a linear displacement field gives
its exact strain at the centroid of any distorted element), a numerical
derivative of the element interpolation, hand-computed asymptote and
approximation values, the KKT conditions themselves, and a restart that must
reproduce an uninterrupted run bit for bit."""
import numpy as np
import pytest


def _vm_closed_form(G, Emax, nu, x):
    """von Mises of the uniform stress of displacement gradient G (sqrt(3 J2)
    of the deviatoric Cauchy stress), scaled by sqrt(x)."""
    lam = nu / ((1 + nu) * (1 - 2 * nu))
    mu = 1 / (2 * (1 + nu))
    eps = 0.5 * (G + G.T)
    sig = Emax * (lam * np.trace(eps) * np.eye(3) + 2 * mu * eps) * np.sqrt(x)
    s = sig - np.trace(sig) / 3 * np.eye(3)
    return np.sqrt(1.5 * np.sum(s * s))


@pytest.mark.parametrize("etype,dims", [(8, (3, 2, 2)), (4, (2, 2, 1))])
def test_stress_patch_linear_field(O, etype, dims):
    P = O.cantilever(etype, *dims, jitter=0.25, seed=4, Emax=2.5, nu=0.3)
    rng = np.random.default_rng(7)
    G = rng.standard_normal((3, 3))
    t = rng.standard_normal(3)
    u = (P.xyz @ G.T + t).ravel()
    x = rng.uniform(0.05, 1.0, P.ne)
    got = P.stress(x, u)
    want = np.array([_vm_closed_form(G, 2.5, 0.3, xe) for xe in x])
    assert np.allclose(got, want, rtol=1e-12, atol=0)


def test_stress_rigid_motion_is_zero(O):
    P = O.cantilever(8, 3, 2, 2, jitter=0.25, seed=5, Emax=1.0)
    W = np.array([[0.0, -0.3, 0.2], [0.3, 0.0, -0.1], [-0.2, 0.1, 0.0]])  # infinitesimal rotation
    u = (P.xyz @ W.T + np.array([1.0, -2.0, 0.5])).ravel()
    assert np.abs(P.stress(np.ones(P.ne), u)).max() < 1e-12


def test_stress_centroid_numerical_derivative(O):
    """Random nodal displacements: the strain is the centroid gradient of the
    trilinear interpolation, here from central differences in the reference
    coordinates and the chain rule (an independent evaluation)."""
    P = O.cantilever(8, 2, 1, 1, jitter=0.3, seed=11, Emax=1.0, nu=0.25)
    rng = np.random.default_rng(3)
    u = rng.standard_normal(3 * P.nn)
    got = P.stress(np.ones(P.ne), u)
    corners = np.array([[(a & 1) * 2 - 1, (a >> 1 & 1) * 2 - 1, (a >> 2 & 1) * 2 - 1] for a in range(8)], float)

    def interp(vals, xi):
        N = np.prod(1 + corners * xi, axis=1) / 8
        return N @ vals

    h = 1e-5
    for e in range(P.ne):
        nodes = P.elem[e]
        X, U = P.xyz[nodes], u.reshape(-1, 3)[nodes]
        dX = np.zeros((3, 3))
        dU = np.zeros((3, 3))
        for r in range(3):
            d = np.zeros(3)
            d[r] = h
            dX[r] = (interp(X, d) - interp(X, -d)) / (2 * h)
            dU[r] = (interp(U, d) - interp(U, -d)) / (2 * h)
        G = np.linalg.solve(dX, dU).T  # G[i][j] = du_i / dx_j
        assert abs(got[e] - _vm_closed_form(G, 1.0, 0.25, 1.0)) < 1e-8 * max(1.0, got[e])


def _kkt_expected(x, df, dg, xmin, xmax, lam):
    """KKT terms from the optimality conditions: at an active lower (upper)
    bound a positive (negative) stationarity residual is absorbed by the
    bound multiplier, which then appears in the complementarity term."""
    r = df + lam @ dg
    terms = []
    for i in range(len(x)):
        if x[i] < xmin[i] + 1e-5 and r[i] > 0:
            terms += [0.0, r[i] * (x[i] - xmin[i])]
        elif x[i] > xmax[i] - 1e-5 and r[i] < 0:
            terms += [0.0, -r[i] * (xmax[i] - x[i])]
        else:
            terms += [r[i]]
    return np.array(terms)


def test_kkt_residual_semantics(O):
    P = O.cantilever(8, 4, 2, 1, Emax=1.0, volfrac=0.3)
    L = O.lib()
    n = P.ne
    X = P.vec("X")
    rng = np.random.default_rng(2)
    X[:] = rng.uniform(0.2, 0.6, n)
    xmin, xmax = np.zeros(n), np.zeros(n)
    df = -rng.uniform(1, 5, n)
    dg = rng.uniform(0.5, 1.5, (1, n)) / n
    gx = np.array([0.02])
    L.orc_mma_movelimit(P.h, X, xmin, xmax)
    L.orc_mma_update(P.h, X, df, gx, dg.ravel(), xmin, xmax)
    lam = np.array([L.orc_mma_lam(P.h, 0)])
    # scripted point: two variables at a bound, the rest interior
    x = X.copy()
    x[0] = xmin[0]           # lower bound
    x[1] = xmax[1] - 2e-6    # within 1e-5 of the upper bound
    df2 = df.copy()
    df2[0] = 3.0 - lam[0] * dg[0, 0]   # r_0 = 3 > 0: absorbed, x - xmin = 0
    df2[1] = -2.0 - lam[0] * dg[0, 1]  # r_1 = -2 < 0: absorbed, complementarity 2 * 2e-6
    n2, ni = P.mma_kkt(x, df2, gx, dg, xmin, xmax)
    terms = _kkt_expected(x, df2, dg, xmin, xmax, lam)
    k, xo1, xo2, U, Lw = P.mma_state()
    # the constraint term lam (a z + y - g) with a = 0 and y = max(0, lam - c)
    rl = lam[0] * (max(0.0, lam[0] - 1000.0) - gx[0])
    want2 = np.sqrt(np.sum(terms ** 2) + rl ** 2)
    assert abs(n2 - want2) <= 1e-12 * want2
    assert abs(ni - max(np.abs(terms).max(), abs(rl))) <= 1e-14 * ni


def test_robust_asymptotes_and_constraint_modification_kat(O):
    """Hand-computed GenSub values with RobustAsymptotesType 1 and
    constraintModification, from a restart state with tight asymptotes."""
    P = O.cantilever(8, 3, 1, 1, Emax=1.0)
    n = P.ne
    P.mma_settings(0.5, 0.7, 1.2, robust=1, conmod=1)
    xo1 = np.array([0.4, 0.4, 0.4])
    xo2 = np.array([0.5, 0.5, 0.5])
    Lr, Ur = xo1 - 0.001, xo1 + 0.001
    P.mma_set_state(3, xo1, xo2, Ur, Lr)
    x = np.array([0.45, 0.3, 0.45])     # oscillating, monotone, oscillating
    xmin = np.array([0.25, 0.35, 0.25])  # element 1 lies below its box
    xmax = np.array([0.65, 0.6, 0.65])
    df = np.array([-1.0, 2.0, -3.0])
    dg = np.array([[0.5, -0.25, 1.0]])
    X = P.vec("X")
    X[:] = x
    O.lib().orc_mma_update(P.h, X, df, np.array([0.0]), dg.ravel(), xmin, xmax)
    Lv, Uv = P.vec("L"), P.vec("Uas")
    # element 0: gamma 0.7, L = 0.45 - 0.7*0.001, kept (>= x - 1e-4 w, w = 0.4)
    assert abs(Lv[0] - (0.45 - 0.0007)) < 1e-15 and abs(Uv[0] - (0.45 + 0.0007)) < 1e-15
    # element 1: x < xmin - 1e-5, so L, U = x -/+ (xmax + 1e-5 - x) / 0.9
    w1 = (0.6 + 1e-5 - 0.3) / 0.9
    assert abs(Lv[1] - (0.3 - w1)) < 1e-15 and abs(Uv[1] - (0.3 + w1)) < 1e-15
    # constraint approximation with the objective's regularisation
    ul = Uv - Lv
    reg = 0.001 * np.abs(dg[0]) + 0.5e-6 / ul
    assert np.allclose(P.vec("pij"), (Uv - x) ** 2 * (np.maximum(dg[0], 0) + reg), rtol=1e-14, atol=0)
    assert np.allclose(P.vec("qij"), (x - Lv) ** 2 * (np.maximum(-dg[0], 0) + reg), rtol=1e-14, atol=0)


def test_restart_reproduces_uninterrupted_run(O):
    """Export the MMA state after 3 iterations, continue to 6; then rewind the
    design vectors and the filter warm start, import the state and redo
    iterations 4-6: identical results."""
    P = O.cantilever(8, 6, 3, 3, jitter=0.1, seed=8, Emax=1.0, volfrac=0.3, rmin=0.3, rtol_el=1e-10, rtol_f=1e-10)
    P.init_design()
    for it in (1, 2, 3):
        P.design_iteration(it)
    st = P.mma_state()
    saved = {k: P.vec(k).copy() for k in ("X", "XFiled", "XPhys", "Xold", "RX")}
    first = [P.design_iteration(it)[:3] for it in (4, 5, 6)]
    for k, v in saved.items():
        P.vec(k)[:] = v
    P.mma_set_state(*st)
    again = [P.design_iteration(it)[:3] for it in (4, 5, 6)]
    assert first == again
