"""Pins the oracle's global operators, solves, filter and MMA (CPU).

- assembly + Dirichlet elimination against an independent dense assembly
  from the element routines (rows/columns zeroed, unit diagonal,
  SolveFem.cpp:316-333), and the solve against a direct solve;
- patch test: a linear displacement is in the null space of the interior rows;
- sensitivities against central finite differences of the compliance;
- filter: the dense operator TFX^T KF^-1 TF assembled independently, constant
  preservation, the non-adjoint sensitivity filter (PDEFilter.cpp:167-183);
- MMA: the subproblem solution against a brute-force dual maximisation whose
  inner minimisations use a generic scalar minimiser (no closed form), and
  hand-computed asymptote updates (MMA.cpp:516-578).
"""
import numpy as np
import pytest
import scipy.optimize as so
import scipy.sparse as sp
import scipy.sparse.linalg as spla


def dense_K(O, P, xphys, E0=1.0, Emin=1e-9, penal=3.0):
    n = 3 * P.nn
    K = np.zeros((n, n))
    npe = P.npe
    for e in range(P.ne):
        X = P.xyz[P.elem[e]]
        KE = (O.hex8_ke if npe == 8 else O.tet4_ke)(X)[0]
        dof = np.array([3 * P.elem[e][k % npe] + k // npe for k in range(3 * npe)])
        K[np.ix_(dof, dof)] += (Emin + xphys[e] ** penal * (E0 - Emin)) * KE
    return K


def eliminate(K, fixed):
    K = K.copy()
    K[fixed, :] = 0
    K[:, fixed] = 0
    K[fixed, fixed] = 1
    return K


@pytest.mark.parametrize("npe", [8, 4])
def test_assembly_and_solve(O, npe):
    P = O.cantilever(npe, 4, 2, 2, jitter=0.15, seed=3, load_kind=0 if npe == 8 else 1, Emax=1.0, rtol_el=1e-13)
    rng = np.random.default_rng(0)
    x = rng.uniform(0.01, 1, P.ne)
    fixed, ld, lv = O.box_bcs(4, 2, 2, 0 if npe == 8 else 1)
    K = eliminate(dense_K(O, P, x), fixed)
    n = 3 * P.nn
    KO = np.column_stack([P.apply_K(x, np.eye(n)[i]) for i in range(n)])
    assert np.max(np.abs(KO - K)) <= 1e-14 * np.max(np.abs(K))
    F = np.zeros(n)
    F[ld] = lv
    F[fixed] = 0
    U = np.linalg.solve(K, F)
    fx, gx, df, dg, rc = P.fem_obj_sens(x)
    assert rc == 0
    assert np.linalg.norm(P.vec("U") - U) <= 1e-10 * np.linalg.norm(U)
    assert abs(fx - F @ U) <= 1e-11 * abs(F @ U)   # sum E u^T KE u = u^T K u = F^T u
    V = P.vec("Vol")
    assert abs(gx[0] - (x @ V / V.sum() - P.params.volfrac)) < 1e-15
    assert np.allclose(dg[0], V / V.sum(), rtol=1e-15)


def test_solve_vs_sparse_direct(O):
    nx, ny, nz = 12, 6, 6
    P = O.cantilever(8, nx, ny, nz, jitter=0.15, seed=5, Emax=1.0, rtol_el=1e-13)
    rng = np.random.default_rng(2)
    x = rng.uniform(0.01, 1, P.ne)
    fixed, ld, lv = O.box_bcs(nx, ny, nz)
    rows, cols, vals = [], [], []
    for e in range(P.ne):
        KE = O.hex8_ke(P.xyz[P.elem[e]])[0] * (1e-9 + x[e] ** 3 * (1 - 1e-9))
        dof = np.array([3 * P.elem[e][k % 8] + k // 8 for k in range(24)])
        rows.append(np.repeat(dof, 24))
        cols.append(np.tile(dof, 24))
        vals.append(KE.ravel())
    n = 3 * P.nn
    K = sp.csr_matrix((np.concatenate(vals), (np.concatenate(rows), np.concatenate(cols))), shape=(n, n)).tolil()
    free = np.ones(n, bool)
    free[fixed] = False
    F = np.zeros(n)
    F[ld] = lv
    Kff = K.tocsr()[free][:, free]
    U = np.zeros(n)
    U[free] = spla.spsolve(Kff.tocsc(), F[free])
    fx, *_ , rc = P.fem_obj_sens(x)
    assert rc == 0
    assert np.linalg.norm(P.vec("U") - U) <= 1e-9 * np.linalg.norm(U)
    assert abs(fx - F @ U) <= 1e-10 * abs(F @ U)


def test_patch_test(O):
    """Undistorted grid, no Dirichlet: K u_linear vanishes at interior nodes."""
    xyz, elem = O.gen_box(8, 4, 3, 3)
    P = O.Problem(elem, xyz, np.zeros(0, np.int32), np.array([1], np.int32), np.array([0.0]),
                  O.default_params(Emax=1.0))
    A = np.array([[0.1, 0.3, -0.2], [0.05, -0.1, 0.2], [0.3, 0.0, 0.1]])
    u = (xyz @ A.T + [0.1, -0.2, 0.3]).ravel()
    y = P.apply_K(np.ones(P.ne), u).reshape(-1, 3)
    ijk = np.array([[i, j, k] for k in range(4) for j in range(4) for i in range(5)])
    interior = (ijk[:, 0] > 0) & (ijk[:, 0] < 4) & (ijk[:, 1] > 0) & (ijk[:, 1] < 3) & (ijk[:, 2] > 0) & (ijk[:, 2] < 3)
    assert np.max(np.abs(y[interior])) < 1e-13
    assert np.max(np.abs(y[~interior])) > 1e-3


def test_sensitivity_finite_difference(O):
    P = O.cantilever(8, 3, 2, 2, jitter=0.1, seed=9, Emax=1.0, rtol_el=1e-14)
    rng = np.random.default_rng(4)
    x = rng.uniform(0.2, 0.9, P.ne)
    fx, gx, df, dg, _ = P.fem_obj_sens(x)
    df = df.copy()
    for j in [0, 3, P.ne - 1]:
        h = 1e-5
        xp, xm = x.copy(), x.copy()
        xp[j] += h
        xm[j] -= h
        fd = (P.fem_obj_sens(xp)[0] - P.fem_obj_sens(xm)[0]) / (2 * h)
        assert abs(fd - df[j]) <= 1e-6 * abs(df[j]), (j, fd, df[j])


def filter_dense(O, P, rmin):
    n, ne, npe = P.nn, P.ne, P.npe
    KF = np.zeros((n, n))
    TF = np.zeros((n, ne))
    TFX = np.zeros((n, ne))
    for e in range(ne):
        X = P.xyz[P.elem[e]]
        KFe, Te = (O.hex8_kehel if npe == 8 else O.tet4_kehel)(X, rmin)
        nodes = P.elem[e]
        KF[np.ix_(nodes, nodes)] += KFe
        TF[nodes, e] += Te
        TFX[nodes, e] += 1.0 / npe
    return TFX.T @ np.linalg.solve(KF, TF)


@pytest.mark.parametrize("npe", [8, 4])
def test_filter_operator(O, npe):
    rmin = 0.6
    P = O.cantilever(npe, 5, 3, 3, jitter=0.15, seed=21, rmin=rmin, rtol_f=1e-14, Xmin=0.0)
    W = filter_dense(O, P, rmin)
    rng = np.random.default_rng(6)
    x = rng.uniform(0.0, 1.0, P.ne)
    xf, rc = P.filter(x)
    assert rc == 0
    assert np.max(np.abs(xf - np.clip(W @ x, 0.0, 1.0))) < 1e-11
    c = np.full(P.ne, 0.37)
    assert np.max(np.abs(P.filter(c)[0] - 0.37)) < 1e-12
    v = rng.standard_normal(P.ne)
    fdf, fdg, rc = P.filter_sens(v, np.abs(v)[None, :])
    assert np.max(np.abs(fdf - W @ v)) < 1e-11 * np.max(np.abs(v))
    assert np.max(np.abs(fdg[0] - W @ np.abs(v))) < 1e-11 * np.max(np.abs(v))
    # the reference applies the forward operator, not its adjoint
    assert np.max(np.abs(W - W.T)) > 1e-4 * np.max(np.abs(W))
    assert np.max(np.abs(fdf - W.T @ v)) > 1e-6


def _mma_problem(O, n_side=(4, 2, 1)):
    P = O.cantilever(8, *n_side, Emax=1.0, volfrac=0.3)
    return P


def test_mma_subproblem_brute_force(O):
    P = _mma_problem(O)
    n = P.ne
    rng = np.random.default_rng(12)
    X = P.vec("X")
    X[:] = rng.uniform(0.2, 0.6, n)
    x0 = X.copy()
    df = -rng.uniform(1.0, 50.0, n)
    dg = np.full((1, n), 1.0 / n) * rng.uniform(0.5, 1.5, n)
    gx = np.array([0.05])
    xmin, xmax = np.zeros(n), np.zeros(n)
    L = O.lib()
    L.orc_mma_movelimit(P.h, X, xmin, xmax)
    L.orc_mma_update(P.h, X, df, gx, dg.ravel(), xmin, xmax)
    xo = X.copy()
    # the subproblem, rebuilt from the documented first-iteration rules (k < 3)
    Lo, Uo = x0 - 0.5 * (xmax - xmin), x0 + 0.5 * (xmax - xmin)
    assert np.allclose(P.vec("L"), Lo, rtol=0, atol=1e-15) and np.allclose(P.vec("Uas"), Uo, rtol=0, atol=1e-15)
    al = np.maximum(xmin, 0.9 * Lo + 0.1 * x0)
    be = np.minimum(xmax, 0.9 * Uo + 0.1 * x0)
    p0 = (Uo - x0) ** 2 * (np.maximum(df, 0) + 0.001 * np.abs(df) + 0.5e-6 / (Uo - Lo))
    q0 = (x0 - Lo) ** 2 * (np.maximum(-df, 0) + 0.001 * np.abs(df) + 0.5e-6 / (Uo - Lo))
    p1 = (Uo - x0) ** 2 * np.maximum(dg[0], 0)
    q1 = (x0 - Lo) ** 2 * np.maximum(-dg[0], 0)
    b = np.sum(p1 / (Uo - x0) + q1 / (x0 - Lo)) - gx[0]
    c = 1000.0

    def xstar(lam):
        out = np.zeros(n)
        for i in range(n):
            f = lambda t: (p0[i] + lam * p1[i]) / (Uo[i] - t) + (q0[i] + lam * q1[i]) / (t - Lo[i])
            out[i] = so.minimize_scalar(f, bounds=(al[i], be[i]), method="bounded",
                                        options={"xatol": 1e-13}).x
        return out

    def dual_grad(lam):  # d/dlam of the dual with y = max(0, lam - c)
        xs = xstar(lam)
        return np.sum(p1 / (Uo - xs) + q1 / (xs - Lo)) - b - max(0.0, lam - c)

    lo, hi = 0.0, 1.0
    while dual_grad(hi) > 0:
        hi *= 2
    if dual_grad(lo) <= 0:
        lam = 0.0
    else:
        for _ in range(200):
            mid = 0.5 * (lo + hi)
            if dual_grad(mid) > 0:
                lo = mid
            else:
                hi = mid
        lam = 0.5 * (lo + hi)
    assert abs(L.orc_mma_lam(P.h, 0) - lam) <= 1e-6 * max(1.0, lam)
    assert np.max(np.abs(xo - xstar(lam))) < 1e-7


def test_mma_asymptote_rules(O):
    """Hand-computed third-iteration asymptotes (gamma 1.2 monotone, 0.7 oscillating)."""
    P = _mma_problem(O, (2, 1, 1))
    L = O.lib()
    X = P.vec("X")
    n = P.ne
    xmin, xmax = np.zeros(n), np.zeros(n)
    df = np.full(n, -1.0)
    dg = np.full(n, 0.5)
    gx = np.array([0.0])
    for xin in ([0.5, 0.5], [0.4, 0.4], [0.3, 0.45]):
        X[:] = xin
        L.orc_mma_movelimit(P.h, X, xmin, xmax)
        L.orc_mma_update(P.h, X, df, gx, dg, xmin, xmax)
    assert L.orc_mma_k(P.h) == 3
    # element 0: 0.5 -> 0.4 -> 0.3 (gamma 1.2); element 1: 0.5 -> 0.4 -> 0.45 (gamma 0.7)
    assert np.allclose(P.vec("L"), [0.06, 0.31], atol=1e-15)
    assert np.allclose(P.vec("Uas"), [0.54, 0.59], atol=1e-15)


def test_design_change_and_movelimit(O):
    P = _mma_problem(O, (2, 1, 1))
    L = O.lib()
    x = np.array([0.95, 0.05])
    xmin, xmax = np.zeros(2), np.zeros(2)
    L.orc_mma_movelimit(P.h, x, xmin, xmax)
    assert np.allclose(xmax, [1.0, 0.25]) and np.allclose(xmin, [0.75, 0.01])
    xold = np.array([0.9, 0.2])
    assert abs(L.orc_design_change(P.h, x, xold) - 0.15) < 1e-15 and np.array_equal(xold, x)
