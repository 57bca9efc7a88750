"""Oracle: the multi-load-case form (separate right-hand sides, summed
compliance).  Pins: linearity (the case solutions add up to the superposed
solution), the compliance identity sum_e E u_k^T KE u_k = F_k . U_k, central
finite differences of the summed compliance, and the eight-patch load
generator's layout."""
import numpy as np
import pytest

PRM = dict(Emax=1.0, volfrac=0.3, rtol_el=1e-13, rtol_f=1e-12)


def test_eight_patch_generator(O):
    nx, ny, nz = 8, 4, 4
    fixed, ld, lv = O.box_bcs(nx, ny, nz, load_kind=2)
    assert len(ld) == 72 and np.all(ld % 3 == 1)          # y-DOFs
    nodes = ld // 3
    assert np.all(nodes % (nx + 1) == nx)                  # on the x = Lx face
    assert np.allclose(lv.reshape(8, 9).sum(axis=1), -1.0)  # unit total per case
    centres = {tuple(sorted(set(c))) for c in nodes.reshape(8, 9)}
    assert len(centres) > 1                                 # seeded, not all alike


def test_case_solutions_superpose(O):
    """Linearity: the case solutions add up to the solution for the summed
    load (built as one case; the reference's INSERT would overwrite the
    DOFs that overlapping patches share, so the sum is formed here)."""
    sep = O.cantilever(8, 6, 3, 3, jitter=0.1, seed=2, load_kind=2, **PRM)
    xyz, elem = O.gen_box(8, 6, 3, 3, 2.0, 1.0, 1.0, 0.1, 2)
    fixed, ld, lv = O.box_bcs(6, 3, 3, load_kind=2)
    dofs = np.unique(ld)
    vals = np.array([lv[ld == d].sum() for d in dofs])
    sup = O.Problem(elem, xyz, fixed, dofs.astype(np.int32), vals, O.default_params(**PRM))
    sep.set_load_separate(True)
    x = np.random.default_rng(4).uniform(0.1, 1.0, sep.ne)
    sep.fem_obj_sens(x)
    sup.fem_obj_sens(x)
    Usum = sum(sep.case_u(k) for k in range(8))
    assert np.abs(Usum - sup.vec("U")).max() <= 1e-9 * np.abs(sup.vec("U")).max()


def test_summed_compliance_identity_and_fd(O):
    P = O.cantilever(8, 6, 3, 3, jitter=0.1, seed=5, load_kind=2, **PRM)
    P.set_load_separate(True)
    x = np.random.default_rng(6).uniform(0.2, 1.0, P.ne)
    fx, gx, df, dg, rc = P.fem_obj_sens(x)
    _, ld, lv = O.box_bcs(6, 3, 3, load_kind=2)
    FU = 0.0
    for k in range(8):
        F = np.zeros(3 * P.nn)
        F[ld[9 * k:9 * k + 9]] = lv[9 * k:9 * k + 9]
        FU += F @ P.case_u(k)
    assert abs(fx - FU) <= 1e-9 * abs(fx)
    h = 1e-6
    for e in (3, 40):
        xp, xm = x.copy(), x.copy()
        xp[e] += h
        xm[e] -= h
        fd = (P.fem_obj_sens(xp)[0] - P.fem_obj_sens(xm)[0]) / (2 * h)
        assert abs(fd - df[e]) <= 1e-5 * abs(df[e])
