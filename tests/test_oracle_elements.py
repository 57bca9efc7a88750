"""Pins the oracle's element routines (CPU, no GPU needed).

Checks are chosen so a plausible mistake fails one of them: closed forms of
box and unit-tet elements, invariants (symmetry, rigid-body null space, rank,
scaling, rotation covariance), 40-digit golden vectors computed through a
different but equivalent formula (tests/golden/make_golden.py), and a
negative test that the truncated Gauss rule of SolveFem.cpp:364-365 is used
rather than the exact one.
"""
import os

import numpy as np
import pytest

GOLD = np.load(os.path.join(os.path.dirname(__file__), "golden", "elements.npz"))
NU = 0.3
LAM = NU / ((1 + NU) * (1 - 2 * NU))
MU = 1 / (2 * (1 + NU))


def rbm(X):
    npe = len(X)
    B = np.zeros((3 * npe, 6))
    for a in range(npe):
        x, y, z = X[a]
        for c in range(3):
            B[c * npe + a, c] = 1
        B[0 * npe + a, 3], B[1 * npe + a, 3] = -y, x
        B[1 * npe + a, 4], B[2 * npe + a, 4] = -z, y
        B[2 * npe + a, 5], B[0 * npe + a, 5] = -x, z
    return B


def unit_hex():
    return np.array([[(a & 1), (a >> 1) & 1, (a >> 2) & 1] for a in range(8)], float)


def box_ke_closed(a, b, c, gp, gw):
    """KE of an a x b x c box from the isotropic bilinear form with the
    quadrature rule (gp, gw), evaluated by hand-derived gradients (J diagonal)."""
    K = np.zeros((24, 24))
    for i, xi in enumerate(gp):
        for j, eta in enumerate(gp):
            for k, ze in enumerate(gp):
                w = gw[i] * gw[j] * gw[k] * (a * b * c / 8)
                g = np.zeros((8, 3))
                for n in range(8):
                    sx, sy, sz = (1 if n & 1 else -1), (1 if n & 2 else -1), (1 if n & 4 else -1)
                    g[n] = [sx * (1 + sy * eta) * (1 + sz * ze) / 8 * 2 / a,
                            sy * (1 + sx * xi) * (1 + sz * ze) / 8 * 2 / b,
                            sz * (1 + sx * xi) * (1 + sy * eta) / 8 * 2 / c]
                for p in range(8):
                    for q in range(8):
                        gg = g[p] @ g[q]
                        for r in range(3):
                            for s in range(3):
                                K[r * 8 + p, s * 8 + q] += w * (LAM * g[p, r] * g[q, s] + MU * g[p, s] * g[q, r]
                                                                + (MU * gg if r == s else 0))
    return K


TRUNC = ([-0.775, 0.0, 0.775], [0.556, 0.889, 0.556])
EXACT = ([-np.sqrt(0.6), 0.0, np.sqrt(0.6)], [5 / 9, 8 / 9, 5 / 9])


def test_hex8_box_closed_form(O):
    for (a, b, c) in [(1, 1, 1), (0.5, 0.25, 0.25), (2.0, 0.3, 0.7)]:
        X = unit_hex() * [a, b, c]
        KE, mindet = O.hex8_ke(X)
        ref = box_ke_closed(a, b, c, *TRUNC)
        assert np.max(np.abs(KE - ref)) <= 1e-14 * np.max(np.abs(ref))
        assert abs(mindet - a * b * c / 8) < 1e-15


def test_hex8_truncated_rule_is_used(O):
    """Negative test: the exact 3-point Gauss rule gives a different KE."""
    KE, _ = O.hex8_ke(unit_hex())
    exact = box_ke_closed(1, 1, 1, *EXACT)
    rel = np.max(np.abs(KE - exact)) / np.max(np.abs(exact))
    assert 1e-4 < rel < 1e-2


def test_hex8_invariants(O):
    rng = np.random.default_rng(1)
    for _ in range(5):
        X = unit_hex() + rng.uniform(-0.15, 0.15, (8, 3))
        KE, mindet = O.hex8_ke(X)
        assert mindet > 0
        assert np.max(np.abs(KE - KE.T)) <= 1e-14 * np.max(np.abs(KE))
        assert np.max(np.abs(KE @ rbm(X))) <= 1e-13 * np.max(np.abs(KE))
        ev = np.linalg.eigvalsh(KE)
        assert np.sum(ev > 1e-10 * ev.max()) == 18 and ev.min() > -1e-12 * ev.max()
        # 3D elasticity: KE(alpha X) = alpha KE(X)
        KE2, _ = O.hex8_ke(2.5 * X)
        assert np.max(np.abs(KE2 - 2.5 * KE)) <= 1e-13 * np.max(np.abs(KE2))
        # rotation covariance with component-blocked DOFs
        th = rng.uniform(0, 2 * np.pi, 3)
        Rz = np.array([[np.cos(th[0]), -np.sin(th[0]), 0], [np.sin(th[0]), np.cos(th[0]), 0], [0, 0, 1]])
        Rx = np.array([[1, 0, 0], [0, np.cos(th[1]), -np.sin(th[1])], [0, np.sin(th[1]), np.cos(th[1])]])
        R = Rz @ Rx
        KR, _ = O.hex8_ke(X @ R.T)
        T = np.kron(R, np.eye(8))  # acts on [u1..u8, v1..v8, w1..w8]
        assert np.max(np.abs(KR - T @ KE @ T.T)) <= 1e-13 * np.max(np.abs(KE))


def test_hex8_golden(O):
    for X, ref in zip(GOLD["hex_X"], GOLD["hex_KE"]):
        KE, _ = O.hex8_ke(X)
        assert np.max(np.abs(KE - ref)) <= 1e-14 * np.max(np.abs(ref))


def test_tet4_closed_form_and_golden(O):
    X = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0], [0, 0, 1]], float)
    KE, det = O.tet4_ke(X)
    assert det == 1.0
    # unit right tet: g_0 = (-1,-1,-1), g_k = e_k, V = 1/6
    g = np.array([[-1, -1, -1], [1, 0, 0], [0, 1, 0], [0, 0, 1]], float)
    ref = np.zeros((12, 12))
    for p in range(4):
        for q in range(4):
            for r in range(3):
                for s in range(3):
                    ref[r * 4 + p, s * 4 + q] = (LAM * g[p, r] * g[q, s] + MU * g[p, s] * g[q, r]
                                                 + (MU * g[p] @ g[q] if r == s else 0)) / 6
    assert np.max(np.abs(KE - ref)) <= 1e-15
    for X, ref in zip(GOLD["tet_X"], GOLD["tet_KE"]):
        KE, _ = O.tet4_ke(X)
        assert np.max(np.abs(KE - ref)) <= 1e-14 * np.max(np.abs(ref))
        assert np.max(np.abs(KE @ rbm(X))) <= 1e-13 * np.max(np.abs(KE))
        assert np.sum(np.linalg.eigvalsh(KE) > 1e-10 * np.abs(KE).max()) == 6


def test_volumes(O):
    # parallelepiped: |det|
    A = np.array([[1.0, 0.2, 0.1], [0.0, 0.8, 0.3], [0.1, 0.0, 1.2]])
    X = unit_hex() @ A.T + [0.3, -0.2, 0.5]
    assert abs(O.hex8_vol(X) - abs(np.linalg.det(A))) < 1e-15
    # frustum of a square pyramid (planar faces): h/3 (a^2 + ab + b^2)
    a, b, h = 1.0, 0.6, 0.8
    F = []
    for n in range(8):
        s = a if n < 4 else b
        off = (a - s) / 2
        F.append([off + s * (n & 1), off + s * ((n >> 1) & 1), h * (n >> 2)])
    assert abs(O.hex8_vol(np.array(F)) - h / 3 * (a * a + a * b + b * b)) < 1e-15
    for X, ref in zip(GOLD["hex_X"], GOLD["hex_vol"]):
        assert abs(O.hex8_vol(X) - ref) <= 1e-15 * ref
    Xt = np.array([[0, 0, 0], [2, 0, 0], [0, 1, 0], [0, 0, 3]], float)
    assert abs(O.tet4_vol(Xt) - 1.0) < 1e-15


def test_helmholtz_element(O):
    rmin = float(GOLD["rmin"])
    for X, rKF, rT in zip(GOLD["hex_X"], GOLD["hex_KF"], GOLD["hex_T"]):
        KF, T = O.hex8_kehel(X, rmin)
        assert np.max(np.abs(KF - rKF)) <= 1e-14 * np.max(np.abs(rKF))
        assert np.max(np.abs(T - rT)) <= 1e-14 * np.max(np.abs(rT))
        # grad 1 = 0 and sum N = 1:  KF 1 = T
        assert np.max(np.abs(KF.sum(1) - T)) <= 1e-15
    # parallelepiped: sum T = (sum w)^3 / 8 * V with the truncated weights
    A = np.array([[0.5, 0.1, 0.0], [0.0, 0.4, 0.1], [0.05, 0.0, 0.6]])
    X = unit_hex() @ A.T
    _, T = O.hex8_kehel(X, 0.2)
    assert abs(T.sum() - 2.001 ** 3 / 8 * np.linalg.det(A)) < 1e-15
    # tet4: Hammer 4-point rule is exact for the mass matrix V/20 (1 + delta) and T = V/4
    Xt = np.array([[0, 0, 0], [1.0, 0.1, 0], [0.2, 0.9, 0.1], [0.1, 0.0, 1.1]])
    V = O.tet4_vol(Xt)
    KF, T = O.tet4_kehel(Xt, 0.0)
    assert np.max(np.abs(KF - V / 20 * (np.ones((4, 4)) + np.eye(4)))) < 1e-15
    assert np.max(np.abs(T - V / 4)) < 1e-15
    for X, rKF, rT in zip(GOLD["tet_X"], GOLD["tet_KF"], GOLD["tet_T"]):
        KF, T = O.tet4_kehel(X, rmin)
        assert np.max(np.abs(KF - rKF)) <= 1e-14 * np.max(np.abs(rKF))
        assert np.max(np.abs(T - rT)) <= 1e-14 * np.max(np.abs(rT))


def test_dofmap(O):
    elem = np.array([[0, 1, 2, 3, 4, 5, 6, 7], [1, 8, 3, 9, 5, 10, 7, 11]], np.int32)
    d = O.dofmap(elem)
    assert list(d[0]) == [0, 3, 6, 9, 12, 15, 18, 21, 1, 4, 7, 10, 13, 16, 19, 22, 2, 5, 8, 11, 14, 17, 20, 23]
    assert list(d[1][:8]) == [3, 24, 9, 27, 15, 30, 21, 33]
    assert list(d[1][8:16]) == [4, 25, 10, 28, 16, 31, 22, 34]
    t = O.dofmap(np.array([[5, 2, 7, 1]], np.int32))
    assert list(t[0]) == [15, 6, 21, 3, 16, 7, 22, 4, 17, 8, 23, 5]


def test_generator_and_bcs(O):
    xyz, elem = O.gen_box(8, 2, 1, 1)
    assert elem.tolist() == [[0, 1, 3, 4, 6, 7, 9, 10], [1, 2, 4, 5, 7, 8, 10, 11]]
    assert np.allclose(xyz[[0, 2, 11]], [[0, 0, 0], [2, 0, 0], [2, 1, 1]])
    fixed, ld, lv = O.box_bcs(140, 70, 40)
    # the reference's compile-time sizes (OptPara.cpp:5-11): 8733 fixed DOFs, 41 load DOFs
    assert len(fixed) == 8733 == 3 * 71 * 41 and len(ld) == 41
    assert np.all(np.diff(fixed) > 0) and np.all(fixed.reshape(-1, 3) % 3 == [0, 1, 2])
    assert np.all(ld % 3 == 1) and np.allclose(lv, -1 / 41)
    xyz, elem = O.gen_box(8, 140, 70, 40, jitter=0.15, seed=392000)
    assert xyz.shape == (410451, 3) and elem.shape == (392000, 8)
    # boundary nodes are not jittered: the clamped face is exactly x = 0
    assert np.all(xyz[fixed[::3] // 3, 0] == 0.0)
    xt, et = O.gen_box(4, 60, 30, 30, jitter=0.15, seed=324000)
    assert xt.shape == (58621, 3) and et.shape == (324000, 4)
    f2, l2, v2 = O.box_bcs(60, 30, 30, 1)
    assert len(f2) == 2883 and len(l2) == 9 and np.allclose(v2, -1 / 9)
    # every tet positively oriented
    Xs = xt[et]
    det = np.linalg.det(Xs[:, 1:] - Xs[:, :1])
    assert det.min() > 0


REF_MESH = "/root/reference/mesh"


@pytest.mark.skipif(not os.path.isdir(REF_MESH), reason="reference BC fixtures not present")
def test_reference_bc_fixture_facts():
    """The reference's BC files: counts, ordering, sums (facts pinned from the
    files themselves; the build's reader semantics must match these)."""
    fixed = np.array(open(os.path.join(REF_MESH, "fixeddofs.txt")).read().split(), dtype=np.int64)
    ld = np.array(open(os.path.join(REF_MESH, "loaddof_1.txt")).read().split(), dtype=np.int64)
    lv = np.array(open(os.path.join(REF_MESH, "load_1.txt")).read().split(), dtype=np.float64)
    assert len(fixed) == 8733 and fixed.sum() == 1073477334 and np.all(np.diff(fixed) > 0)
    assert np.all(fixed.reshape(-1, 3) - 3 * (fixed.reshape(-1, 3)[:, :1] // 3) == [0, 1, 2])
    assert len(ld) == 41 and ld.sum() == 37262276 and np.all(ld % 3 == 1)
    assert len(lv) == 41 and np.all(lv == 1.0)
