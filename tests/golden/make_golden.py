"""Writes tests/golden/elements.npz: element matrices of seeded jittered
hex8/tet4 elements evaluated in 40-digit arithmetic (mpmath), independently
of the FP64 oracle.

The formulas are the reference's element routines as documented:
  hex8  SolveFem.cpp:339-408  KE = sum_{27 GP, truncated rule} W^3 B^T D B detJ
  tet4  tet4/Tet4Isoparametric.cpp:1-41  KE = (1/6) B^T D B sqrt(det(J J^T))
  Helmholtz  PDEFilter.cpp:185-240 / tet4/KeHel.cpp:1-47
  volume  SolveFem.cpp:423-467 (five signed tets, absolute values)
Here they are evaluated through the isotropic bilinear form
  K[(a,i),(b,j)] = lam g_a[i] g_b[j] + mu g_a[j] g_b[i] + mu delta_ij g_a.g_b
which equals B^T D B for the reference's D and B, so a transposed B row or a
misplaced shear term in the oracle would not be reproduced here.

Run:  python tests/golden/make_golden.py
"""
import os

import mpmath as mp
import numpy as np

mp.mp.dps = 40
HERE = os.path.dirname(os.path.abspath(__file__))
NU = mp.mpf("0.3")
LAM = NU / ((1 + NU) * (1 - 2 * NU))
MU = 1 / (2 * (1 + NU))
GP = [mp.mpf("-0.775"), mp.mpf(0), mp.mpf("0.775")]
GW = [mp.mpf("0.556"), mp.mpf("0.889"), mp.mpf("0.556")]


def split64(seed):
    s = seed
    while True:
        s = (s + 0x9E3779B97F4A7C15) & (2**64 - 1)
        z = s
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & (2**64 - 1)
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & (2**64 - 1)
        yield (z ^ (z >> 31)) >> 11


def rand_hex(seed, amp=0.15):
    g = split64(seed)
    X = []
    for a in range(8):
        base = [a & 1, (a >> 1) & 1, (a >> 2) & 1]
        X.append([base[c] * 0.7 + (2 * next(g) / 2**53 - 1) * amp for c in range(3)])
    return np.array(X)


def rand_tet(seed, amp=0.15):
    g = split64(seed)
    base = [[0, 0, 0], [1, 0, 0], [0, 1, 0], [0, 0, 1]]
    return np.array([[base[a][c] * 0.9 + (2 * next(g) / 2**53 - 1) * amp for c in range(3)] for a in range(4)])


def inv3(J):
    return mp.inverse(J), mp.det(J)


def hex_grad(X, xi, eta, zeta):
    dN = mp.matrix(3, 8)
    N = []
    for a in range(8):
        sx, sy, sz = (1 if a & 1 else -1), (1 if a & 2 else -1), (1 if a & 4 else -1)
        fx, fy, fz = 1 + sx * xi, 1 + sy * eta, 1 + sz * zeta
        N.append(fx * fy * fz / 8)
        dN[0, a], dN[1, a], dN[2, a] = sx * fy * fz / 8, sy * fx * fz / 8, sz * fx * fy / 8
    Xm = mp.matrix([[mp.mpf(float(v)) for v in row] for row in X])
    J = dN * Xm
    Ji, det = inv3(J)
    G = Ji * dN
    return N, G, det


def iso_ke(G, npe, w):
    nd = 3 * npe
    K = [[mp.mpf(0)] * nd for _ in range(nd)]
    for a in range(npe):
        for b in range(npe):
            gg = sum(G[r, a] * G[r, b] for r in range(3))
            for i in range(3):
                for j in range(3):
                    v = LAM * G[i, a] * G[j, b] + MU * G[j, a] * G[i, b] + (MU * gg if i == j else 0)
                    K[i * npe + a][j * npe + b] += w * v
    return K


def hex_ke(X):
    K = None
    for i in range(3):
        for j in range(3):
            for k in range(3):
                _, G, det = hex_grad(X, GP[i], GP[j], GP[k])
                Kg = iso_ke(G, 8, GW[i] * GW[j] * GW[k] * det)
                K = Kg if K is None else [[K[r][c] + Kg[r][c] for c in range(24)] for r in range(24)]
    return np.array([[float(v) for v in row] for row in K])


def hex_kehel(X, rmin):
    r2 = (mp.mpf(rmin) / (2 * mp.sqrt(3))) ** 2
    KF = [[mp.mpf(0)] * 8 for _ in range(8)]
    T = [mp.mpf(0)] * 8
    for i in range(3):
        for j in range(3):
            for k in range(3):
                N, G, det = hex_grad(X, GP[i], GP[j], GP[k])
                w = GW[i] * GW[j] * GW[k] * det
                for a in range(8):
                    T[a] += w * N[a]
                    for b in range(8):
                        KF[a][b] += w * (r2 * sum(G[q, a] * G[q, b] for q in range(3)) + N[a] * N[b])
    return np.array([[float(v) for v in row] for row in KF]), np.array([float(v) for v in T])


def tet_parts(X):
    Xm = [[mp.mpf(float(v)) for v in row] for row in X]
    J = mp.matrix([[Xm[r + 1][c] - Xm[0][c] for c in range(3)] for r in range(3)])
    dN = mp.matrix([[-1, 1, 0, 0], [-1, 0, 1, 0], [-1, 0, 0, 1]])
    Ji, det = inv3(J)
    return Ji * dN, det


def tet_ke(X):
    G, det = tet_parts(X)
    K = iso_ke(G, 4, abs(det) / 6)
    return np.array([[float(v) for v in row] for row in K])


def tet_kehel(X, rmin):
    r2 = (mp.mpf(rmin) / (2 * mp.sqrt(3))) ** 2
    G, det = tet_parts(X)
    a = (5 + 3 * mp.sqrt(5)) / 20
    b = (5 - mp.sqrt(5)) / 20
    pts = [(a, b, b), (b, a, b), (b, b, a), (b, b, b)]
    KF = [[mp.mpf(0)] * 4 for _ in range(4)]
    T = [mp.mpf(0)] * 4
    for s, t, v in pts:
        N = [1 - s - t - v, s, t, v]
        w = det / 24
        for p in range(4):
            T[p] += w * N[p]
            for q in range(4):
                KF[p][q] += w * (r2 * sum(G[r, p] * G[r, q] for r in range(3)) + N[p] * N[q])
    return np.array([[float(v) for v in row] for row in KF]), np.array([float(v) for v in T])


def hex_vol(X):
    c = [[mp.mpf(float(v)) for v in row] for row in X]

    def t6(a, b, cc, o):
        u = [o[i] - a[i] for i in range(3)]
        v = [o[i] - b[i] for i in range(3)]
        w = [o[i] - cc[i] for i in range(3)]
        cr = [v[1] * w[2] - v[2] * w[1], v[2] * w[0] - v[0] * w[2], v[0] * w[1] - v[1] * w[0]]
        return (u[0] * cr[0] + u[1] * cr[1] + u[2] * cr[2]) / 6

    p = [t6(c[0], c[1], c[3], c[5]), t6(c[0], c[3], c[2], c[6]), t6(c[5], c[4], c[6], c[0]),
         t6(c[5], c[6], c[7], c[3]), t6(c[0], c[3], c[6], c[5])]
    return float(sum(abs(v) for v in p))


def main():
    hexes = [rand_hex(s) for s in range(1, 7)]
    tets = [rand_tet(s) for s in range(11, 15)]
    rmin = 0.3
    out = dict(
        hex_X=np.array(hexes), hex_KE=np.array([hex_ke(X) for X in hexes]),
        hex_KF=np.array([hex_kehel(X, rmin)[0] for X in hexes]),
        hex_T=np.array([hex_kehel(X, rmin)[1] for X in hexes]),
        hex_vol=np.array([hex_vol(X) for X in hexes]),
        tet_X=np.array(tets), tet_KE=np.array([tet_ke(X) for X in tets]),
        tet_KF=np.array([tet_kehel(X, rmin)[0] for X in tets]),
        tet_T=np.array([tet_kehel(X, rmin)[1] for X in tets]),
        rmin=np.array(rmin), nu=np.array(0.3))
    np.savez_compressed(os.path.join(HERE, "elements.npz"), **out)
    print("wrote", os.path.join(HERE, "elements.npz"))


if __name__ == "__main__":
    main()
