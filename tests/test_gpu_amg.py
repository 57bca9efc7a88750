"""GPU: the multigrid hierarchy behind the elasticity solve.

The coarse operators are checked against explicit Galerkin products
P^T A P built on the CPU from the oracle's assembled stiffness matrix (with
the reference's Dirichlet elimination, SolveFem.cpp:324-333), level by level.
The preconditioner only steers the Krylov iteration, so these checks pin the
hierarchy itself; the solution is pinned by the parity tests."""
import numpy as np
import pytest
import scipy.sparse as sp

pytestmark = pytest.mark.gpu


def _transfer_matrix(parent, P):
    nu, bs, _ = P.shape
    rows, cols, vals = [], [], []
    for u in range(nu):
        if parent[u] < 0:
            continue
        for i in range(bs):
            for j in range(6):
                rows.append(bs * u + i)
                cols.append(6 * parent[u] + j)
                vals.append(P[u, i, j])
    nc = 6 * (parent.max() + 1)
    return sp.csr_matrix((vals, (rows, cols)), shape=(nu * bs, nc))


def _bsr_dense(rp, ci, val):
    nu = len(rp) - 1
    D = np.zeros((6 * nu, 6 * nu))
    for A in range(nu):
        for k in range(rp[A], rp[A + 1]):
            B = ci[k]
            D[6 * A:6 * A + 6, 6 * B:6 * B + 6] = val[k]
    return D


def _unit_diag_rule(D):
    # zero diagonal entries (decoupled rank-deficient coarse columns) become 1
    d = np.diag(D).copy()
    d[d == 0.0] = 1.0
    np.fill_diagonal(D, d)
    return D


def _sa_matrix(pptr, pagg, P, nc):
    nn = len(pptr) - 1
    M = np.zeros((3 * nn, nc))
    for i in range(nn):
        for e in range(pptr[i], pptr[i + 1]):
            M[3 * i:3 * i + 3, 6 * pagg[e]:6 * pagg[e] + 6] = P[e]
    return M


@pytest.mark.parametrize("npe,dims,smooth,precision", [(8, (8, 4, 4), 0, 0), (4, (5, 3, 3), 0, 0),
                                                       (8, (8, 4, 4), 1, 0), (4, (5, 3, 3), 1, 0),
                                                       (8, (8, 4, 4), 1, 1), (8, (8, 4, 4), 1, 2)])
def test_galerkin_operators(O, npe, dims, smooth, precision):
    """Every coarse operator is P^T A P of the level above.  With smoothed
    aggregation (smooth = 1) level 0's prolongator must be the tentative one
    smoothed once with the current operator, (I - omega D^-1 K) P_tent
    (PCGAMG's prolongator smoothing, SolveFem.cpp:129-136), on the whole
    matrix (so its sparsity pattern misses nothing)."""
    from uft.capi import cantilever
    nx, ny, nz = dims
    print(f"[case] galerkin npe={npe} dims={dims} smooth={smooth} precision={precision}", flush=True)
    prm = dict(Emax=1.0, volfrac=0.3, rmin=0.3, rtol_el=1e-10, amg_levels=4, coarse_max=6)
    ctx = cantilever(npe, nx, ny, nz, jitter=0.15, seed=21, prolong_smooth=smooth, precision=precision, **prm)
    ref = O.cantilever(npe, nx, ny, nz, jitter=0.15, seed=21, **prm)
    rng = np.random.default_rng(5)
    x = rng.uniform(0.01, 1.0, ref.ne)
    ctx.set("XPhys", x)
    ctx.fem_obj_sens()
    nlev = len(ctx.solver_info()[0])
    assert nlev >= 3
    ndof = 3 * ref.nn
    K = np.column_stack([ref.apply_K(x, e) for e in np.eye(ndof)])
    A = K
    sa = ctx.amg_prolongator()
    assert (sa is not None) == bool(smooth)
    for lvl in range(1, nlev):
        parent, P = ctx.amg_transfer(lvl - 1)
        Pm = _transfer_matrix(parent, P).toarray()
        if lvl == 1 and sa is not None:
            pptr, pagg, Ps, omega = sa
            assert 0.0 < omega < 4.0 / 3.0 * 2.0
            Psm = _sa_matrix(pptr, pagg, Ps, Pm.shape[1])
            want_p = Pm - omega * (K @ Pm) / np.diag(K)[:, None]
            # the mixed forms assemble the smoothing operator from FP32 element blocks;
            # UFT_MIXED_H keeps the prolongator in FP16 (11-bit significands)
            tol_p = {0: 1e-12, 1: 1e-6, 2: 2e-3}[precision]
            assert np.abs(Psm - want_p).max() <= tol_p * np.abs(want_p).max()
            assert np.abs(Psm - Pm).max() > 1e-3  # really smoothed
            Pm = Psm
        want = _unit_diag_rule(Pm.T @ A @ Pm)
        got = _bsr_dense(*ctx.amg_operator(lvl))
        scale = np.abs(want).max()
        assert np.abs(got - got.T).max() <= 1e-13 * scale, "coarse operator not symmetric"
        tol = 1e-6 if (lvl == 1 and sa is not None and precision != 0) else 1e-12
        assert np.abs(got - want).max() <= tol * scale, lvl
        A = got


def test_transfer_columns_orthonormal():
    """Tentative prolongator: per aggregate the 6 columns are orthonormal
    (local QR of the rigid-body modes), zero rows at fixed DOFs."""
    from uft.capi import cantilever
    import uft
    print("[case] transfer orthonormal", flush=True)
    ctx = cantilever(8, 8, 4, 4, jitter=0.1, seed=3, Emax=1.0, coarse_max=6)
    parent, P = ctx.amg_transfer(0)
    fixed, _, _ = uft.box_bcs(8, 4, 4)
    Pm = _transfer_matrix(parent, P).toarray()
    assert np.all(Pm[fixed] == 0.0)
    G = Pm.T @ Pm
    d = np.diag(G)
    assert np.all((np.abs(d - 1.0) < 1e-12) | (d == 0.0))
    off = G - np.diag(d)
    assert np.abs(off).max() < 1e-12


@pytest.mark.parametrize("dims,coarse_max", [((8, 4, 4), 6), ((24, 12, 12), 600)])
def test_coarse_inverse(dims, coarse_max):
    """The blocked Gauss-Jordan inverse of the coarsest operator."""
    from uft.capi import cantilever
    print(f"[case] coarse inverse {dims}", flush=True)
    nx, ny, nz = dims
    ctx = cantilever(8, nx, ny, nz, jitter=0.15, seed=9, Emax=1.0, volfrac=0.3, coarse_max=coarse_max)
    rng = np.random.default_rng(2)
    ctx.set("XPhys", rng.uniform(0.01, 1.0, nx * ny * nz))
    ctx.fem_obj_sens()
    lvl = len(ctx.solver_info()[0]) - 1
    A = _bsr_dense(*ctx.amg_operator(lvl))
    inv = ctx.amg_coarse_inverse()
    n = A.shape[0]
    cond = np.linalg.cond(A)
    print("coarsest n", n, "cond", cond)
    tol = max(1e-12, 1e-14 * cond)  # backward-stable elimination: error ~ eps * cond
    want = np.linalg.inv(A)
    assert np.abs(inv - want).max() <= tol * np.abs(want).max()
    assert np.abs(inv @ A - np.eye(n)).max() <= tol


@pytest.mark.parametrize("assembled", [1, 0])
def test_level_operators_in_each_precision(assembled):
    """Every multigrid level's stored operator in FP32 and FP16 (UFT_MIXED_H:
    the fine operator assembled into 3x3 node blocks with rows scaled by the
    diagonal, stored as int8 with one FP32 scale per block, or the element
    blocks scaled per element in FP16; coarse levels above the coarsest scaled
    to unit diagonal) against the FP64 one, with a non-smooth random vector."""
    from uft.capi import cantilever
    print(f"[case] level operators FP64/FP32/FP16 fine_assembled={assembled}", flush=True)
    ctx = cantilever(8, 24, 12, 12, jitter=0.15, seed=9, Emax=1.0, volfrac=0.3, rmin=0.2, precision=2,
                     fine_assembled=assembled, coarse_max=600)
    x = np.random.default_rng(4).uniform(0.01, 1.0, ctx.size("ne"))
    ctx.set("XPhys", x)
    ctx.fem_obj_sens()
    n, _ = ctx.solver_info()
    assert len(n) >= 3
    rng = np.random.default_rng(5)
    for lev in range(len(n)):
        u = rng.standard_normal(n[lev])
        y64 = ctx.level_apply(lev, u, 0)
        scale = np.linalg.norm(y64)
        e32 = np.linalg.norm(ctx.level_apply(lev, u, 1) - y64) / scale
        print("level", lev, "FP32 error", e32)
        assert e32 < 1e-6
        if lev < len(n) - 1:  # FP16 blocks: every level but the coarsest (solved densely)
            e16 = np.linalg.norm(ctx.level_apply(lev, u, 2) - y64) / scale
            print("level", lev, "FP16 / int8 error", e16)
            if lev == 0 and assembled:
                # int8 grid of +-127 steps of the block's largest entry: rounding
                # of at most 1/254 of it per entry, a few 1e-3 over the rows
                assert e16 < 1e-2 and e16 > 1e-5
            else:
                assert e16 < 2e-3 and e16 > 1e-7  # FP16 rounding is visible, and only rounding


def test_coarse_sweep_inverse_full_size():
    """The tiled Gauss-Jordan sweep (kernels_dense.hip) that inverts the
    3,840-unknown coarsest level of the 392,000-element cantilever at every
    hierarchy refresh, against numpy's inverse of the same operator (read
    back through uft_amg_operator_get), on a random density field."""
    from uft.capi import cantilever
    print("[case] coarse sweep inverse, 140x70x40", flush=True)
    ctx = cantilever(8, 140, 70, 40, jitter=0.15, seed=392000, Emax=1.0, volfrac=0.3, rmin=1.5 / 70)
    n_lev, _ = ctx.solver_info()
    assert len(n_lev) == 3 and n_lev[-1] >= 512, n_lev
    rng = np.random.default_rng(4)
    ctx.set("XPhys", rng.uniform(0.01, 1.0, 140 * 70 * 40))
    ctx.fem_obj_sens()
    A = _bsr_dense(*ctx.amg_operator(2))
    inv = ctx.amg_coarse_inverse()
    n = A.shape[0]
    cond = np.linalg.cond(A)
    print("coarsest n", n, "cond", cond, flush=True)
    tol = max(1e-12, 1e-14 * cond)
    want = np.linalg.inv(A)
    assert np.array_equal(inv, inv.T)  # the sweep keeps the lower triangle only
    assert np.abs(inv - want).max() <= tol * np.abs(want).max()
    assert np.abs(inv @ A - np.eye(n)).max() <= tol


def test_fp64_node_blocks_match_element_form():
    """The FP64 form's solve operator (the refresh's assembled FP64 node
    blocks, k_fine_ell_op64) against the matrix-free element form of the
    same density (uft_apply_K): the same products summed in another order,
    so they agree to FP64 rounding."""
    from uft.capi import cantilever
    print("[case] FP64 node blocks vs element form, 24x12x12 hex8", flush=True)
    ctx = cantilever(8, 24, 12, 12, jitter=0.15, seed=9, Emax=1.0, volfrac=0.3, rmin=0.2, precision=0,
                     coarse_max=600)
    assert ctx.fine_operator_info()[3] == 72  # the solves apply FP64 node blocks
    x = np.random.default_rng(4).uniform(0.01, 1.0, ctx.size("ne"))
    ctx.set("XPhys", x)
    ctx.fem_obj_sens()  # a hierarchy refresh at this density
    u = np.random.default_rng(6).standard_normal(ctx.size("ndof"))
    yb = ctx.level_apply(0, u, 0)
    ye = ctx.apply_K(u)
    err = np.abs(yb - ye).max() / np.abs(ye).max()
    print("node blocks vs element form:", err)
    assert err <= 1e-13
    ctx.close()


def test_fp32_smoother_form_assembled_matches_element_form():
    """UFT_MIXED (the FP32 smoother of config 3 as written) on the assembled
    FP32 node blocks against its element-block form: the same compliance to
    the solver's tolerance over two design iterations."""
    from uft.capi import cantilever
    print("[case] UFT_MIXED assembled vs element blocks, 16x8x8 hex8", flush=True)
    res = {}
    for fa in (1, 0):
        ctx = cantilever(8, 16, 8, 8, jitter=0.15, seed=5, Emax=1.0, volfrac=0.3, rmin=0.2, rtol_el=1e-10,
                         rtol_f=1e-10, precision=1, fine_assembled=fa)
        assert ctx.fine_operator_info()[3] == (36 if fa else 0)
        ctx.init_design()
        res[fa] = [ctx.design_iteration(it)[0] for it in (1, 2)]
        assert ctx.stats()["relres_el"] <= 1e-10
        ctx.close()
    print("compliances assembled / element form", res[1], res[0])
    for a, b in zip(res[1], res[0]):
        assert abs(a - b) <= 1e-8 * abs(b)
