"""GPU: the multi-load-case form (separate right-hand sides solved together,
summed compliance) against the oracle, and the multi-load-case operator
kernels -- the FP64 vector kernel and the BF16 matrix-core element product --
against the oracle's assembled operator."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

PRM = dict(Emax=1.0, volfrac=0.3, rmin=0.3, rtol_el=1e-12, rtol_f=1e-12, load_cases=1)


def _pair(O, npe, dims, precision, seed=3):
    from uft.capi import cantilever
    ctx = cantilever(npe, *dims, jitter=0.15, seed=seed, load_kind=2, precision=precision, **PRM)
    prm = {k: v for k, v in PRM.items() if k != "load_cases"}
    ref = O.cantilever(npe, *dims, jitter=0.15, seed=seed, load_kind=2, **prm)
    ref.set_load_separate(True)
    return ctx, ref


@pytest.mark.parametrize("npe,dims,precision", [(8, (8, 4, 4), 0), (8, (8, 4, 4), 1), (8, (8, 4, 4), 2),
                                                (4, (5, 3, 3), 1), (4, (5, 3, 3), 2)])
def test_operator_kernels_vs_oracle(O, npe, dims, precision):
    """kind 1: the cycle's fine operator -- BF16 matrix-core element products
    (precision 1) or the assembled FP16 node blocks (precision 2)."""
    print(f"[case] multi-case operator npe={npe} precision={precision}", flush=True)
    ctx, ref = _pair(O, npe, dims, precision)
    rng = np.random.default_rng(11)
    x = rng.uniform(0.05, 1.0, ref.ne)
    ctx.set("XPhys", x)
    U = rng.standard_normal((8, 3 * ref.nn))
    want = np.stack([ref.apply_K(x, U[k]) for k in range(8)])
    got = ctx.apply_K_cases(U, kind=0)
    assert np.abs(got - want).max() <= 1e-12 * np.abs(want).max()
    if precision:
        # BF16 operands (8-bit significands) or FP16 blocks (11-bit), FP32
        # accumulation: entrywise error of a few 2^-8 relative to the row's magnitude
        got = ctx.apply_K_cases(U, kind=1)
        scale = np.abs(want).max()
        err = np.abs(got - want).max() / scale
        print("bf16 relative error", err)
        assert err <= 5e-3
        # and not a layout accident: the product tracks the operator
        assert np.corrcoef(got.ravel(), want.ravel())[0, 1] > 0.9999


@pytest.mark.parametrize("precision", [0, 2])
def test_obj_sens_vs_oracle(O, precision):
    print(f"[case] multi-case obj/sens precision={precision}", flush=True)
    ctx, ref = _pair(O, 8, (8, 4, 4), precision)
    x = np.random.default_rng(5).uniform(0.05, 1.0, ref.ne)
    ctx.set("XPhys", x)
    fx, gx = ctx.fem_obj_sens()
    rfx, rgx, rdf, rdg, _ = ref.fem_obj_sens(x)
    assert abs(fx - rfx) <= 1e-10 * abs(rfx)
    assert abs(gx[0] - rgx[0]) <= 1e-12
    df = ctx.get("dfdx")
    assert np.abs(df - rdf).max() <= 1e-8 * np.abs(rdf).max()
    stats = ctx.load_case_stats()
    assert len(stats) == 8 and all(rr <= 1e-12 for _, rr in stats)
    for k in (0, 5):
        uk = ctx.load_case_u(k)
        assert np.abs(uk - ref.case_u(k)).max() <= 1e-8 * np.abs(ref.case_u(k)).max()


def test_design_loop_vs_oracle_mixed(O):
    print("[case] multi-case 10-iteration loop, mixed", flush=True)
    ctx, ref = _pair(O, 8, (10, 5, 5), 2, seed=13)
    ctx.init_design()
    ref.init_design()
    worst = 0.0
    for it in range(1, 11):
        fx, gx, ch = ctx.design_iteration(it)
        rfx, rgx, rch, _ = ref.design_iteration(it)
        worst = max(worst, abs(fx - rfx) / abs(rfx))
    print("worst relative compliance difference", worst)
    assert worst <= 1e-6
    assert np.max(np.abs(ctx.get("X") - ref.vec("X"))) <= 1e-6


@pytest.mark.parametrize("precision", [0, 2])
def test_three_cases_padded(precision):
    """Three load cases run in a block of four; the padded case is inert
    (also under the mixed forms' per-case reliable updates)."""
    import uft
    from uft.capi import Context
    print(f"[case] three cases precision={precision}", flush=True)
    nx, ny, nz = 8, 4, 4
    xyz, elem = uft.gen_box(8, nx, ny, nz, jitter=0.1, seed=1)
    fixed, ld, lv = uft.box_bcs(nx, ny, nz, load_kind=2)
    one = Context(0)
    one.set_params(Context.default_params(precision=precision,
                                          **{k: v for k, v in PRM.items() if k != "load_cases"}))
    one.set_mesh(elem, xyz)
    one.set_bcs(fixed, ld[9:18], lv[9:18])
    one.setup()
    blk = Context(0)
    blk.set_params(Context.default_params(precision=precision, **PRM))
    blk.set_mesh(elem, xyz)
    blk.set_bcs(fixed, ld[:27], lv[:27], nlc=3)
    blk.setup()
    x = np.random.default_rng(2).uniform(0.1, 1.0, nx * ny * nz)
    one.set("XPhys", x)
    blk.set("XPhys", x)
    one.fem_obj_sens()
    blk.fem_obj_sens()
    stats = blk.load_case_stats()
    assert len(stats) == 3 and all(rr <= 1e-12 for _, rr in stats)
    u1 = blk.load_case_u(1)
    assert np.abs(u1 - one.get("U")).max() <= 1e-9 * np.abs(u1).max()


def test_node_tiles_match_node_order():
    """The eight-case fine products on node tiles (halo entries staged in
    LDS, fine_assembled = 1) against the node-ordered kernels
    (fine_assembled = 2) on a mesh of many tiles: the same FP16 / FP32 block
    products per node in the same slot order, so the operator agrees to FP32
    rounding and a whole mixed solve to its tolerance."""
    from uft.capi import cantilever
    print("[case] node tiles vs node order, 24x12x12 hex8, 8 cases", flush=True)
    prm = dict(PRM, precision=2, rtol_el=1e-10, rtol_f=1e-10)
    tiled = cantilever(8, 24, 12, 12, jitter=0.15, seed=21, load_kind=2, fine_assembled=1, **prm)
    plain = cantilever(8, 24, 12, 12, jitter=0.15, seed=21, load_kind=2, fine_assembled=2, **prm)
    ne = 24 * 12 * 12
    rng = np.random.default_rng(4)
    x = rng.uniform(0.05, 1.0, ne)
    tiled.set("XPhys", x)
    plain.set("XPhys", x)
    U = rng.standard_normal((8, tiled.size("ndof")))
    a = tiled.apply_K_cases(U, kind=1)
    b = plain.apply_K_cases(U, kind=1)
    err = np.abs(a - b).max() / np.abs(b).max()
    print("tiled vs node-ordered FP16 smoother product:", err)
    assert err <= 1e-6
    fa, ga = tiled.fem_obj_sens()
    fb, gb = plain.fem_obj_sens()
    print("compliance", fa, fb)
    assert abs(fa - fb) <= 1e-9 * abs(fb)
    da, db = tiled.get("dfdx"), plain.get("dfdx")
    assert np.abs(da - db).max() <= 1e-7 * np.abs(db).max()
    tiled.close()
    plain.close()
