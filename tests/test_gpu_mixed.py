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


@pytest.mark.parametrize("shuffle", [False, True])
def test_assembled_fine_operator_any_numbering(shuffle):
    """The assembled fine operators (FP16 smoother, FP32 outer product) give
    the element form's compliance with the lexicographic node numbering (16-bit
    column offsets) and with a random one (neighbours more than 32767 nodes
    apart: 32-bit columns)."""
    import uft
    from uft.capi import Context
    nx, ny, nz = 60, 24, 24
    print(f"[case] assembled fine operator, shuffled numbering={shuffle}", flush=True)
    xyz, elem = uft.gen_box(8, nx, ny, nz, jitter=0.1, seed=4)
    fixed, ld, lv = uft.box_bcs(nx, ny, nz)
    if shuffle:  # node n becomes perm[n]
        perm = np.random.default_rng(8).permutation(len(xyz)).astype(np.int32)
        xyz2 = np.empty_like(xyz)
        xyz2[perm] = xyz
        xyz, elem = xyz2, perm[elem]
        remap = lambda d: 3 * perm[np.asarray(d) // 3] + np.asarray(d) % 3
        fixed, ld = np.sort(remap(fixed)), remap(ld)
    res = {}
    for fa in (1, 0):
        ctx = Context(0)
        ctx.set_params(Context.default_params(Emax=1.0, volfrac=0.3, rmin=2.0 * 1.5 / nx, rtol_el=1e-10,
                                              rtol_f=1e-10, precision=2, fine_assembled=fa))
        ctx.set_mesh(elem, xyz)
        ctx.set_bcs(fixed, ld, lv)
        ctx.setup()
        info = ctx.fine_operator_info()
        if fa:
            assert info[2] == (4 if shuffle else 2) and info[1] > 0 and info[0] >= info[1] and info[3] == 12
        else:
            assert info == (0, 0, 0, 0)
        ctx.init_design()
        res[fa] = [ctx.design_iteration(it)[0] for it in (1, 2)]
        assert ctx.stats()["relres_el"] <= 1e-10
        ctx.close()
    print("compliances assembled / element form", res[1], res[0])
    for a, b in zip(res[1], res[0]):
        assert abs(a - b) <= 1e-8 * abs(b)


def test_mixed_tet4_loop_vs_oracle(O):
    """UFT_MIXED_H on a tet4 mesh (assembled blocks from 12x12 element
    matrices): compliance within 1e-6 of the FP64 oracle over 6 iterations."""
    from uft.capi import cantilever
    args = dict(nx=6, ny=3, nz=3, jitter=0.15, seed=17)
    prm = dict(Emax=1.0, volfrac=0.25, rmin=0.5, rtol_el=1e-12, rtol_f=1e-12)
    print("[case] mixed tet4 6x3x3 vs oracle", flush=True)
    ctx = cantilever(4, **args, precision=2, load_kind=1, **prm)
    ref = O.cantilever(4, **args, load_kind=1, **prm)
    assert ctx.fine_operator_info()[1] > 0
    ctx.init_design()
    ref.init_design()
    for it in range(1, 7):
        fx = ctx.design_iteration(it)[0]
        rfx = ref.design_iteration(it)[0]
        assert abs(fx - rfx) <= 1e-6 * abs(rfx), it
    assert ctx.stats()["relres_el"] <= 1e-12
    assert np.max(np.abs(ctx.get("X") - ref.vec("X"))) <= 1e-6


def test_time_fine_operator_kinds():
    """The back-to-back timing entry point accepts exactly the kinds of the
    context's form, rejects a call before the first refresh, and returns a
    positive duration; the solve is unchanged by it (its vectors are
    rebuilt by every solve)."""
    from uft.capi import cantilever, UFTError
    args = dict(nx=12, ny=6, nz=6, jitter=0.1, seed=5)
    prm = dict(Emax=1.0, volfrac=0.3, rmin=0.25, rtol_el=1e-12, rtol_f=1e-12)
    # UFT_MIXED's smoother runs on the assembled FP32 blocks, so kind 2 times its Chebyshev step
    for precision, ok, bad in ((0, (0,), (1, 2)), (1, (1, 2), (0,)), (2, (1, 2), (0,))):
        print(f"[case] time_fine_operator precision={precision}", flush=True)
        ctx = cantilever(8, **args, precision=precision, **prm)
        ctx.init_design()
        with pytest.raises(UFTError):
            ctx.time_fine_operator(ok[0], 3)
        ctx.design_iteration(1)
        f1 = ctx.fem_obj_sens()[0]
        for k in ok + ((3,) if precision == 0 else ()):
            assert ctx.time_fine_operator(k, 5) > 0.0
        for k in bad:
            with pytest.raises(UFTError):
                ctx.time_fine_operator(k, 5)
        f2 = ctx.fem_obj_sens()[0]
        assert abs(f2 - f1) <= 1e-9 * abs(f1)
        ctx.close()
