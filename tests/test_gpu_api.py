"""GPU: the C-ABI boundary — reference text formats, error handling, the
drop-in driver binary, determinism and the VTU writer."""
import os
import subprocess

import numpy as np
import pytest

import conftest

pytestmark = pytest.mark.gpu


def _write_ws(path, values, fmt):
    # whitespace separated, no trailing newline (like the reference's BC fixtures)
    with open(path, "w") as f:
        f.write(" ".join(fmt % v for v in values))


def _ctx(**prm):
    from uft.capi import Context
    c = Context(0)
    c.set_params(Context.default_params(**prm))
    return c


def test_text_readers_match_arrays(O, tmp_path):
    import uft
    print("[case] text readers", flush=True)
    nx, ny, nz = 6, 3, 3
    xyz, elem = uft.gen_box(8, nx, ny, nz, jitter=0.1, seed=5)
    fixed, ld, lv = uft.box_bcs(nx, ny, nz)
    _write_ws(tmp_path / "vertex.txt", xyz.ravel(), "%.17g")
    _write_ws(tmp_path / "element.txt", elem.ravel(), "%d")
    _write_ws(tmp_path / "fixeddofs.txt", fixed, "%d")
    _write_ws(tmp_path / "loaddof_1.txt", ld, "%d")
    _write_ws(tmp_path / "load_1.txt", lv, "%.17g")
    prm = dict(Emax=1.0, volfrac=0.3, rmin=0.5, rtol_el=1e-12, rtol_f=1e-12)
    a = _ctx(**prm)
    L = a.L
    assert L.uft_mesh_load_text(a.h, 8, len(xyz), len(elem), str(tmp_path / "vertex.txt").encode(),
                                str(tmp_path / "element.txt").encode()) == 0
    assert L.uft_bc_load_text(a.h, str(tmp_path).encode(), len(fixed), 1, len(ld)) == 0
    a.setup()
    b = _ctx(**prm)
    b.set_mesh(elem, xyz)
    b.set_bcs(fixed, ld, lv)
    b.setup()
    e2 = np.zeros(elem.size, np.int32)
    x2 = np.zeros(xyz.size)
    assert L.uft_mesh_get(a.h, e2, x2) == 0
    assert np.array_equal(e2, elem.ravel()) and np.array_equal(x2, xyz.ravel())
    fa, ga = a.fem_obj_sens()
    fb, gb = b.fem_obj_sens()
    assert fa == fb and ga == gb


def test_io_errors(tmp_path):
    print("[case] io errors", flush=True)
    a = _ctx()
    rc = a.L.uft_mesh_load_text(a.h, 8, 10, 1, str(tmp_path / "nope.txt").encode(), b"x")
    assert rc == 2  # UFT_ERR_IO
    assert b"cannot open" in a.L.uft_last_error(a.h)
    _write_ws(tmp_path / "v.txt", [0.0] * 5, "%g")
    _write_ws(tmp_path / "e.txt", [0] * 8, "%d")
    rc = a.L.uft_mesh_load_text(a.h, 8, 8, 1, str(tmp_path / "v.txt").encode(), str(tmp_path / "e.txt").encode())
    assert rc == 2 and b"expected" in a.L.uft_last_error(a.h)


def test_state_and_argument_errors():
    from uft.capi import UFTError
    print("[case] state errors", flush=True)
    a = _ctx()
    with pytest.raises(UFTError) as e:
        a.fem_obj_sens()
    assert e.value.code == 7  # UFT_ERR_STATE
    elem = np.array([[0, 1, 2, 3, 4, 5, 6, 99]], np.int32)
    with pytest.raises(UFTError) as e:
        a.set_mesh(elem, np.zeros((8, 3)))
    assert e.value.code == 1
    from uft.capi import Context
    for bad in (dict(cheb_degree=0), dict(cheb_degree=17), dict(cheb_degree_coarse=-1), dict(cycle_type=4),
                dict(precision=3), dict(graph=2), dict(esteig_steps=1), dict(cheb_hi=0.05)):
        with pytest.raises(UFTError) as e:
            a.set_params(Context.default_params(**bad))
        assert e.value.code == 1, bad
    # design batches: offsets must cover the mesh, and elements stay inside their design
    from uft.capi import batch_mesh
    ds = [dict(nx=4, ny=2, nz=2, volfrac=0.3, seed=1), dict(nx=4, ny=2, nz=2, volfrac=0.3, seed=2)]
    el, xyz, fixed, ld, lv, eoff, noff = batch_mesh(ds)
    b = _ctx()
    b.set_mesh(el, xyz)
    for eo, no, vf in ((eoff, noff[::-1].copy(), [0.3, 0.3]), ([0, 10, len(el)], noff, [0.3, 0.3]),
                       (eoff, noff, [0.3, 1.5])):
        with pytest.raises(UFTError) as e:
            b.batch_set(eo, no, vf, [0.1, 0.1])
        assert e.value.code == 1


def test_inverted_element_detected():
    import uft
    from uft.capi import UFTError
    print("[case] inverted element", flush=True)
    xyz, elem = uft.gen_box(8, 2, 1, 1)
    elem = elem.copy()
    elem[0] = elem[0][[1, 0, 3, 2, 5, 4, 7, 6]]  # mirrored node order: det J < 0
    fixed, ld, lv = uft.box_bcs(2, 1, 1)
    a = _ctx()
    a.set_mesh(elem, xyz)
    a.set_bcs(fixed, ld, lv)
    with pytest.raises(UFTError) as e:
        a.setup()
    assert e.value.code == 4  # UFT_ERR_DETJ


def test_nonconvergence_is_reported():
    from uft.capi import cantilever, UFTError
    print("[case] maxit 1", flush=True)
    ctx = cantilever(8, 8, 4, 4, jitter=0.1, seed=2, Emax=1.0, rtol_el=1e-12, maxit_el=1)
    with pytest.raises(UFTError) as e:
        ctx.fem_obj_sens()
    assert e.value.code == 3  # UFT_ERR_NOCONV
    fx, gx = ctx.fem_obj_sens(allow_noconv=True)
    assert np.isfinite(fx) and ctx.stats()["it_el"] == 1


def test_bitwise_reproducible():
    from uft.capi import cantilever
    print("[case] determinism", flush=True)
    out = []
    for _ in range(2):
        ctx = cantilever(8, 10, 5, 5, jitter=0.15, seed=4, Emax=1.0, volfrac=0.3, rmin=0.3, rtol_el=1e-8,
                         rtol_f=1e-8)
        ctx.init_design()
        r = [ctx.design_iteration(it) for it in (1, 2, 3)]
        out.append((r, ctx.get("X"), ctx.get("U")))
        ctx.close()
    assert out[0][0] == out[1][0]
    assert np.array_equal(out[0][1], out[1][1]) and np.array_equal(out[0][2], out[1][2])


def test_dropin_driver_runs(O, tmp_path):
    """bin/uft_opt (the reference's main.cpp sequence over the drop-in classes)."""
    exe = os.path.join(conftest.PKG, "bin", "uft_opt")
    print("[case] drop-in driver", flush=True)
    args = [exe, "-box", "8,4,4", "-jitter", "0.15", "-seed", "7", "-Emax", "1", "-volfrac", "0.3", "-rmin", "0.375",
            "-maxItr", "5", "-ksp_rtol", "1e-12", "-vtu_dir", str(tmp_path)]
    res = subprocess.run(args, capture_output=True, text=True, timeout=120)
    assert res.returncode == 0, res.stderr
    lines = [l for l in res.stdout.splitlines() if l.startswith("It.:")]
    assert len(lines) == 5
    ref = O.cantilever(8, 8, 4, 4, jitter=0.15, seed=7, Emax=1.0, volfrac=0.3, rmin=0.375, rtol_el=1e-12,
                       rtol_f=1e-12)
    n, fx, gx, ch = ref.run(5)
    for line, f in zip(lines, fx):
        got = float(line.split("fx:")[1].split(",")[0])
        assert abs(got - f) <= 1e-6 * abs(f) + 1e-6  # the log line prints %f
    vtus = sorted(os.listdir(tmp_path))
    # the reference's names (Write2Vtu.cpp:35-46): itr 0 initial, -1 final
    assert "output_initial.vtu" in vtus and "output_final.vtu" in vtus
    # the same run through the binary mesh cache: written on the first run,
    # read on the second, identical log
    cache = str(tmp_path / "box.uftc")
    runs = [subprocess.run(args + ["-mesh_cache", cache], capture_output=True, text=True, timeout=120)
            for _ in range(2)]
    assert all(r.returncode == 0 for r in runs) and os.path.exists(cache)
    logs = [[l.split("time:")[0] for l in r.stdout.splitlines() if l.startswith("It.:")] for r in runs]
    assert logs[0] == logs[1] == [l.split("time:")[0] for l in lines]
    txt = open(tmp_path / "output_initial.vtu").read()
    assert 'NumberOfCells="128"' in txt and 'Name="xphys"' in txt


def test_binary_cache_roundtrip(O, tmp_path):
    """Text files -> context -> binary cache -> new context: identical mesh,
    BCs and results; corrupt and truncated caches are rejected."""
    import uft
    print("[case] binary cache", flush=True)
    nx, ny, nz = 6, 3, 3
    xyz, elem = uft.gen_box(8, nx, ny, nz, jitter=0.1, seed=9)
    fixed, ld, lv = uft.box_bcs(nx, ny, nz)
    prm = dict(Emax=1.0, volfrac=0.3, rmin=0.5, rtol_el=1e-12, rtol_f=1e-12)
    a = _ctx(**prm)
    a.set_mesh(elem, xyz)
    a.set_bcs(fixed, ld, lv)
    path = str(tmp_path / "mesh.uftc").encode()
    assert a.L.uft_cache_save(a.h, path) == 0
    a.setup()
    b = _ctx(**prm)
    assert b.L.uft_cache_load(b.h, path) == 0
    b.setup()
    e2 = np.zeros(elem.size, np.int32)
    x2 = np.zeros(xyz.size)
    assert b.L.uft_mesh_get(b.h, e2, x2) == 0
    assert np.array_equal(e2, elem.ravel()) and np.array_equal(x2, xyz.ravel())
    assert a.fem_obj_sens() == b.fem_obj_sens()
    raw = bytearray(open(path, "rb").read())
    raw[100] ^= 0x01
    bad = tmp_path / "bad.uftc"
    bad.write_bytes(bytes(raw))
    c = _ctx(**prm)
    assert c.L.uft_cache_load(c.h, str(bad).encode()) == 2 and b"checksum" in c.L.uft_last_error(c.h)
    bad.write_bytes(bytes(raw[:500]))
    assert c.L.uft_cache_load(c.h, str(bad).encode()) == 2 and b"truncated" in c.L.uft_last_error(c.h)
    bad.write_bytes(b"NOTACACHE" + bytes(64))
    assert c.L.uft_cache_load(c.h, str(bad).encode()) == 2 and b"bad header" in c.L.uft_last_error(c.h)


def _parse_vtu(path):
    import xml.etree.ElementTree as ET
    root = ET.parse(path).getroot()
    piece = root.find("UnstructuredGrid/Piece")
    arr = {}
    for da in root.iter("DataArray"):
        arr[da.get("Name")] = np.array(da.text.split(), dtype=float)
    return int(piece.get("NumberOfPoints")), int(piece.get("NumberOfCells")), arr


@pytest.mark.parametrize("npe", [8, 4])
def test_write_vtu_from_device(O, tmp_path, npe):
    """uft_write_vtu (Write2Vtu::WriteResults) from the device vectors after
    a design iteration: points, hex corners in VTK order 0,1,3,2,4,5,7,6
    (Write2Vtu.cpp:89-96), offsets, types 12/10, ux/uy/uz and xphys equal to
    the context's vectors (17 digits round-trip exactly; the reference's 6
    digits to 5e-7 relative)."""
    from uft.capi import cantilever
    print(f"[case] vtu npe={npe}", flush=True)
    dims = (5, 3, 2)
    ctx = cantilever(npe, *dims, jitter=0.15, seed=2, load_kind=0 if npe == 8 else 1, Emax=1.0, volfrac=0.3,
                     rmin=0.5, rtol_el=1e-12, rtol_f=1e-12)
    ctx.init_design()
    ctx.design_iteration(1)
    U, XP = ctx.get("U"), ctx.get("XPhys")
    xyz, elem = O.gen_box(npe, *dims, jitter=0.15, seed=2)
    for digits, tol in ((17, 0.0), (6, 5e-7)):
        path = tmp_path / f"d{digits}.vtu"
        ctx.write_vtu(path, digits=digits)
        nn, ne, a = _parse_vtu(path)
        assert (nn, ne) == (len(xyz), len(elem))
        close = (lambda g, w: np.array_equal(g, w)) if tol == 0 else \
            (lambda g, w: np.all(np.abs(g - w) <= tol * np.abs(w) + 1e-300))
        assert close(a["nodes"], xyz.ravel())
        order = [0, 1, 3, 2, 4, 5, 7, 6] if npe == 8 else [0, 1, 2, 3]
        assert np.array_equal(a["connectivity"].reshape(ne, npe), elem[:, order])
        assert np.array_equal(a["offsets"], npe * np.arange(1, ne + 1))
        assert np.all(a["types"] == (12 if npe == 8 else 10))
        for c, name in enumerate(("ux", "uy", "uz")):
            assert close(a[name], U[c::3]), name
        assert close(a["xphys"], XP)
    ctx.close()
