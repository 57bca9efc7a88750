"""CPU checks of the C-ABI library: it loads, exports every function that
include/uft.h declares, and its host-side workload generator is bit-exact with
the oracle's.  No GPU compute is called here."""
import os
import subprocess

import numpy as np
import pytest

import conftest


@pytest.fixture(scope="module")
def U():
    import uft
    if not os.path.exists(uft.lib_path()):
        subprocess.run(["make", "-s", "-j8", "-C", conftest.PKG, "lib/libuft.so"], check=True)
    uft.lib()
    return uft


def test_exports_every_declared_symbol(U):
    declared = U.header_symbols()
    assert len(declared) >= 30
    missing = [s for s in declared if s not in U.exported_symbols()]
    assert not missing, missing


def test_params_struct_layout_matches_header(U, tmp_path):
    """The ctypes mirror of uft_params has the C compiler's layout."""
    fields = [f for f, _ in U.Params._fields_]
    src = tmp_path / "layout.c"
    src.write_text('#include <stdio.h>\n#include <stddef.h>\n#include "uft.h"\nint main(void){\n'
                   + 'printf("%zu\\n", sizeof(uft_params));\n'
                   + "".join(f'printf("%zu\\n", offsetof(uft_params, {f}));\n' for f in fields) + "return 0;}\n")
    exe = tmp_path / "layout"
    subprocess.run(["gcc", "-I", os.path.join(conftest.ROOT, "include"), str(src), "-o", str(exe)], check=True)
    out = [int(v) for v in subprocess.run([str(exe)], capture_output=True, text=True, check=True).stdout.split()]
    import ctypes
    assert out[0] == ctypes.sizeof(U.Params)
    assert out[1:] == [getattr(U.Params, f).offset for f in fields]


@pytest.mark.parametrize("npe,dims,seed", [(8, (8, 4, 4), 7), (8, (140, 70, 40), 392000), (4, (60, 30, 30), 324000),
                                           (4, (5, 3, 2), 1)])
def test_generator_bit_exact(U, O, npe, dims, seed):
    a = U.gen_box(npe, *dims, jitter=0.15, seed=seed)
    b = O.gen_box(npe, *dims, jitter=0.15, seed=seed)
    assert np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1])


@pytest.mark.parametrize("kind,dims", [(0, (140, 70, 40)), (0, (40, 20, 20)), (1, (60, 30, 30))])
def test_bcs_bit_exact(U, O, kind, dims):
    for x, y in zip(U.box_bcs(*dims, load_kind=kind), O.box_bcs(*dims, load_kind=kind)):
        assert np.array_equal(x, y)


def test_bad_generator_arguments(U):
    L = U.lib()
    xyz = np.zeros(3)
    el = np.zeros(8, np.int32)
    assert L.uft_gen_box(8, 0, 1, 1, 2.0, 1.0, 1.0, 0.0, 0, xyz, el) == 1
    assert L.uft_gen_box(5, 1, 1, 1, 2.0, 1.0, 1.0, 0.0, 0, xyz, el) == 1


def test_dropin_driver_links(U):
    exe = os.path.join(conftest.PKG, "bin", "uft_opt")
    if not os.path.exists(exe):
        subprocess.run(["make", "-s", "-j8", "-C", conftest.PKG], check=True)
    assert os.access(exe, os.X_OK)


def test_dropin_mma_restart_kkt_stress_compile(tmp_path):
    """The drop-in classes expose the reference's MMA restart constructors,
    Restart, SetAsymptotes, ConstraintModification, SetRobustAsymptotesType,
    KKTresidual (MMA.h:8-52) and SolveFem::ComputeStress."""
    src = tmp_path / "use.cpp"
    src.write_text(r'''
#include "uft_dropin.h"
int use(SolveFem &fem, Vec X, Vec dfdx, Vec *dgdx, Vec xmin, Vec xmax, MeshRef e) {
  const long long n = 16;
  Vec xo1 = VecCreateHost(n), xo2 = VecCreateHost(n), U = VecCreateHost(n), L = VecCreateHost(n);
  MMA a(n, 1, X);
  a.Restart(xo1, xo2, U, L);
  MMA b(n, 1, a.RestartIteration(), xo1, xo2, U, L);
  double ac[1] = {0}, cc[1] = {1000}, dc[1] = {0}, gx[1] = {0}, n2 = 0, ni = 0;
  MMA c(n, 1, 5, xo1, xo2, U, L, ac, cc, dc);
  b.SetAsymptotes(0.5, 0.7, 1.2);
  b.ConstraintModification(true);
  b.SetRobustAsymptotesType(1);
  b.KKTresidual(X, dfdx, gx, dgdx, xmin, xmax, &n2, &ni);
  return fem.ComputeStress(X, 1.0, e, e, e, e) + (int)n2 + (int)ni;
}
''')
    inc = [os.path.join(conftest.ROOT, "include"), os.path.join(conftest.PKG, "dropin")]
    subprocess.run(["g++", "-std=c++17", "-fsyntax-only"] + [f"-I{d}" for d in inc] + [str(src)], check=True)


def test_uneven_design_draw(U):
    """The 32 uneven cantilevers (seed 32): every size inside [8000, 64000]
    elements, nx ~ 2 ny ~ 2 nz, volfrac in [0.2, 0.4); the draw is fixed."""
    from uft.capi import uneven_designs, SplitMix64
    ds = uneven_designs()
    assert len(ds) == 32
    ne = [D["nx"] * D["ny"] * D["nz"] for D in ds]
    assert all(8000 <= v <= 64000 for v in ne)
    assert all(abs(D["nx"] - 2 * D["ny"]) <= 2 and abs(D["nz"] - D["ny"]) <= 1 for D in ds)
    assert all(0.2 <= D["volfrac"] < 0.4 for D in ds)
    assert sum(ne) == 868008 and (ds[0]["nx"], ds[0]["ny"], ds[0]["nz"]) == (46, 23, 24)
    # SplitMix64 known answer (seed 0: first output of the published generator)
    assert SplitMix64(0).next() == 0xE220A8397B1DCDAF


def test_batch_mesh_layout(U):
    """Concatenated batch meshes: each design's elements use only its own
    nodes, its clamped DOFs are its x = 0 nodes, its loads its own tip edge."""
    from uft.capi import batch_mesh
    ds = [dict(nx=6, ny=3, nz=3, volfrac=0.3, seed=1), dict(nx=4, ny=2, nz=2, volfrac=0.3, seed=2)]
    elem, xyz, fixed, ld, lv, eoff, noff = batch_mesh(ds)
    assert list(eoff) == [0, 54, 70] and list(noff) == [0, 112, 157]
    for d in range(2):
        e = elem[eoff[d]:eoff[d + 1]]
        assert e.min() >= noff[d] and e.max() < noff[d + 1]
    fn = np.unique(fixed // 3)
    assert np.all(np.abs(xyz[fn, 0]) < 1e-12)
    assert len(fn) == 16 + 9
    assert len(ld) == 4 + 3 and np.all(np.abs(xyz[ld // 3, 0] - 2.0) < 1e-12)


REF_MESH = os.path.join(conftest.REFERENCE, "mesh")


@pytest.mark.skipif(not os.path.isdir(REF_MESH), reason="the reference's BC files are not on this machine")
def test_reader_on_reference_bc_files(U):
    """The library's host-only BC reader (InputMeshData::GetBcs,
    InputMeshData.cpp:228-285) on the reference's own mesh/fixeddofs.txt,
    loaddof_1.txt and load_1.txt (no line terminators): the counts of
    OptPara.cpp:9-10, the sums and orderings of the fixtures, clamped nodes as
    {3n, 3n+1, 3n+2} triples, y-only unit loads; equal to numpy's parse."""
    import hashlib
    from uft.capi import UFTError
    digests = {"fixeddofs.txt": "5d38111853e5aa3380ab51f2f2d18974832124b7b0d96c8fb019c8f95bbac4fe",
               "loaddof_1.txt": "8c8c871a9f4940b7010408f681dc3110618741b12a0802331c9e175eae8011bd",
               "load_1.txt": "fe9c5ba7692571b3e63ee50c2914080c70720f386e1210ba4faa94b514650cec"}
    for name, want in digests.items():
        assert hashlib.sha256(open(os.path.join(REF_MESH, name), "rb").read()).hexdigest() == want, name
    fixed, ld, lv = U.bc_read_text(REF_MESH, 8733, 1, 41)
    assert len(fixed) == 8733 and int(fixed.astype(np.int64).sum()) == 1073477334
    assert np.all(np.diff(fixed) > 0) and fixed[0] == 0 and fixed[-1] == 489047
    tri = fixed.reshape(-1, 3)
    assert np.all(tri[:, 0] % 3 == 0) and np.all(tri[:, 1] == tri[:, 0] + 1) and np.all(tri[:, 2] == tri[:, 0] + 2)
    assert list(tri[:7, 0] // 3) == [0, 2, 3, 7, 8, 9, 16]
    assert ld.shape == (1, 41) and int(ld.astype(np.int64).sum()) == 37262276
    assert np.all(ld % 3 == 1) and np.all(np.diff(ld[0]) > 0)
    assert np.all(lv == 1.0)
    assert np.array_equal(fixed, np.loadtxt(os.path.join(REF_MESH, "fixeddofs.txt"), dtype=np.int32))
    assert np.array_equal(ld[0], np.loadtxt(os.path.join(REF_MESH, "loaddof_1.txt"), dtype=np.int32))
    # one value more than the file holds: an I/O error naming the file, not a silent zero
    with pytest.raises(UFTError) as e:
        U.bc_read_text(REF_MESH, 8734, 1, 41)
    assert e.value.code == 2 and "fixeddofs.txt: expected 8734 values, read 8733" in str(e.value)
    with pytest.raises(UFTError) as e:
        U.bc_read_text(REF_MESH, 8733, 2, 41)  # no loaddof_2.txt
    assert e.value.code == 2 and "loaddof_2.txt" in str(e.value)


def test_mesh_reader_host_only(U, tmp_path):
    """uft_mesh_read_text on files in the reference's formats
    (InputMeshData.cpp:100-149: whitespace-separated, no terminator)."""
    from uft.capi import mesh_read_text, UFTError
    xyz, elem = U.gen_box(8, 5, 3, 2, jitter=0.15, seed=4)
    (tmp_path / "vertex.txt").write_text(" ".join(f"{v:.17g}" for v in xyz.ravel()))
    (tmp_path / "element.txt").write_text("\n".join(" ".join(str(v) for v in r) for r in elem))
    a, b = mesh_read_text(8, len(xyz), len(elem), tmp_path / "vertex.txt", tmp_path / "element.txt")
    assert np.array_equal(a, xyz) and np.array_equal(b, elem)
    with pytest.raises(UFTError) as e:
        mesh_read_text(8, len(xyz) + 1, len(elem), tmp_path / "vertex.txt", tmp_path / "element.txt")
    assert e.value.code == 2
    with pytest.raises(UFTError) as e:
        mesh_read_text(8, len(xyz), len(elem), tmp_path / "missing.txt", tmp_path / "element.txt")
    assert e.value.code == 2 and "cannot open" in str(e.value)


@pytest.mark.skipif(not os.path.exists(os.path.join(conftest.REFERENCE, "main.cpp")),
                    reason="the reference's main.cpp is not on this machine")
def test_reference_main_compiles_against_dropin(tmp_path):
    """The reference's own driver (main.cpp, read at test time, not kept in
    this repository) with its #include lines replaced by the drop-in header
    compiles and links against libuft.so unchanged otherwise."""
    lines = open(os.path.join(conftest.REFERENCE, "main.cpp")).read().splitlines()
    out, done = [], False
    for ln in lines:
        if ln.lstrip().startswith("#include"):
            if not done:
                out.append('#include "uft_dropin.h"')
                done = True
            continue
        out.append(ln)
    assert done and sum(ln.lstrip().startswith("#include") for ln in lines) == 9
    src = tmp_path / "main.cpp"
    src.write_text("\n".join(out) + "\n")
    inc = [os.path.join(conftest.ROOT, "include"), os.path.join(conftest.PKG, "dropin")]
    libdir = os.path.join(conftest.PKG, "lib")
    subprocess.run(["g++", "-std=c++17", "-Wall", "-Wno-unused-variable", "-Wno-unused-but-set-variable",
                    "-Wno-maybe-uninitialized", "-O1"] + [f"-I{d}" for d in inc] +
                   [str(src), "-o", str(tmp_path / "opt"), f"-L{libdir}", "-luft"], check=True)
    assert os.access(tmp_path / "opt", os.X_OK)


def test_vtu_writer_host(U, tmp_path):
    """uft_vtu_write_host (Write2Vtu.cpp:5-173) on a known mesh: the hex8
    corners reordered to VTK's 0,1,3,2,4,5,7,6 (Write2Vtu.cpp:89-96), type 12
    (10 for tet4), the fields in their positions, exact with 17 digits."""
    import xml.etree.ElementTree as ET
    from uft.capi import vtu_write_host, UFTError
    for npe in (8, 4):
        xyz, elem = U.gen_box(npe, 3, 2, 2, jitter=0.1, seed=5)
        rng = np.random.default_rng(npe)
        u = rng.standard_normal(3 * len(xyz))
        x = rng.uniform(0.01, 1.0, len(elem))
        path = tmp_path / f"t{npe}.vtu"
        vtu_write_host(path, xyz, elem, u, x, digits=17)
        root = ET.parse(path).getroot()
        piece = root.find("UnstructuredGrid/Piece")
        assert int(piece.get("NumberOfPoints")) == len(xyz) and int(piece.get("NumberOfCells")) == len(elem)
        a = {d.get("Name"): np.array(d.text.split(), dtype=float) for d in root.iter("DataArray")}
        assert np.array_equal(a["nodes"], xyz.ravel())
        order = [0, 1, 3, 2, 4, 5, 7, 6] if npe == 8 else [0, 1, 2, 3]
        assert np.array_equal(a["connectivity"].reshape(-1, npe), elem[:, order])
        if npe == 8:  # element 0 of the box: corners (0,0,0),(1,0,0),(0,1,0),(1,1,0) in VTK order 0,1,3,2
            nx1 = 4
            assert list(a["connectivity"][:4]) == [0, 1, nx1 + 1, nx1]
        assert np.array_equal(a["offsets"], npe * np.arange(1, len(elem) + 1))
        assert np.all(a["types"] == (12 if npe == 8 else 10))
        assert np.array_equal(a["ux"], u[0::3]) and np.array_equal(a["uy"], u[1::3]) and np.array_equal(a["uz"], u[2::3])
        assert np.array_equal(a["xphys"], x)
    vtu_write_host(tmp_path / "six.vtu", xyz, elem, u, x)  # the reference's setprecision(6)
    six = open(tmp_path / "six.vtu").read()
    assert f"{x[0]:.6e}" in six
    with pytest.raises(UFTError) as e:
        vtu_write_host(tmp_path / "no" / "such" / "dir.vtu", xyz, elem, u, x)
    assert e.value.code == 2


@pytest.mark.parametrize("npe,dims,seed", [(8, (40, 20, 20), 0), (4, (12, 6, 6), 3), (8, (140, 70, 40), 392000)])
def test_node_tiles_host(U, npe, dims, seed):
    """The node tiles of the eight-right-hand-side products (host side,
    uft_node_tiles_host), on the c0 mesh, a tet4 mesh and the full 392,000-
    element c1 mesh: every node in exactly one tile, ascending inside it; a
    tile's halo is exactly the union of its nodes' neighbour sets (the
    columns of their rows, from the elements), ascending; no halo above the
    840-node cap; and the c1 tiles stay compact (halo per tile node)."""
    import scipy.sparse as sp
    xyz, elem = U.gen_box(npe, *dims, jitter=0.15, seed=seed)
    xyz, elem = np.asarray(xyz).reshape(-1, 3), np.asarray(elem).reshape(-1, npe)
    nn = len(xyz)
    node, hoff, hnode, hmax = U.node_tiles_host(elem, xyz)
    assert node.shape[1] == 256 and len(hoff) == node.shape[0] + 1
    real = node[node >= 0]
    assert np.array_equal(np.sort(real), np.arange(nn))  # each node once
    for t in range(node.shape[0]):
        row = node[t][node[t] >= 0]
        assert len(row) > 0 and np.all(np.diff(row) > 0)
    rows = np.repeat(elem, npe, axis=1).ravel()
    cols = np.tile(elem, (1, npe)).ravel()
    A = sp.csr_matrix((np.ones(len(rows), np.int8), (rows, cols)), shape=(nn, nn))
    A.sum_duplicates()
    sizes = np.diff(hoff)
    assert hmax == sizes.max() and hmax <= 840
    for t in range(node.shape[0]):
        row = node[t][node[t] >= 0]
        want = np.unique(A[row].indices)
        got = hnode[hoff[t]:hoff[t + 1]]
        assert np.array_equal(got, want), t
    if nn > 100000:
        # coordinate bisection keeps tiles compact: a tile's halo averages
        # under three times its node count (consecutive-node tiles: ~6x)
        assert sizes.sum() / nn < 3.0
