"""ctypes binding of include/uft.h.

The binding is plain plumbing: every computation happens in libuft.so on the
GPU.  Host arrays are numpy float64/int32.
"""
import ctypes as C
import os
import re

import numpy as np

_PKG = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
_ROOT = os.path.dirname(_PKG)
HEADER = os.path.join(_ROOT, "include", "uft.h")

FP64, MIXED, MIXED_H = 0, 1, 2
VEC = {"X": 0, "XFiled": 1, "XPhys": 2, "Xold": 3, "dfdx": 4, "xmin": 5, "xmax": 6, "Vol": 7, "U": 8,
       "RX": 9, "VMS": 10, "dgdx0": 16}
SIZE = {"ne": 0, "nn": 1, "ndof": 2, "npe": 3}
STATUS = {0: "UFT_OK", 1: "UFT_ERR_ARG", 2: "UFT_ERR_IO", 3: "UFT_ERR_NOCONV", 4: "UFT_ERR_DETJ",
          5: "UFT_ERR_OOM", 6: "UFT_ERR_HIP", 7: "UFT_ERR_STATE"}


class UFTError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"{STATUS.get(code, code)}: {msg}")
        self.code = code


class Params(C.Structure):
    _fields_ = [(n, C.c_double) for n in
                ("Emin", "Emax", "penal", "volfrac", "rmin", "nu", "movlim", "Xmin", "Xmax", "x0", "beta0",
                 "betamax")] + \
               [("proj", C.c_int), ("m", C.c_int), ("maxItr", C.c_int),
                ("rtol_el", C.c_double), ("rtol_f", C.c_double), ("maxit_el", C.c_int), ("maxit_f", C.c_int),
                ("precision", C.c_int), ("amg_levels", C.c_int), ("cheb_degree", C.c_int),
                ("coarse_max", C.c_int), ("cycle_type", C.c_int),
                ("cheb_lo", C.c_double), ("cheb_hi", C.c_double), ("esteig_steps", C.c_int),
                ("load_cases", C.c_int), ("graph", C.c_int), ("cheb_degree_coarse", C.c_int),
                ("fine_assembled", C.c_int), ("coarse_dense_max", C.c_int),
                ("prolong_smooth", C.c_int), ("coarse_lag", C.c_int)]


_LIB = None


def lib_path():
    return os.path.join(_PKG, "lib", "libuft.so")


def lib():
    """Load libuft.so; raises if it has not been built."""
    global _LIB
    if _LIB is not None:
        return _LIB
    path = lib_path()
    if not os.path.exists(path):
        raise FileNotFoundError(f"{path} is missing: run __graft_entry__.build() (make -C the package) first")
    L = C.CDLL(path)
    dp = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
    ip = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")
    lp = np.ctypeslib.ndpointer(dtype=np.int64, flags="C_CONTIGUOUS")
    vp = C.c_void_p
    pi = C.POINTER(C.c_int)
    pd = C.POINTER(C.c_double)
    sig = {
        "uft_create": (C.c_int, [C.POINTER(vp), C.c_int]),
        "uft_destroy": (None, [vp]),
        "uft_last_error": (C.c_char_p, [vp]),
        "uft_default_params": (None, [C.POINTER(Params)]),
        "uft_params_set": (C.c_int, [vp, C.POINTER(Params)]),
        "uft_params_get": (C.c_int, [vp, C.POINTER(Params)]),
        "uft_mesh_set": (C.c_int, [vp, C.c_int, C.c_int64, C.c_int64, ip, dp]),
        "uft_mesh_load_text": (C.c_int, [vp, C.c_int, C.c_int64, C.c_int64, C.c_char_p, C.c_char_p]),
        "uft_bc_set": (C.c_int, [vp, C.c_int, ip, C.c_int, C.c_int, ip, dp]),
        "uft_bc_load_text": (C.c_int, [vp, C.c_char_p, C.c_int, C.c_int, C.c_int]),
        "uft_setup": (C.c_int, [vp]),
        "uft_size": (C.c_int64, [vp, C.c_int]),
        "uft_total_volume": (C.c_double, [vp]),
        "uft_vec_get": (C.c_int, [vp, C.c_int, dp, C.c_int64]),
        "uft_vec_set": (C.c_int, [vp, C.c_int, dp, C.c_int64]),
        "uft_vec_fill": (C.c_int, [vp, C.c_int, C.c_double]),
        "uft_vec_copy": (C.c_int, [vp, C.c_int, C.c_int]),
        "uft_vec_device": (vp, [vp, C.c_int]),
        "uft_fem_obj_sens": (C.c_int, [vp, pd, pd, C.c_int, C.c_int, pi, C.c_int]),
        "uft_filter": (C.c_int, [vp, C.c_int, C.c_int]),
        "uft_filter_sens": (C.c_int, [vp, C.c_int, C.c_int, pi]),
        "uft_heaviside": (C.c_int, [vp, C.c_int, C.c_int, C.c_int]),
        "uft_projected_sens": (C.c_int, [vp, C.c_int, C.c_int, C.c_int, pi]),
        "uft_mma_movelimit": (C.c_int, [vp, C.c_int, C.c_int, C.c_int]),
        "uft_mma_update": (C.c_int, [vp, C.c_int, C.c_int, pd, pi, C.c_int, C.c_int]),
        "uft_mma_design_change": (C.c_int, [vp, C.c_int, C.c_int, pd]),
        "uft_design_iteration": (C.c_int, [vp, C.c_int, pd, pd, pd]),
        "uft_solve_stats": (C.c_int, [vp, pi, pd, pi, pi]),
        "uft_apply_K": (C.c_int, [vp, C.c_int, dp, dp]),
        "uft_element_ke": (C.c_int, [vp, C.c_int64, dp]),
        "uft_element_kf": (C.c_int, [vp, C.c_int64, dp, dp]),
        "uft_sync": (C.c_int, [vp]),
        "uft_profile_enable": (C.c_int, [vp, C.c_int]),
        "uft_profile_read": (C.c_int, [vp, C.c_int, C.POINTER(C.c_int64), pd]),
        "uft_mesh_get": (C.c_int, [vp, ip, dp]),
        "uft_box_nodes": (C.c_int64, [C.c_int] * 3),
        "uft_box_elems": (C.c_int64, [C.c_int] * 4),
        "uft_gen_box": (C.c_int, [C.c_int] * 4 + [C.c_double] * 4 + [C.c_uint64, dp, ip]),
        "uft_box_bcs": (C.c_int, [C.c_int] * 4 + [ip, pi, ip, dp, pi]),
        "uft_solver_info": (C.c_int, [vp, pi, C.POINTER(C.c_int64), C.POINTER(C.c_int64)]),
        "uft_fine_operator_info": (C.c_int, [vp, C.POINTER(C.c_int64), C.POINTER(C.c_int64), pi, pi]),
        "uft_fine_tiles_info": (C.c_int, [vp, C.POINTER(C.c_int64), C.POINTER(C.c_int64), pi]),
        "uft_amg_operator_get": (C.c_int, [vp, C.c_int, ip, ip, dp]),
        "uft_amg_transfer_get": (C.c_int, [vp, C.c_int, ip, dp]),
        "uft_amg_coarse_inverse_get": (C.c_int, [vp, dp]),
        "uft_phase_times": (C.c_int, [vp, pd]),
        "uft_time_fine_operator": (C.c_int, [vp, C.c_int, C.c_int, C.POINTER(C.c_double)]),
        "uft_mma_set_asymptotes": (C.c_int, [vp, C.c_double, C.c_double, C.c_double]),
        "uft_mma_set_robust_asymptotes": (C.c_int, [vp, C.c_int]),
        "uft_mma_constraint_modification": (C.c_int, [vp, C.c_int]),
        "uft_mma_set_penalization": (C.c_int, [vp, dp, dp, dp]),
        "uft_mma_state_get": (C.c_int, [vp, pi, dp, dp, dp, dp]),
        "uft_mma_state_set": (C.c_int, [vp, C.c_int, dp, dp, dp, dp]),
        "uft_mma_kkt_residual": (C.c_int, [vp, C.c_int, C.c_int, pd, pi, C.c_int, C.c_int, pd, pd]),
        "uft_stress": (C.c_int, [vp, C.c_int, C.c_int]),
        "uft_cache_save": (C.c_int, [vp, C.c_char_p]),
        "uft_load_case_stats": (C.c_int, [vp, pi, pd]),
        "uft_load_case_u": (C.c_int, [vp, C.c_int, dp]),
        "uft_apply_K_cases": (C.c_int, [vp, C.c_int, dp, dp]),
        "uft_cache_load": (C.c_int, [vp, C.c_char_p]),
        "uft_autotune": (C.c_int, [vp, C.c_int, pi, pd, C.c_int]),
        "uft_batch_set": (C.c_int, [vp, C.c_int, lp, lp, dp, dp]),
        "uft_autotune_smoother": (C.c_int, [vp, pi, pi, pd]),
        "uft_level_apply": (C.c_int, [vp, C.c_int, C.c_int, dp, dp]),
        "uft_batch_results": (C.c_int, [vp, pd, pd, pd, pi, pd, pi]),
        "uft_mma_design_change_noalloc": (C.c_int, [C.c_int, vp, vp, C.c_int64, vp, vp]),
        "uft_mesh_read_text": (C.c_int, [C.c_int, C.c_int64, C.c_int64, C.c_char_p, C.c_char_p, dp, ip]),
        "uft_node_tiles_host": (C.c_int, [C.c_int, C.c_int64, C.c_int64, ip, dp, C.POINTER(C.c_int64),
                                          C.POINTER(C.c_int64), C.POINTER(C.c_int), C.c_void_p, C.c_void_p,
                                          C.c_void_p]),
        "uft_bc_read_text": (C.c_int, [C.c_char_p, C.c_int, C.c_int, C.c_int, ip, ip, dp]),
        "uft_write_vtu": (C.c_int, [vp, C.c_char_p, C.c_int, C.c_int, C.c_int]),
        "uft_solve_traffic": (C.c_int, [vp, pd, pd, C.POINTER(C.c_int64), C.POINTER(C.c_int64)]),
        "uft_amg_prolongator_get": (C.c_int, [vp, C.POINTER(C.c_int64), vp, vp, vp, pd]),
        "uft_vtu_write_host": (C.c_int, [C.c_char_p, C.c_int, C.c_int64, C.c_int64, dp, ip, dp, dp, C.c_int]),
    }
    for name, (res, args) in sig.items():
        f = getattr(L, name)
        f.restype = res
        f.argtypes = args
    _LIB = L
    return L


def header_symbols():
    """Function names declared in include/uft.h."""
    txt = open(HEADER).read()
    return sorted(set(re.findall(r"\b(uft_[a-z0-9_]+)\s*\(", txt)))


def exported_symbols():
    L = lib()
    return sorted(n for n in header_symbols() if hasattr(L, n))


def gen_box(npe, nx, ny, nz, lx=2.0, ly=1.0, lz=1.0, jitter=0.0, seed=0):
    L = lib()
    nn = L.uft_box_nodes(nx, ny, nz)
    ne = L.uft_box_elems(npe, nx, ny, nz)
    xyz = np.zeros(nn * 3)
    elem = np.zeros(ne * npe, dtype=np.int32)
    rc = L.uft_gen_box(npe, nx, ny, nz, lx, ly, lz, jitter, seed, xyz, elem)
    if rc:
        raise UFTError(rc, "uft_gen_box")
    return xyz.reshape(nn, 3), elem.reshape(ne, npe)


def box_bcs(nx, ny, nz, load_kind=0):
    L = lib()
    nn = (nx + 1) * (ny + 1) * (nz + 1)
    fixed = np.zeros(3 * nn, dtype=np.int32)
    ld = np.zeros(max(nn, 72), dtype=np.int32)
    lv = np.zeros(max(nn, 72))
    nf, nl = C.c_int(), C.c_int()
    rc = L.uft_box_bcs(nx, ny, nz, load_kind, fixed, C.byref(nf), ld, lv, C.byref(nl))
    if rc:
        raise UFTError(rc, "uft_box_bcs")
    return fixed[:nf.value].copy(), ld[:nl.value].copy(), lv[:nl.value].copy()


def bc_read_text(dirpath, nfixed, nlc, nload):
    """Host-only BC text reader (uft_bc_read_text): fixed DOFs, load DOFs and
    load values (nlc x nload) of the reference's files in dirpath."""
    fixed = np.zeros(nfixed, np.int32)
    ld = np.zeros(nlc * nload, np.int32)
    lv = np.zeros(nlc * nload)
    rc = lib().uft_bc_read_text(os.fsencode(dirpath), nfixed, nlc, nload, fixed, ld, lv)
    if rc:
        raise UFTError(rc, lib().uft_last_error(None).decode())
    return fixed, ld.reshape(nlc, nload), lv.reshape(nlc, nload)


def vtu_write_host(path, xyz, elem, u, xphys, digits=6):
    """Host-only VTU writer (uft_vtu_write_host)."""
    f = lambda a, t: np.ascontiguousarray(a, dtype=t).ravel()
    elem = np.asarray(elem)
    rc = lib().uft_vtu_write_host(os.fsencode(path), elem.shape[1], len(xyz), len(elem), f(xyz, np.float64),
                                  f(elem, np.int32), f(u, np.float64), f(xphys, np.float64), digits)
    if rc:
        raise UFTError(rc, lib().uft_last_error(None).decode())


def node_tiles_host(elem, xyz):
    """Host-only node tiles of the eight-right-hand-side products
    (uft_node_tiles_host): (node [ntile, 256] with -1 padding, hoff, hnode, hmax)."""
    elem = np.ascontiguousarray(elem, dtype=np.int32)
    xyz = np.ascontiguousarray(xyz, dtype=np.float64)
    ne, npe = elem.shape
    nn = len(xyz)
    nt, nh, hm = C.c_int64(), C.c_int64(), C.c_int()
    args = (npe, nn, ne, elem.ravel(), xyz.ravel(), C.byref(nt), C.byref(nh), C.byref(hm))
    rc = lib().uft_node_tiles_host(*args, None, None, None)
    if rc:
        raise UFTError(rc, lib().uft_last_error(None).decode())
    node = np.zeros(nt.value * 256, np.int32)
    hoff = np.zeros(nt.value + 1, np.int32)
    hnode = np.zeros(max(nh.value, 1), np.int32)
    rc = lib().uft_node_tiles_host(*args, node.ctypes.data, hoff.ctypes.data, hnode.ctypes.data)
    if rc:
        raise UFTError(rc, lib().uft_last_error(None).decode())
    return node.reshape(-1, 256), hoff, hnode[:nh.value], hm.value


def mesh_read_text(npe, nn, ne, vertex_path, element_path):
    """Host-only mesh text reader (uft_mesh_read_text): (xyz nn x 3, elem ne x npe)."""
    xyz = np.zeros(nn * 3)
    elem = np.zeros(ne * npe, np.int32)
    rc = lib().uft_mesh_read_text(npe, nn, ne, os.fsencode(vertex_path), os.fsencode(element_path), xyz, elem)
    if rc:
        raise UFTError(rc, lib().uft_last_error(None).decode())
    return xyz.reshape(nn, 3), elem.reshape(ne, npe)


class Context:
    """One design problem on one GPU (uft_ctx)."""

    def __init__(self, device=0):
        L = lib()
        h = C.c_void_p()
        rc = L.uft_create(C.byref(h), device)
        if rc:
            raise UFTError(rc, "uft_create failed (no HIP device?)")
        self.h = h
        self.L = L

    def close(self):
        if getattr(self, "h", None):
            self.L.uft_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()

    def _chk(self, rc, what, allow=()):
        if rc and rc not in allow:
            raise UFTError(rc, f"{what}: {self.L.uft_last_error(self.h).decode()}")
        return rc

    # ---- configuration
    @staticmethod
    def default_params(**kw):
        p = Params()
        lib().uft_default_params(C.byref(p))
        for k, v in kw.items():
            if not hasattr(p, k):
                raise AttributeError(k)
            setattr(p, k, v)
        return p

    def set_params(self, p):
        self._chk(self.L.uft_params_set(self.h, C.byref(p)), "uft_params_set")
        self.params = p

    def get_params(self):
        p = Params()
        self._chk(self.L.uft_params_get(self.h, C.byref(p)), "uft_params_get")
        return p

    def set_mesh(self, elem, xyz):
        elem = np.ascontiguousarray(elem, dtype=np.int32)
        xyz = np.ascontiguousarray(xyz, dtype=np.float64)
        ne, npe = elem.shape
        self._chk(self.L.uft_mesh_set(self.h, npe, xyz.shape[0], ne, elem.ravel(), xyz.ravel()), "uft_mesh_set")

    def set_bcs(self, fixed, load_dofs, load_vals, nlc=1):
        ld = np.ascontiguousarray(load_dofs, dtype=np.int32).ravel()
        lv = np.ascontiguousarray(load_vals, dtype=np.float64).ravel()
        fixed = np.ascontiguousarray(fixed, dtype=np.int32)
        self._chk(self.L.uft_bc_set(self.h, len(fixed), fixed, nlc, len(ld) // nlc, ld, lv), "uft_bc_set")

    def setup(self):
        self._chk(self.L.uft_setup(self.h), "uft_setup")

    def autotune_smoother(self):
        """(degree, cycle type, times ms [3][4]) of the smoother-schedule autotuner."""
        d, c = C.c_int(), C.c_int()
        ms = (C.c_double * 12)()
        self._chk(self.L.uft_autotune_smoother(self.h, C.byref(d), C.byref(c), ms), "uft_autotune_smoother")
        return d.value, c.value, np.array(ms[:]).reshape(3, 4)

    # ---- independent designs in one context
    def batch_set(self, elem_off, node_off, volfrac, rmin):
        f = lambda a, t: np.ascontiguousarray(a, dtype=t)
        nd = len(volfrac)
        self._chk(self.L.uft_batch_set(self.h, nd, f(elem_off, np.int64), f(node_off, np.int64),
                                       f(volfrac, np.float64), f(rmin, np.float64)), "uft_batch_set")
        self.ndesign = nd

    def batch_results(self):
        """Per design: fx, gx, ch, elasticity iterations, relative residual,
        filter iterations of the last iteration."""
        nd = self.ndesign
        fx, gx, ch, rr = ((C.c_double * nd)() for _ in range(4))
        it, itf = (C.c_int * nd)(), (C.c_int * nd)()
        self.L.uft_batch_results(self.h, fx, gx, ch, it, rr, itf)
        return {"fx": np.array(fx[:]), "gx": np.array(gx[:]), "ch": np.array(ch[:]), "it_el": np.array(it[:]),
                "relres_el": np.array(rr[:]), "it_f": np.array(itf[:])}

    def size(self, which):
        return self.L.uft_size(self.h, SIZE[which])

    @property
    def total_volume(self):
        return self.L.uft_total_volume(self.h)

    # ---- vectors
    def vid(self, name):
        if isinstance(name, int):
            return name
        if name.startswith("dgdx") and name != "dgdx0":
            return 16 + int(name[4:])
        return VEC[name]

    def vlen(self, name):
        i = self.vid(name)
        if i == VEC["U"]:
            return self.size("ndof")
        if i == VEC["RX"]:
            return self.size("nn")
        return self.size("ne")

    def get(self, name):
        out = np.zeros(self.vlen(name))
        self._chk(self.L.uft_vec_get(self.h, self.vid(name), out, len(out)), "uft_vec_get")
        return out

    def set(self, name, arr):
        arr = np.ascontiguousarray(arr, dtype=np.float64)
        self._chk(self.L.uft_vec_set(self.h, self.vid(name), arr, len(arr)), "uft_vec_set")

    # ---- the design iteration
    def fem_obj_sens(self, m=1, allow_noconv=False):
        fx = C.c_double()
        gx = (C.c_double * 8)()
        ids = (C.c_int * m)(*[16 + j for j in range(m)])
        rc = self.L.uft_fem_obj_sens(self.h, C.byref(fx), gx, VEC["dfdx"], m, ids, VEC["XPhys"])
        self._chk(rc, "uft_fem_obj_sens", allow=(3,) if allow_noconv else ())
        return fx.value, [gx[j] for j in range(m)]

    def filter(self, src="X", dst="XFiled"):
        self._chk(self.L.uft_filter(self.h, self.vid(src), self.vid(dst)), "uft_filter")

    def filter_sens(self, m=1):
        ids = (C.c_int * m)(*[16 + j for j in range(m)])
        self._chk(self.L.uft_filter_sens(self.h, VEC["dfdx"], m, ids), "uft_filter_sens")

    def heaviside(self, itr):
        self._chk(self.L.uft_heaviside(self.h, VEC["XFiled"], VEC["XPhys"], itr), "uft_heaviside")

    def projected_sens(self, m=1):
        ids = (C.c_int * m)(*[16 + j for j in range(m)])
        self._chk(self.L.uft_projected_sens(self.h, VEC["XFiled"], VEC["dfdx"], m, ids), "uft_projected_sens")

    def mma_movelimit(self):
        self._chk(self.L.uft_mma_movelimit(self.h, VEC["X"], VEC["xmin"], VEC["xmax"]), "uft_mma_movelimit")

    def mma_update(self, gx):
        m = len(gx)
        g = (C.c_double * m)(*gx)
        ids = (C.c_int * m)(*[16 + j for j in range(m)])
        self._chk(self.L.uft_mma_update(self.h, VEC["X"], VEC["dfdx"], g, ids, VEC["xmin"], VEC["xmax"]),
                  "uft_mma_update")

    def design_change(self):
        ch = C.c_double()
        self._chk(self.L.uft_mma_design_change(self.h, VEC["X"], VEC["Xold"], C.byref(ch)), "uft_mma_design_change")
        return ch.value

    # ---- MMA settings, restart state, KKT residual
    def mma_settings(self, init=0.5, dec=0.7, inc=1.2, robust=0, conmod=0):
        self._chk(self.L.uft_mma_set_asymptotes(self.h, init, dec, inc), "uft_mma_set_asymptotes")
        self._chk(self.L.uft_mma_set_robust_asymptotes(self.h, robust), "uft_mma_set_robust_asymptotes")
        self._chk(self.L.uft_mma_constraint_modification(self.h, conmod), "uft_mma_constraint_modification")

    def mma_set_penalization(self, a, c, d):
        f = lambda v: np.ascontiguousarray(v, dtype=np.float64)
        self._chk(self.L.uft_mma_set_penalization(self.h, f(a), f(c), f(d)), "uft_mma_set_penalization")

    def mma_state(self):
        ne = self.size("ne")
        k = C.c_int()
        v = [np.zeros(ne) for _ in range(4)]
        self._chk(self.L.uft_mma_state_get(self.h, C.byref(k), *v), "uft_mma_state_get")
        return k.value, v[0], v[1], v[2], v[3]

    def mma_set_state(self, k, xo1, xo2, U, L):
        f = lambda a: np.ascontiguousarray(a, dtype=np.float64)
        self._chk(self.L.uft_mma_state_set(self.h, k, f(xo1), f(xo2), f(U), f(L)), "uft_mma_state_set")

    def mma_kkt(self, gx):
        m = len(gx)
        g = (C.c_double * m)(*gx)
        ids = (C.c_int * m)(*[16 + j for j in range(m)])
        n2, ni = C.c_double(), C.c_double()
        self._chk(self.L.uft_mma_kkt_residual(self.h, VEC["X"], VEC["dfdx"], g, ids, VEC["xmin"], VEC["xmax"],
                                              C.byref(n2), C.byref(ni)), "uft_mma_kkt_residual")
        return n2.value, ni.value

    def write_vtu(self, path, xphys="XPhys", u="U", digits=6):
        """Write2Vtu::WriteResults of this context (uft_write_vtu)."""
        self._chk(self.L.uft_write_vtu(self.h, os.fsencode(path), self.vid(xphys), self.vid(u), digits),
                  "uft_write_vtu")

    def stress(self, xphys="XPhys"):
        self._chk(self.L.uft_stress(self.h, self.vid(xphys), VEC["VMS"]), "uft_stress")
        return self.get("VMS")

    def design_iteration(self, itr, allow_noconv=False):
        fx, gx, ch = C.c_double(), (C.c_double * 8)(), C.c_double()
        rc = self.L.uft_design_iteration(self.h, itr, C.byref(fx), gx, C.byref(ch))
        self._chk(rc, "uft_design_iteration", allow=(3,) if allow_noconv else ())
        m = self.get_params().m
        return fx.value, [gx[j] for j in range(m)], ch.value

    def init_design(self):
        """Initial filter and projection (main.cpp:64-66)."""
        self.filter("X", "XFiled")
        self.heaviside(0)

    def run(self, maxitr):
        """while (itr < maxItr && ch > 0.01) (main.cpp:70)."""
        self.init_design()
        fxs, gxs, chs = [], [], []
        ch, itr = 1.0, 0
        while itr < maxitr and ch > 0.01:
            itr += 1
            fx, gx, ch = self.design_iteration(itr)
            fxs.append(fx)
            gxs.append(gx[0])
            chs.append(ch)
        return itr, np.array(fxs), np.array(gxs), np.array(chs)

    def load_case_stats(self):
        """Per load case (multi-load-case form): Krylov iterations and final
        relative residual."""
        it = (C.c_int * 8)()
        rr = (C.c_double * 8)()
        n = self.L.uft_load_case_stats(self.h, it, rr)
        return [(it[k], rr[k]) for k in range(max(n, 0))]

    def apply_K_cases(self, U, kind=0):
        U = np.ascontiguousarray(U, dtype=np.float64)
        Y = np.zeros_like(U)
        self._chk(self.L.uft_apply_K_cases(self.h, kind, U.ravel(), Y.ravel()), "uft_apply_K_cases")
        return Y

    def load_case_u(self, k):
        out = np.zeros(self.size("ndof"))
        self._chk(self.L.uft_load_case_u(self.h, k, out), "uft_load_case_u")
        return out

    def stats(self):
        a, b, c, d = C.c_int(), C.c_double(), C.c_int(), C.c_int()
        self._chk(self.L.uft_solve_stats(self.h, C.byref(a), C.byref(b), C.byref(c), C.byref(d)), "uft_solve_stats")
        return {"it_el": a.value, "relres_el": b.value, "it_f": c.value, "outer_el": d.value}

    def apply_K(self, u, xphys_name="XPhys"):
        u = np.ascontiguousarray(u, dtype=np.float64)
        y = np.zeros_like(u)
        self._chk(self.L.uft_apply_K(self.h, self.vid(xphys_name), u, y), "uft_apply_K")
        return y

    def element_ke(self, e):
        nd = 3 * self.size("npe")
        out = np.zeros(nd * nd)
        self._chk(self.L.uft_element_ke(self.h, e, out), "uft_element_ke")
        return out.reshape(nd, nd)

    def element_kf(self, e):
        npe = self.size("npe")
        kf, t = np.zeros(npe * npe), np.zeros(npe)
        self._chk(self.L.uft_element_kf(self.h, e, kf, t), "uft_element_kf")
        return kf.reshape(npe, npe), t

    def sync(self):
        self._chk(self.L.uft_sync(self.h), "uft_sync")

    # ---- solver inspection
    def solver_info(self):
        """(unknowns per level, element count / 6x6 block count per level)."""
        nl = C.c_int()
        n = (C.c_int64 * 8)()
        b = (C.c_int64 * 8)()
        self._chk(self.L.uft_solver_info(self.h, C.byref(nl), n, b), "uft_solver_info")
        return [n[i] for i in range(nl.value)], [b[i] for i in range(nl.value)]

    def fine_tiles_info(self):
        """(tiles, tile-ordered slots, largest halo) of the eight-right-hand-side
        node tiles; all 0 when the node-ordered kernels run."""
        t, k, h = C.c_int64(), C.c_int64(), C.c_int()
        self._chk(self.L.uft_fine_tiles_info(self.h, C.byref(t), C.byref(k), C.byref(h)), "uft_fine_tiles_info")
        return t.value, k.value, h.value

    def fine_operator_info(self):
        """(sliced-ELL slots, nonzero node blocks, bytes per column index,
        bytes per smoother block) of the assembled fine operators; all 0 with
        the element form."""
        s, b, w, v = C.c_int64(), C.c_int64(), C.c_int(), C.c_int()
        self._chk(self.L.uft_fine_operator_info(self.h, C.byref(s), C.byref(b), C.byref(w), C.byref(v)),
                  "uft_fine_operator_info")
        return s.value, b.value, w.value, v.value

    def amg_operator(self, level):
        """BSR (rp, ci, val[nblk, 6, 6]) of multigrid level >= 1."""
        n, b = self.solver_info()
        nu = n[level] // 6
        rp = np.zeros(nu + 1, np.int32)
        ci = np.zeros(b[level], np.int32)
        val = np.zeros(b[level] * 36)
        self._chk(self.L.uft_amg_operator_get(self.h, level, rp, ci, val), "uft_amg_operator_get")
        return rp, ci, val.reshape(-1, 6, 6)

    def amg_transfer(self, level):
        """(parent, P[units, bs, 6]) of the transfer from level to level + 1."""
        n, _ = self.solver_info()
        bs = 3 if level == 0 else 6
        nu = n[level] // bs
        parent = np.zeros(nu, np.int32)
        P = np.zeros(nu * bs * 6)
        self._chk(self.L.uft_amg_transfer_get(self.h, level, parent, P), "uft_amg_transfer_get")
        return parent, P.reshape(nu, bs, 6)

    def amg_prolongator(self):
        """Smoothed prolongator of level 0 (pptr, pagg, P[nent, 3, 6], omega), or None
        with the tentative one."""
        ne_, om = C.c_int64(), C.c_double()
        self._chk(self.L.uft_amg_prolongator_get(self.h, C.byref(ne_), None, None, None, C.byref(om)),
                  "uft_amg_prolongator_get")
        if ne_.value == 0:
            return None
        nn = self.size("nn")
        pptr = np.zeros(nn + 1, np.int32)
        pagg = np.zeros(ne_.value, np.int32)
        P = np.zeros(ne_.value * 18)
        self._chk(self.L.uft_amg_prolongator_get(self.h, C.byref(ne_), pptr.ctypes.data_as(C.c_void_p),
                                                 pagg.ctypes.data_as(C.c_void_p), P.ctypes.data_as(C.c_void_p),
                                                 C.byref(om)), "uft_amg_prolongator_get")
        return pptr, pagg, P.reshape(-1, 3, 6), om.value

    def amg_coarse_inverse(self):
        n = self.solver_info()[0][-1]
        inv = np.zeros(n * n)
        self._chk(self.L.uft_amg_coarse_inverse_get(self.h, inv), "uft_amg_coarse_inverse_get")
        return inv.reshape(n, n)

    def level_apply(self, level, x, kind=0):
        x = np.ascontiguousarray(x, dtype=np.float64)
        y = np.zeros_like(x)
        self._chk(self.L.uft_level_apply(self.h, level, kind, x, y), "uft_level_apply")
        return y

    def solve_traffic(self):
        """(algorithmic bytes of the solves' operator applies, solve ms, applies,
        solves) since the last uft_profile_enable (uft_solve_traffic)."""
        b, ms = C.c_double(), C.c_double()
        a, n = C.c_int64(), C.c_int64()
        self._chk(self.L.uft_solve_traffic(self.h, C.byref(b), C.byref(ms), C.byref(a), C.byref(n)),
                  "uft_solve_traffic")
        return b.value, ms.value, a.value, n.value

    def time_fine_operator(self, kind, reps=20):
        """Average microseconds of reps back-to-back launches of a fine operator
        apply (uft_time_fine_operator: 0 FP64, 1 FP32 outer, 2 smoother step)."""
        us = C.c_double()
        self._chk(self.L.uft_time_fine_operator(self.h, kind, reps, C.byref(us)), "uft_time_fine_operator")
        return us.value

    def phase_times(self):
        ms = (C.c_double * 8)()
        self._chk(self.L.uft_phase_times(self.h, ms), "uft_phase_times")
        return [ms[i] for i in range(7)]


def cantilever(npe=8, nx=8, ny=4, nz=4, jitter=0.0, seed=0, load_kind=0, device=0, **params):
    """Box [0,2]x[0,1]x[0,1] cantilever, clamped at x=0 (uft_box_bcs)."""
    xyz, elem = gen_box(npe, nx, ny, nz, 2.0, 1.0, 1.0, jitter, seed)
    fixed, ld, lv = box_bcs(nx, ny, nz, load_kind)
    ctx = Context(device)
    ctx.set_params(Context.default_params(**params))
    ctx.set_mesh(elem, xyz)
    ctx.set_bcs(fixed, ld, lv, nlc=8 if load_kind == 2 else 1)  # kind 2: eight cases of nine loads
    ctx.setup()
    return ctx


class SplitMix64:
    """SplitMix64 stream (Steele, Lea, Flood 2014): u = (r >> 11) 2^-53."""
    M = (1 << 64) - 1

    def __init__(self, seed):
        self.s = seed & self.M

    def next(self):
        self.s = (self.s + 0x9E3779B97F4A7C15) & self.M
        z = self.s
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & self.M
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & self.M
        return z ^ (z >> 31)

    def uniform(self):
        return (self.next() >> 11) * 2.0 ** -53


def uneven_designs(n=32, seed=32, ne_min=8000, ne_max=64000):
    """Cantilevers of uneven size: per design ny = nz +- 1 in [10, 33] and
    nx = 2 ny +- 2 drawn from a SplitMix64 stream, kept when
    ne_min <= nx ny nz <= ne_max; volfrac ~ U(0.2, 0.4); jitter seed
    1000 seed + index."""
    rng = SplitMix64(seed)
    out = []
    while len(out) < n:
        ny = 10 + int(rng.uniform() * 24)
        nz = ny + int(rng.uniform() * 3) - 1
        nx = 2 * ny + int(rng.uniform() * 5) - 2
        vf = 0.2 + 0.2 * rng.uniform()
        if ne_min <= nx * ny * nz <= ne_max:
            out.append(dict(nx=nx, ny=ny, nz=nz, volfrac=vf, seed=1000 * seed + len(out)))
    return out


def batch_mesh(designs, npe=8, jitter=0.15, load_kind=0):
    """Concatenated mesh, BCs and offsets of independent box cantilevers
    (each [0,2]x[0,1]x[0,1], clamped at x=0)."""
    xyzs, elems, fixed, ld, lv = [], [], [], [], []
    eoff, noff = [0], [0]
    for D in designs:
        xyz, elem = gen_box(npe, D["nx"], D["ny"], D["nz"], 2.0, 1.0, 1.0, jitter, D["seed"])
        f, l, v = box_bcs(D["nx"], D["ny"], D["nz"], load_kind)
        base = noff[-1]
        xyzs.append(xyz)
        elems.append(elem + base)
        fixed.append(f + 3 * base)
        ld.append(l + 3 * base)
        lv.append(v)
        eoff.append(eoff[-1] + len(elem))
        noff.append(base + len(xyz))
    return (np.concatenate(elems), np.concatenate(xyzs), np.concatenate(fixed), np.concatenate(ld),
            np.concatenate(lv), np.array(eoff, np.int64), np.array(noff, np.int64))


def batch_cantilever(designs, npe=8, jitter=0.15, device=0, rmin_h=1.5, **params):
    """One context holding independent cantilevers (uft_batch_set); design d
    uses its own volfrac and rmin = rmin_h * 2 / nx."""
    elem, xyz, fixed, ld, lv, eoff, noff = batch_mesh(designs, npe, jitter)
    ctx = Context(device)
    ctx.set_params(Context.default_params(**params))
    ctx.set_mesh(elem, xyz)
    ctx.set_bcs(fixed, ld, lv, nlc=1)
    ctx.batch_set(eoff, noff, [D["volfrac"] for D in designs], [rmin_h * 2.0 / D["nx"] for D in designs])
    ctx.setup()
    return ctx
