"""ctypes wrapper around liboracle.so, the CPU FP64 restatement of the
reference SIMP design iteration.

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py.  The product library never calls it.
"""
import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None

VEC = {"X": 0, "XFiled": 1, "XPhys": 2, "Xold": 3, "dfdx": 4, "dgdx": 5, "xmin": 6,
       "xmax": 7, "Vol": 8, "U": 9, "RX": 10, "F": 11, "L": 12, "Uas": 13, "xo1": 14,
       "xo2": 15, "Tel": 16, "KFdiag": 17, "alpha": 18, "beta": 19, "p0": 20, "q0": 21, "pij": 22, "qij": 23}


class Params(C.Structure):
    """orc_params; defaults follow OptPara.cpp:66-78."""
    _fields_ = [(n, C.c_double) for n in
                ("Emin", "Emax", "penal", "volfrac", "rmin", "nu", "movlim", "Xmin", "Xmax",
                 "x0", "beta0", "betamax")] + \
               [("proj", C.c_int), ("m", C.c_int), ("maxItr", C.c_int),
                ("rtol_el", C.c_double), ("rtol_f", C.c_double),
                ("maxit_el", C.c_int), ("maxit_f", C.c_int), ("ke_cache", C.c_int)]


def build():
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(_HERE, "liboracle.so")
        if not os.path.exists(path):
            build()
        L = C.CDLL(path)
        dp = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
        ip = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")
        vp = C.c_void_p
        sig = {
            "orc_box_nodes": (C.c_int64, [C.c_int] * 3),
            "orc_box_elems": (C.c_int64, [C.c_int] * 4),
            "orc_gen_box": (C.c_int, [C.c_int] * 4 + [C.c_double] * 4 + [C.c_uint64, dp, ip]),
            "orc_box_bcs": (C.c_int, [C.c_int] * 4 + [ip, C.POINTER(C.c_int), ip, dp, C.POINTER(C.c_int)]),
            "orc_hex8_ke": (C.c_double, [dp, C.c_double, dp]),
            "orc_tet4_ke": (C.c_double, [dp, C.c_double, dp]),
            "orc_hex8_vol": (C.c_double, [dp]),
            "orc_tet4_vol": (C.c_double, [dp]),
            "orc_hex8_kehel": (None, [dp, C.c_double, dp, dp]),
            "orc_tet4_kehel": (None, [dp, C.c_double, dp, dp]),
            "orc_dofmap": (None, [ip, C.c_int64, C.c_int, ip]),
            "orc_default_params": (None, [C.POINTER(Params)]),
            "orc_params_size": (C.c_size_t, []),
            "orc_create": (vp, [C.c_int, C.c_int64, C.c_int64, ip, dp, C.c_int, ip, C.c_int, C.c_int,
                                ip, dp, C.POINTER(Params)]),
            "orc_destroy": (None, [vp]),
            "orc_vec": (C.POINTER(C.c_double), [vp, C.c_int]),
            "orc_total_volume": (C.c_double, [vp]),
            "orc_diag": (None, [vp, C.POINTER(C.c_int), C.POINTER(C.c_int), C.POINTER(C.c_double),
                                C.POINTER(C.c_int)]),
            "orc_set_params": (None, [vp, C.POINTER(Params)]),
            "orc_mma_k": (C.c_int, [vp]),
            "orc_mma_lam": (C.c_double, [vp, C.c_int]),
            "orc_assemble": (None, [vp, dp]),
            "orc_apply_K": (None, [vp, dp, dp]),
            "orc_fem_obj_sens": (C.c_int, [vp, dp, dp, dp, dp, dp]),
            "orc_filter": (C.c_int, [vp, dp, dp]),
            "orc_filter_sens": (C.c_int, [vp, dp, dp]),
            "orc_heaviside": (None, [vp, dp, dp, C.c_int]),
            "orc_projected_sens": (None, [vp, dp, dp, dp]),
            "orc_beta": (C.c_double, [vp]),
            "orc_mma_movelimit": (None, [vp, dp, dp, dp]),
            "orc_mma_update": (None, [vp, dp, dp, dp, dp, dp, dp]),
            "orc_design_change": (C.c_double, [vp, dp, dp]),
            "orc_design_iteration": (C.c_int, [vp, C.c_int, C.POINTER(C.c_double),
                                               C.POINTER(C.c_double), C.POINTER(C.c_double)]),
            "orc_init_design": (C.c_int, [vp]),
            "orc_run": (C.c_int, [vp, C.c_int, dp, dp, dp]),
            "orc_sample_costs": (C.c_int, [vp, C.c_int64, C.c_int, dp]),
            "orc_mma_settings": (None, [vp, C.c_double, C.c_double, C.c_double, C.c_int, C.c_int]),
            "orc_mma_set_acd": (None, [vp, dp, dp, dp]),
            "orc_mma_state_get": (None, [vp, C.POINTER(C.c_int), dp, dp, dp, dp]),
            "orc_mma_state_set": (None, [vp, C.c_int, dp, dp, dp, dp]),
            "orc_mma_kkt": (None, [vp, dp, dp, dp, dp, dp, dp, C.POINTER(C.c_double), C.POINTER(C.c_double)]),
            "orc_stress": (None, [vp, dp, dp, dp]),
            "orc_set_load_separate": (None, [vp, C.c_int]),
            "orc_case_u": (C.POINTER(C.c_double), [vp, C.c_int]),
            "crf_create": (vp, [vp, C.c_int]),
            "crf_destroy": (None, [vp]),
            "crf_threads": (C.c_int, [vp]),
            "crf_time_ke": (C.c_double, [vp]),
            "crf_design_iteration": (C.c_int, [vp, C.c_int, C.POINTER(C.c_double), C.POINTER(C.c_double),
                                               C.POINTER(C.c_double), C.POINTER(C.c_double), dp,
                                               np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")]),
        }
        for name, (res, args) in sig.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        assert L.orc_params_size() == C.sizeof(Params)
        _LIB = L
    return _LIB


def default_params(**kw):
    p = Params()
    lib().orc_default_params(C.byref(p))
    for k, v in kw.items():
        setattr(p, k, v)
    return p


# ---------------------------------------------------------------- mesh / BCs
def gen_box(etype, nx, ny, nz, lx=2.0, ly=1.0, lz=1.0, jitter=0.0, seed=0):
    L = lib()
    nn = L.orc_box_nodes(nx, ny, nz)
    ne = L.orc_box_elems(etype, nx, ny, nz)
    xyz = np.zeros(nn * 3)
    elem = np.zeros(ne * etype, dtype=np.int32)
    rc = L.orc_gen_box(etype, nx, ny, nz, lx, ly, lz, jitter, seed, xyz, elem)
    assert rc == 0
    return xyz.reshape(nn, 3), elem.reshape(ne, etype)


def box_bcs(nx, ny, nz, load_kind=0):
    L = lib()
    nn = (nx + 1) * (ny + 1) * (nz + 1)
    fixed = np.zeros(3 * nn, dtype=np.int32)
    ld = np.zeros(max(nn, 72), dtype=np.int32)
    lv = np.zeros(max(nn, 72))
    nf, nl = C.c_int(), C.c_int()
    L.orc_box_bcs(nx, ny, nz, load_kind, fixed, C.byref(nf), ld, lv, C.byref(nl))
    return fixed[:nf.value].copy(), ld[:nl.value].copy(), lv[:nl.value].copy()


# ---------------------------------------------------------------- elements
def hex8_ke(X, nu=0.3):
    KE = np.zeros(576)
    mindet = lib().orc_hex8_ke(np.ascontiguousarray(X, dtype=np.float64).ravel(), nu, KE)
    return KE.reshape(24, 24), mindet


def tet4_ke(X, nu=0.3):
    KE = np.zeros(144)
    det = lib().orc_tet4_ke(np.ascontiguousarray(X, dtype=np.float64).ravel(), nu, KE)
    return KE.reshape(12, 12), det


def hex8_vol(X):
    return lib().orc_hex8_vol(np.ascontiguousarray(X, dtype=np.float64).ravel())


def tet4_vol(X):
    return lib().orc_tet4_vol(np.ascontiguousarray(X, dtype=np.float64).ravel())


def hex8_kehel(X, rmin):
    KF, T = np.zeros(64), np.zeros(8)
    lib().orc_hex8_kehel(np.ascontiguousarray(X, dtype=np.float64).ravel(), rmin, KF, T)
    return KF.reshape(8, 8), T


def tet4_kehel(X, rmin):
    KF, T = np.zeros(16), np.zeros(4)
    lib().orc_tet4_kehel(np.ascontiguousarray(X, dtype=np.float64).ravel(), rmin, KF, T)
    return KF.reshape(4, 4), T


def dofmap(elem):
    elem = np.ascontiguousarray(elem, dtype=np.int32)
    ne, npe = elem.shape
    out = np.zeros(ne * 3 * npe, dtype=np.int32)
    lib().orc_dofmap(elem.ravel(), ne, npe, out)
    return out.reshape(ne, 3 * npe)


# ---------------------------------------------------------------- problem
class Problem:
    """One oracle context: mesh + BCs + parameters + design state."""

    def __init__(self, elem, xyz, fixed, load_dofs, load_vals, params=None, nlc=1):
        L = lib()
        self.elem = np.ascontiguousarray(elem, dtype=np.int32)
        self.xyz = np.ascontiguousarray(xyz, dtype=np.float64)
        self.ne, self.npe = self.elem.shape
        self.nn = self.xyz.shape[0]
        self.params = params if params is not None else default_params()
        self.m = self.params.m
        ld = np.ascontiguousarray(load_dofs, dtype=np.int32).ravel()
        lv = np.ascontiguousarray(load_vals, dtype=np.float64).ravel()
        self.h = L.orc_create(self.npe, self.nn, self.ne, self.elem.ravel(), self.xyz.ravel(),
                              len(fixed), np.ascontiguousarray(fixed, dtype=np.int32), nlc,
                              len(ld) // nlc, ld, lv, C.byref(self.params))
        if not self.h:
            raise ValueError("orc_create failed")

    def __del__(self):
        if getattr(self, "h", None):
            lib().orc_destroy(self.h)
            self.h = None

    def vec(self, name):
        n = {"U": 3 * self.nn, "F": 3 * self.nn, "RX": self.nn, "KFdiag": self.nn,
             "dgdx": self.m * self.ne, "pij": self.m * self.ne, "qij": self.m * self.ne,
             "Tel": self.npe * self.ne}.get(name, self.ne)
        p = lib().orc_vec(self.h, VEC[name])
        return np.ctypeslib.as_array(p, shape=(n,))

    def set_params(self, params):
        self.params = params
        lib().orc_set_params(self.h, C.byref(params))

    @property
    def total_volume(self):
        return lib().orc_total_volume(self.h)

    def diag(self):
        a, b, c, d = C.c_int(), C.c_double(), C.c_int(), C.c_int()
        it_el, it_f = C.c_int(), C.c_int()
        lib().orc_diag(self.h, C.byref(it_el), C.byref(it_f), C.byref(b), C.byref(d))
        return {"it_el": it_el.value, "it_f": it_f.value, "relres_el": b.value, "dip_steps": d.value}

    def apply_K(self, xphys, u):
        L = lib()
        L.orc_assemble(self.h, np.ascontiguousarray(xphys, dtype=np.float64))
        y = np.zeros(3 * self.nn)
        L.orc_apply_K(self.h, np.ascontiguousarray(u, dtype=np.float64), y)
        return y

    def fem_obj_sens(self, xphys):
        fx = np.zeros(1)
        gx = np.zeros(self.m)
        df = np.zeros(self.ne)
        dg = np.zeros(self.ne * self.m)
        rc = lib().orc_fem_obj_sens(self.h, np.ascontiguousarray(xphys, dtype=np.float64), fx, gx, df, dg)
        return fx[0], gx, df, dg.reshape(self.m, self.ne), rc

    def filter(self, x):
        out = np.zeros(self.ne)
        rc = lib().orc_filter(self.h, np.ascontiguousarray(x, dtype=np.float64), out)
        return out, rc

    def filter_sens(self, df, dg):
        df = np.array(df, dtype=np.float64)
        dg = np.array(dg, dtype=np.float64).ravel()
        rc = lib().orc_filter_sens(self.h, df, dg)
        return df, dg.reshape(self.m, self.ne), rc

    def heaviside(self, xf, itr):
        """PDEFilter::HeavisideFilter on xf at iteration itr; returns (xp, beta)."""
        xp = np.zeros(self.ne)
        lib().orc_heaviside(self.h, np.ascontiguousarray(xf, dtype=np.float64), xp, itr)
        return xp, lib().orc_beta(self.h)

    def projected_sens(self, xf, df, dg):
        """PDEFilter::ComputeProjectedSens with the beta of the last heaviside call."""
        df = np.array(df, dtype=np.float64)
        dg = np.array(dg, dtype=np.float64).ravel()
        lib().orc_projected_sens(self.h, np.ascontiguousarray(xf, dtype=np.float64), df, dg)
        return df, dg.reshape(self.m, self.ne)

    def init_design(self):
        return lib().orc_init_design(self.h)

    def design_iteration(self, itr):
        fx, gx, ch = C.c_double(), (C.c_double * 8)(), C.c_double()
        rc = lib().orc_design_iteration(self.h, itr, C.byref(fx), gx, C.byref(ch))
        return fx.value, [gx[j] for j in range(self.m)], ch.value, rc

    def sample_costs(self, ne_sample=4000, n_it=10):
        out = np.zeros(8)
        lib().orc_sample_costs(self.h, ne_sample, n_it, out)
        return {"ke_per_elem": out[0], "asm_per_elem": out[1], "uku_per_elem": out[2], "pcg_iter": out[3],
                "filter_iter": out[4], "mma_update": out[5]}

    # ---- MMA settings, restart state and KKT residual
    def mma_settings(self, init=0.5, dec=0.7, inc=1.2, robust=0, conmod=0):
        lib().orc_mma_settings(self.h, init, dec, inc, robust, conmod)

    def mma_set_acd(self, a, c, d):
        f = lambda v: np.ascontiguousarray(v, dtype=np.float64)
        lib().orc_mma_set_acd(self.h, f(a), f(c), f(d))

    def mma_state(self):
        k = C.c_int()
        v = [np.zeros(self.ne) for _ in range(4)]
        lib().orc_mma_state_get(self.h, C.byref(k), *v)
        return k.value, v[0], v[1], v[2], v[3]

    def mma_set_state(self, k, xo1, xo2, U, L):
        f = lambda a: np.ascontiguousarray(a, dtype=np.float64)
        lib().orc_mma_state_set(self.h, k, f(xo1), f(xo2), f(U), f(L))

    def mma_kkt(self, x, dfdx, gx, dgdx, xmin, xmax):
        f = lambda a: np.ascontiguousarray(a, dtype=np.float64).ravel()
        n2, ni = C.c_double(), C.c_double()
        lib().orc_mma_kkt(self.h, f(x), f(dfdx), f(gx), f(dgdx), f(xmin), f(xmax), C.byref(n2), C.byref(ni))
        return n2.value, ni.value

    def set_load_separate(self, on=True):
        lib().orc_set_load_separate(self.h, int(on))

    def case_u(self, k):
        p = lib().orc_case_u(self.h, k)
        return np.ctypeslib.as_array(p, shape=(3 * self.nn,)).copy() if p else None

    def stress(self, xphys, u):
        out = np.zeros(self.ne)
        lib().orc_stress(self.h, np.ascontiguousarray(xphys, dtype=np.float64),
                         np.ascontiguousarray(u, dtype=np.float64), out)
        return out

    def cpu_ref(self, nthreads=0):
        """The threaded reference-structure iteration (cpu_ref.c) on this
        problem's state; nthreads 0 = OpenMP's default."""
        return CpuRef(self, nthreads)

    def run(self, maxitr):
        fx = np.zeros(maxitr)
        gx = np.zeros(maxitr)
        ch = np.zeros(maxitr)
        n = lib().orc_run(self.h, maxitr, fx, gx, ch)
        return n, fx[:n], gx[:n], ch[:n]


def cantilever(etype=8, nx=8, ny=4, nz=4, jitter=0.0, seed=0, load_kind=0, **params):
    """Box [0,2]x[0,1]x[0,1] cantilever clamped at x=0 (BCs of orc_box_bcs)."""
    xyz, elem = gen_box(etype, nx, ny, nz, 2.0, 1.0, 1.0, jitter, seed)
    fixed, ld, lv = box_bcs(nx, ny, nz, load_kind)
    p = default_params(**params)
    return Problem(elem, xyz, fixed, ld, lv, p, nlc=8 if load_kind == 2 else 1)


class CpuRef:
    """cpu_ref.c on a Problem: one design iteration in the reference's
    structure (KE recomputed twice unless the problem caches KE, K
    re-assembled, smoothed-aggregation multigrid rebuilt, FGMRES(30), Jacobi-CG
    filter solves, the oracle's MMA) with OpenMP threads."""

    PHASES = ("assembly", "amg_setup", "solve", "sensitivities", "filter_sens", "mma", "filter")

    def __init__(self, prob, nthreads=0):
        self.prob = prob
        self.h = lib().crf_create(prob.h, nthreads)
        self.threads = lib().crf_threads(self.h)

    def __del__(self):
        if getattr(self, "h", None):
            lib().crf_destroy(self.h)
            self.h = None

    def ke_seconds(self):
        """Seconds to compute every element matrix once with these threads."""
        return lib().crf_time_ke(self.h)

    def design_iteration(self, itr):
        fx, gx, ch, rr = C.c_double(), (C.c_double * 8)(), C.c_double(), C.c_double()
        t = np.zeros(8)
        it = np.zeros(4, dtype=np.int32)
        rc = lib().crf_design_iteration(self.h, itr, C.byref(fx), gx, C.byref(ch), C.byref(rr), t, it)
        return {"fx": fx.value, "gx": gx[0], "ch": ch.value, "relres": rr.value, "rc": rc,
                "times": dict(zip(self.PHASES, t[:7].tolist())), "krylov": int(it[0]), "filter_cg": int(it[1]),
                "levels": int(it[2])}
