"""Writes tests/golden/configs/<name>.npz: the CPU oracle's results on the
full-size cantilever workloads, which the GPU parity tests compare against.

Workloads (box [0,2]x[0,1]x[0,1], x=0 clamped, interior nodes jittered
+-0.15 h, E0 = 1, Emin = 1e-9, nu = 0.3, penal = 3, both Krylov solvers at
rtol 1e-12 so the iterate sequences of the two solvers no longer matter):

  c0   40x20x20 hex8 (16,000), seed 0, volfrac 0.3, rmin 1.5 h, 10 iterations
  c1   140x70x40 hex8 (392,000), seed 392000, volfrac 0.3, rmin 1.5 h, 2 iterations
  c2   60x30x30 x 6 tet4 (324,000), seed 324000, 3x3 mid-height load patch,
       volfrac 0.25, rmin 2 h, 2 iterations
  c4a  the c1 mesh with the eight seeded 3x3 load patches solved as separate
       right-hand sides (compliance and sensitivities summed), 1 iteration
  c4b  the 32 uneven designs (8,000-64,000 hex8 each, sizes and volfrac from
       seed 32, rmin 1.5 * 2 / nx), each run alone, 2 iterations

Per iteration the script stores fx, gx, ch and the solvers' iteration
counts; for the full-size vectors (X, XPhys, dfdx) a fixed sample (every
stride-th entry, stride = ceil(ne / 4096)) together with the sum, the 2-norm,
the minimum and the maximum of the whole vector (c0 stores X in full).

The loop is main.cpp:64-119 (initial Filter + HeavisideFilter, then
ComputeObjectiveAndSensitivities, ComputeFilteredSens, SetOuterMovelimit,
Update, DesignChange, Filter, HeavisideFilter), as restated by the oracle.

Run (CPU, in the build container; c1 and c4a take tens of minutes each on
one core, so run them as separate processes):
    python tests/golden/make_configs.py c0 c2 c4b
    python tests/golden/make_configs.py c1
    python tests/golden/make_configs.py c4a
"""
import math
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, os.path.join(ROOT, "oracle"))
OUT = os.path.join(HERE, "configs")

TIGHT = dict(Emax=1.0, Emin=1e-9, nu=0.3, penal=3.0, rtol_el=1e-12, rtol_f=1e-12, ke_cache=1)

# name: (npe, nx, ny, nz, seed, load_kind, volfrac, rmin in units of h = 2/nx, iterations)
CONFIGS = {
    "c0": (8, 40, 20, 20, 0, 0, 0.30, 1.5, 10),
    "c1": (8, 140, 70, 40, 392000, 0, 0.30, 1.5, 2),
    "c2": (4, 60, 30, 30, 324000, 1, 0.25, 2.0, 2),
    "c4a": (8, 140, 70, 40, 392000, 2, 0.30, 1.5, 1),
}
NSAMPLE = 4096


def uneven_designs(n=32, seed=32, ne_min=8000, ne_max=64000):
    """The 32 uneven cantilevers: per design ny in [10, 33], nz = ny + {-1, 0, 1},
    nx = 2 ny + {-2..2} from a SplitMix64 stream (u = (r >> 11) 2^-53), kept
    when ne_min <= nx ny nz <= ne_max; volfrac = 0.2 + 0.2 u; jitter seed
    1000 seed + index."""
    M = (1 << 64) - 1
    s = seed & M

    def u():
        nonlocal s
        s = (s + 0x9E3779B97F4A7C15) & M
        z = s
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & M
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & M
        return ((z ^ (z >> 31)) >> 11) * 2.0 ** -53

    out = []
    while len(out) < n:
        ny = 10 + int(u() * 24)
        nz = ny + int(u() * 3) - 1
        nx = 2 * ny + int(u() * 5) - 2
        vf = 0.2 + 0.2 * u()
        if ne_min <= nx * ny * nz <= ne_max:
            out.append(dict(nx=nx, ny=ny, nz=nz, volfrac=vf, seed=1000 * seed + len(out)))
    return out


def summary(prefix, v, out):
    """Sample and whole-vector statistics of v under keys prefix_*."""
    v = np.asarray(v, dtype=np.float64)
    stride = max(1, math.ceil(len(v) / NSAMPLE))
    out[prefix + "_stride"] = np.int64(stride)
    out[prefix + "_sample"] = v[::stride].copy()
    out[prefix + "_stats"] = np.array([v.sum(), np.linalg.norm(v), v.min(), v.max()])


def run_single(name):
    import oracle as O
    npe, nx, ny, nz, seed, lk, vf, rf, nit = CONFIGS[name]
    prm = dict(TIGHT, volfrac=vf, rmin=rf * 2.0 / nx)
    t0 = time.time()
    P = O.cantilever(npe, nx, ny, nz, jitter=0.15, seed=seed, load_kind=lk, **prm)
    if lk == 2:
        P.set_load_separate(True)
    P.init_design()
    out = {"dims": np.array([npe, nx, ny, nz, seed, lk]), "volfrac": vf, "rmin": prm["rmin"],
           "rtol": 1e-12, "iterations": np.int64(nit)}
    fx, gx, ch, it_el, it_f, relres = [], [], [], [], [], []
    for itr in range(1, nit + 1):
        f, g, c, rc = P.design_iteration(itr)
        assert rc == 0, (name, itr, rc)
        d = P.diag()
        fx.append(f); gx.append(g[0]); ch.append(c)
        it_el.append(d["it_el"]); it_f.append(d["it_f"]); relres.append(d["relres_el"])
        print(f"[{name}] it {itr}: fx {f:.17g} gx {g[0]:.17g} ch {c:.17g} pcg {d['it_el']} "
              f"relres {d['relres_el']:.3e} filter {d['it_f']} ({time.time() - t0:.0f} s)", flush=True)
        if itr == 1:
            summary("dfdx1", P.vec("dfdx"), out)
    out.update(fx=np.array(fx), gx=np.array(gx), ch=np.array(ch), it_el=np.array(it_el),
               it_f=np.array(it_f), relres_el=np.array(relres))
    if name == "c0":
        out["X"] = P.vec("X").copy()
        out["XPhys"] = P.vec("XPhys").copy()
    summary("X", P.vec("X"), out)
    summary("XPhys", P.vec("XPhys"), out)
    out["seconds"] = time.time() - t0
    return out


def run_c4b():
    import oracle as O
    ds = uneven_designs()
    nit = 2
    fx = np.zeros((len(ds), nit))
    gx = np.zeros((len(ds), nit))
    ch = np.zeros((len(ds), nit))
    xs = np.zeros((len(ds), 4))
    t0 = time.time()
    for k, D in enumerate(ds):
        P = O.cantilever(8, D["nx"], D["ny"], D["nz"], jitter=0.15, seed=D["seed"], load_kind=0,
                         **dict(TIGHT, volfrac=D["volfrac"], rmin=1.5 * 2.0 / D["nx"]))
        P.init_design()
        for itr in range(1, nit + 1):
            f, g, c, rc = P.design_iteration(itr)
            assert rc == 0, (k, itr, rc)
            fx[k, itr - 1], gx[k, itr - 1], ch[k, itr - 1] = f, g[0], c
        x = P.vec("X")
        xs[k] = [x.sum(), np.linalg.norm(x), x.min(), x.max()]
        print(f"[c4b] design {k} {D['nx']}x{D['ny']}x{D['nz']}: fx {fx[k].tolist()} ({time.time() - t0:.0f} s)",
              flush=True)
        del P
    return {"designs": np.array([[D["nx"], D["ny"], D["nz"], D["seed"]] for D in ds]),
            "volfrac": np.array([D["volfrac"] for D in ds]), "fx": fx, "gx": gx, "ch": ch, "X_stats": xs,
            "iterations": np.int64(nit), "rtol": 1e-12, "seconds": time.time() - t0}


def main(names):
    os.makedirs(OUT, exist_ok=True)
    for name in names:
        res = run_c4b() if name == "c4b" else run_single(name)
        path = os.path.join(OUT, f"{name}.npz")
        np.savez_compressed(path, **res)
        print(f"wrote {path} ({os.path.getsize(path)} B, {res['seconds']:.0f} s)", flush=True)


if __name__ == "__main__":
    main(sys.argv[1:] or ["c0", "c2", "c4b", "c1", "c4a"])
