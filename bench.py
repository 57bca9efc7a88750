#!/usr/bin/env python3
"""Throughput of the SIMP design iteration on one MI355X.

A step is one design iteration (the body of the reference's main.cpp:72-107:
elasticity solve + objective/sensitivities, sensitivity filter, MMA update,
design change, density filter, projection) on the 392,000-element jittered
hex8 cantilever (140x70x40 on [0,2]x[0,1]x[0,1], interior nodes jittered
+-0.15h from seed 392000, E0=1, Emin=1e-9, nu=0.3, penal=3, volfrac=0.3,
rmin=1.5h, x=0 clamped, -1/41 on the y-DOFs of the bottom edge of x=2),
Krylov rtol 1e-8 for both solvers.  All inputs are resident in HBM
before the timed region; setup (element operators, hierarchy) is reported
separately.

Prints one JSON line.  Extra fields: roofline (the dominant fine operator
kernel: 20 back-to-back launches after the timed steps on the last refresh's
operators, HIP events on its launch stream; the events around one in every
16 launches inside the timed steps are kept beside it) and cpu_baseline (one full
design iteration of the same workload in the reference's structure on the
host cores: oracle/cpu_ref.c with OpenMP threads).

  python bench.py [--steps K] [--warmup W] [--precision both|fp64|mixed|mixed_h] [--no-cpu-baseline]

By default three forms are timed on the same workload: FP64; UFT_MIXED (FP32
outer operator, FP32 multigrid cycle and FP32 smoother blocks inside the FP64
flexible outer loop, the precision form as specified); and UFT_MIXED_H (as
UFT_MIXED with the smoother's fine operator assembled into int8 node blocks
with an FP32 scale per block, FP16 level-1 smoother blocks).  value is the
UFT_MIXED_H rate; config carries the other rates, the ratios and the largest
relative compliance difference against the FP64 form over the timed steps.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
PKG = os.path.join(ROOT, "unstructured-fea-based-topology-optimization-in-petsc-and-eigen_amd")
sys.path.insert(0, PKG)

HBM_PEAK_GBS = 8000.0  # MI355X HBM3E

WORKLOADS = {
    # name: (npe, nx, ny, nz, seed, load_kind, volfrac, rmin_h_factor)
    "c1": (8, 140, 70, 40, 392000, 0, 0.30, 1.5),
    "c0": (8, 40, 20, 20, 0, 0, 0.30, 1.5),
    "c2": (4, 60, 30, 30, 324000, 1, 0.25, 2.0),
    # the c1 mesh with eight separate load cases (seeded 3x3 patches on x=2,
    # unit total each), compliance summed over the cases
    "c4a": (8, 140, 70, 40, 392000, 2, 0.30, 1.5),
    # 32 independent cantilevers of uneven size (8,000-64,000 hex8 each,
    # 868,008 in total; sizes and volfrac ~ U(0.2, 0.4) drawn from seed 32),
    # one context: a step advances every design by one iteration
    "c4b": (8, None, None, None, 32, 0, None, 1.5),
}

def workload(name, precision):
    npe, nx, ny, nz, seed, lk, vf, rf = WORKLOADS[name]
    h = 2.0 / nx
    params = dict(Emax=1.0, Emin=1e-9, nu=0.3, penal=3.0, volfrac=vf, rmin=rf * h, rtol_el=1e-8, rtol_f=1e-8,
                  precision=precision, load_cases=1 if lk == 2 else 0)
    return npe, (nx, ny, nz), seed, lk, params


def algorithmic_bytes_elem_phase(npe, ne, ndof, w=8, nrhs=1):
    """Element phase of one fine stiffness apply: the packed unique KE entries
    (w bytes each), the element's node ids and FP64 density per element, plus
    one read of u (FP64 for w = 8, else FP32); FP16 blocks (w = 2) add the
    FP32 per-element scale."""
    S = 300 if npe == 8 else 78
    wv = 8 if w == 8 else 4
    if nrhs > 1:  # BF16 blocks (w = 2, no scale) or FP64 blocks, nrhs interleaved vectors
        return ne * (S * w + 4 * npe + 8) + nrhs * ndof * wv
    return ne * (S * w + 4 * npe + 8 + (4 if w == 2 else 0)) + ndof * wv


def algorithmic_bytes_fine_ell(blocks, ndof, col_bytes, value_bytes, scaled=True, nrhs=1, step=False):
    """One apply of an assembled fine operator: every nonzero 3x3 node block
    once (value_bytes: 12 for int8 entries with a block scale, 18 for FP16, 36
    for FP32, and its column index) for all nrhs right-hand sides, u read once
    and one output vector per right-hand side (FP32), the row scales of the
    scaled smoother forms, and the fixed-DOF byte per node (padding slots not
    counted).  step: the launch is a fused Chebyshev step (r -= A u;
    d = c1 u + c2 D^-1 r; x += d), whose epilogue reads and writes r and x
    and writes d instead of an output vector, and reads D^-1: per DOF
    4 nrhs (u) + 4 (row scale) + 4 (D^-1) + 20 nrhs (r, x, d) bytes."""
    if step:
        return blocks * (value_bytes + col_bytes) + ndof * (24 * nrhs + (8 if scaled else 4)) + ndof // 3
    return blocks * (value_bytes + col_bytes) + ndof * 4 * (2 * nrhs + (1 if scaled else 0)) + ndof // 3


def algorithmic_bytes_apply(npe, ne, ndof, w=8):
    S = 300 if npe == 8 else 78
    return ne * (S * w + 4 * npe + w) + 2 * ndof * w


def cpu_threads():
    """Host threads for the CPU baseline: the cores this process may run on,
    at most 16 (one GPU's share of the host)."""
    try:
        n = len(os.sched_getaffinity(0))
    except AttributeError:
        n = os.cpu_count() or 1
    return max(1, min(16, n))


def cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(name):
    """One full design iteration of the reference's structure on the host
    cores (oracle/cpu_ref.c through the oracle's wrapper): KE computed twice,
    K re-assembled, the smoothed-aggregation multigrid rebuilt, FGMRES(30) to
    rtol 1e-8 counting its real iterations, Jacobi-CG filter solves, MMA."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O
    npe, (nx, ny, nz), seed, lk, prm = workload(name, 0)
    prm.pop("precision")
    prm.pop("load_cases")
    threads = cpu_threads()
    t0 = time.time()
    P = O.cantilever(npe, nx, ny, nz, jitter=0.15, seed=seed, load_kind=lk, ke_cache=0, **prm)
    P.init_design()
    R = P.cpu_ref(threads)
    t_setup = time.time() - t0
    r = R.design_iteration(1)
    t_iter = sum(r["times"].values())
    t_ke = R.ke_seconds()
    sample = (f"one full design iteration (the first, from the uniform initial design) of the {name} workload in "
              f"the reference's structure on {R.threads} host threads (oracle/cpu_ref.c, OpenMP): KE computed "
              f"twice, K re-assembled, a 4-level smoothed-aggregation multigrid rebuilt, FGMRES(30) to rtol 1e-8 "
              f"({r['krylov']} iterations, {r['levels']} levels), Jacobi-CG filter solves ({r['filter_cg']} "
              f"iterations), the oracle's MMA; setup {t_setup:.1f}s excluded")
    return {"value": 1.0 / t_iter, "unit": "design iters/s", "cores": R.threads, "kind": "port", "sample": sample,
            "cpu_model": cpu_model(), "seconds_per_iter": t_iter, "phases_s": r["times"],
            "krylov_iters": r["krylov"], "compliance": r["fx"], "relres": r["relres"],
            "ke_all_elements_s": t_ke,
            "ke_cached_value": 1.0 / max(t_iter - 2.0 * t_ke, 1e-9),
            "ke_cached_note": "the same iteration less two computations of every element matrix (a CPU code "
                              "that caches KE at setup)"}


PREC = {"fp64": 0, "mixed": 1, "mixed_h": 2}


def run_form(args, precision):
    """Set up the workload in one precision form and time args.steps design
    iterations after args.warmup untimed ones."""
    import ctypes
    import uft
    from uft.capi import Context

    ndesign = 1
    if args.workload == "c4b":
        from uft.capi import uneven_designs, batch_cantilever, batch_mesh
        ds = uneven_designs()
        ndesign = len(ds)
        npe, dims, seed = 8, None, 32
        t0 = time.time()
        ctx = batch_cantilever(ds, jitter=0.15, rmin_h=1.5, precision=PREC[precision], Emax=1.0, Emin=1e-9, nu=0.3,
                               penal=3.0, rtol_el=1e-8, rtol_f=1e-8, graph=args.graph)
        ctx.sync()
        setup_s = time.time() - t0
    else:
        npe, dims, seed, lk, prm = workload(args.workload, PREC[precision])
        prm["graph"] = args.graph
        xyz, elem = uft.gen_box(npe, *dims, jitter=0.15, seed=seed)
        fixed, ld, lv = uft.box_bcs(*dims, load_kind=lk)
        ctx = Context(0)
        ctx.set_params(Context.default_params(**prm))
        ctx.set_mesh(elem, xyz)
        ctx.set_bcs(fixed, ld, lv, nlc=8 if lk == 2 else 1)
        t0 = time.time()
        ctx.setup()
        ctx.sync()
        setup_s = time.time() - t0
    ne, ndof = ctx.size("ne"), ctx.size("ndof")
    ell = ctx.fine_operator_info()
    tiles = ctx.fine_tiles_info()
    ctx.init_design()
    itr = 0
    for _ in range(args.warmup):
        itr += 1
        ctx.design_iteration(itr)
    L = ctx.L
    L.uft_profile_enable(ctx.h, args.timers)
    its, fxs, per_design = [], [], []
    ctx.sync()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        itr += 1
        fx, gx, ch = ctx.design_iteration(itr)
        its.append(ctx.stats()["it_el"])
        fxs.append(fx)
        if ndesign > 1:
            br = ctx.batch_results()
            per_design.append({"fx": br["fx"].tolist(), "it_el": br["it_el"].tolist()})
    ctx.sync()
    t1 = time.perf_counter()
    prof = {}
    for kind in range(6):
        cnt, ms = ctypes.c_int64(), ctypes.c_double()
        L.uft_profile_read(ctx.h, kind, ctypes.byref(cnt), ctypes.byref(ms))
        prof[kind] = (cnt.value, ms.value)
    solve_bytes, solve_ms, applies, solves = ctx.solve_traffic()
    L.uft_profile_enable(ctx.h, 0)
    # back-to-back launches of the fine operators on the last refresh's data
    # (HIP events on the context's stream; one design; with load cases all
    # right-hand sides at once, and the FP64 element phase alone not timed)
    b2b = {}
    if ndesign == 1:
        if WORKLOADS[args.workload][5] == 2:
            kinds = {"fp64": (), "mixed": (1,), "mixed_h": (2, 1)}[precision]
        else:
            kinds = {"fp64": (0, 3), "mixed": (2, 1) if ell[0] else (1,), "mixed_h": (2, 1)}[precision]
        for k in kinds:
            b2b[k] = ctx.time_fine_operator(k, 20)
    ctx.close()
    return {"npe": npe, "dims": dims, "seed": seed, "ne": ne, "ndof": ndof, "setup_s": setup_s, "dt": t1 - t0,
            "its": its, "fxs": fxs, "prof": prof, "b2b": b2b, "ndesign": ndesign, "per_design": per_design, "ell": ell,
            "tiles": tiles,
            "solve": {"algorithmic_bytes": solve_bytes, "ms": solve_ms, "operator_applies": applies,
                      "solves": solves, "gbs": solve_bytes / solve_ms / 1e6 if solve_ms > 0 else None,
                      "ms_per_step": solve_ms / max(solves, 1)}}


FULL_RUN_ITERS = {"c0": 10, "c1": 50, "c2": 50, "c4a": 50}


def full_run(args, precision):
    """The whole loop of main.cpp:64-119 on the workload: the initial filter
    and projection, then design iterations while itr < maxItr and ch > 0.01
    (main.cpp:70), with a device synchronisation at every iteration boundary.
    Returns per-iteration compliance, design change, Krylov iterations and
    wall time, and why the loop stopped."""
    import uft
    from uft.capi import Context
    npe, dims, seed, lk, prm = workload(args.workload, PREC[precision])
    maxitr = FULL_RUN_ITERS[args.workload]
    xyz, elem = uft.gen_box(npe, *dims, jitter=0.15, seed=seed)
    fixed, ld, lv = uft.box_bcs(*dims, load_kind=lk)
    ctx = Context(0)
    ctx.set_params(Context.default_params(maxItr=maxitr, **prm))
    ctx.set_mesh(elem, xyz)
    ctx.set_bcs(fixed, ld, lv, nlc=8 if lk == 2 else 1)
    t0 = time.perf_counter()
    ctx.setup()
    ctx.sync()
    setup_s = time.perf_counter() - t0
    ctx.init_design()
    ctx.sync()
    fxs, chs, its, ms = [], [], [], []
    ch, itr = 1.0, 0
    while itr < maxitr and ch > 0.01:
        itr += 1
        t1 = time.perf_counter()
        fx, gx, ch = ctx.design_iteration(itr)
        ctx.sync()
        ms.append((time.perf_counter() - t1) * 1e3)
        fxs.append(fx)
        chs.append(ch)
        its.append(ctx.stats()["it_el"])
        print(f"[{args.workload} {precision}] itr {itr} fx {fx:.12g} ch {ch:.4g} krylov {its[-1]} {ms[-1]:.1f} ms",
              file=sys.stderr, flush=True)
    ctx.close()
    stop = f"ch = {ch:.4g} <= 0.01 at itr {itr}" if ch <= 0.01 else f"itr = maxItr = {maxitr} (ch = {ch:.4g})"
    return {"precision": precision, "iterations": itr, "stop": stop, "setup_s": setup_s,
            "design_iters_per_s": itr / (sum(ms) / 1e3), "total_s": sum(ms) / 1e3,
            "krylov_iters_per_solve": sum(its) / len(its), "krylov_iters": its, "fx": fxs, "ch": chs,
            "ms_per_iter": [round(v, 2) for v in ms]}


def elem_roofline(name, npe, ne, ndof, prof_elem, prof_node, word, kernel, traffic_key, nrhs=1, algo=None, b2b_us=None):
    """HBM roofline of an element phase (or of the assembled operator's
    apply, algo given) from its HIP-event timings: the back-to-back
    measurement after the timed steps when given (b2b_us), else the events
    around one in every TIMERS launches inside the timed steps (those include
    each timed launch's dispatch)."""
    n, ms = prof_elem
    if algo is None:
        algo = algorithmic_bytes_elem_phase(npe, ne, ndof, word, nrhs)
    avg_s = ms / max(n, 1) / 1e3
    in_run = algo / avg_s / 1e9 if n else None
    if b2b_us:
        avg_s = b2b_us * 1e-6
    achieved = algo / avg_s / 1e9 if (n or b2b_us) else None
    traffic = None
    tfile = os.path.join(ROOT, "profiles", "traffic.json")
    if os.path.exists(tfile):
        traffic = json.load(open(tfile)).get(traffic_key)
    return {"bound": "hbm", "role": name, "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS if achieved else None, "traffic": traffic,
            "traffic_source": ("not measured in this run: HBM bytes per dispatch of the same kernel from "
                               "rocprofv3 --pmc FETCH_SIZE and WRITE_SIZE runs, scaled by a plain copy's, "
                               f"profiles/traffic.json[{traffic_key}]") if traffic else None,
            "kernel": kernel,
            "algorithmic_bytes_per_launch": algo, "launches": n, "avg_launch_us": avg_s * 1e6,
            "timing": ("20 back-to-back launches after the timed steps, HIP events on the launch stream" if b2b_us
                       else "HIP events around one in every TIMERS launches of the timed steps"),
            "achieved_in_run": in_run, "avg_launch_us_in_run": ms / max(n, 1) * 1e3 if n else None,
            "total_ms_timed": ms, "node_phase_avg_us": prof_node[1] / max(prof_node[0], 1) * 1e3}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--workload", default="c1", choices=sorted(WORKLOADS))
    ap.add_argument("--precision", default="both", choices=["both", "fp64", "mixed", "mixed_h"],
                    help="both (default): time the FP64, UFT_MIXED and UFT_MIXED_H forms; the value is the "
                         "UFT_MIXED_H rate")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--timers", type=int, default=16,
                    help="HIP-event timers around one in every TIMERS fine applies of each kind in the timed "
                         "region (the roofline figures; default 16, 1 = every apply), 0 = none")
    ap.add_argument("--graph", type=int, default=0, choices=[0, 1],
                    help="replay each Krylov iteration from a captured hipGraph (1) or launch it kernel by "
                         "kernel (0, default; with timers on the solve launches kernel by kernel anyway)")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="only run a plain 1 GiB device copy three times (reference for rocprof PMC scaling)")
    ap.add_argument("--full-run", action="store_true",
                    help="instead of the timed steps, run the workload's whole optimisation loop (c0: 10, c1 and "
                         "c2: 50 iterations, stopping early on ch <= 0.01 as main.cpp:70) in each form and print "
                         "per-iteration compliance, design change, Krylov iterations and times")
    args = ap.parse_args()
    if args.copy_calibration:
        import torch
        a = torch.ones(2 ** 27, dtype=torch.float64, device="cuda")
        b = torch.empty_like(a)
        for _ in range(3):
            b.copy_(a)
        torch.cuda.synchronize()
        print(json.dumps({"copy_bytes": a.numel() * 8, "copies": 3}))
        return
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one MI355X: the path runs on one GPU (it would shard by element "
              "partition with halo exchanges, not implemented). Run with --gpus 1.", file=sys.stderr)
        sys.exit(2)

    forms = ["fp64", "mixed", "mixed_h"] if args.precision == "both" else [args.precision]
    if args.full_run:
        if args.workload not in FULL_RUN_ITERS:
            sys.exit(f"--full-run covers {sorted(FULL_RUN_ITERS)}")
        runs = {f: full_run(args, f) for f in forms}
        out = {"mode": "full_run", "workload": args.workload, "max_itr": FULL_RUN_ITERS[args.workload],
               "runs": runs}
        if "fp64" in runs:
            ref = runs["fp64"]["fx"]
            for f, r in runs.items():
                k = min(len(ref), len(r["fx"]))
                r["max_rel_compliance_diff_vs_fp64"] = max(abs(a - b) / abs(b) for a, b in zip(r["fx"][:k], ref[:k]))
        print(json.dumps(out))
        return
    res = {f: run_form(args, f) for f in forms}
    main_form = forms[-1]
    r = res[main_form]
    npe, dims, seed, ne, ndof = r["npe"], r["dims"], r["seed"], r["ne"], r["ndof"]
    value = args.steps * r["ndesign"] / r["dt"]
    ename = "hex8" if npe == 8 else "tet4"
    rooflines = []
    multi = WORKLOADS[args.workload][5] == 2
    nrhs = 8 if multi else 1
    k64 = f"k_mr_elem<{npe},double,8>" if multi else f"k_elem_apply_v<{npe},double>"
    if "fp64" in res and res["fp64"]["ell"][3] == 72:
        # the FP64 form's solves apply the assembled FP64 node blocks: blocks
        # (72 bytes and the column) once, u read and the output written in FP64
        q = res["fp64"]
        algo = q["ell"][1] * (72 + q["ell"][2]) + ndof * 8 * 2 + ndof // 3
        rooflines.append(elem_roofline("fp64", npe, ne, ndof, q["prof"][0], q["prof"][1], 8,
                                       "k_fine_ell_op64 (assembled FP64 3x3 node blocks of the refresh, the FP64 "
                                       "form's whole K u in one launch)", f"{args.workload}_fp64_fine_ell", nrhs,
                                       algo=algo, b2b_us=q["b2b"].get(0)))
        if q["b2b"].get(3):
            rooflines.append({"bound": "hbm", "role": "fp64_elem_phase_unused", "kernel": f"{k64} (the element "
                              "phase of the matrix-free apply, which the FP64 solves no longer use; timed for "
                              "comparison)", "avg_launch_us": q["b2b"][3],
                              "algorithmic_bytes_per_launch": algorithmic_bytes_elem_phase(npe, ne, ndof, 8, nrhs),
                              "achieved": algorithmic_bytes_elem_phase(npe, ne, ndof, 8, nrhs) / (q["b2b"][3] * 1e-6)
                              / 1e9, "peak": HBM_PEAK_GBS, "unit": "GB/s", "launches": 0, "total_ms_timed": 0.0})
    elif "fp64" in res:
        q = res["fp64"]
        rooflines.append(elem_roofline("fp64", npe, ne, ndof, q["prof"][0], q["prof"][1], 8,
                                       f"{k64} (FP64 element phase of the fine stiffness apply, FP64 form)",
                                       f"{args.workload}_fp64_elem_apply", nrhs, b2b_us=q["b2b"].get(3)))
        if not multi:
            n, ms = q["prof"][0]
            nn_, msn = q["prof"][1]
            if n and nn_:
                us = q["b2b"].get(0) or (ms / n + msn / nn_) * 1e3
                algo = algorithmic_bytes_apply(npe, ne, ndof, 8)
                rooflines.append({
                    "bound": "hbm", "role": "fp64_k_apply", "achieved": algo / (us * 1e-6) / 1e9,
                    "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": algo / (us * 1e-6) / 1e9 / HBM_PEAK_GBS,
                    "traffic": None, "kernel": f"k_elem_apply_v<{npe},double> + k_node_gather_op (whole FP64 "
                    "K u: element phase plus node phase, FP64 form)", "algorithmic_bytes_per_launch": algo,
                    "launches": n, "avg_launch_us": us, "total_ms_timed": ms + msn,
                    "timing": ("20 back-to-back applies after the timed steps, HIP events on the launch stream"
                               if q["b2b"].get(0) else "HIP events in the timed steps")})
    if main_form != "fp64":
        if multi:
            w, kname = 2, (f"k_mr_elem_mfma<{npe}> (BF16 v_mfma_f32_16x16x32 element products of the "
                           f"preconditioner, 8 load cases, mixed form)")
        elif main_form == "mixed_h":
            w, kname = 2, f"k_elem_apply_h<{npe}> (FP16-block element phase of the smoother, mixed form)"
        else:
            w, kname = 4, f"k_elem_apply_v<{npe},float> (FP32 element phase, mixed form)"
        if r["ell"][0]:  # assembled fine smoother operator: one kernel per apply
            if multi:
                kn = ("k_mr_tile_op<FP16> (8 load cases on node tiles, neighbour entries staged in LDS)"
                      if r["tiles"][0] else f"k_mr_ell_op<{nrhs},FP16> (8 load cases)")
            else:
                kn = "k_fine_ell_op32" if r["ell"][3] == 36 else "k_fine_ell_op_q"
            fmt = {18: "FP16", 12: "int8 (FP32 block scale)", 36: "FP32"}.get(r["ell"][3], "low-precision")
            sc = r["ell"][3] != 36  # the low-precision smoother blocks are scaled by row
            op_bytes = algorithmic_bytes_fine_ell(r["ell"][1], ndof, r["ell"][2], r["ell"][3], sc, nrhs)
            step_bytes = algorithmic_bytes_fine_ell(r["ell"][1], ndof, r["ell"][2], r["ell"][3], sc, nrhs, step=True)
            b2b = r["b2b"].get(2)
            rl = elem_roofline(
                main_form, npe, ne, ndof, r["prof"][2], r["prof"][3], 2,
                f"{kn} (assembled {fmt} 3x3 node blocks of the smoother, sliced ELL, a fused Chebyshev step: "
                "the apply with its row epilogue, mixed form)", f"{args.workload}_{main_form}_fine_ell", nrhs,
                algo=step_bytes if b2b else op_bytes, b2b_us=b2b)
            # the operator alone (K u: blocks, u, one output, row scales), and
            # the in-run events (every epilogue mode of the smoother's launches)
            # priced on the operator alone
            rl["operator_bytes_per_launch"] = op_bytes
            if rl["avg_launch_us"]:
                rl["operator_frac"] = op_bytes / (rl["avg_launch_us"] * 1e-6) / 1e9 / HBM_PEAK_GBS
            n_in = r["prof"][2][0]
            if n_in:
                rl["achieved_in_run"] = op_bytes / (r["prof"][2][1] / n_in / 1e3) / 1e9
            if b2b:
                rl["algorithmic_bytes_note"] = ("the Chebyshev step's operator bytes plus its epilogue's vectors, "
                                                "each read or written once")
            rooflines.append(rl)
        else:
            rooflines.append(elem_roofline(main_form, npe, ne, ndof, r["prof"][2], r["prof"][3], w, kname,
                                           f"{args.workload}_{main_form}_elem_apply", nrhs))
        if r["prof"][4][0] and r["ell"][0]:  # FP32 outer operator from the assembled blocks
            if multi:
                kn = ("k_mr_tile_op<FP32> (8 load cases on node tiles, neighbour entries staged in LDS)"
                      if r["tiles"][0] else f"k_mr_ell_node<{nrhs},FP32> (8 load cases)")
            else:
                kn = "k_fine_ell_op32"
            rooflines.append(elem_roofline(
                "outer", npe, ne, ndof, r["prof"][4], r["prof"][5], 4,
                f"{kn} (assembled FP32 3x3 node blocks of the outer operator, whole apply, mixed form)",
                f"{args.workload}_fp32_fine_ell", nrhs,
                algo=algorithmic_bytes_fine_ell(r["ell"][1], ndof, r["ell"][2], 36, False, nrhs),
                b2b_us=r["b2b"].get(1)))
        elif r["prof"][4][0]:  # FP32 outer operator, FP64 residual at reliable updates
            k32 = f"k_mr_elem<{npe},float,8>" if multi else f"k_elem_apply_v<{npe},float>"
            rooflines.append(elem_roofline("outer", npe, ne, ndof, r["prof"][4], r["prof"][5], 4,
                                           f"{k32} (FP32 element phase of the outer operator, mixed form)",
                                           f"{args.workload}_fp32_elem_apply", nrhs))
        if r["prof"][0][0]:
            rooflines.append(elem_roofline("outer64", npe, ne, ndof, r["prof"][0], r["prof"][1], 8,
                                           f"{k64} (FP64 element phase of the outer loop's residuals, mixed "
                                           f"form)", f"{args.workload}_fp64_elem_apply", nrhs))
    if "mixed" in res and main_form == "mixed_h" and multi:
        q = res["mixed"]
        rooflines.append(elem_roofline("bf16_mfma", npe, ne, ndof, q["prof"][2], q["prof"][3], 2,
                                       f"k_mr_elem_mfma<{npe}> (BF16 v_mfma_f32_16x16x32 element products of the "
                                       f"preconditioner, 8 load cases, UFT_MIXED form)",
                                       f"{args.workload}_mixed_elem_apply", nrhs))
    # the dominant element phase of the timed form: the one with the most time
    cand = [x for x in rooflines if x["launches"] and x["role"] != "fp64_k_apply"] if main_form == "fp64" else \
        [x for x in rooflines if x["launches"] and x["role"] not in ("fp64", "bf16_mfma", "fp64_k_apply")]
    dom = max(cand, key=lambda x: x["total_ms_timed"]) if cand else rooflines[-1]
    if r["ndesign"] > 1:
        wl = (f"{args.workload}: {r['ndesign']} independent {ename} jittered cantilevers of uneven size "
              f"({ne} elements in total, 8,000-64,000 each, sizes and volfrac drawn from seed {seed}), one context, "
              f"rtol 1e-8, {main_form} form; value counts design iterations summed over the designs")
    else:
        wl = (f"{args.workload}: {ne} {ename} jittered cantilever {dims[0]}x{dims[1]}x{dims[2]}, "
              f"seed {seed}, rtol 1e-8, {main_form} form")
    config = {"workload": wl,
              "global_batch": r["ndesign"], "graph": args.graph, "seq_len": None, "parallelism": "single GPU", "precision": main_form,
              "setup_s": r["setup_s"], "krylov_iters_per_solve": sum(r["its"]) / len(r["its"]),
              "final_fx": r["fxs"][-1]}
    if "fp64" in res and main_form != "fp64":
        f64 = res["fp64"]
        config["fp64_design_iters_per_s"] = args.steps * f64["ndesign"] / f64["dt"]
        config["mixed_over_fp64"] = value / config["fp64_design_iters_per_s"]
        config["fp64_krylov_iters_per_solve"] = sum(f64["its"]) / len(f64["its"])
        config["max_rel_compliance_diff_vs_fp64"] = max(abs(a - b) / abs(b) for a, b in zip(r["fxs"], f64["fxs"]))
    config["krylov_iters_per_step"] = r["its"]
    # Solve GB/s: algorithmic bytes of every operator apply inside the timed solves (fine and coarse
    # levels, the coarsest dense inverse; transfers and vector updates not counted) over the solves'
    # wall time (hierarchy refresh excluded)
    config["solve_gbs"] = {f: res[f]["solve"] for f in forms}
    if r["tiles"][0]:
        config["node_tiles"] = {"tiles": r["tiles"][0], "slots": r["tiles"][1], "max_halo_nodes": r["tiles"][2]}
    if "fp64" in res and main_form != "fp64":
        config["fp64_krylov_iters_per_step"] = res["fp64"]["its"]
    if "mixed" in res and main_form == "mixed_h":
        q = res["mixed"]
        key = "bf16_mfma_form" if multi else "fp32_smoother_form"
        config[f"{key}_design_iters_per_s"] = args.steps * q["ndesign"] / q["dt"]
        config[f"{key}_krylov_iters_per_solve"] = sum(q["its"]) / len(q["its"])
        if "fp64" in res:
            config[f"{key}_over_fp64"] = config[f"{key}_design_iters_per_s"] / config["fp64_design_iters_per_s"]
            config[f"{key}_max_rel_compliance_diff_vs_fp64"] = max(
                abs(a - b) / abs(b) for a, b in zip(q["fxs"], res["fp64"]["fxs"]))
    out = {
        "metric": "design iters/s @392k hex8 cantilever; mixed vs FP64 GPU at 1e-6 rel compliance",
        "value": value, "unit": "design iters/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": r["dt"] / args.steps * 1e3, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f64" if main_form == "fp64" else "f32+f64",
        "data": "synthetic",
        "config": config,
        "roofline": {k: v for k, v in dom.items()},
        "rooflines": rooflines,
    }
    if not args.no_cpu_baseline and args.workload in ("c1", "c0", "c2"):
        try:
            cb = cpu_baseline(args.workload)
            out["cpu_baseline"] = cb
            out["vs_cpu_baseline"] = value / cb["value"]
            if "fp64" in res:
                out["fp64_vs_cpu_baseline"] = config.get("fp64_design_iters_per_s", value) / cb["value"]
        except Exception as e:  # the baseline is reported, never required
            out["cpu_baseline"] = {"value": None, "error": repr(e)}
    print(json.dumps(out))


if __name__ == "__main__":
    main()
