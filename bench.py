"""Benchmark: preconditioned solves/s on one MI355X.

Default workload (--config C2, the configuration the metric is quoted on):
the 3-D convection-diffusion operator on a 160^3 grid (n = 4,096,000,
28.5 M nonzeros), INVK preconditioner (ILU(1) factors, inverse fill level
1), BiCGSTAB to 1e-8 relative residual with FP64 vectors, dots and
residuals.  Right-hand sides b_j = A x_j: x_0 = 1 (the reference driver's
convention), x_j ~ N(0,1) from a fixed seed for j >= 1.  All right-hand
sides are resident in HBM before any timed region.

A step is one pass of the served form over a queue of --queue right-hand
sides: FP32-stored factors, --slots column slots of one captured Krylov
iteration, a slot refilled from the queue as soon as its column converges
(continuous batching).  value = right-hand sides solved per second.

Also reported, on right-hand sides of the same kind: the plain FP64
single-RHS GPU solve (one right-hand side at a time, FP64 factors) and the
speedup over it; per-solve latency p50 / p99 in the served form; the
roofline of the dominant kernel (the first apply pass t = d o (W x) at the
served block width; its average duration measured with HIP events on the
solver's stream); the CPU baseline (the oracle's Krylov iterations, single
thread, same factors).

C5 (--config C5) measures the serving engine instead: a pool of 24
matrices (12 pde3d with N ~ U{20..120}, INVK(1,1); 12 random as C4 with
n ~ logU[1e4, --c5-nmax], AINV-LLK 5e-2; FP32 factors, BiCGSTAB 1e-8)
registered with one server; requests of k ~ U{1..64} N(0,1) columns on a
uniformly drawn matrix.  After one warm-up request per matrix (first-touch
builds), a burst (all requests at once) and a closed loop keeping
--c5-window requests outstanding run; the closed loop's sustained rate is
the capacity, and requests then arrive as a Poisson process at 0.8x that
capacity; value = right-hand sides solved per second in the open-loop
phase, latency = submission to completion per request.

Other configurations (--config): C1 pde3d 40^3 with AINV-LLK (fill 8,
thresh 0.1), BiCGSTAB 1e-9, b = 1; C3 2-D anisotropic diffusion 2048^2
(eps = 1e-3) with AINV SYM-LLK (fill 8, thresh 5e-2), CG 1e-8, 32 slots,
N(0,1) right-hand sides (seed 3); C4 random nonsymmetric n = 10^6
(LogNormal(ln 12, 1.0) off-diagonals capped at 400, band 5000, 2
long-range columns, diagonal 1.1 x row sum) with AINV-LLK (fill 8, thresh
5e-2), BiCGSTAB 1e-8, N(0,1) right-hand sides (seed 4).

C1, C2 and C3 also report, as `multilevel`, the multilevel preconditioner
(smoothed-aggregation V-cycle with factors of the configuration's kind as
smoother and coarsest-level solver; ML_CONFIGS below) on the same matrix and
right-hand sides: its FP64 single-RHS solves, its served form with --prec
level factors and one cycle's duration at the served block width
(--no-ml skips it; --ml-cycle times the cycle alone).

Every configuration reports, as `layout`, the sliced layouts of the system
matrix and the factors: stored entries with the rows sorted by length
within windows (the default where the rows in order pad by more than a
quarter; --slice-sort sets the window, 0 keeps rows in order) and in row
order, the window, and the slice widths whose spread is the load balance of
the slices' waves.  C4 also reports, as `bf16`, the served run with
BF16-value / FP32-accumulate factors.
"""
import argparse
import ctypes as C
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
# HBM traffic of the dominant kernel from rocprofv3 --pmc FETCH_SIZE and
# WRITE_SIZE runs, scaled by the same counters on a copy kernel of known size
TRAFFIC_FILES = {
    "C2": os.path.join(ROOT, "profiles", "traffic_apply_pde3d160.json"),
    "C3": os.path.join(ROOT, "profiles", "traffic_apply_aniso2048.json"),
    "C4": os.path.join(ROOT, "profiles", "traffic_apply_random1m.json")}

# name: (description, solver kind, parameters, method, rtol, slots (0: the
#        served form times 8, 16, 32 and 64 slots on its first call and keeps
#        the fastest), queue,
#        right-hand sides ("Ax": x_0 = 1, x_j ~ N(0,1); "ones+N": b_0 = 1,
#        b_j ~ N(0,1); "N": all N(0,1)), seed)
CONFIGS = {
    "C1": ("pde3d 40^3 convection-diffusion, AINV-LLK fill 8 thresh 0.1, "
           "BiCGSTAB rtol 1e-9", "AINV",
           dict(AINV_ALG="LLK", SUB_FILLIN=8, SUB_ILUTHRS=0.1), "BICGSTAB",
           1e-9, 0, 192, "ones+N", 1),
    "C2": ("pde3d 160^3 convection-diffusion, INVK(1,1), BiCGSTAB rtol 1e-8",
           "INVK", dict(SUB_FILLIN=1, INV_FILLIN=1), "BICGSTAB", 1e-8, 0,
           192, "Ax", 5),
    "C3": ("aniso2d 2048^2 eps 1e-3, AINV SYM-LLK fill 8 thresh 5e-2, "
           "CG rtol 1e-8", "AINV",
           dict(AINV_ALG="SYM-LLK", SUB_FILLIN=8, SUB_ILUTHRS=5e-2), "CG",
           1e-8, 0, 96, "N", 3),
    "C4": ("random nonsymmetric n=10^6 (LogNormal sigma 1.0), AINV-LLK fill 8 "
           "thresh 5e-2, BiCGSTAB rtol 1e-8", "AINV",
           dict(AINV_ALG="LLK", SUB_FILLIN=8, SUB_ILUTHRS=5e-2), "BICGSTAB",
           1e-8, 0, 96, "N", 4),
}


# Multilevel leg (C1, C2): the shape of the reference's deck
# tests/pdegen/runs/ppde.inp (ML, 3 levels, SMOOTHED DEC aggregation,
# MULT / POST, one smoother sweep) with factors of the configuration's kind
# as smoother and coarsest-level solver.  Threshold 0 instead of the deck's
# 0.10: the smoothed Galerkin operator's relative couplings fall below 0.10,
# which stops coarsening after one step, and a threshold halved per level
# leaves a coarsest level five times larger for the same iteration count.
# The coarsest level gets 2 sweeps (more do not reduce the iteration count).
ML_PARAMS = dict(N_PREC_LEVS=3, AGGR_KIND="SMOOTHED", AGGR_THRESH=0.0,
                 AGGR_SCALE=1.0, ML_TYPE="MULT", SMOOTHER_POS="POST",
                 SMOOTHER_SWEEPS=1, COARSE_SWEEPS=2)
# factors of every level: INVK with ILU(0) and inverse fill 1 (the Galerkin
# levels are denser than A, and ILU(1) on them fills heavily)
ML_FACTORS = dict(SUB_FILLIN=0, INV_FILLIN=1)
# per configuration: (cycle parameters, factor parameters or None for the
# configuration's own, Krylov method or None for the configuration's).
# C3, the anisotropic problem: threshold 0.10 as the deck (threshold 0
# aggregates across the weak direction and the cycle stops converging), 6
# levels (aggregates of about 3 rows: the coarsest is then ~17,000 rows),
# and BiCGSTAB: the POST cycle is not symmetric and CG fails with it, while
# the symmetric TWOSIDE cycle makes a poor preconditioner with these
# undamped smoothers (it fails under CG too)
ML_CONFIGS = {
    "C1": (ML_PARAMS, None, None),
    "C2": (ML_PARAMS, ML_FACTORS, None),
    "C3": (dict(ML_PARAMS, N_PREC_LEVS=6, AGGR_THRESH=0.1), None,
           "BICGSTAB"),
}


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def make_matrix(m, name, grid):
    if name == "C1":
        return m.gen_pde3d(grid or 40)
    if name == "C2":
        return m.gen_pde3d(grid or 160)
    if name == "C3":
        return m.gen_aniso2d(grid or 2048, 1e-3)
    return m.gen_random(grid or 1000000, seed=40, median=12.0, sigma=1.0,
                        cap=400, band=5000, nlong=2, dom=1.1)


def apply_pass_bytes(lay, k, first, xb=8, yb=8):
    """Algorithmic bytes of one apply pass for k columns (padding
    excluded): factor words (value + int32 column), one row-extent word per
    row, and for the first pass d, the x read (xb bytes per value: 8, or 4
    for the FP32 copy BiCGSTAB and CG pass with FP32 / BF16 factors) and the t
    write; for the second pass the t read and the y write (yb bytes per
    value: 8, or 4 for the FP32 preconditioned vectors of BiCGSTAB and CG;
    beta = 0).
    t is FP32 for FP32 / BF16 factors, FP64 otherwise."""
    n, vb = lay["n"], lay["value_bytes"]
    st = 8 if vb == 8 else 4
    if first:
        return (lay["nnz_w"] * (vb + 4) + n * 4 + n * st + k * n * xb
                + k * n * st)
    return lay["nnz_z"] * (vb + 4) + n * 4 + k * n * st + k * n * yb


def ml_cycle_bytes(levels, k, sweeps=1, csweeps=2, eb=8):
    """Algorithmic bytes of one POST V-cycle over k columns: per product the
    matrix (value + int32 column per entry, one row-extent word per row) and
    each vector row read or written once (eb bytes per vector value: 8, or 4
    for the FP32 cycle, whose level matrices are FP32; the smoother's
    intermediate t in the factors' intermediate type)."""
    tot = 0
    mb = 4 if eb == 4 else 8  # A, P, R value bytes

    def prod(nnz, rows, reads, writes):
        return nnz * (mb + 4) + rows * 4 + (reads + writes) * rows * k * eb

    def smoother(L, beta):
        vb = L["value_bytes"]
        st = 8 if vb == 8 else 4
        n = L["n"]
        w = L["nnz_w"] * (vb + 4) + n * 4 + n * st + n * k * eb + n * k * st
        z = (L["nnz_z"] * (vb + 4) + n * 4 + n * k * st
             + (2 if beta else 1) * n * k * eb)
        return w + z

    nl = len(levels)
    for i, L in enumerate(levels):
        n = L["n"]
        resid = prod(L["nnz"], n, 2, 1)
        if i == nl - 1:
            tot += smoother(L, False) + (csweeps - 1) * (resid +
                                                         smoother(L, True))
            break
        nc = levels[i + 1]["n"]
        tot += L["nnz_r"] * (mb + 4) + nc * 4 + (n + nc) * k * eb  # R b
        tot += L["nnz_p"] * (mb + 4) + n * 4 + (nc + n) * k * eb  # P u_c
        tot += sweeps * (resid + smoother(L, True))
    return tot


ML_LEVELS = 0  # --ml-levels


def ml_setup(config, kind, params, method):
    """(cycle parameters, factor parameters, Krylov method) of the
    multilevel leg of a configuration."""
    mp, fp, meth = ML_CONFIGS[config]
    mp = dict(mp)
    if ML_LEVELS:
        mp["N_PREC_LEVS"] = ML_LEVELS
    return mp, dict(fp if fp is not None else params), meth or method


def new_ml(m, config, kind, params, method):
    mp, fp, _ = ml_setup(config, kind, params, method)
    pv = m.MldMlPrec()
    for key, v in fp.items():
        pv.set(key, v)
    pv.set("SUB_SOLVE", kind)
    pv.set("COARSE_SUBSOLVE", kind)
    for key, v in mp.items():
        pv.set(key, float(v) if key in ("AGGR_THRESH", "AGGR_SCALE") else v)
    return pv


def ml_cycle_only(m, args, a, A, kind, params, method, slots):
    """--ml-cycle: the cycle's kernels without graph replays (for kernel
    traces): build, a two-iteration eager solve at the served width that
    tunes the mappings on the vectors the served form uses, 10 cycles."""
    n = a.shape[0]
    mp, _, method = ml_setup(args.config, kind, params, method)
    pv = new_ml(m, args.config, kind, params, method)
    t0 = time.time()
    pv.build(a)
    t_build = time.time() - t0
    pv.cnv(args.prec)
    B = m.gen_normal(13, (n, slots))
    m.krylov(method, A, pv, B, eps=1e-8, itmax=2, eager=True)
    x32 = method in ("BICGSTAB", "CG") and args.prec != "F64"
    ms, _ = profile_apply(m, pv, n, slots, 10, x32, x32)
    lv = pv.report()["multilevel"]["levels"]
    byt = ml_cycle_bytes(lv, slots, 1, mp["COARSE_SWEEPS"],
                         4 if x32 else 8)
    print(json.dumps({"ml_cycle": True, "config": args.config, "k": slots,
                      "maps": [e for e in pv.report()["multilevel"]["maps"]
                               if e["K"] == slots],
                      "factor_prec": args.prec,
                      "vectors": "FP32" if x32 else "FP64",
                      "build_s": t_build, "ms": ms, "bytes": byt,
                      "GBs": byt / (ms * 1e-3) / 1e9,
                      "levels": [(L["n"], L["nnz"], L["nnz_w"], L["nnz_p"])
                                 for L in lv]}))


def served_warm(m, method, A, pv, Bh, slots, M, eps, times=1):
    """Warm the served form on the first right-hand sides; with slots 0 its
    first call times 8, 16, 32 and 64 slots.  Returns the slot count it runs
    and the ms per column-iteration of each width tried (None when given)."""
    warm = m.DeviceArray.from_numpy(Bh[:min(slots or 64, M)])
    for _ in range(times):
        m.krylov_stream(method, A, pv, warm, slots=slots, eps=eps,
                        itmax=5000)
    m.synchronize()
    warm.free()
    if slots:
        return slots, None
    e = [t for t in pv.report().get("served_slots", [])
         if t["method"] == method][0]
    log(f"served block width {e['slots']} (ms per column-iteration "
        f"{e['ms_per_col_iter']})")
    return e["slots"], e["ms_per_col_iter"]


def ml_leg(m, args, a, A, dB, Bh, sample, kind, params, method, eps, slots,
           M, fp64_rate, it64):
    """The multilevel preconditioner on the same problem and right-hand
    sides: single-RHS FP64 solves, the served form with --prec level
    factors, and one cycle's duration at the served block width."""
    n = a.shape[0]
    mp, fp, method = ml_setup(args.config, kind, params, method)
    pv = new_ml(m, args.config, kind, params, method)
    t0 = time.time()
    pv.build(a)
    t_build = time.time() - t0
    levels = pv.report()["multilevel"]
    log(f"multilevel build {t_build:.2f}s: "
        f"{[(L['n'], L['nnz']) for L in levels['levels']]}")
    dx1 = m.DeviceArray((n,))

    def single(j):
        db1 = m.DeviceArray.from_numpy(Bh[j])
        dx1.zero()
        m.synchronize()
        ts = time.time()
        _, it, err = m.krylov(method, A, pv, db1, dx1, eps=eps, itmax=5000)
        m.synchronize()
        dt = time.time() - ts
        db1.free()
        return dt, int(it[0]), float(err[0])

    single(sample[0])
    ts, its = [], []
    for j in sample:
        dt, it, _ = single(j)
        ts.append(dt)
        its.append(it)
    x = dx1.to_numpy()
    true1 = float(np.linalg.norm(Bh[sample[-1]] - a @ x) /
                  np.linalg.norm(Bh[sample[-1]]))
    rate1 = len(ts) / sum(ts)
    log(f"multilevel FP64 single: {rate1:.2f} solves/s, iterations {its}")
    # served form with --prec level factors (tuned, captured graphs)
    pv.cnv(args.prec)
    dX = m.DeviceArray((M, n))
    slots, width_ms = served_warm(m, method, A, pv, Bh, slots, M, eps)
    t_start = time.time()
    _, iters, errs, lat = m.krylov_stream(method, A, pv, dB, dX, slots=slots,
                                          eps=eps, itmax=5000)
    m.synchronize()
    el = time.time() - t_start
    Xs = dX.to_numpy()
    dX.free()
    true = max(float(np.linalg.norm(Bh[j] - a @ Xs[j]) /
                     np.linalg.norm(Bh[j])) for j in range(0, M,
                                                           max(1, M // 16)))
    rate = M / el
    # one cycle at the served block width, on the vectors the served form
    # gives it (FP32 for BiCGSTAB / CG with FP32 / BF16 factors) and with
    # the mappings it tuned, timed with HIP events on the handle's stream
    x32 = method in ("BICGSTAB", "CG") and args.prec != "F64"
    ms, _ = profile_apply(m, pv, n, slots, 10, x32, x32)
    lv = pv.report()["multilevel"]["levels"]
    byt = ml_cycle_bytes(lv, slots, 1, mp["COARSE_SWEEPS"],
                         4 if x32 else 8)
    cyc = {"factor_prec": args.prec, "vectors": "FP32" if x32 else "FP64",
           "ms": ms, "bytes": byt, "GBs": byt / (ms * 1e-3) / 1e9,
           "roofline_frac": byt / (ms * 1e-3) / 1e9 / HBM_PEAK_GBS}
    log(f"multilevel served {args.prec}: {rate:.2f} solves/s, iterations "
        f"mean {float(np.mean(iters)):.1f}; cycle {cyc}")
    cpu = None
    if not args.no_cpu:
        cpu = ml_cpu_baseline(pv, a, Bh[sample[0]], method, mp,
                              float(np.mean(its)), eps=eps)
        log(f"multilevel cpu baseline {cpu}")
    return {
        "params": mp, "method": method,
        "factors": {"kind": kind, **fp},
        "build_s": t_build,
        "levels": [{k: L[k] for k in ("n", "nnz", "aggregates", "nnz_w",
                                      "nnz_z", "theta", "omega")}
                   for L in levels["levels"]],
        "build_phases_s": {k: levels[k] for k in
                           ("t_aggregation", "t_galerkin", "t_factors",
                            "t_upload")},
        "fp64_single_rhs": {"solves_per_s": rate1, "iterations": its,
                            "true_residual": true1,
                            "vs_one_level_fp64_single":
                                rate1 / fp64_rate if fp64_rate else None,
                            "one_level_iterations": it64},
        "served": {"solves_per_s": rate, "factor_prec": args.prec,
                   "slots": slots, "slots_ms_per_col_iter": width_ms,
                   "iterations_mean": float(np.mean(iters)),
                   "iterations_max": int(np.max(iters)),
                   "latency_ms_p50": float(np.percentile(lat, 50)),
                   "latency_ms_p99": float(np.percentile(lat, 99)),
                   "true_residual_max": true,
                   "vs_one_level_fp64_single":
                       rate / fp64_rate if fp64_rate else None},
        "cycle_k": slots, "cycle": cyc,
        "cpu_baseline": cpu,
    }


def main():
    p = argparse.ArgumentParser()
    p.add_argument("--gpus", type=int, default=1)
    p.add_argument("--steps", type=int, default=2)
    p.add_argument("--warmup", type=int, default=1)
    p.add_argument("--config", default="C2", choices=sorted(CONFIGS) + ["C5"])
    p.add_argument("--c5-nmax", type=float, default=2e6,
                   help="largest random matrix of the C5 pool")
    p.add_argument("--c5-requests", type=int, default=96)
    p.add_argument("--c5-slots", type=int, default=0,
                   help="SLOTS of every served matrix (0: server default)")
    p.add_argument("--c5-window", type=int, default=16,
                   help="outstanding requests of the closed-loop capacity run")
    p.add_argument("--workers", type=int, default=4)
    p.add_argument("--grid", type=int, default=0,
                   help="override the grid size (C1-C3) or order (C4)")
    p.add_argument("--queue", type=int, default=0)
    p.add_argument("--slots", type=int, default=0)
    p.add_argument("--single", type=int, default=6,
                   help="right-hand sides 1..single solved one at a time "
                        "in FP64 (the baseline sample)")
    p.add_argument("--prec", default="F32")
    p.add_argument("--no-cpu", action="store_true")
    p.add_argument("--no-ml", action="store_true",
                   help="skip the multilevel leg (C1, C2, C3)")
    p.add_argument("--no-invt", action="store_true",
                   help="skip C2's INVT leg")
    p.add_argument("--no-c5", action="store_true",
                   help="skip the serving-stream leg of the default line")
    p.add_argument("--ml-levels", type=int, default=0,
                   help="levels of the multilevel leg (0: the configuration's)")
    p.add_argument("--ml-only", action="store_true",
                   help="the multilevel leg alone (no one-level solves)")
    p.add_argument("--ml-cycle", action="store_true",
                   help="time the multilevel cycle alone (C1, C2): build, "
                        "tune the mappings with a short eager solve, then "
                        "10 cycles at the served block width; no graphs")
    p.add_argument("--cpu-iters", type=int, default=None,
                   help="oracle iterations per timed run (default: ~2 s)")
    p.add_argument("--apply-only", action="store_true",
                   help="profiling mode: build, then only the two apply "
                        "kernels at the served block width (and the copy "
                        "kernel that calibrates the traffic counters)")
    p.add_argument("--eager-iteration", action="store_true",
                   help="profiling mode: build, then two 12-iteration "
                        "solves of `slots` right-hand sides with the "
                        "iteration's kernels launched one by one (no graph "
                        "replays), for kernel traces")
    p.add_argument("--map", default="AUTO",
                   help="APPLY_MAP for --apply-only (AUTO, SPLIT, ROW, COOP, "
                        "COOP128, COOP_Q, COOP128_Q, RUN)")
    p.add_argument("--order", default="",
                   help="--apply-only: slice-group order of the run mapping "
                        "as PERIOD,BAND (groups of 4 slices; the served "
                        "run's tuned choice is in roofline.slice_order); "
                        "default linear")
    p.add_argument("--slice-sort", type=int, default=-1,
                   help="row sort of the system matrix and factor layouts: "
                        "-1 the window picked from the row lengths "
                        "(default), 0 rows in order, else a window of that "
                        "many rows (a multiple of 64)")
    args = p.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU only (asked for {args.gpus}): the "
              "sharded path (a row block per GPU, halo exchange and "
              "all-reduced dots in the Krylov loop, no communication in the "
              "apply) is not built", file=sys.stderr)
        sys.exit(2)

    global ML_LEVELS
    ML_LEVELS = args.ml_levels
    import importlib
    m = importlib.import_module("mld-ainv_amd")
    m.init(0)
    lib = m.lib()
    if args.config == "C5":
        serve_bench(m, args)
        return
    desc, kind, params, method, eps, slots, M, rhs_kind, seed = \
        CONFIGS[args.config]
    slots = args.slots or slots
    # the modes that time one pass or cycle take a fixed block width
    kfix = slots or 32
    slots_cfg = slots
    M = args.queue or M
    copy_gbs, _ = m.stream_copy_gbs(1 << 30, 10)
    t0 = time.time()
    a = make_matrix(m, args.config, args.grid)
    n = a.shape[0]
    log(f"{args.config}: n={n} nnz={a.nnz} ({time.time() - t0:.1f}s); "
        f"copy kernel {copy_gbs:.0f} GB/s")

    def new_solver(prec):
        sv = m.SOLVERS[kind]()
        for key, v in params.items():
            sv.set(key, v)
        sv.set("FACTOR_PREC", prec)
        sv.set("SLICE_SORT", args.slice_sort)
        return sv

    if args.apply_only:
        sv = new_solver(args.prec)
        sv.set("APPLY_MAP", args.map)
        if args.order:
            op, ob = (int(v) for v in args.order.split(","))
            sv.set("ORDER_PERIOD", op)
            sv.set("ORDER_BAND", ob)
        else:
            sv.set("ORDER_PERIOD", -1)
        sv.build(a)
        lay = sv.layout()
        x32 = method in ("BICGSTAB", "CG") and args.prec != "F64"
        y32 = method in ("BICGSTAB", "CG") and args.prec != "F64"
        ms_w, ms_z = profile_apply(m, sv, n, kfix, 10, x32, y32)
        bw = apply_pass_bytes(lay, kfix, True, 4 if x32 else 8)
        print(json.dumps({"apply_only": True, "config": args.config,
                          "prec": args.prec, "map": args.map, "k": kfix,
                          "order": sv.report()["slice_order"]["w"],
                          "x": "FP32" if x32 else "FP64",
                          "ms_w": ms_w, "ms_z": ms_z, "bytes_w": bw,
                          "GBs_w": bw / (ms_w * 1e-3) / 1e9}))
        return
    A = m.SystemMatrix(a)
    if args.slice_sort != -1:
        A.set_slice_sort(args.slice_sort)
    if args.eager_iteration:
        sv = new_solver(args.prec)
        sv.build(a)
        B = m.DeviceArray.from_numpy(np.ascontiguousarray(
            m.gen_normal(3, (n, kfix))))
        for _ in range(2):
            t0 = time.time()
            m.krylov(method, A, sv, B, eps=1e-30, itmax=12, eager=True)
            m.synchronize()
            log(f"eager: 12 iterations at k={kfix} in "
                f"{time.time() - t0:.3f}s")
        print(json.dumps({"eager_iteration": True, "config": args.config,
                          "k": kfix, "prec": args.prec}))
        return
    if args.ml_cycle:
        ml_cycle_only(m, args, a, A, kind, params, method, kfix)
        return

    # ---- right-hand sides, resident on the device: M n-vectors
    t0 = time.time()
    if rhs_kind == "Ax":
        X = m.gen_normal(seed, (M, n))
        X[0] = 1.0
        dXt = m.DeviceArray.from_numpy(X)
        del X
        dB = m.DeviceArray((M, n))
        for j in range(M):
            off = j * n * 8
            m._ck(lib.ainv_mat_spmm(A._h, 1.0,
                                    C.c_void_p(dXt.ptr.value + off), 0.0,
                                    C.c_void_p(dB.ptr.value + off), 1, 1, 1,
                                    None))
        m.synchronize()
        dXt.free()
    else:
        Bn = m.gen_normal(seed, (M, n))
        if rhs_kind == "ones+N":
            Bn[0] = 1.0
        dB = m.DeviceArray.from_numpy(Bn)
        del Bn
    log(f"{M} right-hand sides on the device ({time.time() - t0:.1f}s)")

    if args.ml_only:
        Bh = dB.to_numpy()
        sample = list(range(1, min(args.single, M - 1) + 1))
        print(json.dumps({"config": args.config, "multilevel": ml_leg(
            m, args, a, A, dB, Bh, sample, kind, params, method, eps, slots,
            M, None, None)}))
        return

    sv = new_solver("F64")
    t0 = time.time()
    sv.build(a)
    t_build = time.time() - t0
    log(f"{kind} build {t_build:.2f}s")

    # ---- plain FP64 single-RHS GPU solves of a sample of the queue
    dx1 = m.DeviceArray((n,))
    Bh = dB.to_numpy()
    # the sample skips b_0 (A*1 or 1, the reference driver's right-hand
    # side; it appears once in the queue)
    sample = list(range(1, min(args.single, M - 1) + 1))

    def single(j):
        db1 = m.DeviceArray.from_numpy(Bh[j])
        dx1.zero()
        m.synchronize()
        ts = time.time()
        _, it, err = m.krylov(method, A, sv, db1, dx1, eps=eps, itmax=5000)
        m.synchronize()
        dt = time.time() - ts
        db1.free()
        return dt, int(it[0]), float(err[0])

    single(sample[0])  # warm-up (plan, graph capture)
    t64, it64 = [], []
    for j in sample:
        dt, it, err = single(j)
        t64.append(dt)
        it64.append(it)
    x64 = dx1.to_numpy()
    true64 = float(np.linalg.norm(Bh[sample[-1]] - a @ x64) /
                   np.linalg.norm(Bh[sample[-1]]))
    lay64 = sv.layout()
    ms_w64, ms_z64 = profile_apply(m, sv, n, 1, 20)
    gbs64 = ((apply_pass_bytes(lay64, 1, True) +
              apply_pass_bytes(lay64, 1, False)) /
             ((ms_w64 + ms_z64) * 1e-3) / 1e9)
    fp64_rate = len(t64) / sum(t64)
    log(f"FP64 single: {fp64_rate:.2f} solves/s, iterations {it64}, "
        f"true residual {true64:.2e}, apply {ms_w64:.3f}+{ms_z64:.3f} ms "
        f"({gbs64:.0f} GB/s)")

    # ---- served form with reduced-precision factors (the timed workload)
    sv.cnv(args.prec)
    lay = sv.layout()
    dX = m.DeviceArray((M, n))
    slots, width_ms = served_warm(m, method, A, sv, Bh, slots, M, eps,
                                  max(1, args.warmup))
    lat, iters, errs = [], None, None
    t_start = time.time()
    for _ in range(args.steps):
        _, iters, errs, l = m.krylov_stream(method, A, sv, dB, dX,
                                            slots=slots, eps=eps, itmax=5000)
        lat.extend(l.tolist())
    m.synchronize()
    elapsed = time.time() - t_start
    rate = args.steps * M / elapsed
    Xs = dX.to_numpy()
    true = max(float(np.linalg.norm(Bh[j] - a @ Xs[j]) / np.linalg.norm(Bh[j]))
               for j in range(0, M, max(1, M // 16)))
    # the served solution of the last sampled right-hand side against its
    # FP64 single-RHS solve (both at the same tolerance)
    j64 = sample[-1]
    xdiff = float(np.linalg.norm(Xs[j64] - x64) / np.linalg.norm(x64))
    rep = sv.report()
    tuned = [t for t in rep.get("tuning", []) if t["K"] == slots]
    kmap_w = tuned[0]["map_w"] if tuned else "split"
    kmap_z = tuned[0]["map_z"] if tuned else "split"
    so = rep.get("slice_order", {})
    order_w = [so.get("w", {}).get("period", 0), so.get("w", {}).get("band", 0)]
    order_z = [so.get("z", {}).get("period", 0), so.get("z", {}).get("band", 0)]
    it_red = [int(iters[j]) for j in sample]
    log(f"served {args.prec} slots={slots}: {rate:.2f} solves/s, "
        f"{elapsed / args.steps:.2f} s/step, iterations mean "
        f"{float(np.mean(iters)):.1f} max {int(np.max(iters))}, tuning {tuned}")

    # the two apply passes as the served form runs them (BiCGSTAB and CG with
    # FP32 / BF16 factors give the first pass an FP32 x); the dominant kernel
    # is the longer of the two
    x32 = method in ("BICGSTAB", "CG") and args.prec != "F64"
    y32 = method in ("BICGSTAB", "CG") and args.prec != "F64"
    ms_w, ms_z = profile_apply(m, sv, n, slots, 10, x32, y32)
    bw = apply_pass_bytes(lay, slots, True, 4 if x32 else 8)
    bz = apply_pass_bytes(lay, slots, False, 8, 4 if y32 else 8)
    vt = {"F64": "double", "F32": "float", "BF16": "unsigned short"}[args.prec]
    tt = "double" if args.prec == "F64" else "float"
    xt = "float" if x32 else "double"

    def kname(kmap, first):
        kernel, extra = {"split": ("k_hll_spmm", ""),
                         "row": ("k_hll_spmm_row", ""),
                         "coop": ("k_hll_spmm_coop", " 256,"),
                         "coop128": ("k_hll_spmm_coop", " 128,"),
                         "coop_q": ("k_hll_spmm_coopq", " 256,"),
                         "coop128_q": ("k_hll_spmm_coopq", " 128,"),
                         "run": ("k_hll_spmm_run", " 128,")}[kmap]
        if first:
            return f"{kernel}<{vt}, {xt}, {tt}, {tt}, {slots},{extra} 0>"
        yt = "float" if y32 else "double"
        return f"{kernel}<{vt}, {tt}, {tt}, {yt}, {slots},{extra} 1>"

    # slice-group order ([period, band] in groups of 4 slices; [0, 0]:
    # linear) of the run mapping's passes, as the tuner left them
    passes = [{"pass": "first: t = d o (W x)", "kernel": kname(kmap_w, True),
               "bytes_per_launch": bw, "ms_per_launch": ms_w,
               "slice_order": order_w if kmap_w == "run" else [0, 0]},
              {"pass": "second: y = Z t", "kernel": kname(kmap_z, False),
               "bytes_per_launch": bz, "ms_per_launch": ms_z,
               "slice_order": order_z if kmap_z == "run" else [0, 0]}]
    traffic = {}
    tfile = TRAFFIC_FILES.get(args.config, "")
    if os.path.exists(tfile):
        with open(tfile) as f:
            tr = json.load(f)
        if tr.get("config") == args.config and tr.get("n") == n:
            traffic = {(e["kernel"], tuple(e.get("slice_order", [0, 0]))):
                       e["bytes_per_launch"] for e in tr.get("kernels", [])}
    for ps in passes:
        ps["achieved_GBs"] = ps["bytes_per_launch"] / (ps["ms_per_launch"]
                                                       * 1e-3) / 1e9
        ps["traffic"] = traffic.get((ps["kernel"], tuple(ps["slice_order"])))
    dom = max(passes, key=lambda q: q["ms_per_launch"])
    other = passes[1] if dom is passes[0] else passes[0]
    log(f"apply at k={slots}: W {ms_w:.3f} ms "
        f"({passes[0]['achieved_GBs']:.0f} GB/s), Z {ms_z:.3f} ms "
        f"({passes[1]['achieved_GBs']:.0f} GB/s)")

    bf16 = None
    if args.config == "C4" and args.prec == "F32":
        bf16 = bf16_leg(m, method, A, sv, a, dB, dX, Bh, M, eps, args.steps,
                        x64, j64)

    ml = None
    if not args.no_ml and args.config in ML_CONFIGS:
        ml = ml_leg(m, args, a, A, dB, Bh, sample, kind, params, method, eps,
                    slots_cfg, M, fp64_rate, it64)
    invt = None
    if args.config == "C2" and not args.no_invt:
        invt = invt_leg(m, args, a, A, dB, Bh, sample, method, eps,
                        slots_cfg, M)

    cpu = None
    if not args.no_cpu:
        cpu = cpu_baseline(sv, a, Bh[sample[0]], method, args.cpu_iters,
                           float(np.mean(it64)), eps=eps)
        log(f"cpu baseline {cpu}")

    # the serving stream over the 24-matrix pool (its own matrices; the
    # queue above is released first)
    c5 = None
    if args.config == "C2" and not args.no_c5:
        dX.free()
        dB.free()
        c5 = c5_compact(c5_run(m, args, burst=False))

    lat_ms = np.array(lat)
    out = {
        "metric": "preconditioned solves/s + apply HBM GB/s (% roofline); "
                  "p50/p99 solve latency",
        "value": rate,
        "unit": "solves/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": elapsed / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {
            "workload": f"{args.config}: {desc}; served form: queue of {M} "
                        f"right-hand sides through {slots} slots, "
                        f"{args.prec}-stored factors, FP64 Krylov",
            "name": args.config, "n": n, "nnz": int(a.nnz), "queue": M,
            "slots": slots, "slots_ms_per_col_iter": width_ms,
            "factor_prec": args.prec, "parallelism": "1 GPU",
        },
        "latency_ms": {"p50": float(np.percentile(lat_ms, 50)),
                       "p99": float(np.percentile(lat_ms, 99)),
                       "note": "admission to completion in the served form, "
                               "queue fully loaded"},
        "iterations": {"served_mean": float(np.mean(iters)),
                       "served_max": int(np.max(iters)),
                       "served_sample": it_red,
                       "fp64_single_sample": it64,
                       "sample": f"right-hand sides {sample[0]}.."
                                 f"{sample[-1]}"},
        "true_residual_max": true,
        "error_vs_fp64": {"rel_diff": xdiff, "rhs": j64,
                          "iterations_fp64": it64[-1],
                          "iterations_served": int(iters[j64]),
                          "note": "||x_served - x_fp64|| / ||x_fp64|| for one "
                                  "right-hand side, both solves at the "
                                  "stopping tolerance"},
        "fp64_single_rhs": {
            "solves_per_s": fp64_rate, "ms_mean": 1e3 * float(np.mean(t64)),
            "iterations": it64, "true_residual": true64,
            "apply_GBs": gbs64, "apply_roofline_frac": gbs64 / HBM_PEAK_GBS,
        },
        "speedup_vs_fp64_single": rate / fp64_rate,
        "roofline": {
            "bound": "hbm", "achieved": dom["achieved_GBs"],
            "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": dom["achieved_GBs"] / HBM_PEAK_GBS,
            "traffic": dom["traffic"],
            "kernel": dom["kernel"], "pass": dom["pass"],
            "slice_order": dom["slice_order"],
            "slice_order_trial": so.get("trial"),
            "bytes_per_launch": dom["bytes_per_launch"],
            "ms_per_launch": dom["ms_per_launch"],
            "achievable_copy_GBs": copy_gbs,
            "achievable_copy_runtime_GBs": m.runtime_copy_gbs(1 << 30, 10),
            "other_pass": other,
        },
        "build": {"seconds": t_build},
        "layout": layout_record(A, rep),
        "bf16": bf16,
        "cpu_baseline": cpu,
        "multilevel": ml,
        "invt": invt,
        "c5": c5,
    }
    print(json.dumps(out))


def bf16_leg(m, method, A, sv, a, dB, dX, Bh, M, eps, steps, x64, j64):
    """The served run again with BF16-value / FP32-accumulate factors (the
    same build's factors rounded to BF16), its own block-width trial, for
    the comparison with FP32-stored factors."""
    n = a.shape[0]
    sv.cnv("BF16")
    slots, width_ms = served_warm(m, method, A, sv, Bh, 0, M, eps, 1)
    lat, iters = [], None
    t0 = time.time()
    for _ in range(steps):
        _, iters, _, l = m.krylov_stream(method, A, sv, dB, dX, slots=slots,
                                         eps=eps, itmax=5000)
        lat.extend(l.tolist())
    m.synchronize()
    el = time.time() - t0
    Xs = dX.to_numpy()
    true = max(float(np.linalg.norm(Bh[j] - a @ Xs[j]) / np.linalg.norm(Bh[j]))
               for j in range(0, M, max(1, M // 16)))
    ms_w, ms_z = profile_apply(m, sv, n, slots, 10, True, True)
    lay = sv.layout()
    out = {"solves_per_s": steps * M / el, "slots": slots,
           "slots_ms_per_col_iter": width_ms,
           "iterations_mean": float(np.mean(iters)),
           "iterations_max": int(np.max(iters)),
           "true_residual_max": true,
           "error_vs_fp64": float(np.linalg.norm(Xs[j64] - x64) /
                                  np.linalg.norm(x64)),
           "latency_ms": {"p50": float(np.percentile(lat, 50)),
                          "p99": float(np.percentile(lat, 99))},
           "apply_ms": {"first": ms_w, "second": ms_z},
           "apply_GBs": {
               "first": apply_pass_bytes(lay, slots, True, 4) /
               (ms_w * 1e-3) / 1e9,
               "second": apply_pass_bytes(lay, slots, False, 8, 4) /
               (ms_z * 1e-3) / 1e9}}
    log(f"served BF16 slots={slots}: {out['solves_per_s']:.1f} solves/s, "
        f"apply {ms_w:.3f}+{ms_z:.3f} ms")
    return out


def layout_record(A, rep):
    """Sliced layouts of the system matrix and the factors: stored entries
    with the row sort and with the rows in order, the sort window, and the
    slice widths (entries per row slot) whose spread is the load balance of
    the slices' waves."""
    out = {"a": A.report()["a"]}
    for f in ("w", "z"):
        if f in rep:
            out[f] = rep[f]
    for v in out.values():
        v["padding_ratio"] = v["padded"] / max(v["nnz"], 1)
        v["unsorted_padding_ratio"] = v["unsorted_padded"] / max(v["nnz"], 1)
        v["max_over_mean_width"] = v["max_width"] / max(v["width_mean"], 1e-9)
    return out


def serve_bench(m, args):
    print(json.dumps(c5_run(m, args, burst=True)))


def c5_run(m, args, burst=True):
    """The serving stream: first-touch builds of the 24-matrix pool, an
    optional burst, the closed-loop capacity and the open loop (Poisson
    arrivals at 0.8x that capacity); returns the record."""
    import threading
    rng = np.random.default_rng(5)
    pool = []
    for i in range(12):
        N = int(rng.integers(20, 121))
        pool.append((f"pde3d N={N}", m.gen_pde3d(N), "INVK",
                     dict(SUB_FILLIN=1, INV_FILLIN=1, FACTOR_PREC="F32",
                          METHOD="BICGSTAB")))
    for i in range(12):
        n = int(np.exp(rng.uniform(np.log(1e4), np.log(args.c5_nmax))))
        pool.append((f"random n={n}",
                     m.gen_random(n, seed=100 + i, median=12.0, sigma=1.0,
                                  cap=400, band=5000, nlong=2, dom=1.1),
                     "AINV", dict(AINV_ALG="LLK", SUB_FILLIN=8,
                                  SUB_ILUTHRS=5e-2, FACTOR_PREC="F32",
                                  METHOD="BICGSTAB")))
    sizes = [a.shape[0] for _, a, _, _ in pool]
    log(f"C5 pool: {len(pool)} matrices, n from {min(sizes)} to {max(sizes)}")
    srv = m.Server(hbm_budget_bytes=200 << 30, workers=args.workers)
    if args.c5_slots:
        for _, _, _, prm in pool:
            prm["SLOTS"] = args.c5_slots
    ids = [srv.add_matrix(a, kind, **prm) for _, a, kind, prm in pool]
    blocks = [m.gen_normal(500 + j, (sizes[j], 64)) for j in range(len(pool))]
    t0 = time.time()
    warm = [srv.submit(ids[j], np.ascontiguousarray(blocks[j][:, :1]),
                       eps=1e-8) for j in range(len(pool))]
    for t in warm:
        srv.wait(t)
    t_warm = time.time() - t0
    st_warm = srv.stats()
    # builds run on the server's builder threads while the workers keep
    # their batches on resident matrices running on the device
    first_touch = {
        "seconds": t_warm, "builds": st_warm["builds"],
        "build_seconds": st_warm["build_seconds"],
        "build_overlap_seconds": st_warm["build_overlap_seconds"],
        "overlap_frac": st_warm["build_overlap_seconds"] /
                        max(st_warm["build_seconds"], 1e-9),
        "builders": st_warm.get("builders"),
        "note": "build_seconds summed over the builder threads; overlap: "
                "the build seconds during which a worker ran a slot batch"}
    log(f"first-touch builds + one solve per matrix: {t_warm:.1f}s "
        f"({first_touch})")

    def requests(count, seed):
        rr = np.random.default_rng(seed)
        out = []
        for _ in range(count):
            j = int(rr.integers(0, len(pool)))
            k = int(rr.integers(1, 65))
            out.append((j, k, np.ascontiguousarray(blocks[j][:, :k])))
        return out

    def check(res):
        worst = max(float(np.max(e)) for _, _, e, _ in res)
        if worst > 1e-8:
            raise SystemExit(f"C5: a request ended at ||r||/||b|| = {worst}")

    # burst: every request submitted at once (the most batching there is)
    burst_req = burst_cols = None
    nburst = 0
    if burst:
        bl = requests(max(24, args.c5_requests // 2), 6)
        nburst = len(bl)
        t0 = time.time()
        tk = [srv.submit(ids[j], b, eps=1e-8, itmax=2000) for j, _, b in bl]
        res = [srv.wait(t) for t in tk]
        t_burst = time.time() - t0
        check(res)
        burst_req = len(bl) / t_burst
        burst_cols = sum(k for _, k, _ in bl) / t_burst
        log(f"burst: {burst_req:.2f} requests/s, {burst_cols:.1f} columns/s")
    st_burst = srv.stats()

    # capacity: closed loop keeping `window` requests outstanding, each new
    # one submitted when one completes (sustained throughput), over the same
    # request list the open loop replays
    closed = requests(args.c5_requests, 7)
    window = args.c5_window
    lock = threading.Lock()
    nxt = [0]
    done_res = []

    def client():
        while True:
            with lock:
                i = nxt[0]
                nxt[0] += 1
            if i >= len(closed):
                return
            j, _, b = closed[i]
            r = srv.wait(srv.submit(ids[j], b, eps=1e-8, itmax=2000))
            with lock:
                done_res.append(r)

    t0 = time.time()
    cl = [threading.Thread(target=client) for _ in range(window)]
    for t in cl:
        t.start()
    for t in cl:
        t.join()
    t_closed = time.time() - t0
    check(done_res)
    cap_req = len(closed) / t_closed
    cap_cols = sum(k for _, k, _ in closed) / t_closed
    log(f"capacity (closed loop, {window} outstanding): {cap_req:.2f} "
        f"requests/s, {cap_cols:.1f} columns/s")
    st_closed = srv.stats()

    # open loop: Poisson arrivals at 0.8x the sustained capacity
    reqs = requests(args.c5_requests, 7)
    lam = 0.8 * cap_req
    arr = np.cumsum(np.random.default_rng(8).exponential(1.0 / lam,
                                                         len(reqs)))
    tickets, results = [], []
    cv = threading.Condition()

    def waiter():
        for i in range(len(reqs)):
            with cv:
                while len(tickets) <= i:
                    cv.wait()
                t = tickets[i]
            results.append(srv.wait(t))

    th = threading.Thread(target=waiter)
    th.start()
    t_start = time.time()
    for (j, _, b), ta in zip(reqs, arr):
        delay = t_start + ta - time.time()
        if delay > 0:
            time.sleep(delay)
        t = srv.submit(ids[j], b, eps=1e-8, itmax=2000)
        with cv:
            tickets.append(t)
            cv.notify()
    th.join()
    elapsed = time.time() - t_start
    check(results)
    cols = sum(k for _, k, _ in reqs)
    lat = np.array([r[3] for r in results])
    st = srv.stats()
    srv.free()

    def phase(a, b):
        return {k: (round(b[k] - a[k], 6) if isinstance(b[k], float)
                    else b[k] - a[k])
                for k in ("batches", "steps", "admitted_midflight",
                          "open_seconds", "admit_seconds", "step_seconds",
                          "extract_seconds", "queue_seconds",
                          "service_seconds", "completed")}
    log(f"burst phase {phase(st_warm, st_burst)}")
    log(f"closed-loop phase {phase(st_burst, st_closed)}")
    log(f"open-loop phase {phase(st_closed, st)}")
    out = {
        "metric": "preconditioned solves/s + apply HBM GB/s (% roofline); "
                  "p50/p99 solve latency",
        "value": cols / elapsed,
        "unit": "solves/s",
        "n_gpus": 1, "steps": 1, "warmup": 1,
        "ms_per_step": elapsed * 1e3,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f64", "data": "synthetic",
        "config": {
            "workload": "C5: serving pool of 24 matrices (12 pde3d N~U{20..120} "
                        "INVK(1,1); 12 random n~logU[1e4, "
                        f"{args.c5_nmax:.0e}] AINV-LLK 5e-2), FP32 factors, "
                        "BiCGSTAB 1e-8, requests of U{1..64} columns, Poisson "
                        "arrivals at 0.8x the closed-loop capacity",
            "name": "C5", "requests": len(reqs), "columns": cols,
            "workers": args.workers, "slots": args.c5_slots or "adaptive",
            "parallelism": "1 GPU"},
        "latency_ms": {"p50": float(np.percentile(lat, 50)),
                       "p99": float(np.percentile(lat, 99)),
                       "note": "submission to completion per request "
                               "(host buffers, queueing included)"},
        "capacity": {"requests_per_s": cap_req, "columns_per_s": cap_cols,
                     "note": f"closed loop, {window} requests outstanding, "
                             f"{len(closed)} requests"},
        "burst": {"requests_per_s": burst_req, "columns_per_s": burst_cols,
                  "requests": nburst},
        "arrival_rate_per_s": lam,
        "first_touch_seconds": t_warm,
        "first_touch": first_touch,
        "server_stats": st,
        "phases": {"burst": phase(st_warm, st_burst),
                   "closed_loop": phase(st_burst, st_closed),
                   "open_loop": phase(st_closed, st)},
        "roofline": None, "cpu_baseline": None,
        "true_residual_max": max(float(np.max(e)) for _, _, e, _ in results),
    }
    return out


def c5_compact(rec):
    """The serving stream's figures for the default benchmark line."""
    return {
        "workload": rec["config"]["workload"],
        "requests": rec["config"]["requests"],
        "columns": rec["config"]["columns"],
        "first_touch_seconds": rec["first_touch_seconds"],
        "first_touch": rec["first_touch"],
        "capacity": rec["capacity"],
        "open_loop": {"solves_per_s": rec["value"],
                      "unit": "right-hand sides per second",
                      "arrival_rate_requests_per_s":
                          rec["arrival_rate_per_s"]},
        "latency_ms": rec["latency_ms"],
        "true_residual_max": rec["true_residual_max"],
        "peak_resident_bytes": rec["server_stats"]["peak_resident_bytes"],
        "evictions": rec["server_stats"]["evictions"],
    }


INVT_PARAMS = dict(SUB_FILLIN=10, SUB_ILUTHRS=1e-2, INV_FILLIN=10,
                   INV_THRESH=1e-2)


def invt_leg(m, args, a, A, dB, Bh, sample, method, eps, slots, M):
    """C2's INVT preconditioner (ILUT fill 10, drop 1e-2 with maxval
    scaling; inverse fill 10, inverse drop 1e-2) on the same matrix and
    queue: FP64 single-RHS solves, the served form with --prec factors and
    its apply passes against the roofline."""
    n = a.shape[0]
    sv = m.SOLVERS["INVT"]()
    for key, v in INVT_PARAMS.items():
        sv.set(key, v)
    sv.set("FACTOR_PREC", "F64")
    t0 = time.time()
    sv.build(a)
    t_build = time.time() - t0
    dx1 = m.DeviceArray((n,))

    def single(j):
        db1 = m.DeviceArray.from_numpy(Bh[j])
        dx1.zero()
        m.synchronize()
        ts = time.time()
        _, it, _ = m.krylov(method, A, sv, db1, dx1, eps=eps, itmax=5000)
        m.synchronize()
        dt = time.time() - ts
        db1.free()
        return dt, int(it[0])

    sub = sample[:3]
    single(sub[0])
    ts, its = [], []
    for j in sub:
        dt, it = single(j)
        ts.append(dt)
        its.append(it)
    x64 = dx1.to_numpy()
    dx1.free()
    rate64 = len(ts) / sum(ts)
    true64 = float(np.linalg.norm(Bh[sub[-1]] - a @ x64) /
                   np.linalg.norm(Bh[sub[-1]]))
    sv.cnv(args.prec)
    lay = sv.layout()
    dX = m.DeviceArray((M, n))
    slots, width_ms = served_warm(m, method, A, sv, Bh, slots, M, eps)
    t_start = time.time()
    _, iters, _, lat = m.krylov_stream(method, A, sv, dB, dX, slots=slots,
                                       eps=eps, itmax=5000)
    m.synchronize()
    el = time.time() - t_start
    Xs = dX.to_numpy()
    dX.free()
    rate = M / el
    true = max(float(np.linalg.norm(Bh[j] - a @ Xs[j]) /
                     np.linalg.norm(Bh[j])) for j in range(0, M,
                                                           max(1, M // 16)))
    xdiff = float(np.linalg.norm(Xs[sub[-1]] - x64) / np.linalg.norm(x64))
    x32 = method in ("BICGSTAB", "CG") and args.prec != "F64"
    ms_w, ms_z = profile_apply(m, sv, n, slots, 10, x32, x32)
    bw = apply_pass_bytes(lay, slots, True, 4 if x32 else 8)
    bz = apply_pass_bytes(lay, slots, False, 8, 4 if x32 else 8)
    rep = sv.report()
    tuned = [t for t in rep.get("tuning", []) if t["K"] == slots]
    dom_w = ms_w >= ms_z
    gbs = (bw / (ms_w * 1e-3) if dom_w else bz / (ms_z * 1e-3)) / 1e9
    log(f"INVT: build {t_build:.2f}s, FP64 single {rate64:.2f} solves/s "
        f"({its}), served {rate:.2f} solves/s, apply {ms_w:.3f}+{ms_z:.3f}"
        f" ms")
    return {
        "params": INVT_PARAMS, "method": method, "build_s": t_build,
        "nnz_w": lay["nnz_w"], "nnz_z": lay["nnz_z"],
        "fp64_single_rhs": {"solves_per_s": rate64, "iterations": its,
                            "true_residual": true64},
        "served": {"solves_per_s": rate, "factor_prec": args.prec,
                   "slots": slots, "slots_ms_per_col_iter": width_ms,
                   "queue": M, "iterations_mean": float(np.mean(iters)),
                   "iterations_max": int(np.max(iters)),
                   "latency_ms_p50": float(np.percentile(lat, 50)),
                   "latency_ms_p99": float(np.percentile(lat, 99)),
                   "true_residual_max": true,
                   "error_vs_fp64": xdiff},
        "speedup_vs_fp64_single": rate / rate64,
        "apply": {"ms_first": ms_w, "ms_second": ms_z,
                  "bytes_first": bw, "bytes_second": bz,
                  "maps": tuned[0] if tuned else None},
        "roofline": {"bound": "hbm", "achieved": gbs, "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": gbs / HBM_PEAK_GBS,
                     "pass": "first" if dom_w else "second"},
    }


def profile_apply(m, sv, n, k, reps, x32=False, y32=False):
    """Average durations (ms) of the two apply kernels on k columns; x32:
    FP32 x (the input the served BiCGSTAB and CG give FP32 / BF16 factors),
    y32: FP32 y (BiCGSTAB's and CG's preconditioned vectors)."""
    xh = m.gen_normal(11, (n, k)) if k > 1 else m.gen_normal(11, (n,))
    x = m.DeviceArray.from_numpy(xh.astype(np.float32) if x32 else xh)
    y = m.DeviceArray((n, k) if k > 1 else (n,))
    a, b = C.c_double(), C.c_double()
    m._ck(m.lib().ainv_apply_profile(sv._h, x.ptr, y.ptr, k, reps,
                                     (1 if x32 else 0) | (2 if y32 else 0),
                                     C.byref(a), C.byref(b)))
    x.free()
    y.free()
    return a.value, b.value


def _host_info():
    model = "unknown"
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return model, len(os.sched_getaffinity(0))


# a whole CPU solve is timed beside the per-iteration sample when it is
# expected to take at most this long
CPU_FULL_SOLVE_MAX_S = 30.0


def _full_solve(one, full, gpu_iters, O):
    """full: (label, callable running one whole single-thread solve and
    returning (iterations, final relative residual)), or None.  Timed when
    the per-iteration sample in `one` predicts at most CPU_FULL_SOLVE_MAX_S."""
    if full is None:
        return
    label, solve = full
    expect = one["seconds_per_iteration"] * gpu_iters
    if expect > CPU_FULL_SOLVE_MAX_S:
        one["full_solve"] = {"skipped": f"predicted {expect:.0f} s > "
                                        f"{CPU_FULL_SOLVE_MAX_S:.0f} s"}
        return
    O.set_threads(1)
    t0 = time.perf_counter()
    it, err = solve()
    dt = time.perf_counter() - t0
    one["full_solve"] = {"seconds": dt, "solves_per_s": 1.0 / dt,
                         "iterations": int(it), "relative_residual": err,
                         "cores": 1, "note": label}


def _timed_legs(run, label, gpu_iters, threads, full=None):
    """One warm-up iteration (which sizes the runs to about 2 s), then 5
    timed runs; median seconds per iteration extrapolated to one solve of
    gpu_iters iterations; single-thread figure with the threaded one beside
    it (and one whole solve, see _full_solve)."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O
    model, nproc = _host_info()
    if threads is None:
        threads = min(16, nproc)

    def timed(nt, mm):
        O.set_threads(nt)
        t0 = time.perf_counter()
        it = run(mm)
        return (time.perf_counter() - t0) / max(it, 1), it

    def leg(nt):
        per0, _ = timed(nt, 1)
        mm = max(1, int(round(2.0 / max(per0, 1e-6))))
        runs = [timed(nt, mm) for _ in range(5)]
        per = float(np.median([r[0] for r in runs]))
        return {"value": 1.0 / (per * gpu_iters), "unit": "solves/s",
                "cores": nt, "kind": "port",
                "sample": f"5 runs of {runs[0][1]} {label}, after one "
                          f"warm-up; median {per:.4f} s per iteration, "
                          f"extrapolated to {gpu_iters:.1f} iterations per "
                          f"solve",
                "seconds_per_iteration": per}

    try:
        one = leg(1)
        one["host"] = {"cpu_model": model, "nproc": nproc,
                       "compiler": O.build_info()}
        _full_solve(one, full, gpu_iters, O)
        if threads > 1:
            mt = leg(threads)
            mt["note"] = ("the same loops row-parallel with OpenMP")
            one["threaded"] = mt
    finally:
        O.set_threads(1)
    return one


def ml_cpu_baseline(pv, a, b, method, mp, gpu_iters, threads=None,
                    eps=None):
    """The oracle's cycle inside its Krylov method on the same hierarchy
    (every level's A, P, R and FP64 factors exported from the library, which
    builds them bit-identical to the oracle's own build), timed as
    cpu_baseline."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O
    nl = pv.levels()
    H = O.MlHierarchy.from_levels(
        [pv.level(l, "A") for l in range(nl)],
        [pv.level(l, "P") for l in range(nl - 1)],
        [pv.level(l, "R") for l in range(nl - 1)],
        [(pv.level(l, "W"), pv.level(l, "Z"), pv.level(l, "D"))
         for l in range(nl)],
        ml_type=mp["ML_TYPE"], pos=mp["SMOOTHER_POS"],
        sweeps=mp["SMOOTHER_SWEEPS"], csweeps=mp["COARSE_SWEEPS"])
    A = O.wrap(a)

    def run(mm):
        _, it, _ = O.ml_krylov(method, A, H, b, 1e-30, itmax=mm)
        return it

    def whole():
        _, it, err = O.ml_krylov(method, A, H, b, eps, itmax=100000)
        return it, err

    full = None if eps is None else (
        f"one whole {method} solve with the oracle's cycle, same right-hand "
        f"side, to {eps:g} on one thread", whole)
    return _timed_legs(run, f"{method} iterations of the oracle's cycle on "
                       "the same hierarchy (FP64 factors) and one right-hand "
                       "side of the queue", gpu_iters, threads, full)


def cpu_baseline(sv, a, b, method, iters, gpu_iters, threads=None, eps=None):
    """The oracle's Krylov method on the same system with the same FP64
    factors: one warm-up, then 5 timed runs of `iters` iterations each
    (`iters` = None: as many as take about 2 s), median time per iteration,
    extrapolated to one solve at the FP64 GPU path's mean iteration count.
    Returns the single-thread figure (the reference's one rank, no threads)
    and, beside it, the same loops row-parallel on `threads` host cores.
    With eps, one whole single-thread solve to that tolerance is timed too
    (`full_solve`) when the sample predicts at most CPU_FULL_SOLVE_MAX_S."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O
    sv.cnv("F64")
    w, z, d = sv.export("W"), sv.export("Z"), sv.export("D")
    A, W, Z = O.wrap(a), O.wrap(w), O.wrap(z)
    model, nproc = _host_info()
    if threads is None:
        threads = min(16, nproc)

    def timed(nt, m):
        O.set_threads(nt)
        t0 = time.perf_counter()
        _, it, _ = O.krylov(method, A, W, Z, d, b, 1e-30, itmax=m)
        return (time.perf_counter() - t0) / max(it, 1), it

    def leg(nt):
        per0, _ = timed(nt, 1)  # warm-up, also sizes the runs
        m = iters or max(1, int(round(2.0 / max(per0, 1e-6))))
        runs = [timed(nt, m) for _ in range(5)]
        per = float(np.median([r[0] for r in runs]))
        return {"value": 1.0 / (per * gpu_iters), "unit": "solves/s",
                "cores": nt, "kind": "port",
                "sample": f"5 runs of {runs[0][1]} {method} iterations of the "
                          f"oracle on the same matrix, one right-hand side of "
                          f"the queue and the FP64 factors, after one warm-up; "
                          f"median {per:.4f} s per iteration, extrapolated "
                          f"to {gpu_iters:.1f} iterations per solve",
                "seconds_per_iteration": per}

    def whole():
        _, it, err = O.krylov(method, A, W, Z, d, b, eps, itmax=100000)
        return it, err

    full = None if eps is None else (
        f"one whole {method} solve of the same right-hand side to {eps:g} "
        f"on one thread", whole)
    try:
        one = leg(1)
        one["host"] = {"cpu_model": model, "nproc": nproc,
                       "compiler": O.build_info()}
        _full_solve(one, full, gpu_iters, O)
        if threads > 1:
            mt = leg(threads)
            mt["note"] = ("beyond the reference (one rank, no threads): the "
                          "same loops row-parallel with OpenMP")
            one["threaded"] = mt
    finally:
        O.set_threads(1)
    return one

if __name__ == "__main__":
    main()
