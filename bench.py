"""Benchmark: preconditioned solves/s on one MI355X.

Workload (the configuration the metric is quoted on): the 3-D
convection-diffusion operator on a 160^3 grid (n = 4,096,000, 28.5 M
nonzeros), INVK preconditioner (ILU(1) factors, inverse fill level 1),
BiCGSTAB to 1e-8 relative residual in FP64, right-hand side A*1 (first
column; further columns A*x_c with x_c ~ N(0,1) from a fixed seed), x0 = 0.

A step is one batched preconditioned solve of --k right-hand sides with
FP32-stored factors (FP64 Krylov recurrence, FP64 dot products and
residuals).  Inputs are resident in HBM before the timed region.  Also
reported: the plain FP64 single-RHS GPU solve, the roofline of the
dominant apply kernel (HIP events on the stream it runs on), and the CPU
baseline (the oracle's BiCGSTAB iterations, single thread, on the same
matrix and factors).
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def apply_bytes(lay, k, st):
    """Algorithmic bytes of the two apply kernels for k columns:
    factor words (value + int32 column), one row length per row, d, the x
    read, the t write / read and the y write (beta = 0)."""
    n, vb = lay["n"], lay["value_bytes"]
    sd = 8 if vb == 8 else 4
    first = lay["nnz_w"] * (vb + 4) + n * 4 + n * sd + k * n * 8 + k * n * st
    second = lay["nnz_z"] * (vb + 4) + n * 4 + k * n * st + k * n * 8
    return first, second


def main():
    p = argparse.ArgumentParser()
    p.add_argument("--gpus", type=int, default=1)
    p.add_argument("--steps", type=int, default=3)
    p.add_argument("--warmup", type=int, default=1)
    p.add_argument("--grid", type=int, default=160)
    p.add_argument("--k", type=int, default=32)
    p.add_argument("--eps", type=float, default=1e-8)
    p.add_argument("--no-cpu", action="store_true")
    p.add_argument("--cpu-iters", type=int, default=4)
    args = p.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for {args.gpus}); the "
              "path shards by row blocks with no communication in the "
              "apply, see DESIGN.md", file=sys.stderr)
        sys.exit(2)

    import importlib
    m = importlib.import_module("mld-ainv_amd")
    m.init(0)
    N, k = args.grid, args.k
    t0 = time.time()
    a = m.gen_pde3d(N)
    n = a.shape[0]
    log(f"matrix {N}^3: n={n} nnz={a.nnz} ({time.time() - t0:.1f}s)")
    X = m.gen_normal(5, (n, k))
    X[:, 0] = 1.0
    B = np.ascontiguousarray(np.column_stack([a @ X[:, c] for c in range(k)]))
    b1 = np.ascontiguousarray(B[:, 0])
    A = m.SystemMatrix(a)
    sv = m.MldDInvkSolver()
    sv.set("SUB_FILLIN", 1)
    sv.set("INV_FILLIN", 1)
    sv.set("FACTOR_PREC", "F64")
    t0 = time.time()
    sv.build(a)
    t_build = time.time() - t0
    rep = sv.report()
    log(f"INVK build {t_build:.2f}s  report {json.dumps(rep)}")

    dB = m.DeviceArray.from_numpy(B)
    dX = m.DeviceArray((n, k))
    db1 = m.DeviceArray.from_numpy(b1)
    dx1 = m.DeviceArray((n,))

    def solve(bvec, xvec, kk):
        xvec.zero()
        x, it, err = m.krylov("BICGSTAB", A, sv, bvec, xvec, eps=args.eps,
                              itmax=2000)
        return it, err

    # ---- plain FP64 single-RHS GPU solve (the comparison point)
    solve(db1, dx1, 1)
    t0 = time.time()
    it64, err64 = solve(db1, dx1, 1)
    t_fp64 = time.time() - t0
    x1 = dx1.to_numpy()
    true64 = float(np.linalg.norm(b1 - a @ x1) / np.linalg.norm(b1))
    lay64 = sv.layout()
    ms_w64, ms_z64 = profile_apply(m, sv, n, 1, 20)
    log(f"FP64 1-RHS: {it64[0]} its, {t_fp64 * 1e3:.1f} ms, true res "
        f"{true64:.2e}; apply kernels {ms_w64:.3f}+{ms_z64:.3f} ms")

    # ---- FP32 factors, batched k RHS (the timed workload)
    sv.cnv("F32")
    lay = sv.layout()
    for _ in range(args.warmup):
        solve(dB, dX, k)
    m.synchronize()
    lat = []
    iters = None
    t_start = time.time()
    for _ in range(args.steps):
        ts = time.time()
        iters, errs = solve(dB, dX, k)
        m.synchronize()
        lat.append(time.time() - ts)
    m.synchronize()
    elapsed = time.time() - t_start
    Xs = dX.to_numpy()
    true = [float(np.linalg.norm(B[:, c] - a @ Xs[:, c]) /
                  np.linalg.norm(B[:, c])) for c in range(k)]
    solves_per_s = args.steps * k / elapsed
    # single-RHS FP32 for the iteration delta against FP64
    solve(db1, dx1, 1)
    it32, _ = solve(db1, dx1, 1)

    ms_w, ms_z = profile_apply(m, sv, n, k, 10)
    bw, bz = apply_bytes(lay, k, 4)
    gbs_w = bw / (ms_w * 1e-3) / 1e9
    gbs_apply = (bw + bz) / ((ms_w + ms_z) * 1e-3) / 1e9
    bw64, bz64 = apply_bytes(lay64, 1, 8)
    gbs64 = (bw64 + bz64) / ((ms_w64 + ms_z64) * 1e-3) / 1e9
    log(f"FP32 k={k}: {elapsed / args.steps * 1e3:.1f} ms/step, "
        f"{solves_per_s:.1f} solves/s, iters {list(iters)}, apply W "
        f"{ms_w:.3f} ms ({gbs_w:.0f} GB/s) Z {ms_z:.3f} ms")

    cpu = None
    if not args.no_cpu:
        cpu = cpu_baseline(sv, a, b1, args.cpu_iters, int(it64[0]))
        log(f"cpu baseline {cpu}")

    lat_ms = np.array(lat) * 1e3
    out = {
        "metric": "preconditioned solves/s + apply HBM GB/s (% roofline); "
                  "p50/p99 solve latency",
        "value": solves_per_s,
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
            "workload": f"pde3d-{N}^3 convection-diffusion, INVK(1,1), "
                        f"BiCGSTAB eps={args.eps:g}, FP32-stored factors, "
                        f"FP64 Krylov, batch of {k} RHS per solve step",
            "n": n, "nnz": int(a.nnz), "rhs_per_step": k,
            "parallelism": "1 GPU",
        },
        "latency_ms": {"p50": float(np.percentile(lat_ms, 50)),
                       "p99": float(np.percentile(lat_ms, 99))},
        "iterations": {"fp32_batch": [int(v) for v in iters],
                       "fp32_single": int(it32[0]),
                       "fp64_single": int(it64[0])},
        "true_residual_max": max(true),
        "fp64_single_rhs": {
            "solves_per_s": 1.0 / t_fp64, "ms": t_fp64 * 1e3,
            "iterations": int(it64[0]), "true_residual": true64,
            "apply_GBs": gbs64, "apply_roofline_frac": gbs64 / HBM_PEAK_GBS,
        },
        "speedup_vs_fp64_single": solves_per_s * t_fp64,
        "roofline": {
            "kernel": "k_hll_spmm (first apply pass: t = d o (W x), FP32 "
                      f"factors, k={k})",
            "bound": "hbm", "achieved": gbs_w, "peak": HBM_PEAK_GBS,
            "unit": "GB/s", "frac": gbs_w / HBM_PEAK_GBS, "traffic": None,
            "bytes_per_launch": bw, "ms_per_launch": ms_w,
            "apply_both_kernels_GBs": gbs_apply,
        },
        "build": {"seconds": t_build, "report": rep},
        "cpu_baseline": cpu,
    }
    print(json.dumps(out))


def profile_apply(m, sv, n, k, reps):
    import ctypes as C
    x = m.DeviceArray.from_numpy(m.gen_normal(11, (n, k)) if k > 1
                                 else m.gen_normal(11, (n,)))
    y = m.DeviceArray((n, k) if k > 1 else (n,))
    a, b = C.c_double(), C.c_double()
    m._ck(m.lib().ainv_apply_profile(sv._h, x.ptr, y.ptr, k, reps,
                                     C.byref(a), C.byref(b)))
    return a.value, b.value


def cpu_baseline(sv, a, b, iters, gpu_iters):
    """Oracle BiCGSTAB (single thread) for `iters` iterations on the same
    160^3 system with the same FP64 INVK factors, extrapolated to one full
    solve at the GPU's iteration count."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O
    sv.cnv("F64")
    w, z, d = sv.export("W"), sv.export("Z"), sv.export("D")
    t0 = time.time()
    _, it, _ = O.krylov("BICGSTAB", a, w, z, d, b, 1e-30, itmax=iters)
    dt = time.time() - t0
    per_iter = dt / max(it, 1)
    return {"value": 1.0 / (per_iter * gpu_iters), "unit": "solves/s",
            "cores": 1, "kind": "port",
            "sample": f"{it} BiCGSTAB iterations of the oracle on the same "
                      f"matrix and FP64 factors ({dt:.1f} s), extrapolated "
                      f"to {gpu_iters} iterations per solve",
            "seconds_per_iteration": per_iter}


if __name__ == "__main__":
    main()
