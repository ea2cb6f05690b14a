"""Benchmark of the HPCG PCG iteration phase on one MI355X.

A step is one CG set: 50 MG-preconditioned CG iterations from x = 0 on the
HPCG problem (27-point stencil, 4 levels, FP64), with the matrix and vectors
resident in HBM.  value = HPCG GFLOP/s by the reference's FLOP model
(ReportResults.cpp:81-99), with the convergence correction 50/optMaxIters
(:99) when the device needs more than 50 iterations to reach the reference's
50-iteration residual reduction.  Also reported: algorithmic HBM TB/s, the
roofline of the dominant kernel (the fine-level Gauss-Seidel colour step,
timed with HIP events on the stream it runs on), time to relative residual
1e-9 for the FP64 and mixed-precision solvers, and the reference CPU build
timed on this host on a bounded sample.

    python bench.py [--steps K] [--warmup W] [--grid N] [--gpus 1]
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "HPCG GFLOP/s + HBM TB/s; time to rel-res 1e-9: mixed vs FP64 MG-PCG"
HBM_PEAK_GBS = 8000.0
# the reference's 50-iteration scaled residual (CG_ref, natural-order
# Gauss-Seidel), measured from the reference build at these grids
REF_REDUCTION = {16: 3.66234e-41, 192: 5.31511e-05}


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def fine_gs_bytes(g):
    """Algorithmic bytes of the 15 fine-level full colour steps of one
    standalone symmetric Gauss-Seidel call (forward blocks of classes 7..0,
    backward 1..7): sum of m_c (8+4) + 3 n_c 8."""
    from hpcg_amd.models import colour_nnz, colour_rows
    m, n = colour_nnz(g, g, g), colour_rows(g, g, g)
    classes = [7, 6, 5, 4, 3, 2, 1, 0, 1, 2, 3, 4, 5, 6, 7]
    return sum(m[c] * 12 + 3 * n[c] * 8 for c in classes), len(classes)


def cpu_baseline(g, iters):
    """The reference's CG_ref (OpenMP build, setup/Make.GCC_OMP flags) on the
    same grid for a few iterations, rated with the same FLOP model."""
    from hpcg_amd.models import hpcg_flops, level_shapes
    from oracle import pyoracle as po
    if not po.ref_available(omp=True):
        return None
    R = po.RefProblem(g, g, g, omp=True)
    t = time.perf_counter()
    out = R.cg(max_iter=iters)
    el = time.perf_counter() - t
    fl = hpcg_flops(level_shapes(g, g, g), out["niters"], sets=1)["total"]
    cores = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    return {"value": round(fl / el / 1e9, 4), "unit": "GFLOP/s", "cores": cores, "kind": "reference",
            "sample": f"CG_ref, {out['niters']} iterations + preamble at {g}^3 after setup "
                      f"({el:.1f} s), OpenMP build with setup/Make.GCC_OMP flags; its Gauss-Seidel "
                      f"is serial (ComputeSYMGS_ref.cpp has no OpenMP)"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--grid", type=int, default=192)
    ap.add_argument("--iters", type=int, default=50)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-tol", action="store_true", help="skip the time-to-1e-9 solves")
    ap.add_argument("--cpu-iters", type=int, default=5)
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only: this build has no multi-GPU path "
              f"(--gpus {args.gpus} requested)", file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import torch

    import hpcg_amd
    from hpcg_amd.models import algorithmic_bytes_per_iteration, hpcg_flops, hpcg_rhs, level_shapes

    g, it = args.grid, args.iters
    torch.cuda.init()
    stream = torch.cuda.Stream()  # default priority; every kernel below runs on it
    ctx = hpcg_amd.Context(0)
    ctx.set_stream(stream.cuda_stream)
    t0 = time.perf_counter()
    ctx.generate(g, g, g, levels=4)
    ctx.sync()
    setup_s = time.perf_counter() - t0
    shapes = level_shapes(g, g, g)
    b = hpcg_rhs(g, g, g)
    db, dx = ctx.vec(0, b), ctx.vec(0)
    log(f"[bench] {g}^3: {shapes[0][0]} rows, {shapes[0][1]} nnz, setup {setup_s:.2f} s, "
        f"device bytes {ctx.memory_bytes() / 1e9:.2f} GB")

    # iterations needed for the reference's 50-iteration reduction (main.cpp:281-293)
    opt_iters = it
    red = REF_REDUCTION.get(g)
    if red is not None:
        ctx.fill(dx, 0.0)
        v = ctx.cg(db, dx, max_iter=10 * it, tol=red)
        opt_iters = max(it, v["niters"])
        log(f"[bench] reference reduction {red:.5e} reached in {v['niters']} iterations")

    def step():
        ctx.fill(dx, 0.0)
        return ctx.cg(db, dx, max_iter=it, tol=0.0)

    for _ in range(args.warmup):
        step()
    ctx.sync()
    torch.cuda.synchronize()
    t = time.perf_counter()
    for _ in range(args.steps):
        res = step()
    ctx.sync()
    torch.cuda.synchronize()
    el = time.perf_counter() - t
    ms = el / args.steps * 1e3
    fl = hpcg_flops(shapes, it, sets=args.steps)["total"] * (it / opt_iters)
    gflops = fl / el / 1e9
    bytes_it = algorithmic_bytes_per_iteration(g, g, g)
    tbs = bytes_it * it * args.steps / el / 1e12
    log(f"[bench] {args.steps} sets: {ms:.2f} ms/set, {gflops:.1f} GFLOP/s, {tbs:.3f} TB/s "
        f"(algorithmic), scaled residual {res['normr'] / res['normr0']:.4e}")

    # roofline of the dominant kernel: the fine-level full colour step
    dr = ctx.vec(0, np.random.default_rng(0).random(shapes[0][0]))
    reps = 20
    ctx.symgs(0, dr, dx)
    ev0, ev1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    ev0.record(stream)
    for _ in range(reps):
        ctx.symgs(0, dr, dx)
    ev1.record(stream)
    ev1.synchronize()
    gs_bytes, nlaunch = fine_gs_bytes(g)
    t_launch = ev0.elapsed_time(ev1) * 1e-3 / (reps * nlaunch)
    achieved = gs_bytes / nlaunch / t_launch / 1e9
    roof = {"kernel": "k_gs<double,double,double,-1,true> (fine-level Gauss-Seidel colour step)",
            "bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": None,
            "avg_launch_us": round(t_launch * 1e6, 2), "bytes_per_launch": int(gs_bytes / nlaunch)}
    log(f"[bench] fine GS step: {t_launch * 1e6:.1f} us/launch, {achieved:.0f} GB/s")

    # time to relative residual 1e-9 (true residual checked in fp64)
    tol_res = None
    if not args.no_tol:
        tol_res = {}
        dAx = ctx.vec(0)
        bn = float(np.linalg.norm(b))
        for name, variant, prec in (("fp64", hpcg_amd.HPCGX_CG_FP64, None),
                                    ("mixed_fp32", hpcg_amd.HPCGX_CG_MIXED_FP32, hpcg_amd.HPCGX_PREC_FP32),
                                    ("mixed_fp16", hpcg_amd.HPCGX_CG_MIXED_FP16, hpcg_amd.HPCGX_PREC_FP16)):
            if prec is not None:
                ctx.prepare(prec)
            best = None
            for _ in range(2):
                ctx.fill(dx, 0.0)
                ctx.sync()
                t = time.perf_counter()
                o = ctx.cg(db, dx, max_iter=2000, tol=1e-9, variant=variant)
                ctx.sync()
                el2 = time.perf_counter() - t
                best = el2 if best is None else min(best, el2)
            ctx.spmv(0, dx, dAx)
            rel = ctx.waxpby_dot(1.0, db, -1.0, dAx, dAx) ** 0.5 / bn
            tol_res[name] = {"seconds": round(best, 4), "iterations": o["niters"],
                             "true_rel_residual": float(f"{rel:.4e}")}
            log(f"[bench] {name}: {o['niters']} it, {best * 1e3:.1f} ms, true rel res {rel:.3e}")
        tol_res["speedup_fp32_vs_fp64"] = round(tol_res["fp64"]["seconds"] / tol_res["mixed_fp32"]["seconds"], 3)
        tol_res["speedup_fp16_vs_fp64"] = round(tol_res["fp64"]["seconds"] / tol_res["mixed_fp16"]["seconds"], 3)

    cpu = None
    if not args.no_cpu_baseline:
        try:
            cpu = cpu_baseline(g, args.cpu_iters)
            if cpu:
                log(f"[bench] cpu baseline {cpu['value']} GFLOP/s on {cpu['cores']} threads")
        except Exception as e:  # the baseline never decides the bench
            log(f"[bench] cpu baseline failed: {e}")

    out = {
        "metric": METRIC, "value": round(gflops, 2), "unit": "GFLOP/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(ms, 3),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": f"HPCG FP64 MG-PCG, {g}^3 grid, 4 levels, {it} CG iterations per set",
                   "grid": [g, g, g], "levels": 4, "iterations_per_set": it, "opt_max_iters": opt_iters},
        "hbm_tb_s": round(tbs, 4),
        "roofline": roof,
        "cpu_baseline": cpu,
        "time_to_rel_res_1e-9": tol_res,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
