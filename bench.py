"""Benchmark of the HPCG PCG iteration phase on one MI355X.

A step is one CG set: 50 MG-preconditioned CG iterations from x = 0 on the
HPCG problem (27-point stencil, 4 levels, FP64), with the matrix and vectors
resident in HBM, at 192^3 (the drop-in workload).  value = HPCG GFLOP/s by
the reference's FLOP model (ReportResults.cpp:81-99), with the convergence
correction 50/optMaxIters (:99) when the device needs more than 50
iterations to reach the reference's 50-iteration residual reduction.

The same JSON line also carries:
  * the rating with the setup charged as ReportResults.cpp:353 charges it
    (device generation and packing, amortised /10 per set);
  * HBM rates of the iteration on the model's bytes and on the bytes the
    kernels load as built (zero-start masks, elided step, column forms),
    and the reference report's own "Raw B/W" figures (ReportResults.cpp:
    103-135,339-341) for comparison;
  * the roofline of the dominant kernel (the fine-level Gauss-Seidel colour
    step, timed with HIP events on the stream it runs on);
  * the same set with the explicit 32-bit column indices (the reference's
    data structure) beside the default compact columns;
  * time to relative residual 1e-9 for the FP64 and mixed-precision solvers
    at 192^3 and at 512^3 (true residual and forward error checked in FP64),
    with HPCG-MxP-style rates for the mixed solvers;
  * the fine step's roofline and an FP64 set at 512^3;
  * the general layout (a randomly renumbered 128^3 hierarchy, which the
    reference's rule for general matrices, GenerateProblem.cpp:47-49, asks
    to work) against the stencil layout of the same matrices;
  * the 384^3 variable-coefficient operator: FP16 / FP32 / FP64 MG to 1e-9
    (FP16 also to 1e-12), forward error against an FP64 solve at 1e-12;
  * the 128^3 adjoint through torch.autograd: backward / forward time,
    finite differences on 64 kappa entries, the FP64 gradient;
  * the reference CPU build timed on this host on a bounded sample, with the
    host's CPU model, threads and NUMA nodes.

    python bench.py [--steps K] [--warmup W] [--grid N] [--gpus 1]
                    [--no-512] [--no-general] [--no-kappa] [--no-adjoint]
                    [--no-tol] [--no-cpu-baseline]

Other workloads print their own JSON line:
    --case mixed     512^3 leg alone
    --case kappa     384^3 leg alone
    --case adjoint   128^3 adjoint leg alone, plus the oracle's FP64 CPU gradient
    --case general   the general-layout leg alone
    --case cpu       the reference CPU build to relative residual 1e-9 (no GPU)
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
# HBM bytes per launch of the fine-level colour step from rocprofv3 --pmc
# runs (FETCH_SIZE + WRITE_SIZE per dispatch, each calibrated against a plain
# 1 GiB copy), keyed by grid and column format.  These are profile constants
# quoted with their source, not measured by this run.
PROFILED_TRAFFIC = {
    (192, "compact"): (269637643, "profiles/k_gs_fine_192_fp64_pmc_traffic.md"),
    (192, "explicit"): (361616960, "profiles/k_gs_fine_192_fp64_pmc_traffic.md"),
    (512, "compact"): (5050622898, "profiles/k_gs_fine_512_fp64_pmc_traffic.md"),
    (512, "explicit"): (6795489152, "profiles/k_gs_fine_512_fp64_pmc_traffic.md"),
}


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def fine_gs_bytes(g, ib=4.0):
    """Bytes of the 15 fine-level full colour steps of one standalone
    symmetric Gauss-Seidel call (forward blocks of classes 7..0, backward
    1..7).  Returns (model, as_built, launches): model = sum of m_c (8+4) +
    3 n_c 8 (the algorithmic bytes: stored nonzeros, 4-byte columns); as
    built = all 27 ELL slots of each row at 8 + ib bytes (ib: column-index
    bytes per slot of the level's column forms) + 3 n_c 8."""
    from hpcg_amd.models import colour_nnz, colour_rows
    m, n = colour_nnz(g, g, g), colour_rows(g, g, g)
    classes = [7, 6, 5, 4, 3, 2, 1, 0, 1, 2, 3, 4, 5, 6, 7]
    model = sum(m[c] * 12 + 3 * n[c] * 8 for c in classes)
    built = sum(n[c] * 27 * (8 + ib) + 3 * n[c] * 8 for c in classes)
    return model, built, len(classes)


def time_fine_step(ctx, stream, dr, dx, reps):
    """Average duration of one fine Gauss-Seidel colour step: `reps`
    symmetric Gauss-Seidel calls (15 launches each, all on `stream`)
    between two HIP events on that stream."""
    import torch
    ctx.symgs(0, dr, dx)
    ev0, ev1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    ev0.record(stream)
    for _ in range(reps):
        ctx.symgs(0, dr, dx)
    ev1.record(stream)
    ev1.synchronize()
    return ev0.elapsed_time(ev1) * 1e-3 / (reps * 15)


def level_forms(ctx, levels=4):
    return [ctx.column_forms(l) for l in range(levels)]


def per_kernel_rates(ctx, stream, g, dr, dx):
    """Each hook's kernel on the finest level, timed with HIP events on the
    bench stream (10 calls each after one untimed), against its algorithmic
    bytes (hpcg_amd.models.kernel_bytes): achieved GB/s and the fraction of
    the 8 TB/s HBM peak."""
    import torch
    from hpcg_amd.models import kernel_bytes, level_shapes
    sh = level_shapes(g, g, g)
    n, m = sh[0][0], sh[0][1]
    nc = sh[1][0]
    dy, dw, dxc, drc = ctx.vec(0), ctx.vec(0), ctx.vec(1), ctx.vec(1)
    ctx.fill(dxc, 0.5)
    cases = {
        "spmv": (lambda: ctx.spmv(0, dx, dy), kernel_bytes("spmv", n, m)),
        "symgs": (lambda: ctx.symgs(0, dr, dx), kernel_bytes("symgs", n, m)),
        "restriction (fused residual)": (lambda: ctx.restrict(0, dr, dx, drc),
                                         kernel_bytes("restrict", n, m, nc=nc, mc=m // 8)),
        "prolongation": (lambda: ctx.prolong(0, dxc, dx), kernel_bytes("prolong", n, m, nc=nc)),
        "waxpby": (lambda: ctx.waxpby(1.0, dx, -0.5, dr, dw), kernel_bytes("waxpby", n, m)),
        # these two return the scalar to the host: each call includes that read
        "dot (incl. host read)": (lambda: ctx.dot(dx, dr), kernel_bytes("dot", n, m)),
        "waxpby+dot (incl. host read)": (lambda: ctx.waxpby_dot(1.0, dx, -0.5, dr, dw),
                                         kernel_bytes("waxpby_dot", n, m)),
    }
    out = {}
    for name, (call, by) in cases.items():
        call()
        ev0, ev1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        ev0.record(stream)
        for _ in range(10):
            call()
        ev1.record(stream)
        ev1.synchronize()
        t = ev0.elapsed_time(ev1) * 1e-3 / 10
        gbs = by / t / 1e9
        out[name] = {"us": round(t * 1e6, 2), "gb_s": round(gbs, 1), "frac": round(gbs / HBM_PEAK_GBS, 4)}
        log(f"[bench] {name}: {t * 1e6:.1f} us, {gbs:.0f} GB/s algorithmic")
    for v in (dy, dw, dxc, drc):
        v.free()
    return out


def host_cpu():
    """The host the CPU baseline ran on: CPU model, hardware threads this
    process may use, NUMA nodes."""
    model = None
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    try:
        threads = len(os.sched_getaffinity(0))
    except (AttributeError, OSError):
        threads = os.cpu_count()
    try:
        numa = len([d for d in os.listdir("/sys/devices/system/node") if d.startswith("node") and d[4:].isdigit()])
    except OSError:
        numa = None
    return {"model": model, "threads_available": threads, "numa_nodes": numa}


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
            "host": host_cpu(),
            "sample": f"CG_ref, {out['niters']} iterations + preamble at {g}^3, timed after the problem "
                      f"setup: {el:.1f} s of CG; OpenMP build with setup/Make.GCC_OMP flags; its "
                      f"Gauss-Seidel is serial (ComputeSYMGS_ref.cpp has no OpenMP)"}


def run_sets(ctx, db, dx, it, steps, warmup):
    """`warmup` untimed then `steps` timed CG sets of `it` iterations from
    x = 0, tolerance 0 (main.cpp:329); returns (seconds, last result)."""
    import torch

    def step():
        ctx.fill(dx, 0.0)
        return ctx.cg(db, dx, max_iter=it, tol=0.0)

    for _ in range(warmup):
        step()
    ctx.sync()
    torch.cuda.synchronize()
    t = time.perf_counter()
    res = None
    for _ in range(steps):
        res = step()
    ctx.sync()
    torch.cuda.synchronize()
    return time.perf_counter() - t, res


def solve_timed(ctx, db, dx, variant, tol, reps=1, max_iter=5000):
    best, out = None, None
    for _ in range(reps):
        ctx.fill(dx, 0.0)
        ctx.sync()
        t = time.perf_counter()
        out = ctx.cg(db, dx, max_iter=max_iter, tol=tol, variant=variant)
        ctx.sync()
        el = time.perf_counter() - t
        best = el if best is None else min(best, el)
    out["seconds"] = best
    return out


def forward_error_ones(ctx, dx, n):
    """||x - 1|| / ||1||: the HPCG problem's exact solution is all ones
    (b = A 1 exactly, GenerateProblem_ref.cpp:171-180)."""
    e = ctx.vec(0)
    ctx.fill(e, 1.0)
    err2 = ctx.waxpby_dot(1.0, dx, -1.0, e, e)
    e.free()
    return (err2 / n) ** 0.5


def true_relres(ctx, db, dx, bnorm):
    r = ctx.vec(0)
    ctx.spmv(0, dx, r)
    rr = ctx.waxpby_dot(1.0, db, -1.0, r, r) ** 0.5
    r.free()
    return rr / bnorm


def tolerance_solves(ctx, db, dx, g, reps, tol=1e-9):
    """FP64 MG-PCG and the FP64-outer flexible PCG with FP32 and scaled-FP16
    MG to relative residual `tol`; each result's true residual ||b - A x|| /
    ||b|| and forward error ||x - 1|| / ||1|| in FP64.  The mixed solvers'
    rate is reported HPCG-MxP style: the HPCG model's FP64 operation count of
    the iterations performed over the solve time (not an HPCG rating: the
    mixed solvers stand outside the reference's rules)."""
    import numpy as np

    import hpcg_amd
    from hpcg_amd.models import algorithmic_bytes_per_iteration, hpcg_rhs, level_shapes, mxp_rate
    shapes = level_shapes(g, g, g)
    bn = float(np.linalg.norm(hpcg_rhs(g, g, g)))
    res = {}
    for name, variant, prec, vmg, smg in (("fp64", hpcg_amd.HPCGX_CG_FP64, None, 8, 8),
                                          ("mixed_fp32", hpcg_amd.HPCGX_CG_MIXED_FP32, hpcg_amd.HPCGX_PREC_FP32, 4, 4),
                                          ("mixed_fp16", hpcg_amd.HPCGX_CG_MIXED_FP16, hpcg_amd.HPCGX_PREC_FP16, 2, 4)):
        if prec is not None:
            ctx.prepare(prec)
        o = solve_timed(ctx, db, dx, variant, tol, reps=reps)
        rel = true_relres(ctx, db, dx, bn)
        fwd = forward_error_ones(ctx, dx, shapes[0][0])
        per_it = o["seconds"] / max(1, o["niters"])
        by = algorithmic_bytes_per_iteration(g, g, g, v_mg=vmg, s_mg=smg)
        mxp = mxp_rate(shapes, o["niters"], o["seconds"])["gflops"]
        res[name] = {"seconds": round(o["seconds"], 4), "iterations": o["niters"],
                     "true_rel_residual": float(f"{rel:.4e}"), "forward_error": float(f"{fwd:.4e}"),
                     "ms_per_iteration": round(per_it * 1e3, 3),
                     "algorithmic_tb_s": round(by / per_it / 1e12, 3),
                     "hpcg_model_gflops": round(mxp, 1)}
        log(f"[bench {g}^3] {name}: {o['niters']} it, {o['seconds'] * 1e3:.1f} ms, true rel res {rel:.3e}, "
            f"fwd err {fwd:.2e}, {per_it * 1e3:.2f} ms/it")
    res["speedup_fp32_vs_fp64"] = round(res["fp64"]["seconds"] / res["mixed_fp32"]["seconds"], 3)
    res["speedup_fp16_vs_fp64"] = round(res["fp64"]["seconds"] / res["mixed_fp16"]["seconds"], 3)
    return res


def leg_512(stream, g=512, reps=1):
    """The full-memory FP64 vs mixed-precision case: 512^3 (134M rows, 3.6G
    nonzeros on the fine level).  An FP64 50-iteration set rated like the
    headline, the fine step's roofline, the iteration's HBM rates, and the
    three solvers to 1e-9."""
    import numpy as np

    import hpcg_amd
    from hpcg_amd.models import (algorithmic_bytes_per_iteration, hpcg_flops, hpcg_rhs,
                                 implementation_bytes_per_iteration, index_bytes_per_slot, level_shapes)
    ctx = hpcg_amd.Context(0)
    ctx.set_stream(stream.cuda_stream)
    try:
        t = time.perf_counter()
        ctx.generate(g, g, g, levels=4)
        ctx.sync()
        setup = time.perf_counter() - t
        shapes = level_shapes(g, g, g)
        b = hpcg_rhs(g, g, g)
        db, dx = ctx.vec(0, b), ctx.vec(0)
        del b
        log(f"[bench {g}^3] setup {setup:.2f} s, device {ctx.memory_bytes() / 1e9:.1f} GB")
        el, _ = run_sets(ctx, db, dx, 50, 1, 1)
        fl = hpcg_flops(shapes, 50, sets=1)["total"]
        forms = level_forms(ctx)
        model_it = algorithmic_bytes_per_iteration(g, g, g)
        built_it = implementation_bytes_per_iteration(g, g, g, forms=forms)
        dr = ctx.vec(0, np.random.default_rng(0).random(shapes[0][0]))
        tl = time_fine_step(ctx, stream, dr, dx, 3)
        mb, bb, nl = fine_gs_bytes(g, index_bytes_per_slot(forms[0]))
        dr.free()
        out = {"grid": [g, g, g], "setup_s": round(setup, 2), "device_gb": round(ctx.memory_bytes() / 1e9, 2),
               "fp64_set": {"ms_per_set": round(el * 1e3, 2), "gflops": round(fl / el / 1e9, 1),
                            "rating_with_setup": round(fl / (el + setup / 10.0) / 1e9, 1),
                            "model_tb_s": round(model_it * 50 / el / 1e12, 3),
                            "as_built_tb_s": round(built_it * 50 / el / 1e12, 3)},
               "fine_gs_step": {"avg_launch_us": round(tl * 1e6, 1),
                                "achieved_gb_s": round(mb / nl / tl / 1e9, 1),
                                "frac": round(mb / nl / tl / (HBM_PEAK_GBS * 1e9), 4),
                                "as_built_gb_s": round(bb / nl / tl / 1e9, 1),
                                "as_built_frac": round(bb / nl / tl / (HBM_PEAK_GBS * 1e9), 4),
                                "traffic": PROFILED_TRAFFIC[(g, "compact")][0] if (g, "compact") in PROFILED_TRAFFIC
                                else None,
                                "traffic_source": PROFILED_TRAFFIC[(g, "compact")][1]
                                if (g, "compact") in PROFILED_TRAFFIC else None},
               "column_forms_level0": forms[0]}
        log(f"[bench {g}^3] FP64 set {el * 1e3:.1f} ms ({fl / el / 1e9:.1f} GFLOP/s); fine step {tl * 1e6:.1f} us")
        out["time_to_rel_res_1e-9"] = tolerance_solves(ctx, db, dx, g, reps)
        return out
    finally:
        ctx.close()


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--grid", type=int, default=192)
    ap.add_argument("--iters", type=int, default=50)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-tol", action="store_true", help="skip the time-to-1e-9 solves")
    ap.add_argument("--no-512", action="store_true", help="skip the 512^3 leg")
    ap.add_argument("--no-general", action="store_true", help="skip the general-layout leg")
    ap.add_argument("--no-kappa", action="store_true", help="skip the 384^3 variable-coefficient leg")
    ap.add_argument("--no-adjoint", action="store_true", help="skip the 128^3 adjoint leg")
    ap.add_argument("--cpu-iters", type=int, default=20)
    ap.add_argument("--case", default="hpcg", choices=("hpcg", "mixed", "kappa", "adjoint", "general", "cpu"))
    ap.add_argument("--pmc-calibration", action="store_true",
                    help="first copy 1 GiB with a plain copy kernel (scale for FETCH_SIZE/WRITE_SIZE)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only: this build has no multi-GPU path "
              f"(--gpus {args.gpus} requested)", file=sys.stderr)
        sys.exit(2)

    if args.case == "mixed":
        return case_mixed_512(args)
    if args.case == "kappa":
        return case_kappa_fp16(args)
    if args.case == "adjoint":
        return case_adjoint(args)
    if args.case == "general":
        return case_general(args)
    if args.case == "cpu":
        return case_cpu(args)

    import numpy as np
    import torch

    import hpcg_amd
    from hpcg_amd.models import (algorithmic_bytes_per_iteration, hpcg_bytes, hpcg_flops, hpcg_rhs,
                                 implementation_bytes_per_iteration, index_bytes_per_slot, level_shapes)

    g, it = args.grid, args.iters
    torch.cuda.init()
    if args.pmc_calibration:
        src = torch.ones(1 << 27, dtype=torch.float64, device="cuda")  # 1 GiB
        dst = torch.empty_like(src)
        for _ in range(3):
            dst.copy_(src)
        torch.cuda.synchronize()
        del src, dst
        log("[bench] calibration: 3 copies of 1 GiB (1048576 KiB read + 1048576 KiB written each)")
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

    el, res = run_sets(ctx, db, dx, it, args.steps, args.warmup)
    ms = el / args.steps * 1e3
    fl = hpcg_flops(shapes, it, sets=args.steps)["total"] * (it / opt_iters)
    gflops = fl / el / 1e9
    # ReportResults.cpp:353: setup (t9) and optimisation (t7) charged /10 per
    # set; here both are the device generation and packing (setup_s)
    rating = fl / (el + args.steps * setup_s / 10.0) / 1e9
    forms = level_forms(ctx)
    bytes_it = algorithmic_bytes_per_iteration(g, g, g)
    built_it = implementation_bytes_per_iteration(g, g, g, forms=forms)
    tbs = bytes_it * it * args.steps / el / 1e12
    built_tbs = built_it * it * args.steps / el / 1e12
    # the reference report's "Raw Read / Write / Total B/W" (ReportResults.cpp:339-341)
    rb = hpcg_bytes(shapes, it, sets=args.steps)
    ref_bw = {"read_gb_s": round(rb["reads"] / el / 1e9, 1), "write_gb_s": round(rb["writes"] / el / 1e9, 1),
              "total_gb_s": round(rb["total"] / el / 1e9, 1),
              "rule": "ReportResults.cpp:103-135,339-341: the reference's read/write model over the CG time; "
                      "it counts residual and post-smoother writes per nonzero and every sweep in full, "
                      "so it runs above the HBM peak and is no roofline"}
    log(f"[bench] {args.steps} sets: {ms:.2f} ms/set, {gflops:.1f} GFLOP/s (rating with setup {rating:.1f}), "
        f"{tbs:.3f} TB/s model, {built_tbs:.3f} TB/s as built, scaled residual {res['normr'] / res['normr0']:.4e}")

    # roofline of the dominant kernel: the fine-level full colour step
    dr = ctx.vec(0, np.random.default_rng(0).random(shapes[0][0]))
    t_launch = time_fine_step(ctx, stream, dr, dx, 20)
    gs_model, gs_built, nlaunch = fine_gs_bytes(g, index_bytes_per_slot(forms[0]))
    achieved = gs_model / nlaunch / t_launch / 1e9
    traffic = PROFILED_TRAFFIC.get((g, "compact"))
    roof = {"kernel": "k_gs<double,double,double,-1,true,true> (fine-level Gauss-Seidel colour step, compact columns)",
            "bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4),
            "traffic": traffic[0] if traffic else None,
            "traffic_source": traffic[1] if traffic else None,
            "avg_launch_us": round(t_launch * 1e6, 2), "bytes_per_launch": int(gs_model / nlaunch),
            "as_built_bytes_per_launch": int(gs_built / nlaunch),
            "as_built_gb_s": round(gs_built / nlaunch / t_launch / 1e9, 1),
            "as_built_frac": round(gs_built / nlaunch / t_launch / 1e9 / HBM_PEAK_GBS, 4)}
    log(f"[bench] fine GS step: {t_launch * 1e6:.1f} us/launch, {achieved:.0f} GB/s model, "
        f"{roof['as_built_gb_s']:.0f} GB/s as built")
    kernels = per_kernel_rates(ctx, stream, g, dr, dx)

    # the reference's data structure: explicit 32-bit column indices
    ctx.set_column_format(hpcg_amd.HPCGX_COLS_EXPLICIT)
    el_x, res_x = run_sets(ctx, db, dx, it, max(2, args.steps // 2), 1)
    t_launch_x = time_fine_step(ctx, stream, dr, dx, 10)
    ctx.set_column_format(hpcg_amd.HPCGX_COLS_COMPACT)
    sets_x = max(2, args.steps // 2)
    explicit = {"ms_per_set": round(el_x / sets_x * 1e3, 3),
                "gflops": round(hpcg_flops(shapes, it, sets=sets_x)["total"] * (it / opt_iters) / el_x / 1e9, 2),
                "fine_step_us": round(t_launch_x * 1e6, 2),
                "fine_step_frac": round(gs_model / nlaunch / t_launch_x / 1e9 / HBM_PEAK_GBS, 4),
                "same_residual": res_x["normr"] == res["normr"]}
    log(f"[bench] explicit columns: {explicit}")
    dr.free()

    tol_res = None
    if not args.no_tol:
        tol_res = tolerance_solves(ctx, db, dx, g, reps=2)

    ctx.close()
    torch.cuda.synchronize()

    big = None
    if not args.no_512:
        try:
            big = leg_512(stream)
        except Exception as e:  # reported, never silently dropped
            log(f"[bench] 512^3 leg failed: {e!r}")
            big = {"error": repr(e)}

    general = None
    if not args.no_general:
        try:
            general = leg_general(stream, 128)
        except Exception as e:
            log(f"[bench] general leg failed: {e!r}")
            general = {"error": repr(e)}

    kap = None
    if not args.no_kappa:
        try:
            kap = leg_kappa(stream, 384)
        except Exception as e:
            log(f"[bench] kappa leg failed: {e!r}")
            kap = {"error": repr(e)}

    adj = None
    if not args.no_adjoint:
        try:
            adj = leg_adjoint(128)
        except Exception as e:
            log(f"[bench] adjoint leg failed: {e!r}")
            adj = {"error": repr(e)}

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
                   "grid": [g, g, g], "levels": 4, "iterations_per_set": it, "opt_max_iters": opt_iters,
                   "column_format": "compact"},
        "rating_with_setup": {"gflops": round(rating, 2), "setup_s": round(setup_s, 3),
                              "rule": "ReportResults.cpp:353, setup and optimisation /10 per set"},
        "hbm_tb_s": round(tbs, 4),
        "hbm_tb_s_as_built": round(built_tbs, 4),
        "reference_raw_bw": ref_bw,
        "column_forms_level0": forms[0],
        "roofline": roof,
        "explicit_columns": explicit,
        "cpu_baseline": cpu,
        "kernels": kernels,
        "time_to_rel_res_1e-9": tol_res,
        "grid_512": big,
        "general_layout_128": general,
        "kappa_384": kap,
        "adjoint_128": adj,
    }
    print(json.dumps(out), flush=True)


def case_mixed_512(args):
    """512^3 HPCG problem, FP64 MG-PCG vs FP64-outer + FP32 (and
    scaled FP16) MG to relative residual 1e-9, verified in FP64."""
    import torch
    g = args.grid if args.grid != 192 else 512
    stream = torch.cuda.Stream()
    r = leg_512(stream, g)
    t = r["time_to_rel_res_1e-9"]
    out = {"metric": METRIC, "config": {"workload": f"HPCG {g}^3, FP64 MG-PCG vs FP64-outer "
                                                    f"flexible PCG with FP32 / scaled-FP16 MG, to rel-res 1e-9",
                                        "grid": [g, g, g], "levels": 4},
           "value": t["speedup_fp32_vs_fp64"], "unit": "x (time to 1e-9, FP64 / mixed FP32)", **r}
    print(json.dumps(out), flush=True)


def leg_general(stream, g=128):
    """The reference requires optimisations to work for general sparse
    matrices (GenerateProblem.cpp:47-49).  The g^3 HPCG hierarchy, each level
    renumbered by a seeded random permutation (P A P^T, injection mapped
    along, entries kept in their row order), is uploaded through
    hpcgx_level_upload_rows and runs in the general layout; the same
    matrices in natural numbering run in the stencil layout.  Reports the
    FP64 50-iteration set of both and the colours."""
    import numpy as np

    import hpcg_amd
    from hpcg_amd.models import hpcg_flops, hpcg_rhs, level_shapes
    st = hpcg_amd.Context(0)
    st.set_stream(stream.cuda_stream)
    st.generate(g, g, g, levels=4)
    rng = np.random.default_rng(4242)
    levels = []
    for k in range(4):
        nr, cols, vals = st.export_level(k)
        levels.append((nr, cols, vals, g >> k))
    perms = [rng.permutation(len(lv[0])) for lv in levels]  # new index of old row
    gen = hpcg_amd.Context(0)
    gen.set_stream(stream.cuda_stream)
    t = time.perf_counter()
    for k, (nr, cols, vals, d) in enumerate(levels):
        p = perms[k]
        inv = np.argsort(p)
        c2 = cols[inv]
        c2 = np.where(c2 >= 0, p[np.maximum(c2, 0)], -1).astype(np.int32)
        f2c = None
        if k + 1 < len(levels):
            dc = d // 2
            i = np.arange(dc ** 3)
            fx, fy, fz = 2 * (i % dc), 2 * ((i // dc) % dc), 2 * (i // (dc * dc))
            f_old = (fz * d + fy) * d + fx  # GenerateCoarseProblem.cpp:71-82
            f2c = np.empty(dc ** 3, np.int32)
            f2c[perms[k + 1]] = p[f_old]
        gen.upload_rows(k, nr[inv], c2, vals[inv], f2c, 0 if f2c is None else len(f2c))
    gen.sync()
    upload_s = time.perf_counter() - t
    shapes = level_shapes(g, g, g)
    b = hpcg_rhs(g, g, g)
    res = {}
    try:
        for name, ctx, bb, fmt in (("stencil layout", st, b, hpcg_amd.HPCGX_COLS_COMPACT),
                                   ("general layout", gen, b[np.argsort(perms[0])], hpcg_amd.HPCGX_COLS_COMPACT),
                                   ("general layout, explicit columns", gen, b[np.argsort(perms[0])],
                                    hpcg_amd.HPCGX_COLS_EXPLICIT)):
            ctx.set_column_format(fmt)
            db, dx = ctx.vec(0, bb), ctx.vec(0)
            el, o = run_sets(ctx, db, dx, 50, 3, 1)
            el /= 3
            gfl = hpcg_flops(shapes, 50, sets=1)["total"] / el / 1e9
            cols_ = [ctx.colouring(k)[1] for k in range(4)]
            res[name] = {"ms_per_set": round(el * 1e3, 2), "gflops": round(gfl, 1),
                         "scaled_residual_50": float(f"{o['normr'] / o['normr0']:.4e}"), "colours": cols_,
                         "column_forms_level0": ctx.column_forms(0),
                         "columns_read": [ctx.column_reads(k) for k in range(4)]}
            log(f"[bench general] {name}: {el * 1e3:.1f} ms/set, {gfl:.1f} GFLOP/s, colours {cols_}, "
                f"residual {o['normr'] / o['normr0']:.3e}")
    finally:
        st.close()
        gen.close()
    res["general_over_stencil"] = round(res["general layout"]["gflops"] / res["stencil layout"]["gflops"], 4)
    res["general_upload_s"] = round(upload_s, 2)
    res["grid"] = [g, g, g]
    return res


def case_cpu(args):
    """The reference CPU build alone, to relative residual 1e-9 as the
    reference's own CG_ref measures it (CG_ref.cpp:100, recursive residual;
    natural-order Gauss-Seidel, so more iterations than the device's
    colouring needs): the CPU side of the time-to-solution comparison.  Takes
    tens of seconds at 192^3, so it is not part of the default line."""
    from hpcg_amd.models import hpcg_flops, level_shapes
    from oracle import pyoracle as po
    g = args.grid
    if not po.ref_available(omp=True):
        print(json.dumps({"metric": METRIC, "error": "reference CPU build not present (make -C oracle ref)"}))
        return 1
    R = po.RefProblem(g, g, g, omp=True)
    t = time.perf_counter()
    out = R.cg(max_iter=500, tol=1e-9)
    el = time.perf_counter() - t
    cores = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    fl = hpcg_flops(level_shapes(g, g, g), out["niters"], sets=1)["total"]
    res = {"metric": METRIC, "config": {"workload": f"reference CG_ref (OpenMP build, setup/Make.GCC_OMP flags) "
                                                    f"to relative residual 1e-9 at {g}^3", "grid": [g, g, g]},
           "value": round(el, 3), "unit": "s", "higher_is_better": False, "iterations": out["niters"],
           "relative_residual": out["normr"] / out["normr0"], "gflops": round(fl / el / 1e9, 4),
           "cores": cores, "host": host_cpu()}
    log(f"[bench cpu {g}^3] CG_ref to 1e-9: {out['niters']} iterations, {el:.2f} s on {cores} threads")
    print(json.dumps(res), flush=True)
    return 0


def case_general(args):
    import torch
    g = args.grid if args.grid != 192 else 128
    res = leg_general(torch.cuda.Stream(), g)
    out = {"metric": METRIC, "config": {"workload": f"HPCG {g}^3 hierarchy renumbered by a random permutation per "
                                                    f"level, FP64 MG-PCG 50-iteration set, general vs stencil layout",
                                        "grid": [g, g, g]},
           "value": res["general layout"]["gflops"], "unit": "GFLOP/s (general layout)", "layouts": res}
    print(json.dumps(out), flush=True)


def kappa_field(g, seed=2026):
    import numpy as np
    rng = np.random.default_rng(seed)
    kappa = np.exp(0.5 * rng.standard_normal(g ** 3))
    b = rng.uniform(-1.0, 1.0, g ** 3)
    return kappa, b


def leg_kappa(stream, g=384):
    """The variable-coefficient operator at g^3 (log kappa ~ N(0, 0.5^2),
    b ~ U(-1,1), seeded): FP64 MG-PCG to 1e-12 is the reference solve; the
    FP64-outer flexible PCG with the scaled-FP16 MG (values D^-1/2 A D^-1/2
    in FP16, FP32 arithmetic) and with the FP32 MG, and FP64 MG-PCG, each to
    1e-9 (FP16 also to 1e-12): time, iterations, true residual in FP64 and
    forward error against the reference solve."""
    import numpy as np

    import hpcg_amd
    kappa, b = kappa_field(g)
    ctx = hpcg_amd.Context(0)
    ctx.set_stream(stream.cuda_stream)
    try:
        t = time.perf_counter()
        ctx.generate(g, g, g, levels=4, kappa=kappa)
        ctx.prepare(hpcg_amd.HPCGX_PREC_FP16)
        ctx.prepare(hpcg_amd.HPCGX_PREC_FP32)
        ctx.sync()
        setup = time.perf_counter() - t
        del kappa
        bn = float(np.linalg.norm(b))
        db, dx, dref = ctx.vec(0, b), ctx.vec(0), ctx.vec(0)
        ref = solve_timed(ctx, db, dref, hpcg_amd.HPCGX_CG_FP64, 1e-12)
        xref = ctx.download(dref)
        xn = float(np.linalg.norm(xref))
        rows = {"fp64_1e-12": {"seconds": round(ref["seconds"], 3), "iterations": ref["niters"],
                               "true_rel_residual": float(f"{true_relres(ctx, db, dref, bn):.4e}")}}
        log(f"[bench kappa {g}^3] fp64 reference: {ref['niters']} it {ref['seconds']:.2f} s")
        for name, variant, tol in (("mixed_fp16", hpcg_amd.HPCGX_CG_MIXED_FP16, 1e-9),
                                   ("mixed_fp16", hpcg_amd.HPCGX_CG_MIXED_FP16, 1e-12),
                                   ("mixed_fp32", hpcg_amd.HPCGX_CG_MIXED_FP32, 1e-9),
                                   ("fp64", hpcg_amd.HPCGX_CG_FP64, 1e-9)):
            o = solve_timed(ctx, db, dx, variant, tol)
            fe = float(np.linalg.norm(ctx.download(dx) - xref) / xn)
            rel = true_relres(ctx, db, dx, bn)
            rows[f"{name}_{tol:g}"] = {"seconds": round(o["seconds"], 3), "iterations": o["niters"],
                                       "true_rel_residual": float(f"{rel:.4e}"), "forward_error": float(f"{fe:.4e}")}
            log(f"[bench kappa {g}^3] {name} tol {tol:g}: {o['niters']} it {o['seconds']:.2f} s rel {rel:.2e} "
                f"fwd err {fe:.2e}")
    finally:
        ctx.close()
    return {"grid": [g, g, g], "setup_s": round(setup, 2), "solves": rows,
            "speedup_fp16_vs_fp64_1e-9": round(rows["fp64_1e-09"]["seconds"] / rows["mixed_fp16_1e-09"]["seconds"], 3),
            "speedup_fp32_vs_fp64_1e-9": round(rows["fp64_1e-09"]["seconds"] / rows["mixed_fp32_1e-09"]["seconds"], 3)}


def case_kappa_fp16(args):
    import torch
    g = args.grid if args.grid != 192 else 384
    res = leg_kappa(torch.cuda.Stream(), g)
    out = {"metric": METRIC, "config": {"workload": f"variable-coefficient 27-point operator {g}^3, "
                                                    f"log kappa ~ N(0, 0.25), b ~ U(-1,1); MG in scaled FP16 values, "
                                                    f"FP32 arithmetic, FP64 outer PCG", "grid": [g, g, g]},
           "value": res["solves"]["mixed_fp16_1e-09"]["forward_error"],
           "unit": "forward error at rel-res 1e-9 (FP16 MG)", "higher_is_better": False, **res}
    print(json.dumps(out), flush=True)


def leg_adjoint(g=128, cpu_check=False):
    """Adjoint through the solve at g^3 (kappa, b drawn as for the kappa
    leg); L = 1/2 |x - xt|^2 with xt a seeded N(0,1) field smoothed by 3
    SYMGS sweeps.  Forward and adjoint solves: mixed FP32-MG / FP64-outer to
    1e-10.  Checks dL/dkappa against central finite differences on 64 kappa
    entries (h = 1e-6, FP64 solves to 1e-13) and against the FP64 gradient
    (FP64 MG-PCG to 1e-13 forward and adjoint, on the device; with
    cpu_check also the oracle's CPU gradient)."""
    import numpy as np
    import torch

    from hpcg_amd.adjoint import KappaOperator, KappaSolve, kappa_solve
    kappa, b = kappa_field(g)
    rng = np.random.default_rng(77)
    # target: seeded N(0,1) field, 3 symmetric Gauss-Seidel sweeps with r = 0 on the operator
    op = KappaOperator(kappa, (g, g, g), variant="fp64")
    xt = torch.tensor(rng.standard_normal(g ** 3), device="cuda")
    xd, zr = op._dev(None), torch.zeros(op.npad, dtype=torch.float64, device="cuda")
    op.ctx.from_natural(xd, xt, level=0)
    for _ in range(3):
        op.ctx.symgs(0, zr, xd)
    op.ctx.to_natural(xd, xt, level=0)
    op.ctx.sync()
    op.close()

    def grad(variant, tol):
        k = torch.tensor(kappa, device="cuda", requires_grad=True)
        bt = torch.tensor(b, device="cuda", requires_grad=True)
        torch.cuda.synchronize()
        t = time.perf_counter()
        x = kappa_solve(k, bt, (g, g, g), tol, variant)
        loss = 0.5 * ((x - xt) ** 2).sum()
        torch.cuda.synchronize()
        t_fwd = time.perf_counter() - t
        t = time.perf_counter()
        loss.backward()
        torch.cuda.synchronize()
        return k.grad.detach(), bt.grad.detach(), t_fwd, time.perf_counter() - t

    for rep in range(2):  # the first pass warms torch's kernels up; the second is timed
        gk, gb, t_fwd, t_bwd = grad("mixed_fp32", 1e-10)
    tim = dict(KappaSolve.last_timing)
    log(f"[bench adjoint {g}^3] forward {t_fwd:.3f} s backward {t_bwd:.3f} s timing {tim}")
    gk64, gb64, _, _ = grad("fp64", 1e-13)
    dev_rel = float(torch.linalg.norm(gk - gk64) / torch.linalg.norm(gk64))
    devb_rel = float(torch.linalg.norm(gb - gb64) / torch.linalg.norm(gb64))
    log(f"[bench adjoint {g}^3] vs FP64 gradient (1e-13): dL/dkappa rel {dev_rel:.2e}, dL/db rel {devb_rel:.2e}")
    # finite differences on 64 kappa entries, fp64 solves to 1e-13
    bd = torch.tensor(b, device="cuda")
    idx = rng.choice(g ** 3, 64, replace=False)
    h = 1e-6
    fd = []
    for j in idx:
        xs = []
        for sgn in (1.0, -1.0):
            kp = kappa.copy()
            kp[j] += sgn * h
            o = KappaOperator(kp, (g, g, g), variant="fp64")
            xs.append(o.solve(bd, tol=1e-13))
            o.close()
        fd.append(0.5 * torch.dot(xs[0] - xs[1], xs[0] + xs[1] - 2 * xt).item() / (2 * h))
    fd = np.array(fd)
    ga = gk[torch.tensor(idx, device="cuda")].cpu().numpy()
    fd_rel = float(np.max(np.abs(fd - ga)) / np.max(np.abs(ga)))
    log(f"[bench adjoint {g}^3] FD max rel err (vs max |g|) {fd_rel:.2e}")
    out = {"grid": [g, g, g], "backward_over_forward": round(t_bwd / t_fwd, 3),
           "solve_ratio_adjoint_over_forward": round(tim["adjoint_solve_s"] / tim["forward_solve_s"], 3),
           "forward_s": round(t_fwd, 4), "backward_s": round(t_bwd, 4), "timing": tim,
           "fd_64_entries_max_rel_err": fd_rel, "fp64_gradient_rel_err": dev_rel, "fp64_db_rel_err": devb_rel}
    if cpu_check:
        # fp64 CPU reference gradient (the oracle, fp64 CG to 1e-13): the checker
        try:
            from oracle import pyoracle as po
            P = po.Problem(g, g, g, kappa=kappa)
            xtn = xt.cpu().numpy()
            xo = P.cg(b, max_iter=2000, tol=1e-13, mode=2)["x"]
            lo = P.cg(xo - xtn, max_iter=2000, tol=1e-13, mode=2)["x"]
            go = P.kappa_grad(lo, xo)
            out["fp64_cpu_gradient_rel_err"] = float(np.linalg.norm(gk.cpu().numpy() - go) / np.linalg.norm(go))
            out["fp64_cpu_db_rel_err"] = float(np.linalg.norm(gb.cpu().numpy() - lo) / np.linalg.norm(lo))
            log(f"[bench adjoint {g}^3] vs fp64 CPU gradient: dL/dkappa rel {out['fp64_cpu_gradient_rel_err']:.2e}, "
                f"dL/db rel {out['fp64_cpu_db_rel_err']:.2e}")
        except Exception as e:
            log(f"[bench adjoint {g}^3] CPU gradient failed: {e}")
    return out


def case_adjoint(args):
    g = args.grid if args.grid != 192 else 128
    res = leg_adjoint(g, cpu_check=True)
    out = {"metric": METRIC, "config": {"workload": f"adjoint of the {g}^3 variable-coefficient solve "
                                                    f"through torch.autograd.Function, mixed FP32-MG / FP64-outer "
                                                    f"PCG to 1e-10", "grid": [g, g, g]},
           "value": res["backward_over_forward"], "unit": "backward / forward time (autograd calls, incl. setup)",
           "higher_is_better": False, **res}
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
