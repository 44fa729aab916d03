"""Benchmark: bk_sugi box full-mode signal stage at 512^3 (split-bf16 MFMA).

Workload (the metric's configuration): box 2000 Mpc/h, nmesh 512^3, PCS
interlaced, 30 k-bins over 0.01-0.31 h/Mpc (465 k1 <= k2 pairs),
angu_config [0,0,0] and [2,0,2], tracer aaa, 5e6 lognormal + RSD particles
(seeded synthetic catalog).

One step = the signal stage of both angular configurations: from the
compensated delta_k resident in HBM (FP32) through G_00, Ylm-weighted shell
filtering, the pruned shell synthesis (x and y passes) and the split-bf16
MFMA contraction with the z pass fused in, to the two 30 x 30 signal
matrices.  bins/step = 2 x 465.

Also reported (outside the timed steps): the same stage with the real shells
written to HBM and contracted apart, the split-fp16 and split-fp8 variants,
the plain FP64 pairwise path and the FP64 MFMA GEMM path on the same
catalog, each split variant's error relative to max|B| against the FP64
result, per-kernel roofline timings, the batched-mocks comparison (64 box
mocks at 256^3, batched driver against the per-mock loop) and the CPU
baseline (the FP64 numpy oracle timed on one full angular configuration).
"""
import argparse
import json
import os
import platform
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
MFMA_BF16_PEAK_TFLOPS = 2500.0


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def parse():
    p = argparse.ArgumentParser()
    p.add_argument("--gpus", type=int, default=1)
    p.add_argument("--steps", type=int, default=20)
    p.add_argument("--warmup", type=int, default=3)
    p.add_argument("--nmesh", type=int, default=512)
    p.add_argument("--nparticles", type=int, default=5_000_000)
    p.add_argument("--k-bins", type=int, default=30)
    p.add_argument("--no-fp64", action="store_true", help="skip the FP64 reference paths")
    p.add_argument("--no-cpu", action="store_true", help="skip the CPU oracle baseline")
    p.add_argument("--no-graph", action="store_true", help="time the eager launches only (no hipGraph replay)")
    p.add_argument("--no-mocks", action="store_true", help="skip the batched-mocks comparison of the default run")
    p.add_argument("--roofline-only", action="store_true")
    p.add_argument("--workload", choices=("metric", "mocks", "kernels", "survey"), default="metric",
                   help="mocks: 64 box mocks (pk + bk_sugi diagonal at 256^3), batched driver vs per-mock loop; "
                        "kernels: one launch of each hot kernel at 512^3 plus a 1 GiB device copy (for counters)")
    p.add_argument("--mocks", type=int, default=64)
    p.add_argument("--mock-batch", type=int, default=16)
    p.add_argument("--loop-mocks", type=int, default=16, help="mocks timed in the per-mock run_stats loop")
    return p.parse_args()


def main():
    args = parse()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (got --gpus {args.gpus}); the path is not sharded across GPUs.",
              file=sys.stderr)
        sys.exit(2)
    if args.workload == "mocks":
        return bench_mocks(args)
    if args.workload == "kernels":
        return bench_kernels(args)
    if args.workload == "survey":
        return bench_survey(args)
    # the mock catalogs are drawn in worker processes before anything touches the GPU
    mock_cats = None if args.no_mocks or args.mocks < 1 else mock_catalogs(args.mocks)
    import numpy as np
    import torch

    from cosmonpc_amd import angular as A
    from cosmonpc_amd import synth
    from cosmonpc_amd.device import get_context
    from cosmonpc_amd.estimators import bk_box_g00, bk_box_signal_auto, bk_box_signal_multi, contraction_grid
    from cosmonpc_amd.mesh import paint_field

    n, L = args.nmesh, 2000.0
    kmin, kmax, K = 0.01, 0.31, args.k_bins
    k_edge = np.linspace(kmin, kmax, K + 1)
    sampler = "pcs"
    angus = [(0, 0, 0), (2, 0, 2)]
    npairs = K * (K + 1) // 2
    vpc = L ** 3 / n ** 3

    t = time.time()
    pos = synth.lognormal_box(args.nparticles, boxsize=L, ngrid=256, seed=2024)
    w = np.ones(len(pos))
    log(f"catalog {len(pos)} particles in {time.time() - t:.1f}s")
    ctx = get_context(0)
    dev = ctx.device
    pos_d = torch.from_numpy(pos).to(dev)
    w_d = torch.from_numpy(w).to(dev)
    NZ = w.sum() / L ** 3

    def delta_of(field):
        return field.delta_k(ctx, 1, scale=1.0 / vpc, sub0=NZ)

    # ---------------------------------------------------------- end to end paint
    # (one untimed pass first: rocFFT builds its 512^3 plans on first use)
    paint_field(ctx, pos_d, w_d, n, L, sampler, True, 1)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    f32 = paint_field(ctx, pos_d, w_d, n, L, sampler, True, 1)
    d32 = delta_of(f32)
    torch.cuda.synchronize()
    t_paint32 = time.perf_counter() - t0
    del f32
    log(f"fp32 paint+r2c+prepare {t_paint32 * 1e3:.1f} ms")

    def enqueue_step(delta, prec, mode, zsynth="fused", shared=False):
        """G_00 + both configurations' signal matrices, on the contraction
        grid the estimator uses (reduced for the GEMM modes): the device work
        of every configuration is enqueued before any result is read; the
        returned function reads them.  ``shared``: one multi-configuration
        call that synthesises the unweighted shells once for both
        configurations (they are [0,0,0]'s legs and [2,0,2]'s second leg)."""
        m, nmax = contraction_grid(n, L, kmax, mode)
        dm = ctx.resample(delta, n, m, prec, nmax) if m != n else delta
        if shared:
            # G_00 made inside, from the resampled spectrum once the shells
            # have read it (no copy of it)
            fin = bk_box_signal_multi(ctx, dm, None, m, L, prec, k_edge, angus, mode, zsynth=zsynth, defer=True,
                                      consume_delta=m != n)
        else:
            G = bk_box_g00(ctx, dm, m, prec)
            fins = [bk_box_signal_auto(ctx, dm, G, m, L, prec, k_edge, a, mode, zsynth=zsynth, defer=True)
                    for a in angus]
            fin = lambda: [f() for f in fins]  # noqa: E731
        return fin, (n / m) ** 3

    def finish_step(pending):
        fin, factor = pending
        return [factor * r for r in fin()]

    def signal_step(delta, prec, mode, zsynth="fused", shared=False):
        return finish_step(enqueue_step(delta, prec, mode, zsynth, shared))

    grid_m = contraction_grid(n, L, kmax, "bf16x3")[0]
    log(f"contraction grid {grid_m}^3 for the GEMM modes")
    bins_per_step = npairs * len(angus)

    def timed_steps(fn):
        for _ in range(args.warmup):
            res = fn()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(args.steps):
            res = fn()
        torch.cuda.synchronize()
        return res, (time.perf_counter() - t0) / args.steps

    def timed_pipelined(enq, fin):
        """K steps with the next step's launches queued before the current
        step's results are read (each result lands in pinned host memory
        right behind its kernels), so the device never waits for the host
        between steps."""
        for _ in range(args.warmup):
            res = fin(enq())
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        prev = enq()
        for _ in range(args.steps - 1):
            cur = enq()
            res = fin(prev)
            prev = cur
        res = fin(prev)
        torch.cuda.synchronize()
        return res, (time.perf_counter() - t0) / args.steps

    def rel_diff(a, b):
        return max(float(np.max(np.abs(x - y)) / np.max(np.abs(y))) for x, y in zip(a, b))

    # each configuration on its own (as one run_stats per configuration does)
    res_indep, t_indep = timed_pipelined(lambda: enqueue_step(d32, 1, "bf16x3"), finish_step)
    log(f"bf16x3 signal stage, configurations apart {t_indep * 1e3:.1f} ms/step -> "
        f"{bins_per_step / t_indep:.1f} bins/s")
    # both configurations in one call, the shared leg synthesised once
    res_eager, t_eager = timed_pipelined(lambda: enqueue_step(d32, 1, "bf16x3", shared=True), finish_step)
    share_diff = rel_diff(res_eager, res_indep)
    log(f"bf16x3 signal stage, shared legs {t_eager * 1e3:.1f} ms/step -> {bins_per_step / t_eager:.1f} bins/s "
        f"(max rel diff vs apart {share_diff:.1e})")
    graph_info = {"configs_apart_ms_per_step": t_indep * 1e3, "configs_apart_bins_per_s": bins_per_step / t_indep,
                  "shared_legs_eager_ms_per_step": t_eager * 1e3, "shared_vs_apart_max_rel_diff": share_diff}
    # both are split-bf16 results (~1e-6 of max|B| from FP64); sharing changes
    # only the FP32 summation order of the contraction
    if share_diff > 1e-5:
        res_eager, t_eager = res_indep, t_indep

    # the step recorded once as a hipGraph and replayed (the input spectrum
    # d32 stays where it is; every replay recomputes G_00, the shells and the
    # contractions from it).  Reported beside the eager step, which is the
    # headline: with ~40 long kernels per step the replay gains nothing.
    res_bf16, t_step = res_eager, t_eager
    if not args.no_graph:
        from cosmonpc_amd.graphs import StepGraph

        sg = StepGraph(ctx)
        try:
            pending = sg.record(lambda: enqueue_step(d32, 1, "bf16x3", shared=share_diff <= 1e-5), finish_step)
        except Exception as e:  # recorded nothing: report the eager step
            graph_info["graph_error"] = f"{type(e).__name__}: {e}"
            log(f"graph recording failed: {graph_info['graph_error']}")
            sg = None
        if sg is not None:
            def replay():
                sg.replay()
                return finish_step(pending)

            res_graph, t_graph = timed_steps(replay)
            diff = rel_diff(res_graph, res_eager)
            graph_info.update({"graph_ms_per_step": t_graph * 1e3, "graph_bins_per_s": bins_per_step / t_graph,
                               "graph_vs_eager_max_rel_diff": diff})
            log(f"bf16x3 signal stage, graph replay {t_graph * 1e3:.1f} ms/step -> "
                f"{bins_per_step / t_graph:.1f} bins/s (max rel diff vs eager {diff:.1e})")
            sg.close()
            del pending
    value = bins_per_step / t_step
    log(f"bf16x3 signal stage {t_step * 1e3:.1f} ms/step -> {value:.1f} bins/s")

    # the same stage with the real shells written to HBM and contracted apart
    signal_step(d32, 1, "bf16x3", "separate")
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    res_sep = signal_step(d32, 1, "bf16x3", "separate")
    torch.cuda.synchronize()
    t_sep = time.perf_counter() - t0
    log(f"bf16x3 separate z pass: {t_sep * 1e3:.1f} ms/step -> {bins_per_step / t_sep:.1f} bins/s")
    # the FP8 split variant (three e4m3 pieces, six products; separate z pass)
    signal_step(d32, 1, "fp8x6")
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    res_fp8 = signal_step(d32, 1, "fp8x6")
    torch.cuda.synchronize()
    t_fp8 = time.perf_counter() - t0
    log(f"fp8x6 (separate z pass): {t_fp8 * 1e3:.1f} ms/step -> {bins_per_step / t_fp8:.1f} bins/s")
    # the FP16 split variant (hi + lo halves, three products; separate z pass)
    signal_step(d32, 1, "fp16x3")
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    res_fp16 = signal_step(d32, 1, "fp16x3")
    torch.cuda.synchronize()
    t_fp16 = time.perf_counter() - t0
    log(f"fp16x3 (separate z pass): {t_fp16 * 1e3:.1f} ms/step -> {bins_per_step / t_fp16:.1f} bins/s")

    # ---------------------------------------------------------- per-kernel roofline
    m, nmax = contraction_grid(n, L, kmax, "bf16x3")
    dm32 = ctx.resample(d32, n, m, 1, nmax) if m != n else d32
    roof = kernel_roofline(ctx, dm32, m, L, k_edge, torch, np, A)
    del dm32
    log(f"kernels (grid {m}^3): {json.dumps(roof['stages'])}")

    extra = {**graph_info, "separate_z_ms_per_step": t_sep * 1e3, "separate_z_bins_per_s": bins_per_step / t_sep,
             "fp8x6_ms_per_step": t_fp8 * 1e3, "fp8x6_bins_per_s": bins_per_step / t_fp8,
             "fp16x3_ms_per_step": t_fp16 * 1e3, "fp16x3_bins_per_s": bins_per_step / t_fp16}
    if not args.no_fp64:
        paint_field(ctx, pos_d, w_d, n, L, sampler, True, 0)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        f64 = paint_field(ctx, pos_d, w_d, n, L, sampler, True, 0)
        d64 = delta_of(f64)
        torch.cuda.synchronize()
        extra["paint_fp64_ms"] = (time.perf_counter() - t0) * 1e3
        del f64
        res = {}
        for mode in ("pairwise", "gemm_fp64"):
            signal_step(d64, 0, mode)  # warm-up: FFT plans and work areas
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            res[mode] = signal_step(d64, 0, mode)
            torch.cuda.synchronize()
            dt = time.perf_counter() - t0
            extra[f"{mode}_ms_per_step"] = dt * 1e3
            extra[f"{mode}_bins_per_s"] = bins_per_step / dt
            log(f"fp64 {mode}: {dt * 1e3:.1f} ms/step -> {bins_per_step / dt:.1f} bins/s")
        ref = res["pairwise"]
        extra["speedup_vs_fp64_plain"] = extra["pairwise_ms_per_step"] / (t_step * 1e3)
        for a, r64, rb, rs, r8, r16, rg in zip(angus, ref, res_bf16, res_sep, res_fp8, res_fp16, res["gemm_fp64"]):
            tag = "".join(map(str, a))
            extra[f"err_bf16x3_B{tag}"] = float(np.max(np.abs(rb - r64)) / np.max(np.abs(r64)))
            extra[f"err_bf16x3_separate_B{tag}"] = float(np.max(np.abs(rs - r64)) / np.max(np.abs(r64)))
            extra[f"err_fp8x6_B{tag}"] = float(np.max(np.abs(r8 - r64)) / np.max(np.abs(r64)))
            extra[f"err_fp16x3_B{tag}"] = float(np.max(np.abs(r16 - r64)) / np.max(np.abs(r64)))
            extra[f"err_gemm_fp64_B{tag}"] = float(np.max(np.abs(rg - r64)) / np.max(np.abs(r64)))
        del d64
    del d32
    extra["end_to_end"] = end_to_end(pos, w, n, L, kmax, K, sampler)
    if mock_cats is not None:
        extra["mocks"] = mock_comparison(args, mock_cats)
        del mock_cats
    cpu = None if args.no_cpu else cpu_baseline(n, L, k_edge)
    line = {
        "metric": "bk_sugi full-mode bins/s @512^3, split-bf16 MFMA vs FP64 plain (<=1e-4 rel)",
        "value": value,
        "unit": "bins/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": t_step * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32 shells, split-bf16 MFMA, fp32 accumulate, fp64 partials",
        "data": "synthetic (seeded lognormal + linear RSD catalog)",
        "config": {"workload": f"bk_sugi box full {n}^3 pcs interlaced, {K} k-bins 0.01-0.31, "
                               f"angu [0,0,0]+[2,0,2], {len(pos)} particles, box 2000",
                   "bins_per_step": bins_per_step, "paint_fp32_ms": t_paint32 * 1e3,
                   "step": ("both configurations in one pass (shared shell legs, two-output contraction)"
                            if graph_info.get("shared_vs_apart_max_rel_diff", 1.0) <= 1e-5
                            else "configurations one at a time")
                           + ", eager launches, the next step queued before the current one is read"},
        "roofline": roof["roofline"],
        "cpu_baseline": cpu,
        "extra": dict(extra, autotune=autotune_choices()),
        "kernels": roof["stages"],
    }
    print(json.dumps(line), flush=True)


def kernel_roofline(ctx, delta, n, L, k_edge, torch, np, A):
    """Times each stage of the signal pass with HIP events on the context
    stream and returns the dominant kernel's roofline."""
    from cosmonpc_amd.estimators import bk_box_g00

    K = len(k_edge) - 1
    N3 = n ** 3
    half = n * n * (n // 2 + 1)
    batch = 4
    ev = lambda: torch.cuda.Event(enable_timing=True)  # noqa: E731
    work = ctx.spectrum(n, 1, batch)
    shells = ctx.mesh(n, 1, K).view(K, N3)
    G = bk_box_g00(ctx, delta, n, 1)
    poly = A.shell_components(0, 0)[0][0]
    kmax = float(k_edge[-1])
    R = ctx.synth_radius(L, kmax)
    H = ctx.half_shells(n, R, K)

    def timed(fn, reps=1):
        e0, e1 = ev(), ev()
        e0.record()
        for _ in range(reps):
            fn()
        e1.record()
        e1.synchronize()
        return e0.elapsed_time(e1) / reps

    # production path: x and y synthesis passes into the half-synthesised
    # layout, then the z pass fused into the split-bf16 GEMM
    t_half, kz = [], np.zeros(K, dtype=np.int32)
    for b0 in range(0, K, batch):
        c = min(batch, K - b0)
        t_half.append((timed(lambda: ctx.shell_synth_half(delta, n, L, k_edge, b0, c, poly, work, H, b0)), c))
        kz[b0:b0 + c] = ctx.shell_synth_half(delta, n, L, k_edge, b0, c, poly, work, H, b0)
    ctx.contract_zgemm(n, L, kmax, G, H, K, kz)
    t_zgemm = timed(lambda: ctx.contract_zgemm(n, L, kmax, G, H, K, kz), 3)
    # the launch of the measured step: [2,0,2] (A = Y_20-weighted shells,
    # B = the unweighted ones) with [0,0,0] = B^T G B from the same launch
    H2 = ctx.half_shells(n, R, K)
    kz2 = np.zeros(K, dtype=np.int32)
    poly2 = A.shell_components(2, 0)[0][0]
    for b0 in range(0, K, batch):
        c = min(batch, K - b0)
        kz2[b0:b0 + c] = ctx.shell_synth_half(delta, n, L, k_edge, b0, c, poly2, work, H2, b0)
    ctx.contract_zgemm(n, L, kmax, G, H2, K, kz2, H, K, kz, with_bb=True)
    t_bb = timed(lambda: ctx.contract_zgemm(n, L, kmax, G, H2, K, kz2, H, K, kz, with_bb=True), 3)
    del H2
    # the separate path: real shells (cube, x, y and z FFT kernels) + GEMM
    t_synth = []
    for b0 in range(0, K, batch):
        c = min(batch, K - b0)
        dst = shells[b0:b0 + c].view(c, n, n, n)
        t_synth.append((timed(lambda: ctx.shell_synth(delta, n, L, 1, k_edge, b0, c, poly, work, dst)), c))
    # the full 3-D c2r path it replaces, for comparison
    t_filter, t_c2r = [], []
    for b0 in range(0, K, batch):
        c = min(batch, K - b0)
        dst = shells[b0:b0 + c].view(c, n, n, n)
        t_filter.append((timed(lambda: ctx.shell_filter(delta, n, L, 1, k_edge, b0, c, poly, work)), c))
        t_c2r.append((timed(lambda: ctx.c2r(work[:c], dst, n, 1, c)), c))
    ctx.contract_gemm("bf16x3", G, shells, None)
    t_gemm = timed(lambda: ctx.contract_gemm("bf16x3", G, shells, None), 5)

    def per_launch(ts):
        return float(np.mean([x[0] for x in ts if x[1] == batch]))

    # algorithmic bytes per launch: 4 real FP32 shells written (synthesis);
    # spectra read + meshes written (c2r); delta read + spectra written
    # (filter); every shell and G read once (GEMM)
    b_synth = batch * N3 * 4
    # written / read kz rows only (each shell holds kz < its own radius)
    b_half = int(np.mean([kz[b0:b0 + batch].sum() for b0 in range(0, K - batch + 1, batch)])) * n * n * 8
    b_zgemm = int(kz.sum()) * n * n * 8 + N3 * 4
    b_bb = int(kz.sum() + kz2.sum()) * n * n * 8 + N3 * 4
    b_c2r = batch * (half * 8 + N3 * 4)
    b_filter = half * 8 + batch * half * 8
    b_gemm = (K + 1) * N3 * 4
    stages = {
        "shell_synth_half_batch4": {"ms_per_launch": per_launch(t_half), "launches": len(t_half),
                                    "total_ms": sum(x[0] for x in t_half), "bytes": b_half},
        "zgemm_bf16x3": {"ms_per_launch": t_zgemm, "launches": 1, "total_ms": t_zgemm, "bytes": b_zgemm},
        "zgemm_bf16x3_bb": {"ms_per_launch": t_bb, "launches": 1, "total_ms": t_bb, "bytes": b_bb},
        "shell_synth_batch4": {"ms_per_launch": per_launch(t_synth), "launches": len(t_synth),
                               "total_ms": sum(x[0] for x in t_synth), "bytes": b_synth},
        "gemm_x3_bf16": {"ms_per_launch": t_gemm, "launches": 1, "total_ms": t_gemm, "bytes": b_gemm},
        "ref_shell_filter": {"ms_per_launch": per_launch(t_filter), "launches": len(t_filter),
                             "total_ms": sum(x[0] for x in t_filter), "bytes": b_filter},
        "ref_c2r3d_batch4": {"ms_per_launch": per_launch(t_c2r), "launches": len(t_c2r),
                             "total_ms": sum(x[0] for x in t_c2r), "bytes": b_c2r},
    }
    for v in stages.values():
        v["GBps"] = v["bytes"] / v["ms_per_launch"] / 1e6
    # the measured step runs 15 synthesis launches (both weights) and one
    # two-output contraction.  Algorithmic FLOPs of a contraction: 2 K^2 n^3
    # per output matrix (the G and Ylm products are counted as bytes only)
    stages["zgemm_bf16x3"]["flops"] = 2.0 * K * K * N3
    stages["zgemm_bf16x3_bb"]["flops"] = 4.0 * K * K * N3
    stages["gemm_x3_bf16"]["flops"] = 2.0 * K * K * N3
    dom = max(("shell_synth_half_batch4", "zgemm_bf16x3_bb"), key=lambda k: stages[k]["total_ms"])
    st = stages[dom]
    ach = st["GBps"]
    # the bounding roof: the larger of the two lower bounds on the launch
    # time (bytes at the HBM peak, FLOPs at the dense bf16 MFMA peak)
    t_hbm = st["bytes"] / (HBM_PEAK_GBS * 1e9)
    t_mfma = st.get("flops", 0.0) / (MFMA_BF16_PEAK_TFLOPS * 1e12)
    tflops = st.get("flops", 0.0) / (st["ms_per_launch"] * 1e-3) / 1e12
    if t_mfma > t_hbm:
        roof = {"bound": "mfma", "achieved": tflops, "peak": MFMA_BF16_PEAK_TFLOPS, "unit": "TFLOP/s",
                "frac": tflops / MFMA_BF16_PEAK_TFLOPS}
    else:
        roof = {"bound": "hbm", "achieved": ach, "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": ach / HBM_PEAK_GBS}
    roof.update({"traffic": measured_traffic(dom), "kernel": dom, "algorithmic_bytes_per_launch": st["bytes"],
                 "traffic_source": "rocprofv3 FETCH_SIZE + WRITE_SIZE of this kernel (profiles/"
                                   "kernels_512_hbm_traffic_pmc.txt), not measured in this run"})
    if "flops" in st:
        roof["mfma"] = {"algorithmic_flops_per_launch": st["flops"], "achieved_tflops": tflops,
                        "peak_tflops_bf16_dense": MFMA_BF16_PEAK_TFLOPS, "frac": tflops / MFMA_BF16_PEAK_TFLOPS}
    util = measured_utilisation(dom)
    if util:
        roof["measured_utilisation"] = dict(util, source="rocprofv3 PMC (profiles/kernels_512_pmc_utilisation.txt)")
        # what the counters say limits the kernel: the busiest of its pipes
        busy = {"valu": util.get("valu_busy_pct", 0.0), "lds": util.get("lds_active_pct", 0.0),
                "mfma": util.get("mfma_busy_pct", 0.0)}
        pipe = max(busy, key=busy.get)
        roof["limiter"] = (f"{pipe}: {busy[pipe]:.0f} % busy (" +
                           ", ".join(f"{k} {v:.0f} %" for k, v in busy.items() if k != pipe) +
                           f" busy; the bound's roof at {100 * roof['frac']:.0f} %)")
        if dom == "zgemm_bf16x3_bb":
            roof["limiter"] += "; the in-LDS z FFT and the bf16 hi/lo splits between barriers"
    del work, shells, G, H
    return {"roofline": roof, "stages": stages}


def end_to_end(pos, w, n, L, kmax, K, sampler):
    """The workload end to end as a user runs it: one run_stats per angular
    configuration ([0,0,0], then [2,0,2]) from the catalog in host memory
    to the saved result (paint, FFTs, shells, contraction, shot noise,
    normalisation, file), FP32 bulk with the split-bf16 contraction.  Each
    configuration runs once untimed first (plans, work areas)."""
    import tempfile

    import torch

    from cosmonpc_amd import run_stats

    out = {}
    with tempfile.TemporaryDirectory() as tmp:
        base = dict(statistic="bk_sugi", correlation_mode="auto", geometry="box-like",
                    catalogs={"data_a": "cat.npy", "randoms_a": None, "data_b": None, "randoms_b": None},
                    column_names=["x", "y", "z", "w"], rsd=[0, 0, 1], apply_rsd=False, nmesh=[n] * 3,
                    boxsize=[L] * 3, sampler=sampler, interlaced=True, k_min=0.01, k_max=kmax, k_bins=K,
                    tracer_type="aaa", data_vector_mode="full", shotnoise_mode="ana", compensation=True,
                    cosmology={"h": 0.6766, "Omega0": 0.30966}, redshift_box=0.5, output_dir=tmp,
                    precision="fp32", arrays={"data_a": (pos, w)})
        for angu in ([0, 0, 0], [2, 0, 2]):
            cfg = dict(base, angu_config=angu)
            run_stats(dict(cfg))
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            run_stats(dict(cfg))
            torch.cuda.synchronize()
            out["run_stats_" + "".join(map(str, angu)) + "_s"] = time.perf_counter() - t0
    out["note"] = ("one run_stats per configuration, catalog in host memory -> saved .npy result, fp32 bulk + "
                   "split-bf16 contraction (the reference's way: no legs shared between configurations)")
    log(f"end to end: {out}")
    return out


def autotune_choices():
    from cosmonpc_amd import autotune

    return {k: v for k, v in autotune.choices().items()}


def measured_traffic(kernel):
    """HBM bytes per launch of ``kernel`` as measured by the rocprofv3
    FETCH_SIZE / WRITE_SIZE runs summarised in profiles/ (None if absent)."""
    path = os.path.join(ROOT, "profiles", "kernels_512_hbm_traffic_pmc.txt")
    if not os.path.exists(path):
        return None
    for line in open(path):
        parts = line.split()
        if len(parts) == 3 and parts[0] == "traffic_bytes_per_launch" and parts[1] == kernel:
            return float(parts[2])
    return None


def measured_utilisation(stage):
    """Busy fractions of ``stage``'s kernel from the rocprofv3 PMC runs
    summarised in profiles/ (None if absent)."""
    path = os.path.join(ROOT, "profiles", "kernels_512_pmc_utilisation.txt")
    if not os.path.exists(path):
        return None
    for line in open(path):
        parts = line.split()
        if len(parts) >= 4 and parts[0] == "utilisation" and parts[1] == stage:
            return {parts[i]: float(parts[i + 1]) for i in range(2, len(parts) - 1, 2)}
    return None


def host_info():
    model = platform.processor() or platform.machine()
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    return {"cpu_model": model, "os_cpu_count": os.cpu_count(), "affinity_cpus": len(os.sched_getaffinity(0))}


def cpu_baseline(n, L, k_edge):
    """The FP64 numpy oracle timed on one full angular configuration of the
    workload, [0,0,0]: the K shell inverse FFTs (mask and c2c IFFT on the
    full complex128 grid, real parts kept in host memory: K n^3 doubles, the
    whole block at once), the G_00 transform, and the K (K + 1) / 2
    triple-product sums sum_x G S_i S_j.  Threads: the host's CPUs available
    to this process, at most 16 (the CPU share of one GPU): the elementwise
    work and the sums in x-slabs on a thread pool, the 3-D IFFTs with
    scipy's multi-threaded pocketfft.  The timed operations do not depend on
    the field values, so the input spectrum is seeded noise."""
    from concurrent.futures import ThreadPoolExecutor

    import numpy as np
    import scipy.fft

    from oracle.cnpc_ref import angular as OA
    from oracle.cnpc_ref import conventions as C

    cores = max(1, min(16, len(os.sched_getaffinity(0))))
    K = len(k_edge) - 1
    rng = np.random.default_rng(0)
    d = rng.standard_normal((n, n, n)) + 1j * rng.standard_normal((n, n, n))
    kx, ky, kz = C.k_axes(n, L)
    step = max(1, -(-n // cores))
    slabs = [(i, min(n, i + step)) for i in range(0, n, step)]
    kn = np.empty((n, n, n))

    def knorm(sl):
        kn[sl[0]:sl[1]] = OA.get_kgrid((kx[sl[0]:sl[1]], ky, kz))[1]

    S = np.empty((K, n, n, n))
    w = np.empty_like(d)
    with ThreadPoolExecutor(cores) as ex:
        list(ex.map(knorm, slabs))
        t0 = time.perf_counter()
        for i in range(K):
            def shell(sl, i=i):
                a, b = sl
                w[a:b] = d[a:b] * np.logical_and(kn[a:b] >= k_edge[i], kn[a:b] < k_edge[i + 1])
            list(ex.map(shell, slabs))
            F = scipy.fft.ifftn(w, workers=cores, overwrite_x=False)

            def keep(sl, i=i):
                S[i, sl[0]:sl[1]] = F[sl[0]:sl[1]].real
            list(ex.map(keep, slabs))
            del F
            if i == 0 or (i + 1) % 10 == 0:
                log(f"cpu oracle: {i + 1}/{K} shell IFFTs, {time.perf_counter() - t0:.1f}s")
        F = scipy.fft.ifftn(d, workers=cores)
        G = np.empty((n, n, n))

        def keep_g(sl):
            G[sl[0]:sl[1]] = F[sl[0]:sl[1]].real
        list(ex.map(keep_g, slabs))
        del F
        t_shells = time.perf_counter() - t0
        del w, d, kn
        t0 = time.perf_counter()
        B = np.zeros((K, K))
        GS = np.empty((n, n, n))
        for i in range(K):
            def gs(sl, i=i):
                np.multiply(G[sl[0]:sl[1]], S[i, sl[0]:sl[1]], out=GS[sl[0]:sl[1]])
            list(ex.map(gs, slabs))
            for j in range(i, K):
                B[i, j] = sum(ex.map(lambda sl, j=j: np.dot(GS[sl[0]:sl[1]].ravel(), S[j, sl[0]:sl[1]].ravel()),
                                     slabs))
            if (i + 1) % 10 == 0:
                log(f"cpu oracle: triple sums of {i + 1}/{K} rows, {time.perf_counter() - t0:.1f}s")
        t_pairs = time.perf_counter() - t0
    del S, G, GS
    npairs = K * (K + 1) // 2
    t_cfg = t_shells + t_pairs
    bins_per_s = npairs / t_cfg
    log(f"cpu oracle ({cores} threads): {K} shell IFFTs + G {t_shells:.1f}s, {npairs} triple sums {t_pairs:.1f}s "
        f"-> {bins_per_s:.3f} bins/s (check {abs(B[0, 0]):.3e})")
    return {"value": bins_per_s, "unit": "bins/s", "cores": cores, "kind": "port",
            "sample": f"one full angular configuration ([0,0,0]) of the {n}^3 workload timed end to end with the "
                      f"complex128 numpy oracle on {cores} threads (x-slab thread pool, scipy.fft workers): "
                      f"{K} shell IFFTs + the G_00 transform in {t_shells:.1f}s, {npairs} triple sums in "
                      f"{t_pairs:.1f}s; [2,0,2] would add {K} Ylm-weighted shell IFFTs",
            "host": host_info()}


def bench_kernels(args):
    """One launch of each hot kernel of the 512^3 workload, timed with HIP
    events on the context stream, and a plain 1 GiB device-to-device copy of
    known size (the reference point for hardware byte counters).  Prints one
    JSON line of per-kernel times and algorithmic bytes."""
    import numpy as np
    import torch

    from cosmonpc_amd import angular as A
    from cosmonpc_amd import synth
    from cosmonpc_amd.device import get_context
    from cosmonpc_amd.estimators import bk_box_g00, contraction_grid

    n, L, kmax, K = args.nmesh, 2000.0, 0.31, args.k_bins
    k_edge = np.linspace(0.01, kmax, K + 1)
    pos = synth.lognormal_box(args.nparticles, boxsize=L, ngrid=256, seed=2024)
    ctx = get_context(0)
    dev = ctx.device
    pos_d = torch.from_numpy(pos).to(dev)
    w_d = torch.ones(len(pos), dtype=torch.float64, device=dev)
    half = n * n * (n // 2 + 1)

    def timed(fn):
        fn()  # first call: plans and work areas
        torch.cuda.synchronize()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        fn()
        e1.record()
        e1.synchronize()
        return e0.elapsed_time(e1)

    out = {}
    m0, m1 = ctx.mesh(n, 1), ctx.mesh(n, 1)
    out["paint_pcs_interlaced"] = (timed(lambda: ctx.paint(pos_d, w_d, n, L, "pcs", True, 1, m0, m1)),
                                   len(pos) * 32 + 2 * n ** 3 * 4)
    s1, s2 = ctx.spectrum(n, 1), ctx.spectrum(n, 1)
    out["rocfft_r2c_512"] = (timed(lambda: ctx.r2c(m0, s1, n, 1)), n ** 3 * 4 + half * 8)
    ctx.r2c(m1, s2, n, 1)
    d = ctx.spectrum(n, 1)
    out["prepare"] = (timed(lambda: ctx.prepare(s1, s2, d, n, L, 1, True, 1, "pcs", 1.0 / n ** 3)), 3 * half * 8)
    polys = [A.make_poly(0, [1.0], A.PART_FULL, 1.0, 1.0)] + [A.legendre_poly(p, np.array([0.0, 0.0, 1.0]))
                                                              for p in (2, 4)]
    out["bin_pk_3poles"] = (timed(lambda: ctx.bin(d, d, n, L, 1, np.linspace(0.005, 0.305, 31), polys)), half * 8)
    m, nmax = contraction_grid(n, L, kmax, "bf16x3")
    dm = ctx.resample(d, n, m, 1, nmax)
    G = bk_box_g00(ctx, dm, m, 1)
    # G_00's c2r on the contraction grid (the fixed 20 x 20 plan at 400^3):
    # the spectrum read and the mesh written once
    tmp, Gt = dm.clone(), ctx.mesh(m, 1)
    out["c2r_g00_%d" % m] = (timed(lambda: ctx.c2r(tmp, Gt, m, 1)), m * m * (m // 2 + 1) * 8 + m ** 3 * 4)
    del tmp, Gt
    work = ctx.spectrum(m, 1, 4)
    shells = ctx.mesh(m, 1, K).view(K, m ** 3)
    poly = A.shell_components(0, 0)[0][0]
    out["shell_synth_batch4"] = (timed(lambda: ctx.shell_synth(dm, m, L, 1, k_edge, 8, 4, poly, work,
                                                               shells[8:12].view(4, m, m, m))), 4 * m ** 3 * 4)
    for b0 in range(0, K, 4):
        c = min(4, K - b0)
        ctx.shell_synth(dm, m, L, 1, k_edge, b0, c, poly, work, shells[b0:b0 + c].view(c, m, m, m))
    out["gemm_x3_bf16"] = (timed(lambda: ctx.contract_gemm("bf16x3", G, shells, None)), (K + 1) * m ** 3 * 4)
    del shells
    R = ctx.synth_radius(L, kmax)
    H = ctx.half_shells(m, R, K)
    kz = np.zeros(K, dtype=np.int32)
    for b0 in range(0, K, 4):
        kz[b0:b0 + 4] = ctx.shell_synth_half(dm, m, L, k_edge, b0, min(4, K - b0), poly, work, H, b0)
    out["shell_synth_half_batch4"] = (timed(lambda: ctx.shell_synth_half(dm, m, L, k_edge, 8, 4, poly, work, H, 8)),
                                      int(kz[8:12].sum()) * m * m * 8)
    out["zgemm_bf16x3"] = (timed(lambda: ctx.contract_zgemm(m, L, kmax, G, H, K, kz)),
                           int(kz.sum()) * m * m * 8 + m ** 3 * 4)
    # [2,0,2] beside [0,0,0]: A = Y_20-weighted shells, B = the unweighted
    # ones, and B^T G B ([0,0,0]) from the same launch
    H2 = ctx.half_shells(m, R, K)
    kz2 = np.zeros(K, dtype=np.int32)
    poly2 = A.shell_components(2, 0)[0][0]
    for b0 in range(0, K, 4):
        kz2[b0:b0 + 4] = ctx.shell_synth_half(dm, m, L, k_edge, b0, min(4, K - b0), poly2, work, H2, b0)
    out["zgemm_bf16x3_bb"] = (timed(lambda: ctx.contract_zgemm(m, L, kmax, G, H2, K, kz2, H, K, kz, with_bb=True)),
                              int(kz.sum() + kz2.sum()) * m * m * 8 + m ** 3 * 4)
    del H, H2
    src = torch.empty(1 << 28, dtype=torch.float32, device=dev)
    dst = torch.empty_like(src)
    src.fill_(1.0)
    out["copy_1GiB"] = (timed(lambda: dst.copy_(src)), 2 * (1 << 30))
    torch.cuda.synchronize()
    line = {k: {"ms": t, "bytes": b, "GBps": b / t / 1e6} for k, (t, b) in out.items()}
    for k, v in line.items():
        log(f"{k}: {v['ms']:.3f} ms, {v['bytes'] / 1e9:.3f} GB algorithmic -> {v['GBps']:.0f} GB/s")
    print(json.dumps({"workload": f"kernels of the bk_sugi {n}^3 pcs workload (contraction grid {m}^3)",
                      "kernels": line}), flush=True)


def _mock_catalog(i):
    """Mock i: its own seed, N ~ U[5e5, 2e6], lognormal + RSD on a 256^3
    grid in a 1000 Mpc/h box."""
    import numpy as np

    from cosmonpc_amd import synth

    npart = int(np.random.default_rng(1000 + i).integers(500_000, 2_000_001))
    return synth.lognormal_box(npart, boxsize=1000.0, ngrid=256, seed=5000 + i)


def mock_catalogs(nmocks):
    """The mock catalogs, drawn in worker processes (call before the GPU is
    touched: the workers are forked)."""
    from concurrent.futures import ProcessPoolExecutor

    t = time.time()
    with ProcessPoolExecutor(max_workers=min(16, os.cpu_count() or 1)) as ex:
        cats = list(ex.map(_mock_catalog, range(nmocks)))
    log(f"{nmocks} mock catalogs ({sum(len(c) for c in cats)} particles) in {time.time() - t:.1f}s")
    return cats


def mock_comparison(args, cats):
    """Batched mocks: the box catalogs ``cats``, pk poles [0,2,4] and bk_sugi
    [0,0,0] diagonal (19 bins 0.01-0.20) at 256^3 CIC interlaced, FP64.  One
    step = every mock through both tasks with the batched driver
    (``run_mocks``: one paint launch and one batched r2c per chunk); the
    reference's way, one ``run_stats`` per mock and task, is timed beside it
    on ``args.loop_mocks`` mocks.  Outputs go to a temporary directory."""
    import shutil
    import tempfile

    import numpy as np
    import torch

    from cosmonpc_amd import run_stats
    from cosmonpc_amd.mocks import LAST_TIMINGS, run_mocks

    nm = len(cats)
    tmp = tempfile.mkdtemp(prefix="cnpc_mocks_")
    paths = [os.path.join(tmp, "mocks", f"mock_{i:03d}.npy") for i in range(nm)]
    arrays = {p: (c, np.ones(len(c))) for p, c in zip(paths, cats)}
    base = dict(correlation_mode="auto", geometry="box-like", column_names=["x", "y", "z"], rsd=[0, 0, 1],
                apply_rsd=False, nmesh=[256] * 3, boxsize=[1000.0] * 3, sampler="cic", interlaced=True,
                compensation=True, cosmology={"h": 0.6766, "Omega0": 0.30966}, redshift_box=0.5)
    pk = dict(base, statistic="pk", k_min=0.005, k_max=0.305, k_bins=30, poles=[0, 2, 4])
    bk = dict(base, statistic="bk_sugi", k_min=0.01, k_max=0.20, k_bins=19, angu_config=[0, 0, 0],
              tracer_type="aaa", data_vector_mode="diagonal", shotnoise_mode="ana")

    def batched(out, sel):
        cfgs = [dict(c, output_dir=out) for c in (pk, bk)]
        return run_mocks(cfgs, sel, batch=args.mock_batch, arrays=arrays)

    def loop(out, sel):
        for p in sel:
            for c in (pk, bk):
                run_stats(dict(c, output_dir=out, catalogs={"data_a": p}, arrays={"data_a": arrays[p]}))

    import logging

    logging.disable(logging.INFO)
    try:
        batched(os.path.join(tmp, "warm_b"), paths[:2])  # FFT plans, work areas
        loop(os.path.join(tmp, "warm_l"), paths[:1])
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        batched(os.path.join(tmp, "batched"), paths)
        torch.cuda.synchronize()
        t_b = time.perf_counter() - t0
        timings = dict(LAST_TIMINGS)
        nloop = min(args.loop_mocks, nm)
        t0 = time.perf_counter()
        loop(os.path.join(tmp, "loop"), paths[:nloop])
        torch.cuda.synchronize()
        t_l = time.perf_counter() - t0
        worst = 0.0
        for p in paths[:nloop]:
            for kind in ("pk_res_auto_{}.npy", "bk_sugi_res_diagonal_000_auto_aaa_mocks_{}.npy"):
                name = kind.format(os.path.splitext(os.path.basename(p))[0])
                a = np.load(os.path.join(tmp, "batched", name), allow_pickle=True).item()
                b = np.load(os.path.join(tmp, "loop", name), allow_pickle=True).item()
                for key in ("P0", "P2", "P4", "B_sugi"):
                    if key in b:
                        ref = np.asarray(b[key])
                        worst = max(worst, float(np.max(np.abs(np.asarray(a[key]) - ref)) / np.max(np.abs(ref))))
    finally:
        logging.disable(logging.NOTSET)
        shutil.rmtree(tmp, ignore_errors=True)
    value = nm / t_b
    loop_rate = nloop / t_l
    log(f"mocks batched: {t_b:.2f}s for {nm} mocks -> {value:.2f} mocks/s ({timings}); "
        f"per-mock loop: {t_l:.2f}s for {nloop} -> {loop_rate:.2f} mocks/s; max rel diff {worst:.2e}")
    return {"value": value, "unit": "mocks/s", "mocks": nm, "seconds": t_b, "mock_batch": args.mock_batch,
            "workload": f"{nm} box mocks 256^3 cic interlaced, box 1000, N ~ U[5e5, 2e6], pk [0,2,4] 30 bins + "
                        "bk_sugi [0,0,0] diagonal 19 bins, FP64",
            "per_mock_loop": {"value": loop_rate, "unit": "mocks/s", "mocks": nloop},
            "speedup_vs_loop": value / loop_rate, "max_rel_diff_vs_loop": worst, "stages_s": timings}


def bench_mocks(args):
    """The batched-mocks comparison on its own, as one JSON line."""
    cats = mock_catalogs(args.mocks)
    r = mock_comparison(args, cats)
    line = {
        "metric": "batched box mocks/s (pk [0,2,4] + bk_sugi [0,0,0] diagonal per mock)",
        "value": r["value"], "unit": "mocks/s", "n_gpus": 1, "steps": 1, "warmup": 1,
        "ms_per_step": r["seconds"] * 1e3, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f64", "data": "synthetic (seeded lognormal + linear RSD catalogs, one seed per mock)",
        "config": {"workload": r["workload"], "mock_batch": args.mock_batch},
        "per_mock_loop": r["per_mock_loop"], "speedup_vs_loop": r["speedup_vs_loop"],
        "max_rel_diff_vs_loop": r["max_rel_diff_vs_loop"], "stages_s": r["stages_s"],
    }
    print(json.dumps(line), flush=True)


def survey_catalogs(ndata, nrandoms, seed, omega_m):
    """Survey-shaped catalogs of the 512^3 survey workload, drawn on the GPU
    (the numpy generator of cosmonpc_amd.synth takes minutes at this size):
    RA 150-250 deg, DEC 0-60 deg, 0.4 < z < 0.7 with Gaussian n(z) (mean
    0.55, sigma 0.08, peak 3e-4) times dV/dz; data thinned by a lognormal
    density (128^3, P(k) = 2e4 (k/0.1)^-1.5), randoms uniform; WEIGHT_FKP =
    1 / (1 + 1e4 NZ).  Returns host structured arrays (RA, DEC, Z,
    WEIGHT_FKP, NZ)."""
    import numpy as np
    import torch

    from cosmonpc_amd import synth

    dev = torch.device("cuda", 0)
    g = torch.Generator(device=dev)
    g.manual_seed(seed)
    ra, dec, zr = (150.0, 250.0), (0.0, 60.0), (0.4, 0.7)

    def nz(z):
        return 3.0e-4 * np.exp(-0.5 * ((z - 0.55) / 0.08) ** 2)

    zt, chi = synth._chi_table(omega_m, zr[1] * 1.05)
    m = (zt >= zr[0]) & (zt <= zr[1])
    zg = zt[m]
    pdf = nz(zg) * chi[m] ** 2 * np.gradient(chi, zt)[m]
    cdf = np.concatenate([[0.0], np.cumsum(0.5 * (pdf[1:] + pdf[:-1]) * np.diff(zg))])
    cdf /= cdf[-1]
    cdf_t, zg_t = torch.from_numpy(cdf).to(dev), torch.from_numpy(zg).to(dev)
    zt_t, chi_t = torch.from_numpy(zt).to(dev), torch.from_numpy(chi).to(dev)

    def interp(x, xp, fp):
        i = torch.clamp(torch.searchsorted(xp, x), 1, len(xp) - 1)
        t = (x - xp[i - 1]) / (xp[i] - xp[i - 1])
        return fp[i - 1] + t * (fp[i] - fp[i - 1])

    def draw(count):
        u = torch.rand(count, generator=g, device=dev, dtype=torch.float64)
        z = interp(u, cdf_t, zg_t)
        r = ra[0] + (ra[1] - ra[0]) * torch.rand(count, generator=g, device=dev, dtype=torch.float64)
        s0, s1 = np.sin(np.deg2rad(dec[0])), np.sin(np.deg2rad(dec[1]))
        d = torch.rad2deg(torch.asin(s0 + (s1 - s0) * torch.rand(count, generator=g, device=dev,
                                                                   dtype=torch.float64)))
        c = interp(z, zt_t, chi_t)
        rr, dd = torch.deg2rad(r), torch.deg2rad(d)
        pos = torch.stack([c * torch.cos(dd) * torch.cos(rr), c * torch.cos(dd) * torch.sin(rr), c * torch.sin(dd)], 1)
        return r, d, z, pos

    def rows(r, d, z):
        out = np.zeros(len(z), dtype=[("RA", "f8"), ("DEC", "f8"), ("Z", "f8"), ("WEIGHT_FKP", "f8"), ("NZ", "f8")])
        out["RA"], out["DEC"], out["Z"] = r.cpu().numpy(), d.cpu().numpy(), z.cpu().numpy()
        out["NZ"] = nz(out["Z"])
        out["WEIGHT_FKP"] = 1.0 / (1.0 + out["NZ"] * 1.0e4)
        return out

    r, d, z, pos = draw(int(nrandoms))
    randoms = rows(r, d, z)
    lo, hi = pos.min(0).values, pos.max(0).values
    L = float((hi - lo).max()) * 1.02
    center = 0.5 * (lo + hi)
    del r, d, z, pos
    n = 128
    _, rho = synth._lognormal_density(n, L, np.random.default_rng(seed), 2.0e4, -1.5)
    rho_t = torch.from_numpy(rho).to(dev)
    cmax = float(rho.max())
    parts, need = [], int(ndata)
    while need > 0:
        cnt = max(1024, int(need * cmax * 1.2))
        r, d, z, pos = draw(cnt)
        idx = torch.clamp(((pos - center) / L + 0.5) * n, 0, n - 1).long()
        acc = torch.rand(cnt, generator=g, device=dev, dtype=torch.float64) < rho_t[idx[:, 0], idx[:, 1], idx[:, 2]] / cmax
        sel = torch.nonzero(acc).flatten()[:need]
        parts.append((r[sel], d[sel], z[sel]))
        need -= len(sel)
    data = rows(*(torch.cat([p[i] for p in parts]) for i in range(3)))
    return data, randoms


def bench_survey(args):
    """The survey workload end to end through run_stats (catalog rows in
    host memory -> RA/DEC/Z to Cartesian on the GPU -> paint -> FFTs ->
    estimators -> saved result): pk poles 0/2/4 and bk_sugi [2,0,2]
    diagonal, 512^3 TSC interlaced, box 3000, 2e6 data + 2e7 randoms, FP32
    bulk (alpha, I22, I33 and shot-noise sums in FP64) and FP64.  Each
    configuration runs once untimed (plans, work areas), then timed."""
    import tempfile

    import numpy as np
    import torch

    from cosmonpc_amd import run_stats

    cosmo = {"h": 0.6766, "Omega0": 0.30966}
    t0 = time.perf_counter()
    data, randoms = survey_catalogs(2_000_000, 20_000_000, 2025, cosmo["Omega0"])
    log(f"survey catalogs: {len(data)} data, {len(randoms)} randoms in {time.perf_counter() - t0:.1f}s")
    n, L = 512, 3000.0
    out = {}
    with tempfile.TemporaryDirectory() as tmp:
        base = dict(correlation_mode="auto", geometry="survey-like",
                    catalogs={"data_a": "gal.fits", "randoms_a": "ran.fits"},
                    column_names=["RA", "DEC", "Z", "WEIGHT_FKP", "NZ"], nmesh=[n] * 3, boxsize=[L] * 3,
                    sampler="tsc", interlaced=True, k_min=0.01, k_max=0.20, k_bins=19, compensation=True,
                    normalization_scheme="particle", cosmology=dict(cosmo), z_range=None,
                    comp_weight_plan={"scheme": None, "name_alias": None}, output_dir=tmp,
                    arrays={"data_a": data, "randoms_a": randoms})
        cfgs = {"pk_poles_024": dict(base, statistic="pk", poles=[0, 2, 4]),
                "bk_sugi_202_diagonal": dict(base, statistic="bk_sugi", tracer_type="aaa", angu_config=[2, 0, 2],
                                             data_vector_mode="diagonal", shotnoise_mode="ana")}
        for prec in ("fp32", "fp64"):
            for name, cfg in cfgs.items():
                c = dict(cfg, precision=prec)
                run_stats(dict(c))
                torch.cuda.synchronize()
                t0 = time.perf_counter()
                res = run_stats(dict(c))
                torch.cuda.synchronize()
                dt = time.perf_counter() - t0
                out[f"{name}_{prec}_s"] = dt
                log(f"survey {name} {prec}: {dt:.3f} s end to end")
                if name.startswith("bk"):
                    out[f"{name}_{prec}_B0"] = complex(np.asarray(res["B_sugi"]).ravel()[3])
        b32, b64 = out["bk_sugi_202_diagonal_fp32_B0"], out["bk_sugi_202_diagonal_fp64_B0"]
        out["bk_fp32_vs_fp64_rel_bin3"] = abs(b32 - b64) / abs(b64)
        for k in ("bk_sugi_202_diagonal_fp32_B0", "bk_sugi_202_diagonal_fp64_B0"):
            out[k] = [out[k].real, out[k].imag]
    total = out["pk_poles_024_fp32_s"] + out["bk_sugi_202_diagonal_fp32_s"]
    line = {
        "metric": "survey pk [0,2,4] + bk_sugi [2,0,2] diagonal, end to end through run_stats (s per catalog pair)",
        "value": total, "unit": "s", "n_gpus": 1, "steps": 1, "warmup": 1, "ms_per_step": total * 1e3,
        "higher_is_better": False, "scaling": "weak", "vs_baseline": None,
        "dtype": "fp32 meshes / FFTs / Ylm, fp64 catalog sums", "data": "synthetic (seeded lognormal survey, GPU-drawn)",
        "config": {"workload": f"survey {n}^3 tsc interlaced, box {L:g}, 2e6 data + 2e7 randoms, RA 150-250, "
                               "DEC 0-60, 0.4 < z < 0.7, 19 k-bins 0.01-0.20"},
        "extra": out,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
