#!/usr/bin/env python3
"""Benchmark of the B-spline MSE score+gradient (threading=hip) on one MI355X.

One step = one score+gradient evaluation of the 256x256x128 MSE workload
(1 x 1 x 2.5 mm, 10-voxel control grid, nearest-voxel moving gradient as in
flavor g), with the images, moving gradient, coefficient vectors and
outputs resident in HBM.  The coefficient vector cycles over 20 seeded
N(0, 3 mm) draws.  Prints one JSON line.

The dominant kernel (the fused tile kernel) is timed separately with HIP
events on its launch stream; its roofline uses the algorithmic bytes of the
flavor-g path: fixed 4 B + moving 4 B + moving gradient 12 B per voxel, plus
the coefficient vector.

The same line carries the batched multi-atlas comparison ("batched"): 16
pairs with edges drawn from 96..256 per axis (10-voxel grid, N(0, 3 mm)
fields), images and moving gradient stored as BF16, evaluated (a) as one
batched evaluation replayed from a hipGraph and (b) by looping the
single-pair entry point over the 16 pairs on the caller's stream (and, as
a second figure, on each context's own stream, where the pairs' launches
may overlap); both report aggregate pair-evaluations per second.  The "adjoint" object covers the
differentiable path (torch.autograd.Function) on 192x192x96 with an 8-voxel
grid: forward+backward time and memory of its two modes and a central
finite-difference check of the adjoint.  "warp_regularize" times the
warped moving image and the analytic regularizer on the headline workload,
with their rooflines and the oracle's versions timed beside them.
"optimizer_loop" runs the device L-BFGS (liblbfgs's algorithm) for 20
iterations on the headline workload, with and without regularization,
both loops' cost per evaluation (host-driven and device-resident), and the
reference's liblbfgs on the oracle's score for 3 iterations beside it.
"stage_preprocess" times a stage's preprocessing on the device (both
images subsampled, moving gradient, histogram ranges) against the oracle,
and the V-optimal histograms of both stage images.
"mse_128cube" evaluates the 128^3 pair (15-voxel grid) at zero and at the
true coefficients, single-pair HIP beside the oracle's flavor g.
"mattes_mi" evaluates Mattes MI on the 512x512x160 pair (12-voxel grid,
32x32 bins, fixed-point PV histogram) with its roofline, the oracle's
flavor g timed beside it and the histograms compared bin for bin.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

WORKLOAD = "mse_256x256x128_vpr10"
METRIC = "score+grad evals/s and HBM GB/s: batched vs single-pair HIP at \u22641e-4 rel grad err"
HBM_PEAK_GBS = 8000.0
N_COEFF_VECTORS = 20


def alg_bytes(pair):
    nvox = pair.fixed.size
    return 20 * nvox + 4 * pair.true_coeff.size


def cpu_model():
    """The host CPU's model name, as /proc/cpuinfo gives it."""
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def time_cpu_evals(case, coeffs, threads, budget_s, warmup=3, evals=10):
    """Median seconds per evaluation of the oracle's flavor g: `warmup`
    untimed evaluations, then up to `evals` timed ones (fewer if the budget
    runs out)."""
    for i in range(warmup):
        case.mse(coeffs[i % len(coeffs)], "g", nthreads=threads)
    times = []
    t_start = time.perf_counter()
    for i in range(evals):
        t0 = time.perf_counter()
        case.mse(coeffs[i % len(coeffs)], "g", nthreads=threads)
        times.append(time.perf_counter() - t0)
        if time.perf_counter() - t_start > budget_s:
            break
    return float(np.median(times)), len(times)


def cpu_threads():
    """Host threads for the CPU legs: OMP_NUM_THREADS where the environment
    sets it (the host's share of cores), else every core of this process's
    affinity mask; never more than the mask holds."""
    aff = len(os.sched_getaffinity(0))
    env = os.environ.get("OMP_NUM_THREADS", "")
    return max(1, min(aff, int(env))) if env.isdigit() and int(env) > 0 else aff


def cpu_host_info():
    return {"threads": cpu_threads(), "affinity_cpus": len(os.sched_getaffinity(0)),
            "os_cpu_count": os.cpu_count(), "omp_num_threads": os.environ.get("OMP_NUM_THREADS")}


def cpu_baseline(pair, coeffs, gpu_results, budget_s=12.0):
    """Time the oracle's flavor-g restatement (OpenMP) on a bounded sample of
    the same workload, with the host's threads (cpu_threads) and with one,
    and report the GPU results' error against the FP64 oracle on the
    same coefficient vectors."""
    import oracle
    threads = cpu_threads()
    case = oracle.Case(pair.fixed, pair.moving, pair.spacing, pair.vox_per_rgn)
    _ = case.m_grad
    med, n = time_cpu_evals(case, coeffs, threads, budget_s)
    med1, n1 = time_cpu_evals(case, coeffs, 1, budget_s)
    errs = []
    for k, (score, grad) in gpu_results.items():
        s64, g64, _ = case.mse_f64(coeffs[k], 0, nthreads=threads)
        errs.append((abs(score - s64) / abs(s64), float(np.linalg.norm(grad - g64) / np.linalg.norm(g64))))
    return {
        "value": 1.0 / med,
        "unit": "evals/s",
        "cores": threads,
        "kind": "port",
        "cpu": cpu_model(),
        "host": cpu_host_info(),
        "sample": f"{n} evals of the same workload after 3 warm-ups, median "
                  f"{med * 1e3:.1f} ms/eval, oracle flavor g (OpenMP, float as the reference)",
        "gb_per_s": alg_bytes(pair) / med / 1e9,
        "one_thread": {"value": 1.0 / med1, "unit": "evals/s", "cores": 1,
                       "sample": f"{n1} evals after 3 warm-ups, median {med1 * 1e3:.1f} ms/eval"},
    }, errs


BATCH_PAIRS = 16
BATCH_SEED = 0


def batch_oracle_errors(pairs, coeffs, results):
    """Per-pair relative score and gradient error of the batched GPU results
    against the FP64 oracle on the same (BF16-rounded) images and BF16-stored
    moving gradient."""
    import oracle
    from plastimatch_amd.bspline import bf16_round
    threads = cpu_threads()
    errs = []
    for p, c, (score, grad) in zip(pairs, coeffs, results):
        case = oracle.Case(p.fixed, p.moving, p.spacing, p.vox_per_rgn, origin=p.origin)
        case._m_grad = bf16_round(case.m_grad)
        s64, g64, _ = case.mse_f64(c, 0, nthreads=threads)
        errs.append((abs(score - s64) / abs(s64), float(np.linalg.norm(grad - g64) / np.linalg.norm(g64))))
    return errs


def batched_leg(args, dev, want_errors):
    """Batched (hipGraph replay) vs looped single-pair evaluation of the
    16-pair BF16 multi-atlas workload; aggregate pair-evaluations per second."""
    import torch
    from plastimatch_amd import synth
    from plastimatch_amd.bspline import (BsplineBatch, BsplineHip, Bspline_xform, Volume, bf16_round,
                                         volume_gradient)

    pairs = synth.make_batch(BATCH_PAIRS, seed=BATCH_SEED)
    stages = []
    for p in pairs:
        p.fixed = bf16_round(p.fixed)
        p.moving = bf16_round(p.moving)
        fixed = Volume(torch.from_numpy(p.fixed).to(dev), p.spacing, p.origin)
        moving = Volume(torch.from_numpy(p.moving).to(dev), p.spacing, p.origin)
        mg = volume_gradient(moving)
        bxf = Bspline_xform.for_volume(fixed, p.vox_per_rgn)
        stages.append(BsplineHip(bxf, fixed, moving, mg, storage="bf16"))
        del fixed, moving, mg
    nvox = sum(p.fixed.size for p in pairs)
    coeffs = [synth.random_coeff(p.dim, p.vox_per_rgn, 3.0, seed=200 + i) for i, p in enumerate(pairs)]
    d_coeffs = [torch.from_numpy(c).to(dev) for c in coeffs]
    d_grads = [torch.zeros(c.size, dtype=torch.float64, device=dev) for c in coeffs]
    d_score = torch.zeros(BATCH_PAIRS, dtype=torch.float64, device=dev)
    d_nvox = torch.zeros(BATCH_PAIRS, dtype=torch.int64, device=dev)
    # the caller's stream (a stream handle of its own: the null handle would
    # select each context's own stream)
    caller = torch.cuda.Stream()
    stream = caller.cuda_stream
    steps = max(10, args.steps // 10)
    warm = max(2, args.warmup // 10)

    # (b) looping the single-pair entry point, each stage autotuned on its
    # own: on the caller's stream (the loop runs the pairs one after the
    # other), and on each context's own stream (the pairs' launches may
    # overlap, up to the hardware queues)
    for st, c in zip(stages, d_coeffs):
        st.autotune("mse", "ref_nn", c, reps=3)

    def loop_once(own_streams=False):
        for i, st in enumerate(stages):
            st.score_device(d_coeffs[i], d_score[i:i + 1], d_grads[i], d_nvox[i:i + 1], "mse", "ref_nn",
                            stream=None if own_streams else stream)

    def time_loop(own_streams):
        for _ in range(warm):
            loop_once(own_streams)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(steps):
            loop_once(own_streams)
        torch.cuda.synchronize()
        return (time.perf_counter() - t0) / steps

    t_loop = time_loop(False)
    t_loop_ctx = time_loop(True)

    # (a) one batched evaluation, autotuned, captured once and replayed
    batch = BsplineBatch(stages)
    batch.set_buffers(d_coeffs, d_grads, d_score, d_nvox)
    batch.autotune("ref_nn", reps=3)
    vpt, wgs = batch.get_tuning()
    batch.capture("ref_nn")
    for _ in range(warm):
        batch.launch(stream)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        batch.launch(stream)
    torch.cuda.synchronize()
    t_batch = (time.perf_counter() - t0) / steps
    # the batched tile kernel alone, event-timed on its stream (eager launches)
    batch.set_timing(True)
    for _ in range(steps):
        batch.score_device("ref_nn", stream)
    ms, n = batch.get_timing()
    batch.set_timing(False)
    tile_ms = ms[0] / n
    torch.cuda.synchronize()
    results = [(float(d_score[i].item()), d_grads[i].cpu().numpy()) for i in range(BATCH_PAIRS)]
    bytes_per_eval = 10 * nvox + sum(4 * c.size + 8 * c.size for c in coeffs)   # BF16: 2 + 2 + 6 B/voxel
    out = {
        "workload": f"batch{BATCH_PAIRS}_edges96-256_vpr10_bf16",
        "pairs": BATCH_PAIRS,
        "voxels": int(nvox),
        "dims": [list(p.dim) for p in pairs],
        "storage": "bf16 (FP32 interpolation, FP64 reductions)",
        "vox_per_thread": vpt,
        "workgroups_per_cu": wgs,
        "batched_graph_ms_per_eval": t_batch * 1e3,
        "looped_single_ms_per_eval": t_loop * 1e3,
        "looped_single_context_streams_ms_per_eval": t_loop_ctx * 1e3,
        "batched_pair_evals_per_s": BATCH_PAIRS / t_batch,
        "looped_pair_evals_per_s": BATCH_PAIRS / t_loop,
        "speedup_batched_vs_looped": t_loop / t_batch,
        "speedup_batched_vs_looped_context_streams": t_loop_ctx / t_batch,
        "batched_tile_kernel_ms": tile_ms,
        "batched_gb_per_s": bytes_per_eval / t_batch / 1e9,
        "tile_kernel_gb_per_s": 10 * nvox / (tile_ms * 1e-3) / 1e9,
    }
    if want_errors:
        errs = batch_oracle_errors(pairs, coeffs, results)
        out["rel_err_vs_fp64_per_pair"] = [{"score": e[0], "grad_l2": e[1]} for e in errs]
        out["max_rel_err_vs_fp64"] = {"score": max(e[0] for e in errs), "grad_l2": max(e[1] for e in errs)}
    batch.close()
    for st in stages:
        st.close()
    return out


ADJOINT_WORKLOAD = "adjoint_192x192x96_vpr8"
FD_COMPONENTS = 200
FD_STEP_MM = 1e-3
# the FP64 score resolves a much smaller step, where the piecewise-smooth
# trilinear score is differentiable across nearly every voxel
FD64_STEP_MM = 1e-5


def adjoint_leg(args, dev, want_errors):
    """The differentiable MSE (torch.autograd.Function) on 192x192x96, vpr 8:
    forward+backward time and extra memory of the recompute-in-backward and
    stored-field modes, and the adjoint against central finite differences
    of the GPU score on 200 random coefficients (and, in the CPU leg, against
    the reference flavor's gradient, FP64, and central differences of the
    FP64 score on 200 coefficients whose support holds no clamped voxel)."""
    import torch
    from plastimatch_amd import synth
    from plastimatch_amd.autograd import mse
    from plastimatch_amd.bspline import BsplineHip, Bspline_xform, Volume

    pair = synth.make_config(ADJOINT_WORKLOAD, seed=0)
    fixed = Volume(torch.from_numpy(pair.fixed).to(dev), pair.spacing, pair.origin)
    moving = Volume(torch.from_numpy(pair.moving).to(dev), pair.spacing, pair.origin)
    bxf = Bspline_xform.for_volume(fixed, pair.vox_per_rgn)
    st = BsplineHip(bxf, fixed, moving)
    c = synth.random_coeff(pair.dim, pair.vox_per_rgn, 2.0, seed=300)
    steps = max(10, args.steps // 10)
    out = {"workload": ADJOINT_WORKLOAD, "dim": list(pair.dim), "vox_per_rgn": list(pair.vox_per_rgn),
           "grad_semantics": "exact trilinear derivative (flavor p)"}
    grads = {}
    for mode in ("recompute", "stored"):
        coeff = torch.from_numpy(c).to(dev).requires_grad_(True)

        def fwd_bwd():
            coeff.grad = None
            s = mse(coeff, st, mode)
            s.backward()
            return s

        for _ in range(3):
            fwd_bwd()
        torch.cuda.synchronize()
        base_mem = torch.cuda.memory_allocated(dev)
        torch.cuda.reset_peak_memory_stats(dev)
        t0 = time.perf_counter()
        for _ in range(steps):
            s = fwd_bwd()
        torch.cuda.synchronize()
        dt = (time.perf_counter() - t0) / steps
        peak = torch.cuda.max_memory_allocated(dev) - base_mem
        grads[mode] = (float(s.item()), coeff.grad.detach().cpu().numpy().astype(np.float64))
        out[mode] = {"ms_per_forward_backward": dt * 1e3, "extra_device_bytes": int(peak)}
    out["stored_field_bytes"] = 12 * pair.fixed.size
    g = grads["recompute"][1]
    out["stored_vs_recompute_grad_rel_l2"] = float(np.linalg.norm(grads["stored"][1] - g) / np.linalg.norm(g))
    # central differences of the GPU score on random coefficients
    rng = np.random.default_rng(301)
    idx = rng.choice(c.size, FD_COMPONENTS, replace=False)
    base = torch.from_numpy(c).to(dev)
    fd = np.empty(FD_COMPONENTS)
    with torch.no_grad():
        for n, i in enumerate(idx):
            a = base.clone()
            a[i] += FD_STEP_MM
            b = base.clone()
            b[i] -= FD_STEP_MM
            fd[n] = (mse(a, st).item() - mse(b, st).item()) / (2 * FD_STEP_MM)
    out["fd_check"] = {"components": FD_COMPONENTS, "h_mm": FD_STEP_MM,
                       "rel_l2": float(np.linalg.norm(g[idx] - fd) / np.linalg.norm(fd)),
                       "median_abs_rel": float(np.median(np.abs(g[idx] - fd) / np.maximum(np.abs(fd), 1e-30)))}
    if want_errors:
        import oracle
        threads = cpu_threads()
        case = oracle.Case(pair.fixed, pair.moving, pair.spacing, pair.vox_per_rgn)
        s_ref, g_ref, _, _ = case.mse(c, "p", nthreads=threads)
        s64, g64, _ = case.mse_f64(c, 1, nthreads=threads)
        sc = grads["recompute"][0]
        out["rel_err_vs_reference_flavor_p"] = {
            "score": abs(sc - s_ref) / abs(s_ref),
            "grad_l2": float(np.linalg.norm(g - g_ref) / np.linalg.norm(g_ref))}
        out["rel_err_vs_fp64"] = {"score": abs(sc - s64) / abs(s64),
                                  "grad_l2": float(np.linalg.norm(g - g64) / np.linalg.norm(g64))}
        # central differences of the FP64 score (flavor p) on coefficients
        # whose support holds no clamped voxel
        inner = []
        for i in np.random.default_rng(302).permutation(c.size):
            if case.mse_f64_support(c, int(i), 0.0)[2] == 0:
                inner.append(int(i))
                if len(inner) == FD_COMPONENTS:
                    break
        fd64, crossed = case.fd_mse_f64(c, np.array(inner), FD64_STEP_MM)
        out["fd64_check"] = {"components": len(inner), "h_mm": FD64_STEP_MM, "unclamped_support": True,
                             "border_voxels_in_steps": int(np.sum(crossed)),
                             "rel_l2": float(np.linalg.norm(g[inner] - fd64) / np.linalg.norm(fd64))}
    st.close()
    return out


# The deformation-output and regularization rows, on the headline workload
REG_WEIGHTS = dict(total_displacement_penalty=0.3, diffusion_penalty=0.7, curvature_penalty=1.1,
                   lame_coefficient_1=0.4, lame_coefficient_2=0.9, linear_elastic_multiplier=1.3,
                   third_order_penalty=0.2, curvature_mixed_weight=0.6)
# FP64 operations per knot slot of k_regularize with every term on (FMA = 2):
# per component the x-digit step 6 contractions x 7, the y-digit step 16
# contractions x 7 + 31 for the weighted sums, the z-digit step 4
# contractions x 7 + 3; the elastic couplings 6 x (2 x 7 + 3); the gradient 9
# and the score 11 -- 3 x (42 + 143 + 31) + 102 + 20.
REG_FLOPS_PER_KNOT = 3 * (42 + 143 + 31) + 6 * 17 + 20
FP64_PEAK_TFLOPS = 78.6


def warp_reg_leg(args, dev, want_cpu):
    """The warped moving image (trilinear) and the analytic regularizer (all
    terms on) on 256x256x128, vpr 10: kernel time with HIP events on the
    stream they run on, rooflines (the warp reads a moving value and writes
    a float per voxel, 8 B; the regularizer is FP64 arithmetic), and the
    oracle's warp and regularizer timed beside them with their errors."""
    import torch
    from plastimatch_amd import synth
    from plastimatch_amd.bspline import BsplineHip, Bspline_xform, Volume

    pair = synth.make_config(WORKLOAD, seed=0)
    fixed = Volume(torch.from_numpy(pair.fixed).to(dev), pair.spacing, pair.origin)
    moving = Volume(torch.from_numpy(pair.moving).to(dev), pair.spacing, pair.origin)
    bxf = Bspline_xform.for_volume(fixed, pair.vox_per_rgn)
    st = BsplineHip(bxf, fixed, moving)
    c = synth.random_coeff(pair.dim, pair.vox_per_rgn, 3.0, seed=100)
    d_c = torch.from_numpy(c).to(dev)
    npix = pair.fixed.size
    out_img = torch.empty(npix, dtype=torch.float32, device=dev)
    sc = torch.zeros(1, dtype=torch.float64, device=dev)
    gr = torch.zeros(bxf.num_coeff, dtype=torch.float64, device=dev)
    stream = torch.cuda.current_stream()
    s = stream.cuda_stream
    steps = max(20, args.steps // 10)

    def timed(fn):
        for _ in range(5):
            fn()
        e0 = torch.cuda.Event(enable_timing=True)
        e1 = torch.cuda.Event(enable_timing=True)
        e0.record(stream)
        for _ in range(steps):
            fn()
        e1.record(stream)
        torch.cuda.synchronize()
        return e0.elapsed_time(e1) / steps

    warp_ms = timed(lambda: st.warp(d_c, out_img, -1000.0, True, stream=s))
    reg_ms = timed(lambda: st.regularize(d_c, sc, gr, stream=s, **REG_WEIGHTS))
    warp_bytes = 8 * npix + 4 * bxf.num_coeff
    warp_gbs = warp_bytes / (warp_ms * 1e-3) / 1e9
    reg_flops = REG_FLOPS_PER_KNOT * 64 * st.num_tiles
    reg_tflops = reg_flops / (reg_ms * 1e-3) / 1e12
    out = {
        "workload": WORKLOAD,
        "warp": {"ms": warp_ms, "voxels_per_s": npix / (warp_ms * 1e-3),
                 "roofline": {"bound": "hbm", "achieved": warp_gbs, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                              "frac": warp_gbs / HBM_PEAK_GBS, "alg_bytes_per_launch": warp_bytes,
                              "kernel": "k_tiles<6> (+ no fill: the ROI is the image)"}},
        "regularize": {"ms": reg_ms, "weights": REG_WEIGHTS, "regions": st.num_tiles,
                       "roofline": {"bound": "mfma", "achieved": reg_tflops, "peak": FP64_PEAK_TFLOPS,
                                    "unit": "TFLOP/s", "frac": reg_tflops / FP64_PEAK_TFLOPS,
                                    "flops_per_launch": reg_flops, "dtype": "f64",
                                    "kernel": "k_regularize + k_reg_reduce"}},
    }
    got_img = out_img.cpu().numpy().reshape(pair.fixed.shape)
    got_s, got_g = sc.item(), gr.cpu().numpy()
    st.close()
    if want_cpu:
        import oracle
        threads = cpu_threads()
        g = oracle.geom(pair.dim, pair.spacing, pair.vox_per_rgn, origin=pair.origin)
        t0 = time.perf_counter()
        ref_img = oracle.warp(g, pair.moving, pair.spacing, c, True, -1000.0, m_origin=pair.origin,
                              nthreads=threads)
        t_warp = time.perf_counter() - t0
        both = (got_img != -1000.0) & (ref_img != -1000.0)
        span = float(pair.moving.max() - pair.moving.min())
        t0 = time.perf_counter()
        s_ref, g_ref = oracle.reg_analytic(g, oracle.reg_parms(**REG_WEIGHTS), c, nthreads=threads)
        t_reg = time.perf_counter() - t0
        out["warp"]["cpu_baseline"] = {"value": npix / t_warp, "unit": "voxels/s", "cores": threads,
                                       "kind": "port", "sample": "one full warp (orc_warp, OpenMP)"}
        out["warp"]["rms_err_vs_oracle_rel_range"] = float(
            np.sqrt(((got_img[both].astype(np.float64) - ref_img[both]) ** 2).mean()) / span)
        out["warp"]["default_mismatches"] = int(np.count_nonzero((got_img == -1000.0) != (ref_img == -1000.0)))
        out["regularize"]["cpu_baseline"] = {"value": 1.0 / t_reg, "unit": "evals/s", "cores": threads,
                                             "kind": "port",
                                             "sample": "one evaluation, dense 64x64 forms (orc_reg_analytic)"}
        out["regularize"]["rel_err_vs_oracle"] = {
            "score": abs(got_s - s_ref) / abs(s_ref),
            "grad_max_rel": float(np.abs(got_g - g_ref).max() / np.abs(g_ref).max())}
    return out


OPT_ITS = 20
OPT_CPU_ITS = 3
# regularization of the optimizer leg: the analytic penalties integrate over
# the grid in mm^3, so weights of this size keep them about a tenth of the
# MSE at the synthetic field
OPT_REG_WEIGHTS = {"curvature_penalty": 1e-4, "diffusion_penalty": 1e-6}


def optimizer_leg(args, dev, eval_ms, want_cpu):
    """The optimizer loop on the device: liblbfgs's L-BFGS (More-Thuente,
    defaults) over the headline workload's flavor-g MSE from the identity,
    OPT_ITS iterations (max_feval unbounded), everything resident; wall time
    of the whole call against the bare evaluation.  CPU beside it: the
    reference's own liblbfgs (oracle/_ref) driving the oracle's flavor-g
    score for OPT_CPU_ITS iterations, whose progress reports the device
    run's first iterations are compared with."""
    import torch
    from plastimatch_amd import synth
    from plastimatch_amd.bspline import BsplineHip, Bspline_xform, Volume, volume_gradient

    pair = synth.make_config(WORKLOAD, seed=0)
    fixed = Volume(torch.from_numpy(pair.fixed).to(dev), pair.spacing, pair.origin)
    moving = Volume(torch.from_numpy(pair.moving).to(dev), pair.spacing, pair.origin)
    bxf = Bspline_xform.for_volume(fixed, pair.vox_per_rgn)
    st = BsplineHip(bxf, fixed, moving, volume_gradient(moving))
    st.autotune("mse", "ref_nn", None, reps=5)
    d_x = torch.zeros(bxf.num_coeff, dtype=torch.float64, device=dev)
    d_c = torch.zeros(bxf.num_coeff, dtype=torch.float32, device=dev)
    st.optimize_lbfgs(d_x, d_c, max_its=2, max_feval=-1)                  # warm-ups
    st.optimize_lbfgs(d_x, d_c, max_its=2, max_feval=-1, resident=False)
    out = {"workload": WORKLOAD, "optimizer": "L-BFGS, More-Thuente (liblbfgs 1.9 defaults)",
           "start": "identity (zero coefficients)",
           "loop": "host-driven (the default); mse_loops also times the device-resident loop "
                   "(scalar logic on the device, evaluation chains replayed from a hipGraph, 4 a replay)"}
    runs = {}
    for label, reg in (("mse", None), ("mse_plus_regularization", OPT_REG_WEIGHTS)):
        d_x.zero_()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        r = st.optimize_lbfgs(d_x, d_c, "mse", "ref_nn", reg, record=False, max_its=OPT_ITS, max_feval=-1)
        dt = time.perf_counter() - t0
        if label == "mse":
            d_x.zero_()
            runs[label] = st.optimize_lbfgs(d_x, d_c, "mse", "ref_nn", None, max_its=OPT_CPU_ITS + 1,
                                            max_feval=-1)
        out[label] = {"status": r.status, "iterations": r.iterations, "fevals": r.fevals,
                      "ms_per_iteration": dt * 1e3 / max(r.iterations, 1),
                      "ms_per_evaluation_in_loop": dt * 1e3 / r.fevals,
                      "fx_end": r.fx}
    # the host-driven loop on the same run, and both loops' cost per
    # evaluation from the difference of a long and a short run (the fixed
    # cost of a call cancels)
    per_eval = {}
    for resident in (True, False):
        tt = []
        for its in (OPT_ITS, 3 * OPT_ITS):
            d_x.zero_()
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            r = st.optimize_lbfgs(d_x, d_c, "mse", "ref_nn", None, record=False, resident=resident,
                                  max_its=its, max_feval=-1)
            tt.append((time.perf_counter() - t0, r.fevals))
        per_eval["resident" if resident else "host_loop"] = {
            "ms_per_evaluation_in_loop": tt[0][0] * 1e3 / tt[0][1],
            "marginal_ms_per_evaluation": (tt[1][0] - tt[0][0]) * 1e3 / (tt[1][1] - tt[0][1]),
            "fevals": [tt[0][1], tt[1][1]]}
    out["mse_loops"] = per_eval
    out["mse_plus_regularization"]["weights"] = OPT_REG_WEIGHTS
    out["mse"]["bare_evaluation_ms"] = eval_ms
    out["mse"]["loop_overhead_ms_per_evaluation"] = out["mse"]["ms_per_evaluation_in_loop"] - eval_ms
    for v in per_eval.values():
        v["marginal_overhead_ms_per_evaluation"] = v["marginal_ms_per_evaluation"] - eval_ms
    st.close()
    if want_cpu:
        import oracle
        threads = cpu_threads()
        case = oracle.Case(pair.fixed, pair.moving, pair.spacing, pair.vox_per_rgn)
        _ = case.m_grad

        def fun(x):
            sc, g, _, _ = case.mse(x.astype(np.float32), "g", nthreads=threads)
            return float(sc), g.astype(np.float64)

        impl = "ref" if oracle.ref_lbfgs_lib() is not None else "oracle"
        t0 = time.perf_counter()
        ref = oracle.lbfgs(fun, np.zeros(bxf.num_coeff), impl=impl, max_its=OPT_CPU_ITS, max_feval=1000)
        dt = time.perf_counter() - t0
        dev_rep = runs["mse"].reports[:len(ref.reports)]
        diffs = [abs(a[0] - b[0]) / abs(b[0]) for a, b in zip(dev_rep, ref.reports)]
        out["cpu_baseline"] = {
            "value": ref.feval / dt, "unit": "evaluations/s in the loop", "cores": threads,
            "kind": "reference" if impl == "ref" else "port",
            "ms_per_iteration": dt * 1e3 / max(ref.it, 1),
            "sample": f"{ref.feval} evaluations ({OPT_CPU_ITS} iterations + the stopping trial) of the "
                      "reference's liblbfgs driving the oracle's flavor-g score (OpenMP)"}
        out["vs_cpu_reference_run"] = {
            "reports_compared": len(diffs),
            "same_iteration_record": [a[4:] for a in dev_rep] == [b[4:] for b in ref.reports],
            "max_rel_diff_fx": max(diffs) if diffs else None}
    return out


STAGE_RATE = (2, 2, 1)


def stage_leg(args, dev, want_cpu):
    """Stage preprocessing of the headline workload on the device: both
    256x256x128 images subsampled by (2, 2, 1) with legacy subsampling, the
    moving gradient of the subsampled moving image, and the MI histogram
    ranges of both -- everything registration_data.cxx:350-384 and
    joint_histogram.cxx:115-139 do before a stage's first evaluation.  Wall
    time of the whole sequence from device-resident inputs (each call also
    uploads its small axis tables); the oracle's restatements timed beside it,
    and whether the results are the oracle's bit for bit."""
    import torch
    from plastimatch_amd import synth
    from plastimatch_amd.bspline import Volume, hist_eqsp, hist_vopt, stage_resample, volume_gradient
    from plastimatch_amd._abi import VOPT_KEYS

    pair = synth.make_config(WORKLOAD, seed=0)
    fx = Volume(torch.from_numpy(pair.fixed).to(dev), pair.spacing, pair.origin)
    mv = Volume(torch.from_numpy(pair.moving).to(dev), pair.spacing, pair.origin)

    def run():
        f_ss = stage_resample(fx, "voxel_legacy", STAGE_RATE)
        m_ss = stage_resample(mv, "voxel_legacy", STAGE_RATE)
        mg = volume_gradient(m_ss)
        h = (hist_eqsp(f_ss, 32), hist_eqsp(m_ss, 32))
        return f_ss, m_ss, mg, h

    for _ in range(3):
        run()
    torch.cuda.synchronize()
    reps = 20
    t0 = time.perf_counter()
    for _ in range(reps):
        f_ss, m_ss, mg, h = run()
    torch.cuda.synchronize()
    dt = (time.perf_counter() - t0) / reps
    n_in, n_out = pair.fixed.size, f_ss.npix
    out = {"workload": WORKLOAD, "resample": f"voxel_legacy {STAGE_RATE}, trilinear",
           "in_dim": list(pair.dim), "out_dim": list(f_ss.dim), "ms_per_stage": dt * 1e3,
           "resample_alg_bytes_per_image": 4 * (n_in + n_out)}
    # the histogram_type=vopt alternative to the equal-spacing ranges:
    # V-optimal bins of both stage images
    for _ in range(2):
        hist_vopt(f_ss, 32), hist_vopt(m_ss, 32)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(reps):
        hv = (hist_vopt(f_ss, 32), hist_vopt(m_ss, 32))
    torch.cuda.synchronize()
    out["hist_vopt"] = {"bins": 32, "keys": VOPT_KEYS, "ms_both_images": (time.perf_counter() - t0) / reps * 1e3}
    if want_cpu:
        import oracle
        t0 = time.perf_counter()
        rf, _, _ = oracle.stage_resample(pair.fixed, pair.spacing, pair.origin, "voxel_legacy", STAGE_RATE)
        rm, o, s = oracle.stage_resample(pair.moving, pair.spacing, pair.origin, "voxel_legacy", STAGE_RATE)
        rg = oracle.volume_gradient(rm, s, o)
        rh = (oracle.hist_eqsp(rf, 32), oracle.hist_eqsp(rm, 32))
        t_cpu = time.perf_counter() - t0
        out["cpu_baseline"] = {"value": t_cpu * 1e3, "unit": "ms per stage", "cores": 1, "kind": "port",
                               "sample": "one stage preprocessing (oracle restatements, one thread)"}
        out["bit_exact_vs_oracle"] = {
            "fixed": bool(np.array_equal(f_ss.img.cpu().numpy(), rf)),
            "moving": bool(np.array_equal(m_ss.img.cpu().numpy(), rm)),
            "moving_gradient": bool(np.array_equal(mg.img.cpu().numpy(), rg)),
            "hist_eqsp": h == rh}
        t0 = time.perf_counter()
        rv = (oracle.hist_vopt(rf, 32), oracle.hist_vopt(rm, 32))
        t_cpu = time.perf_counter() - t0
        out["hist_vopt"]["cpu_baseline"] = {"value": t_cpu * 1e3, "unit": "ms both images", "cores": 1,
                                            "kind": "port", "sample": "oracle orc_hist_vopt on both stage images"}
        out["hist_vopt"]["bit_exact_vs_oracle"] = all(
            a[0] == b[0] and a[1] == b[1] and np.array_equal(a[2], b[2]) for a, b in zip(hv, rv))
    return out


MI_WORKLOAD = "mi_512x512x160_vpr12"
MI_BINS = 32


def mi_leg(args, dev, want_cpu):
    """Mattes MI (flavor g) on the 512x512x160 CT-like pair with a 12-voxel
    grid and 32x32 bins: evaluations per second at four coefficient vectors
    (the true field and three N(0, 4 mm) draws), the stage times, and the
    HBM roofline of the whole two-pass evaluation (fixed 4 + moving 4 B per
    voxel and pass, the float-image byte model; the device reads 1-byte bin
    images instead).  With the CPU leg: the oracle's flavor g (serial pass 1 as the
    reference) on one of the vectors, timed, and the device's integer
    histogram, score and gradient against it and the FP64 oracle."""
    import torch
    from plastimatch_amd import synth
    from plastimatch_amd.bspline import stage_for_pair

    pair = synth.make_config(MI_WORKLOAD, seed=0)
    nvox = pair.fixed.size
    st = stage_for_pair(pair, with_grad=False, bins=MI_BINS)
    coeffs = [pair.true_coeff] + [synth.random_coeff(pair.dim, pair.vox_per_rgn, 4.0, seed=400 + i)
                                  for i in range(3)]
    d_coeffs = [torch.from_numpy(np.ascontiguousarray(c, np.float32)).to(dev) for c in coeffs]
    d_score = torch.zeros(1, dtype=torch.float64, device=dev)
    d_grad = torch.zeros(st.bxf.num_coeff, dtype=torch.float64, device=dev)
    d_nvox = torch.zeros(1, dtype=torch.int64, device=dev)
    stream = torch.cuda.current_stream().cuda_stream
    st.autotune("mi", "ref_nn", d_coeffs[0], reps=3)
    vpt, wgs = st.get_tuning()

    def step(i):
        st.score_device(d_coeffs[i % len(d_coeffs)], d_score, d_grad, d_nvox, "mi", stream=stream)

    for i in range(3):
        step(i)
    torch.cuda.synchronize()
    steps = max(8, args.steps // 50)
    t0 = time.perf_counter()
    for i in range(steps):
        step(i)
    torch.cuda.synchronize()
    dt = (time.perf_counter() - t0) / steps
    st.set_timing(True)
    for i in range(steps):
        step(i)
    ms, n = st.get_timing()
    st.set_timing(False)
    stage = [float(x / n) for x in ms[:5]]
    alg = 16 * nvox + 4 * coeffs[0].size + 8 * coeffs[0].size
    achieved = alg / dt / 1e9
    out = {"workload": MI_WORKLOAD, "dim": list(pair.dim), "vox_per_rgn": list(pair.vox_per_rgn),
           "bins": [MI_BINS, MI_BINS], "coeff_vectors": len(coeffs),
           "vox_per_thread": vpt, "workgroups_per_cu": wgs,
           "value": 1.0 / dt, "unit": "evals/s", "ms_per_eval": dt * 1e3,
           "stage_ms": {"zero": stage[0], "k_mi_hist": stage[1], "k_mi_score": stage[2],
                        "k_tiles_pass2": stage[3], "k_reduce": stage[4]},
           "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                        "frac": achieved / HBM_PEAK_GBS, "alg_bytes_per_eval": alg,
                        "scope": "whole evaluation (both passes)"}}
    if want_cpu:
        import oracle
        threads = cpu_threads()
        oc = oracle.Case(pair.fixed, pair.moving, pair.spacing, pair.vox_per_rgn, origin=pair.origin)
        hp = oc.hist(MI_BINS)
        t0 = time.perf_counter()
        o = oc.mi(coeffs[1], hp, nthreads=threads)
        t_cpu = time.perf_counter() - t0
        s64, g64, _ = oc.mi_f64(coeffs[1], hp, nthreads=threads)
        r = st.score(coeffs[1], "mi", want_hist=True)
        nj = MI_BINS * MI_BINS
        out["cpu_baseline"] = {"value": 1.0 / t_cpu, "unit": "evals/s", "cores": threads, "kind": "port",
                               "sample": "one evaluation of the oracle's flavor g (pass 1 serial as the "
                                         "reference, pass 2 OpenMP)"}
        out["identical_histograms"] = bool(np.array_equal(r.hist[:nj], o["fx_j_hist"])
                                           and np.array_equal(r.hist[nj:], o["fx_f_hist"]))
        out["num_vox_equal"] = r.num_vox == o["num_vox"]
        out["rel_err_vs_oracle"] = {"score": abs(r.score - o["score"]) / abs(o["score"]),
                                    "grad_l2": float(np.linalg.norm(r.grad - o["grad"]) / np.linalg.norm(o["grad"]))}
        out["rel_err_vs_fp64"] = {"score": abs(r.score - s64) / abs(s64),
                                  "grad_l2": float(np.linalg.norm(r.grad - g64) / np.linalg.norm(g64))}
        # the reference's float flavor is itself this far from FP64 (MI's PV
        # derivative cancels); the parity bar is 1.5 times this where it
        # exceeds 1e-4
        out["reference_flavor_rel_err_vs_fp64"] = {
            "score": abs(o["score"] - s64) / abs(s64),
            "grad_l2": float(np.linalg.norm(o["grad"] - g64) / np.linalg.norm(g64))}
    st.close()
    return out


CUBE_WORKLOAD = "mse_128cube_vpr15"


def cube_leg(args, dev, want_cpu):
    """MSE flavor g on the 128^3 pair (1 mm, 15-voxel grid, N(0, 2 mm)
    field) at zero and at the true coefficients: evaluations per second of
    the single-pair HIP path per vector, its tile kernel's roofline, and --
    with the CPU leg -- the oracle's flavor g timed on the host's threads
    and on one, with the device results' errors against the FP64 oracle
    and the float flavor, and the overlap counts."""
    import torch
    from plastimatch_amd import synth
    from plastimatch_amd.bspline import stage_for_pair

    pair = synth.make_config(CUBE_WORKLOAD, seed=0)
    st = stage_for_pair(pair)
    vecs = {"zero": np.zeros(pair.true_coeff.size, np.float32), "true": pair.true_coeff}
    d_score = torch.zeros(1, dtype=torch.float64, device=dev)
    d_grad = torch.zeros(st.bxf.num_coeff, dtype=torch.float64, device=dev)
    d_nvox = torch.zeros(1, dtype=torch.int64, device=dev)
    stream = torch.cuda.current_stream().cuda_stream
    d0 = torch.from_numpy(vecs["true"]).to(dev)
    st.autotune("mse", "ref_nn", d0, reps=5)
    vpt, wgs = st.get_tuning()
    steps = max(50, args.steps // 5)
    out = {"workload": CUBE_WORKLOAD, "dim": list(pair.dim), "vox_per_rgn": list(pair.vox_per_rgn),
           "grad_mode": "ref_nn", "vox_per_thread": vpt, "workgroups_per_cu": wgs, "per_vector": {}}
    alg = alg_bytes(pair)
    for name, c in vecs.items():
        dc = torch.from_numpy(c).to(dev)
        for _ in range(5):
            st.score_device(dc, d_score, d_grad, d_nvox, "mse", "ref_nn", stream=stream)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(steps):
            st.score_device(dc, d_score, d_grad, d_nvox, "mse", "ref_nn", stream=stream)
        torch.cuda.synchronize()
        dt = (time.perf_counter() - t0) / steps
        st.set_timing(True)
        for _ in range(steps):
            st.score_device(dc, d_score, d_grad, d_nvox, "mse", "ref_nn", stream=stream)
        ms, n = st.get_timing()
        st.set_timing(False)
        tile_ms = float(ms[0] / n)
        r = st.score(c, "mse", "ref_nn")
        v = {"value": 1.0 / dt, "unit": "evals/s", "ms_per_eval": dt * 1e3,
             "stage_ms": {"k_tiles": tile_ms, "k_reduce": float(ms[1] / n)},
             "roofline": {"bound": "hbm", "achieved": alg / (tile_ms * 1e-3) / 1e9, "peak": HBM_PEAK_GBS,
                          "unit": "GB/s", "frac": alg / (tile_ms * 1e-3) / 1e9 / HBM_PEAK_GBS,
                          "alg_bytes_per_launch": alg, "kernel": "k_tiles"},
             "score": r.score, "num_vox": r.num_vox, "_grad": r.grad}
        out["per_vector"][name] = v
    if want_cpu:
        import oracle
        threads = cpu_threads()
        case = oracle.Case(pair.fixed, pair.moving, pair.spacing, pair.vox_per_rgn)
        mg = case.m_grad
        for name, c in vecs.items():
            v = out["per_vector"][name]
            times = []
            for k in range(4):
                t0 = time.perf_counter()
                s_ref, g_ref, n_ref, _ = case.mse(c, "g", nthreads=threads)
                times.append(time.perf_counter() - t0)
            t1 = time.perf_counter()
            case.mse(c, "g", nthreads=1)
            t_one = time.perf_counter() - t1
            s64, g64, n64 = case.mse_f64(c, 0, nthreads=threads)
            g = v["_grad"]
            v["cpu_baseline"] = {"value": 1.0 / float(np.median(times[1:])), "unit": "evals/s",
                                 "cores": threads, "kind": "port", "host": cpu_host_info(),
                                 "sample": "median of 3 evals after 1 warm-up, oracle flavor g (OpenMP)",
                                 "one_thread": {"value": 1.0 / t_one, "unit": "evals/s", "cores": 1,
                                                "sample": "one eval"}}
            v["num_vox_oracle"] = n_ref
            v["num_vox_fp64"] = n64
            v["rel_err_vs_fp64"] = {"score": abs(v["score"] - s64) / abs(s64) if s64 else abs(v["score"]),
                                    "grad_l2": float(np.linalg.norm(g - g64) / max(np.linalg.norm(g64), 1e-300))}
            v["rel_err_vs_reference_flavor"] = {
                "score": abs(v["score"] - s_ref) / abs(s_ref) if s_ref else abs(v["score"]),
                "grad_l2": float(np.linalg.norm(g - g_ref) / max(np.linalg.norm(g_ref), 1e-300))}
        del mg
    for v in out["per_vector"].values():
        del v["_grad"]
    st.close()
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1000)
    ap.add_argument("--warmup", type=int, default=50)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--no-batch", action="store_true", help="skip the batched multi-atlas leg")
    ap.add_argument("--no-adjoint", action="store_true", help="skip the differentiable-path leg")
    ap.add_argument("--no-warp-reg", action="store_true", help="skip the warp and regularization leg")
    ap.add_argument("--no-optimizer", action="store_true", help="skip the optimizer-loop leg")
    ap.add_argument("--no-stage", action="store_true", help="skip the stage-preprocessing leg")
    ap.add_argument("--no-mi", action="store_true", help="skip the Mattes-MI leg")
    ap.add_argument("--no-cube", action="store_true", help="skip the 128^3 MSE leg")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="only run a plain 256 MiB device-to-device copy kernel 10 times "
                         "(reference for scaling the FETCH_SIZE / WRITE_SIZE counters)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU only; --gpus {args.gpus} is not supported "
              "(across GPUs this path would split the tiles into slabs of z-planes, one slab a GPU, "
              "and add one FP64 all-reduce of the gradient per evaluation)", file=sys.stderr)
        sys.exit(2)

    import torch
    if args.copy_calibration:
        n = 64 * 1024 * 1024
        src = torch.ones(n, dtype=torch.float32, device="cuda:0")
        dst = torch.empty_like(src)
        for _ in range(10):
            dst.copy_(src)
        torch.cuda.synchronize()
        print(json.dumps({"copy_bytes_read": 4 * n, "copy_bytes_written": 4 * n, "launches": 10}))
        return
    from plastimatch_amd import synth
    from plastimatch_amd.bspline import BsplineHip, Bspline_xform, Volume, volume_gradient

    dev = torch.device("cuda:0")
    pair = synth.make_config(WORKLOAD, seed=0)
    fixed = Volume(torch.from_numpy(pair.fixed).to(dev), pair.spacing, pair.origin)
    moving = Volume(torch.from_numpy(pair.moving).to(dev), pair.spacing, pair.origin)
    mgrad = volume_gradient(moving)
    bxf = Bspline_xform.for_volume(fixed, pair.vox_per_rgn)
    st = BsplineHip(bxf, fixed, moving, mgrad)
    coeffs = np.stack([synth.random_coeff(pair.dim, pair.vox_per_rgn, 3.0, seed=100 + i)
                       for i in range(N_COEFF_VECTORS)])
    d_coeffs = torch.from_numpy(coeffs).to(dev)
    d_score = torch.zeros(N_COEFF_VECTORS, dtype=torch.float64, device=dev)
    d_grad = torch.zeros(N_COEFF_VECTORS, bxf.num_coeff, dtype=torch.float64, device=dev)
    d_nvox = torch.zeros(N_COEFF_VECTORS, dtype=torch.int64, device=dev)
    stream = torch.cuda.current_stream().cuda_stream
    st.autotune("mse", "ref_nn", d_coeffs[0], reps=10)
    vpt, wgs = st.get_tuning()

    def step(i):
        k = i % N_COEFF_VECTORS
        st.score_device(d_coeffs[k], d_score[k:k + 1], d_grad[k], d_nvox[k:k + 1], "mse", "ref_nn",
                        stream=stream)

    for i in range(args.warmup):
        step(i)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for i in range(args.steps):
        step(i)
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    ms_per_step = elapsed * 1e3 / args.steps
    value = args.steps / elapsed

    # dominant-kernel timing with HIP events on the launch stream
    st.set_timing(True)
    for i in range(args.steps):
        step(i)
    ms, n = st.get_timing()
    st.set_timing(False)
    tile_ms = ms[0] / n
    stage_ms = [float(x / n) for x in ms[:2]]
    bytes_per_launch = alg_bytes(pair)
    achieved = bytes_per_launch / (tile_ms * 1e-3) / 1e9

    traffic = None
    tpath = os.path.join(ROOT, "profiles", f"k_tiles_{WORKLOAD}_traffic.json")
    if os.path.exists(tpath):
        with open(tpath) as f:
            traffic = json.load(f).get("hbm_bytes_per_launch")

    out = {
        "metric": METRIC,
        "value": value,
        "unit": "evals/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": WORKLOAD, "dim": list(pair.dim), "spacing": list(pair.spacing),
                   "vox_per_rgn": list(pair.vox_per_rgn), "grad_mode": "ref_nn",
                   "coeff_vectors": N_COEFF_VECTORS, "num_tiles": st.num_tiles,
                   "num_coeff": bxf.num_coeff, "vox_per_thread": vpt, "workgroups_per_cu": wgs},
        "eval_gb_per_s": bytes_per_launch / (ms_per_step * 1e-3) / 1e9,
        "stage_ms": {"k_tiles": stage_ms[0], "k_reduce": stage_ms[1]},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": traffic,
                     "kernel": "k_tiles", "alg_bytes_per_launch": bytes_per_launch,
                     "kernel_ms": tile_ms},
        "cpu_baseline": None,
    }
    torch.cuda.synchronize()
    check = {}
    for k in (0, 1):
        check[k] = (float(d_score[k].item()), d_grad[k].cpu().numpy())
    st.close()
    del fixed, moving, mgrad
    if not args.no_cpu:
        base, errs = cpu_baseline(pair, coeffs, check)
        out["cpu_baseline"] = base
        out["rel_err_vs_fp64"] = {"score": max(e[0] for e in errs), "grad_l2": max(e[1] for e in errs)}
    if not args.no_cube:
        out["mse_128cube"] = cube_leg(args, dev, want_cpu=not args.no_cpu)
    if not args.no_batch:
        out["batched"] = batched_leg(args, dev, want_errors=not args.no_cpu)
    if not args.no_adjoint:
        out["adjoint"] = adjoint_leg(args, dev, want_errors=not args.no_cpu)
    if not args.no_warp_reg:
        out["warp_regularize"] = warp_reg_leg(args, dev, want_cpu=not args.no_cpu)
    if not args.no_optimizer:
        out["optimizer_loop"] = optimizer_leg(args, dev, ms_per_step, want_cpu=not args.no_cpu)
    if not args.no_stage:
        out["stage_preprocess"] = stage_leg(args, dev, want_cpu=not args.no_cpu)
    if not args.no_mi:
        out["mattes_mi"] = mi_leg(args, dev, want_cpu=not args.no_cpu)
    print(json.dumps(out))


if __name__ == "__main__":
    main()
