"""Throughput of the PhyloCSF scoring path on one MI355X.

Workload (the configuration the headline metric is quoted on): 58mammals tree
and ECMs, 10,000 synthetic regions with lengths U{50..2000} codons (half
simulated under the coding model, half under the noncoding model), each
species-codon gapped with probability 0.1 and 1% of codons holding an N,
fixed strategy.  One step = one batched forward over all ~10.25 M codon
columns under both models (FP32 MFMA pruning with FP64 P(t) and FP64
log-scaling, per-region FP64 reduction, results copied to the host).
Inputs are resident in HBM before the timed region.

Printed: one JSON line with value = codon-columns/s of the batched MFMA path,
the plain FP64 GPU kernel's rate on the same batch, the accuracy of the
FP32 scores against the FP64 kernel, the roofline of the dominant kernel
(flops from HIP events on the kernel's own stream) and the FP64 CPU oracle's
rate on a bounded sample.
"""
from __future__ import annotations

import argparse
import json
import os
import platform
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "codon-columns/s, batched MFMA vs plain FP64 GPU, at |delta score|<=1e-3 dB"
PEAK_FP32_TFLOPS = 157.3   # MI355X dense FP32 matrix
PEAK_HBM_GBS = 8000.0
PEAK_F16_TFLOPS = 2500.0   # MI355X dense FP16/BF16 matrix
# name -> (C-ABI precision, dtype, peak of the MFMA it runs on, MFMA ops per algorithmic flop)
PRECISIONS = {"fp32": (1, "fp32", PEAK_FP32_TFLOPS, 1.0),
              "fp16s": (2, "fp16x2 split (hi+lo), fp32 accumulate", PEAK_F16_TFLOPS, 3.0),
              "fp16": (3, "fp16, fp32 accumulate", PEAK_F16_TFLOPS, 1.0),
              "bf16": (4, "bf16, fp32 accumulate", PEAK_F16_TFLOPS, 1.0)}


def profiled_traffic(precision, shape):
    """HBM bytes per launch of the fused kernel at this precision and launch
    shape, as measured by the rocprofv3 FETCH_SIZE / WRITE_SIZE passes
    (scaled by a plain copy kernel) committed under profiles/; None when that
    configuration has not been profiled."""
    path = os.path.join(ROOT, "profiles", "prune_mfma_58mammals_c2.md")
    key = "traffic %s shape%d " % (precision, shape)
    try:
        for line in open(path):
            line = line.strip()
            if line.startswith(key):
                return float(line[len(key):])
    except OSError:
        pass
    return None


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return platform.processor()


def _oracle_run(wl, seconds, threads, first):
    """Scores regions first, first+1, ... of the workload with the FP64 oracle on
    `threads` threads (one region at a time per thread, as the reference's
    fork-per-region `-p N`) until `seconds` have passed; returns
    (regions, columns, elapsed).  The oracle's C calls release the GIL."""
    import threading

    from oracle import oracle as O
    from phylocsf_amd.api import paramset_prefix
    prefix = paramset_prefix("58mammals")
    t = O.newick_tree(open(prefix + ".nh").read())
    Sc, pc = O.ecm_read(prefix + "_coding.ECM")
    Sn, pn = O.ecm_read(prefix + "_noncoding.ECM")
    om = O.Model(t, [Sc, Sn], [pc, pn])
    lock = threading.Lock()
    state = {"next": first, "regions": 0, "cols": 0}
    t0 = time.perf_counter()

    def worker():
        while True:
            with lock:
                r = state["next"]
                if r >= wl.n_regions or time.perf_counter() - t0 >= seconds:
                    return
                state["next"] = r + 1
            lo, hi = int(wl.offsets[r]), int(wl.offsets[r + 1])
            om.score_fixed(wl.codes[lo:hi])
            with lock:
                state["regions"] += 1
                state["cols"] += hi - lo

    ts = [threading.Thread(target=worker) for _ in range(threads)]
    for th in ts:
        th.start()
    for th in ts:
        th.join()
    return state["regions"], state["cols"], time.perf_counter() - t0


def cpu_baseline(wl, seconds, threads):
    """FP64 oracle (C restatement of the reference loop) on the same workload:
    `threads` threads over regions for `seconds`, and 1 thread (the reference's
    default -p 1) for a third of that on the following regions."""
    nr, nc, dt = _oracle_run(wl, seconds, threads, 0)
    nr1, nc1, dt1 = _oracle_run(wl, seconds / 3.0, 1, nr)
    return dict(value=nc / dt, unit="codon-columns/s", cores=threads, kind="port",
                sample="regions 0..%d (%d codon columns) of the workload, fixed strategy, "
                       "FP64 oracle, %.1f s on %d threads of %s" % (nr - 1, nc, dt, threads,
                                                                    cpu_model()),
                value_1_thread=nc1 / dt1,
                sample_1_thread="next %d regions (%d codon columns), %.1f s on 1 thread"
                                % (nr1, nc1, dt1))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--fp64-steps", type=int, default=1)
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--cpu-threads", type=int, default=min(16, os.cpu_count() or 1),
                    help="threads of the CPU baseline (the GPU host's share is 16 cores)")
    ap.add_argument("--regions", type=int, default=None, help="override region count")
    ap.add_argument("--precision", choices=sorted(PRECISIONS) + ["fp64", "fp64m"], default=None,
                    help="GEMM precision of the batched MFMA path (fp16s and fp32 meet "
                         "the 1e-3 dB bound; fp16 and bf16 do not); default fp16s, and "
                         "fp32 for C4 (its configuration names FP32)")
    ap.add_argument("--workload", choices=["C2", "C3", "C4", "C5", "omega", "estep"],
                    default="C2",
                    help="C2 (default) is the headline measurement; C3/C4/C5 print their own "
                         "line for the other scoring configurations, omega the omega test's "
                         "and estep PhyloEM's expected substitution counts")
    args = ap.parse_args()
    if args.precision is None:
        args.precision = "fp32" if args.workload == "C4" else "fp16s"
    if args.precision in ("fp64", "fp64m") and args.workload != "omega":
        ap.error("--precision fp64/fp64m are for --workload omega (the other lines time the MFMA "
                 "paths and report the FP64 kernels beside them)")
    if args.gpus != 1:
        print("bench.py runs on one GPU: this build has no multi-GPU path (regions shard "
              "across GPUs by design only); run with --gpus 1", file=sys.stderr)
        sys.exit(2)
    rank = int(os.environ.get("RANK", "0"))
    if args.workload != "C2":
        other_workload(args)
        return

    import phylocsf_amd as P
    from phylocsf_amd import workloads as W

    ctx = P.Context(0)
    ctx.set_timing(True)
    model = W.model_for("C2", ctx)
    log("simulating C2 ...")
    t0 = time.perf_counter()
    wl = W.make("C2", model, n_regions=args.regions, n_threads=16)
    log("simulated %d regions, %d columns in %.1f s" % (wl.n_regions, wl.n_cols,
                                                         time.perf_counter() - t0))
    batch = P.Batch(model, wl.codes, wl.offsets)
    nl = model.tree.nl
    ncols = wl.n_cols

    # ---- batched FP32 MFMA path (the measured value); the autotuner picks the
    # launch shape before the timed region
    prec, dtype, peak, mfma_per_flop = PRECISIONS[args.precision]
    batch.forward(1.0, prec)
    chosen, shape_ms = batch.autotune(reps=2, precision=prec)
    log("autotune: " + ", ".join("v%d %.2f ms" % (i, t) for i, t in enumerate(shape_ms)) +
        " -> v%d" % chosen)
    for _ in range(args.warmup):
        batch.forward(1.0, prec)
    ctx.synchronize()
    kernel_ms = []
    t0 = time.perf_counter()
    for _ in range(args.steps):
        ll32, el32 = batch.forward(1.0, prec)
        kernel_ms.append(batch.timing()["prune_ms"])
    ctx.synchronize()
    elapsed = time.perf_counter() - t0
    value = ncols * args.steps / elapsed
    ms_per_step = 1000 * elapsed / args.steps
    log("%s MFMA: %.3f ms/step, %.3e col/s, kernel %.3f ms" % (args.precision, ms_per_step, value,
                                                              np.mean(kernel_ms)))

    # ---- the same batch at FP32 (the configuration's stated GEMM precision)
    fp32 = None
    if args.precision != "fp32":
        batch.forward(1.0, P.PCSF_FP32)
        c32, _ = batch.autotune(reps=2, precision=P.PCSF_FP32)
        for _ in range(args.warmup):
            batch.forward(1.0, P.PCSF_FP32)
        ctx.synchronize()
        k32 = []
        t0 = time.perf_counter()
        for _ in range(args.steps):
            llf, _ = batch.forward(1.0, P.PCSF_FP32)
            k32.append(batch.timing()["prune_ms"])
        ctx.synchronize()
        el32 = time.perf_counter() - t0
        fp32 = {"value": ncols * args.steps / el32, "unit": "codon-columns/s",
                "ms_per_step": 1000 * el32 / args.steps, "kernel_ms": float(np.mean(k32)),
                "launch_shape": int(c32),
                "achieved_tflops": 16384.0 * (nl - 2) * ncols / (np.mean(k32) / 1000) / 1e12,
                "peak_tflops": PEAK_FP32_TFLOPS}
        log("fp32 MFMA: %.3f ms/step, %.3e col/s" % (fp32["ms_per_step"], fp32["value"]))

    # ---- plain FP64 GPU kernel on the same batch (the 50x denominator)
    fp64_rate, dmax, fp64m = None, None, None
    if args.fp64_steps > 0:
        batch.forward(1.0, P.PCSF_FP64)
        ctx.synchronize()
        t0 = time.perf_counter()
        for _ in range(args.fp64_steps):
            ll64, el64 = batch.forward(1.0, P.PCSF_FP64)
        ctx.synchronize()
        fp64_rate = ncols * args.fp64_steps / (time.perf_counter() - t0)
        log("FP64 plain: %.3e col/s" % fp64_rate)
        s32 = P.DB * (ll32[:, 0] - ll32[:, 1])
        s64 = P.DB * (ll64[:, 0] - ll64[:, 1])
        dmax = float(np.max(np.abs(s32 - s64))) if len(s32) else 0.0
        if fp32 is not None:
            fp32["max_abs_delta_score_db_vs_fp64"] = float(
                np.max(np.abs(P.DB * (llf[:, 0] - llf[:, 1]) - s64)))
        # the FP64 forward on matrix cores (the CLI's default precision)
        llm, _ = batch.forward(1.0, P.PCSF_FP64M)
        ctx.synchronize()
        t0 = time.perf_counter()
        nm_steps = max(2, args.fp64_steps)
        for _ in range(nm_steps):
            batch.forward(1.0, P.PCSF_FP64M)
        ctx.synchronize()
        fp64m = {"value": ncols * nm_steps / (time.perf_counter() - t0),
                 "unit": "codon-columns/s",
                 "max_abs_delta_score_db_vs_fp64": float(
                     np.max(np.abs(P.DB * (llm[:, 0] - llm[:, 1]) - s64))) if len(s64) else 0.0}
        log("FP64 matrix cores: %.3e col/s" % fp64m["value"])

    # ---- roofline of the dominant kernel (fused pruning: MFMA-bound; its HBM
    # traffic is the codes plus per-column results, ~nl + 32 B per column)
    flops_per_col = 16384.0 * (nl - 2)          # 2 models x (L-2) x 2*64^2
    kavg_s = float(np.mean(kernel_ms)) / 1000.0
    achieved = flops_per_col * ncols / kavg_s / 1e12
    # executed matrix-core work: GEMMs left after the cherry tables, times the
    # MFMA products per GEMM (three for split FP16)
    from phylocsf_amd.api import program
    ops, _ = program(model.tree, with_tables=1)
    n_gemm = int(np.sum((np.asarray(ops).reshape(-1, 4)[:, 0] & 0xff) == 2))
    exec_per_col = 2 * n_gemm * 2 * 64 * 64 * mfma_per_flop
    executed = exec_per_col * ncols / kavg_s / 1e12
    roofline = dict(bound="mfma", achieved=achieved, peak=peak, unit="TFLOP/s",
                    frac=achieved / peak, traffic=(profiled_traffic(args.precision, chosen)
                                          if args.regions is None else None),
                    algorithmic_bytes=float(ncols * (nl + 32)),
                    executed_mfma_tflops=executed, executed_frac=executed / peak,
                    gemms_per_column_per_model=n_gemm,
                    # the level-by-level design's HBM roofline (partials written and
                    # read once per tree level: 2*(L-2)*2*64*4 + L + 32 B per column)
                    hbm_roofline_of_level_design={
                        "bytes_per_column": 2 * (nl - 2) * 2 * 64 * 4 + nl + 32,
                        "columns_per_s": PEAK_HBM_GBS * 1e9 / (2 * (nl - 2) * 2 * 64 * 4 + nl + 32),
                        "value_over_it": value / (PEAK_HBM_GBS * 1e9 /
                                                  (2 * (nl - 2) * 2 * 64 * 4 + nl + 32))},
                    kernel="prune_mfma_kernel", kernel_ms=float(np.mean(kernel_ms)),
                    launch_shape=int(chosen),
                    autotune_ms=[round(float(t), 3) for t in shape_ms],
                    flops_per_column=flops_per_col)

    cpu = None
    if rank == 0 and args.cpu_seconds > 0:
        cpu = cpu_baseline(wl, args.cpu_seconds, args.cpu_threads)
        log("CPU oracle: %.1f col/s on %d threads, %.1f col/s on 1" % (
            cpu["value"], cpu["cores"], cpu["value_1_thread"]))

    out = {
        "metric": METRIC,
        "value": value,
        "unit": "codon-columns/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": dtype,
        "data": "synthetic",
        "config": {
            "workload": "58mammals, %d regions, lengths U{50..2000}, 50/50 coding/noncoding, "
                        "gap p=0.1, 1%% N, fixed strategy (seed 2)" % wl.n_regions,
            "codon_columns": ncols,
            "models": 2,
            "precision": "%s MFMA GEMM inputs, FP32 partials, FP64 P(t) and log-scaling"
                         % args.precision,
        },
        "plain_fp64_gpu": {"value": fp64_rate, "unit": "codon-columns/s",
                           "speedup_of_value": value / fp64_rate if fp64_rate else None},
        "fp64_matrix_cores": fp64m,
        "accuracy": {"max_abs_delta_score_db_vs_fp64": dmax, "bound_db": 1e-3},
        "fp32": fp32,
        "roofline": roofline,
        "cpu_baseline": cpu,
    }
    if rank == 0:
        print(json.dumps(out), flush=True)


def _line(metric, value, unit, steps, warmup, ms, dtype, config, **extra):
    out = {"metric": metric, "value": value, "unit": unit, "n_gpus": 1, "steps": steps,
           "warmup": warmup, "ms_per_step": ms, "higher_is_better": True, "scaling": "weak",
           "vs_baseline": None, "dtype": dtype, "data": "synthetic", "config": config}
    out.update(extra)
    print(json.dumps(out), flush=True)


def other_workload(args):
    """The other scoring configurations, one JSON line each (not the headline):
      C3  100vertebrates, 2,000 regions, mle: the reference's maximiser and the
          gradient maximiser (adjoint d lpr/d rho), end-to-end scored columns/s
          and likelihood evaluations/s;
      C4  120mammals, 100,000 regions, 6 reading frames built on the GPU from
          nucleotide rows, FP32 (autotuned) fixed scoring of every frame;
      C5  100vertebrates, 20,000 regions, every GEMM precision against FP64
          (error per precision), and per-branch gradients at the mle estimates.
    """
    import phylocsf_amd as P
    from phylocsf_amd import workloads as W

    ctx = P.Context(0)
    ctx.set_timing(True)
    name = args.workload
    if name in ("omega", "estep"):
        return extra_workload(args, P, W, ctx)
    model = W.model_for(name, ctx)
    prec, dtype, _, _ = PRECISIONS[args.precision]
    steps = max(1, args.steps)
    t0 = time.perf_counter()
    if name == "C4":
        lengths, rmodel, rrho, kw = W.region_params("C4", args.regions)
        nuc, noff = model.simulate_nucleotides(lengths, rmodel, rrho, seed=W.SPECS["C4"]["seed"],
                                               n_threads=16, **kw)
        log("simulated %d regions, %d nucleotides in %.1f s" % (len(lengths), noff[-1],
                                                                 time.perf_counter() - t0))
        t1 = time.perf_counter()
        batch = P.Batch.frames(model, nuc, noff, 6)
        build_s = time.perf_counter() - t1
        del nuc
        log("6-frame batch: %d regions, %d codon columns, built in %.2f s (upload included)"
            % (batch.n_regions, batch.n_cols, build_s))
        batch.forward(1.0, prec)
        chosen, shape_ms = batch.autotune(reps=1, precision=prec)
        for _ in range(args.warmup):
            batch.forward(1.0, prec)
        ctx.synchronize()
        kms = []
        t1 = time.perf_counter()
        for _ in range(steps):
            ll, el = batch.forward(1.0, prec)
            kms.append(batch.timing()["prune_ms"])
        ctx.synchronize()
        el_s = time.perf_counter() - t1
        score = P.DB * (ll[:, 0] - ll[:, 1])
        best_frame = score.reshape(-1, 6).argmax(axis=1)
        _line("codon-columns/s, 6-frame fixed scoring (%s MFMA)" % args.precision.upper(),
              batch.n_cols * steps / el_s,
              "codon-columns/s", steps, args.warmup, 1000 * el_s / steps, dtype,
              {"workload": "120mammals, %d regions U{20..600}, 30/70 coding/noncoding, gap runs "
                           "(mean 5 codons) covering 0.25, 6 reading frames, fixed strategy (seed 4)"
                           % len(lengths), "codon_columns": batch.n_cols},
              kernel_ms=float(np.mean(kms)), launch_shape=int(chosen),
              autotune_ms=[round(float(t), 3) for t in shape_ms],
              frame_batch_build_s=build_s,
              best_frame_histogram=np.bincount(best_frame, minlength=6).tolist())
        return
    wl = W.make(name, model, n_regions=args.regions, n_threads=16)
    log("simulated %d regions, %d columns in %.1f s" % (wl.n_regions, wl.n_cols,
                                                         time.perf_counter() - t0))
    batch = P.Batch(model, wl.codes, wl.offsets)
    if name == "C3":
        res = {}
        for strat in ("mle", "mle_grad"):
            P.score(batch, strat, prec)  # warm-up (builds the aligned layout)
            ctx.synchronize()
            t1 = time.perf_counter()
            r = P.score(batch, strat, prec)
            dt = time.perf_counter() - t1
            res[strat] = (r, dt)
            log("%s: %.2f s, %.3e scored col/s, %d evaluations per region (mean)"
                % (strat, dt, wl.n_cols / dt, r["n_evals"].mean()))
        t1 = time.perf_counter()
        r64 = P.score(batch, "mle", P.PCSF_FP64)
        dtr64 = time.perf_counter() - t1
        t1 = time.perf_counter()
        r64m = P.score(batch, "mle", P.PCSF_FP64M)
        dtr64m = time.perf_counter() - t1
        t1 = time.perf_counter()
        g64 = P.score(batch, "mle_grad", P.PCSF_FP64)
        dt64 = time.perf_counter() - t1
        log("FP64 mle %.2f s, FP64M mle %.2f s, FP64 mle_grad %.2f s" % (dtr64, dtr64m, dt64))
        g, dtg = res["mle_grad"]
        b, dtb = res["mle"]
        # finite-difference check of d lpr / d rho at the gradient estimates
        # (FP64 forwards, central, step 1e-5 relative to rho: estimates reach
        # the 0.01 bound, where an absolute 1e-5 step is not small), and the
        # gradient pass of the measured precision against the FP64 adjoint
        rho = g["rho"]
        batch.forward(rho, P.PCSF_FP64, per_region=True)
        d, _ = batch.backward(np.ones_like(rho))
        h = 1e-5 * rho
        lp, _ = batch.forward(rho + h, P.PCSF_FP64, per_region=True)
        lm, _ = batch.forward(rho - h, P.PCSF_FP64, per_region=True)
        fd = (lp - lm) / (2 * h)
        fd_err = float(np.max(np.abs(d - fd) / np.maximum(1.0, np.abs(fd))))
        batch.forward(rho, prec, per_region=True)
        dp, _ = batch.backward(np.ones_like(rho))
        grad_err = float(np.max(np.abs(dp - d) / np.maximum(1.0, np.abs(d))))
        _line("scored codon-columns/s, mle with the adjoint (gradient) maximiser",
              wl.n_cols / dtg, "codon-columns/s", 1, 1, 1000 * dtg, dtype,
              {"workload": "100vertebrates, %d regions lognormal(300, 0.8) clipped 30..5000, "
                           "coding, tree scale U(0.3, 3.0), mle (seed 3)" % wl.n_regions,
               "codon_columns": wl.n_cols},
              # n_evals per region = per model: 7 shared-grid forwards, the
              # gradient passes, 1 final FP64 forward
              gradient_passes_per_pair=float((g["n_evals"].mean() - 16.0) / 2.0),
              gradient_evaluations_per_s=float(wl.n_cols * (g["n_evals"].mean() - 16.0) / dtg),
              reference_maximiser={"scored_columns_per_s": wl.n_cols / dtb,
                                   "evaluations_per_s": float(wl.n_cols * b["n_evals"].mean() / dtb),
                                   "mean_evaluations": float(b["n_evals"].mean())},
              mean_evaluations_per_region=float(g["n_evals"].mean()),
              max_abs_delta_score_db={
                  "mle_grad_vs_fp64_mle_grad": float(np.max(np.abs(g["score"] - g64["score"]))),
                  "mle_grad_vs_fp64_mle": float(np.max(np.abs(g["score"] - r64["score"]))),
                  "mle_vs_fp64_mle": float(np.max(np.abs(b["score"] - r64["score"])))},
              fp64_gradient_maximiser={"scored_columns_per_s": wl.n_cols / dt64,
                                       "gradient_passes_per_pair":
                                           float((g64["n_evals"].mean() - 16.0) / 2.0)},
              fp64_reference_maximiser={
                  "plain_kernel_scored_columns_per_s": wl.n_cols / dtr64,
                  "matrix_core_kernel_scored_columns_per_s": wl.n_cols / dtr64m,
                  "max_abs_delta_score_db_matrix_core_vs_plain":
                      float(np.max(np.abs(r64m["score"] - r64["score"])))},
              lpr_gain_of_gradient_maximiser=float(np.max(g["lpr"] - r64["lpr"])),
              fd_rel_error_of_d_lpr_d_rho=fd_err,
              gradient_rel_error_vs_fp64_adjoint=grad_err)
        return
    # C5: every precision against FP64, fixed strategy; per-branch gradients at the mle
    ctx.synchronize()
    t1 = time.perf_counter()
    ll64, _ = batch.forward(1.0, P.PCSF_FP64)
    dt64 = time.perf_counter() - t1
    s64 = P.DB * (ll64[:, 0] - ll64[:, 1])
    per = {"fp64": {"codon_columns_per_s": wl.n_cols / dt64}}
    batch.forward(1.0, P.PCSF_FP64M)
    ctx.synchronize()
    t1 = time.perf_counter()
    llm, _ = batch.forward(1.0, P.PCSF_FP64M)
    dtm = time.perf_counter() - t1
    sm = P.DB * (llm[:, 0] - llm[:, 1])
    per["fp64m"] = {"codon_columns_per_s": wl.n_cols / dtm,
                    "max_abs_delta_score_db": float(np.max(np.abs(sm - s64))),
                    "max_rel_loglik_error": float(np.max(np.abs(llm - ll64) /
                                                         np.maximum(1.0, np.abs(ll64))))}
    log("fp64: %.3e col/s; fp64m: %.3e col/s, max |d score| %.3g dB"
        % (wl.n_cols / dt64, wl.n_cols / dtm, per["fp64m"]["max_abs_delta_score_db"]))
    for pname in ("fp32", "fp16s", "fp16", "bf16"):
        pc = PRECISIONS[pname][0]
        batch.forward(1.0, pc)
        batch.autotune(reps=1, precision=pc)
        ctx.synchronize()
        t1 = time.perf_counter()
        for _ in range(steps):
            ll, _ = batch.forward(1.0, pc)
        ctx.synchronize()
        dt = (time.perf_counter() - t1) / steps
        s = P.DB * (ll[:, 0] - ll[:, 1])
        rel = float(np.max(np.abs(ll - ll64) / np.maximum(1.0, np.abs(ll64))))
        per[pname] = {"codon_columns_per_s": wl.n_cols / dt,
                      "max_abs_delta_score_db": float(np.max(np.abs(s - s64))),
                      "mean_abs_delta_score_db": float(np.mean(np.abs(s - s64))),
                      "max_rel_loglik_error": rel}
        log("%s: %.3e col/s, max |d score| %.3g dB, max rel lpr %.3g"
            % (pname, wl.n_cols / dt, per[pname]["max_abs_delta_score_db"], rel))
    # gradient maximiser and per-branch gradients (FP64 outside pass) on the
    # first regions of the workload (about 1.5 M columns)
    nsub = min(wl.n_regions, 1000)
    sub = P.Batch(model, wl.codes[:wl.offsets[nsub]], wl.offsets[:nsub + 1])
    sub_cols = int(wl.offsets[nsub])
    t1 = time.perf_counter()
    g = P.score(sub, "mle_grad", prec)
    dt_mle = time.perf_counter() - t1
    sub.forward(g["rho"], P.PCSF_FP64, per_region=True)
    ctx.synchronize()
    t1 = time.perf_counter()
    _, gbr = sub.backward(np.ones_like(g["rho"]), branches=True)
    dt_b = time.perf_counter() - t1
    best = per[args.precision]
    _line("codon-columns/s per GEMM precision, fixed scoring", best["codon_columns_per_s"],
          "codon-columns/s", steps, 1, 1000 * wl.n_cols / best["codon_columns_per_s"], dtype,
          {"workload": "100vertebrates, %d regions U{100..3000}, 50/50 coding/noncoding, gap p=0.15"
                       " (seed 5)" % wl.n_regions, "codon_columns": wl.n_cols},
          precisions=per,
          gradient_sample={"regions": nsub, "codon_columns": sub_cols,
                           "mle_grad_scored_columns_per_s": sub_cols / dt_mle,
                           "gradient_passes_per_pair": float((g["n_evals"].mean() - 16.0) / 2.0),
                           "per_branch_gradient_columns_per_s": sub_cols / dt_b,
                           "branches": int(gbr.shape[-1])})


def extra_workload(args, P, W, ctx):
    """The rows next to the scoring path, one JSON line each:
      omega  the omega test (OmegaModel.score) on 58mammals regions simulated as
             the C2 workload, lengths U{50..500}: every region's two MAP fits in
             lock step on the GPU; scored codon-columns/s, evaluations per region,
             error against the FP64 run, and the FP64 oracle on a sample;
      estep  PhyloEM's expected substitution counts (collect_sufficient_statistics,
             every branch posterior summed over columns) for the 58mammals coding
             model over C2 columns: codon-columns/s against the FP64 oracle's.
    """
    from oracle import oracle as O
    from phylocsf_amd.api import paramset_prefix
    prefix = paramset_prefix("58mammals")
    tree = P.newick_tree(open(prefix + ".nh").read())
    otree = O.newick_tree(open(prefix + ".nh").read())
    sim = P.PhyloCSFModel.from_paramset(None, "58mammals")
    if args.workload == "omega":
        R = args.regions or 200
        rng = np.random.default_rng(6)
        lengths = rng.integers(50, 501, R)
        codes, off = sim.simulate(lengths, rng.integers(0, 2, R), np.ones(R), gap_p=0.1,
                                  n_p=0.01, seed=6, n_threads=16)
        batch = P.Batch(P.PhyloCSFModel.tree_only(ctx, tree), codes, off)
        prec, dtype, _, _ = (PRECISIONS[args.precision] if args.precision in PRECISIONS else
                             {"fp64": (0, "fp64", None, None),
                              "fp64m": (5, "fp64 (matrix cores)", None, None)}[args.precision])
        t0 = time.perf_counter()
        r = P.score_omega(batch, precision=prec)
        dt = time.perf_counter() - t0
        log("omega (%s): %.2f s, %.3e scored col/s" % (args.precision, dt, off[-1] / dt))
        err = None
        if prec != 0:
            t1 = time.perf_counter()
            r64 = P.score_omega(batch, precision=0)
            log("omega (fp64): %.2f s" % (time.perf_counter() - t1))
            err = float(np.max(np.abs(r["score"] - r64["score"])))
        # FP64 oracle on the first regions, about cpu-seconds of one thread
        t0 = time.perf_counter()
        nr = nc = 0
        while nr < R and time.perf_counter() - t0 < args.cpu_seconds:
            O.omega_score(otree, codes[off[nr]:off[nr + 1]])
            nc += int(off[nr + 1] - off[nr])
            nr += 1
        cdt = time.perf_counter() - t0
        _line("scored codon-columns/s, omega test (two MAP fits per region)", off[-1] / dt,
              "codon-columns/s", 1, 0, 1000 * dt, dtype,
              {"workload": "58mammals tree, %d regions U{50..500} simulated under the C2 "
                           "models (gaps 0.1, 1%% N), omega test, H1 omega 0.2 sigma 0.01 "
                           "(seed 6)" % R, "codon_columns": int(off[-1])},
              mean_evaluations_per_region=float(r["n_evals"].mean()),
              failed_regions=int((r["status"] != 0).sum()),
              max_abs_delta_score_db_vs_fp64=err,
              cpu_baseline={"value": nc / cdt, "unit": "codon-columns/s", "cores": 1,
                            "kind": "port",
                            "sample": "regions 0..%d (%d codon columns), FP64 oracle, %.1f s "
                                      "on 1 thread of %s" % (nr - 1, nc, cdt, cpu_model())})
        return
    # estep
    wl = W.make("C2", sim, n_regions=args.regions or 600, n_threads=16)
    leaves = np.where(wl.codes == 255, -1, wl.codes.astype(np.int32))
    S, pi = O.ecm_read(prefix + "_coding.ECM")
    Q = O.build_q(S, pi)
    Pm = np.stack([O.pt(Q, pi, t) for t in otree.branch[:-1]])
    n = tree.n
    lik = P.PhyloLik(ctx)
    it = lik.prepare(tree, Pm, pi, leaves[:100])
    it.add_branch_posteriors(np.zeros((n - 1, 64, 64)))  # warm-up
    steps = max(1, args.steps)
    t0 = time.perf_counter()
    for _ in range(steps):
        ss = np.zeros((n - 1, 64, 64))
        P.collect_sufficient_statistics(ctx, tree, Pm, pi, leaves, ss)
    dt = (time.perf_counter() - t0) / steps
    ncols = leaves.shape[0]
    log("E-step: %d columns in %.3f s, %.3e col/s" % (ncols, dt, ncols / dt))
    t0 = time.perf_counter()
    nsmp = 0
    while nsmp < ncols and time.perf_counter() - t0 < args.cpu_seconds:
        O.lik_posteriors(otree, Pm, pi, leaves[nsmp:nsmp + 50], posteriors=False)
        nsmp += 50
    cdt = time.perf_counter() - t0
    _z, _p, ec0 = O.lik_posteriors(otree, Pm, pi, leaves[:200], posteriors=False)
    ss = np.zeros((n - 1, 64, 64))
    P.collect_sufficient_statistics(ctx, tree, Pm, pi, leaves[:200], ss)
    err = float(np.max(np.abs(ss - ec0)) / np.max(np.abs(ec0)))
    _line("codon-columns/s, PhyloEM expected substitution counts (every branch)", ncols / dt,
          "codon-columns/s", steps, 1, 1000 * dt, "fp64",
          {"workload": "58mammals coding model, %d columns of the C2 workload (seed 2), "
                       "branch posteriors of all %d branches summed over columns"
                       % (ncols, n - 1), "codon_columns": ncols},
          max_error_vs_oracle_first_200_columns_relative_to_max_count=err,
          cpu_baseline={"value": nsmp / cdt, "unit": "codon-columns/s", "cores": 1,
                        "kind": "port",
                        "sample": "first %d columns, FP64 oracle, %.1f s on 1 thread of %s"
                                  % (nsmp, cdt, cpu_model())})


if __name__ == "__main__":
    main()
