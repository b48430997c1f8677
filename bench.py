#!/usr/bin/env python3
"""Benchmark of the MI355X query engine on one GPU.

Workload: exact L2 nearest-neighbour search over r = 7028 synthetic
log-expression references with d = 60660 coordinates, query batches of
B = 4096 (seeded generator in ultrahigh_ann_amd/synthetic.py). One step is
one batch through the engine's C ABI with queries and references already
resident in HBM. The timed region is bracketed by torch.cuda.synchronize()
and the per-step latency comes from HIP events on the launch stream.

`value` is the mixed-precision strategy (BF16 Gram on MFMA, exact binary32
re-rank of a bounded candidate set, binary64 near ties); the plain binary32
direct path is timed beside it on the same batch, and both are checked
against the binary64 reference CPU path on a query sample.

Prints ONE JSON line (metric, value = queries/s, roofline of the dominant
kernel, CPU baseline timed on this host with the reference's CPU path).
"""

from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "queries/s at r=7028 d=60660, mixed vs FP32 at equal argmin agreement; p99 ms"
PEAK = {"fp32": 157.3, "bf16": 2500.0, "fp8": 5000.0, "hbm": 8000.0}


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def cpu_threads() -> int:
    env = os.environ.get("OMP_NUM_THREADS")
    if env and env.isdigit() and int(env) > 0:
        return int(env)
    return max(1, min(16, os.cpu_count() or 1))


def cpu_baseline(refs_np, queries_np, sample):
    """Reference CPU exact L2 (oracle/_ref, OpenMP parallel_queries) on a
    bounded query sample; also the binary64 argmin for agreement."""
    import ctypes

    import oracle

    threads = cpu_threads()
    os.environ.setdefault("OMP_NUM_THREADS", str(threads))
    q = np.ascontiguousarray(queries_np[:sample])
    r, d = refs_np.shape
    idx = np.zeros(sample, np.uint64)
    kind = "reference"
    t0 = time.perf_counter()
    if oracle.ref_available():
        rc = oracle.ref().ref_exact_l2(oracle._p(refs_np, oracle._f32p), r, d, oracle._p(q, oracle._f32p),
                                       sample, 1, oracle._p(idx, oracle._u64p))
        if rc != 0:
            raise RuntimeError("reference CPU scan failed")
    else:
        kind = "port"
        oracle.lib().orc_l2_fp64_scan(oracle._p(refs_np, oracle._f32p), r, d, oracle._p(q, oracle._f32p),
                                      sample, threads, oracle._p(idx, oracle._u64p),
                                      ctypes.POINTER(ctypes.c_double)())
    elapsed = time.perf_counter() - t0
    return {
        "value": sample / elapsed,
        "unit": "queries/s",
        "cores": threads,
        "kind": kind,
        "sample": f"{sample} queries of the same batch, exact L2 binary64 over r={r} d={d}, "
                  f"OpenMP parallel_queries, {elapsed:.2f} s wall",
    }, idx


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--rows", type=int, default=7028)
    ap.add_argument("--dim", type=int, default=60660)
    ap.add_argument("--batch", type=int, default=4096)
    ap.add_argument("--seed", type=int, default=2026)
    ap.add_argument("--cpu-sample", type=int, default=0, help="queries timed on the CPU (0 = auto)")
    ap.add_argument("--no-cpu", action="store_true")
    args = ap.parse_args()

    if args.gpus != 1:
        print(json.dumps({"error": "this engine runs on one GPU; --gpus must be 1",
                          "n_gpus": args.gpus}))
        sys.exit(2)

    import torch

    import ultrahigh_ann_amd as uh
    from ultrahigh_ann_amd import synthetic

    if not torch.cuda.is_available() or not uh.device_available():
        print(json.dumps({"error": "no GPU visible"}))
        sys.exit(1)

    r, d, B = args.rows, args.dim, args.batch
    log(f"generating synthetic data r={r} d={d} B={B} seed={args.seed}")
    refs, queries = synthetic.log_expression(r, B, d, seed=args.seed, device="cuda")
    refs_np = refs.cpu().numpy()
    log("building index (upload references, BF16 copy)")
    index = uh.HipExactL2Index(refs_np)
    index.enable_mixed()
    ws = index.make_query_workspace(B)
    stream = torch.cuda.current_stream()
    out_idx = torch.zeros(B, dtype=torch.int64, device="cuda")
    out_dist = torch.zeros(B, dtype=torch.float32, device="cuda")

    def run(strategy, steps, warmup, label):
        for i in range(warmup):
            index.query_device_async(queries, B, ws, out_idx, out_dist, strategy=strategy, stream=stream)
            torch.cuda.synchronize()
        events = [torch.cuda.Event(enable_timing=True) for _ in range(steps + 1)]
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        events[0].record(stream)
        for i in range(steps):
            index.query_device_async(queries, B, ws, out_idx, out_dist, strategy=strategy, stream=stream)
            events[i + 1].record(stream)
        torch.cuda.synchronize()
        wall = time.perf_counter() - t0
        ms = np.array([events[i].elapsed_time(events[i + 1]) for i in range(steps)])
        log(f"{label}: {steps} steps {wall:.3f} s wall, per-step ms {np.round(ms, 3).tolist()}")
        return wall, ms, out_idx.cpu().numpy().astype(np.uint64)

    fp32_steps = max(1, min(args.steps, 3))
    fp32_wall, fp32_ms, fp32_idx = run(uh.STRATEGY_DIRECT, fp32_steps, 1, "fp32 direct")
    K = args.steps
    wall, step_ms, mixed_idx = run(uh.STRATEGY_GRAM_BF16, K, args.warmup, "bf16 mixed")
    stats = ws.mixed_stats()
    qps = B * K / wall
    fp32_qps = B * fp32_steps / fp32_wall

    # Per-stage HIP-event times of the mixed strategy (same stream), outside
    # the timed region.
    ws.set_timing(True)
    stage = []
    for _ in range(3):
        index.query_device_async(queries, B, ws, out_idx, out_dist, strategy=uh.STRATEGY_GRAM_BF16,
                                 stream=stream)
        stage.append(ws.stage_times())
    ws.set_timing(False)
    stage = np.array(stage).mean(0)
    gemm_ms = float(stage[1])
    flops = 2.0 * B * r * d
    achieved = flops / (gemm_ms * 1e-3) / 1e12
    roofline = {
        "bound": "mfma",
        "achieved": round(achieved, 2),
        "peak": PEAK["bf16"],
        "unit": "TFLOP/s",
        "frac": round(achieved / PEAK["bf16"], 4),
        "traffic": None,
    }

    cpu = None
    agreement = None
    if not args.no_cpu:
        threads = cpu_threads()
        sample = args.cpu_sample or max(8, 2 * threads)
        log(f"CPU baseline: {sample} queries on {threads} threads")
        cpu, fp64_idx = cpu_baseline(refs_np, queries.cpu().numpy(), sample)
        agreement = {
            "sample": sample,
            "fp32_direct": float(np.mean(fp32_idx[:sample] == fp64_idx)),
            "bf16_mixed": float(np.mean(mixed_idx[:sample] == fp64_idx)),
        }
        log(f"CPU baseline {cpu['value']:.2f} q/s; agreement with binary64 {agreement}")

    line = {
        "metric": METRIC,
        "value": round(qps, 2),
        "unit": "queries/s",
        "n_gpus": 1,
        "steps": K,
        "warmup": args.warmup,
        "ms_per_step": round(wall * 1e3 / K, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "bf16",
        "data": "synthetic",
        "config": {
            "workload": f"exact-l2 r={r} d={d} batch={B}: BF16 Gram on MFMA + exact re-rank",
            "references": r,
            "dim": d,
            "batch": B,
            "strategy": "gram-bf16 (fp32 re-rank, fp64 near ties)",
        },
        "p50_ms": round(float(np.percentile(step_ms, 50)), 3),
        "p99_ms": round(float(np.percentile(step_ms, 99)), 3),
        "fp32_direct": {
            "value": round(fp32_qps, 2),
            "unit": "queries/s",
            "ms_per_step": round(fp32_wall * 1e3 / fp32_steps, 3),
            "speedup_mixed_vs_fp32": round(qps / fp32_qps, 2),
        },
        "mixed_stages_ms": {"bf16_convert": round(float(stage[0]), 3), "gram_gemm": round(gemm_ms, 3),
                            "select_rerank": round(float(stage[2]), 3),
                            "escalation": round(float(stage[3]), 3)},
        "mixed_counters": {"avg_candidates": round(stats["candidates"] / B, 3),
                           "fp64_reranked": stats["fp64_reranked"], "escalated": stats["escalated"]},
        "agreement_fp64": agreement,
        "roofline": roofline,
        "cpu_baseline": cpu,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
