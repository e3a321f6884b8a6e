"""Stitched pairs/s and p50/p99 latency on one MI355X: the plain per-pair GPU
path (the reference's execute_run flow: one pair at a time, per-stage
launches, two host synchronizations) beside the batched form, on the
1920x1080 workload (ORB 5000, cross-checked 5000x5000 Hamming matching,
RANSAC 4096, 6-band multiband blend, 30-50 % overlap), inputs resident in
HBM before the timed region. Prints one JSON line.

    python bench.py [--gpus 1] [--steps K] [--warmup W]

A step is one batch of --pairs synthetic pairs through S1-S6.
"""
from __future__ import annotations

import argparse
import ctypes as C
import json
import os
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
for p in (ROOT, ROOT / "crossstage-stitch_amd", ROOT / "tests"):
    sys.path.insert(0, str(p))

W, H, K, R = 1920, 1080, 5000, 4096
BLEND = "multiband_6"
SEED0 = 20261022


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--pairs", type=int, default=16, help="distinct synthetic pairs per step")
    ap.add_argument("--cpu-pairs", type=int, default=1, help="pairs timed on the CPU oracle (0 = skip)")
    ap.add_argument("--no-cpu", action="store_true")
    return ap.parse_args()


def make_pairs(n):
    from cs_stitch import pipeline as P
    return [P.synth_pair(SEED0 + i, W, H, overlap=(0.3, 0.5)) for i in range(n)]


def upload(pairs):
    import torch
    from cs_stitch import stages as st
    dev = []
    for a, b, _ in pairs:
        ta, tb = st.pitched(H, W, 3), st.pitched(H, W, 3)
        ta.copy_(torch.from_numpy(a).cuda())
        tb.copy_(torch.from_numpy(b).cuda())
        dev.append((ta, tb))
    torch.cuda.synchronize()
    return dev


def plain_path(dev_pairs, steps, warmup):
    """One pair at a time through PipelineRunner.execute_run."""
    from cs_stitch import pipeline as P
    runner = P.PipelineRunner(W, H, max_k=K, max_r=R, canvas_width=4 * W, canvas_height=2 * H,
                              workspace_mode=P.F.BLEND_NAMES[BLEND])
    cfg = P.harness_config(S=1.0, K=K, R=R, blend=BLEND, cross_check=True)
    lat, ok, gpu_span = [], 0, []
    stage = np.zeros(6)
    t_total = 0.0
    for step in range(warmup + steps):
        t0 = time.perf_counter()
        for ta, tb in dev_pairs:
            runner.use_device_pair(ta, tb)
            rec = runner.execute_run(cfg)
            if step >= warmup:
                lat.append(rec.t_e2e)
                gpu_span.append(rec.t_gpu_span)
                ok += rec.status == b"OK"
                stage += [rec.t_s1, rec.t_s2, rec.t_s3, rec.t_s4, rec.t_s5, rec.t_s6]
        if step >= warmup:
            t_total += time.perf_counter() - t0
    n = steps * len(dev_pairs)
    runner.close()
    return dict(pairs_per_s=n / t_total, ms_per_step=1e3 * t_total / steps, p50_ms=float(np.percentile(lat, 50)),
                p99_ms=float(np.percentile(lat, 99)), ok=ok, n=n, gpu_span_ms=float(np.median(gpu_span)),
                stage_ms=dict(zip(["s1", "s2", "s3", "s4", "s5", "s6"], (stage / n).round(4).tolist())))


def cpu_baseline(pairs, n):
    import pipeline_oracle
    t0 = time.perf_counter()
    for a, b, _ in pairs[:n]:
        pipeline_oracle.run(a, b, 1.0, K, R, BLEND, True)
    dt = time.perf_counter() - t0
    return dict(value=n / dt, unit="pairs/s", cores=1, kind="port",
                sample=f"{n} pair(s) of {W}x{H}, K={K}, R={R}, cross-check, {BLEND}, single-thread CPU oracle "
                       f"({dt:.1f} s)")


def main():
    args = parse()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (--gpus {args.gpus} requested): every pair is independent, "
              "so the per-GPU figure is the measurement; see DESIGN.md for how the path shards.", file=sys.stderr)
        sys.exit(2)
    import torch
    assert torch.cuda.is_available(), "bench.py needs a HIP device"
    from cs_stitch import _ffi
    _ffi.load()
    pairs = make_pairs(args.pairs)
    dev = upload(pairs)
    plain = plain_path(dev, args.steps, args.warmup)
    cpu = None if args.no_cpu or args.cpu_pairs <= 0 else cpu_baseline(pairs, args.cpu_pairs)
    line = {
        "metric": "stitched pairs/s + p50/p99 latency, batched/served vs plain per-pair, 1 GPU",
        "value": round(plain["pairs_per_s"], 3),
        "unit": "pairs/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(plain["ms_per_step"], 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u8/int32/f32/f64",
        "data": "synthetic",
        "config": {"workload": f"plain per-pair path, {W}x{H} RGB8 pairs, 30-50% overlap, ORB {K}, "
                               f"cross-checked {K}x{K} Hamming, RANSAC {R}, {BLEND}",
                   "pairs_per_step": args.pairs},
        "plain": plain,
        "cpu_baseline": cpu,
    }
    print(json.dumps(line))


if __name__ == "__main__":
    main()
