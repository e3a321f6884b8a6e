"""Stitched pairs/s and p50/p99 latency of the batched and served forms beside
the plain per-pair GPU path, on one MI355X. Prints one JSON line.

    python bench.py [--gpus 1] [--steps K] [--warmup W]

Workload (value): the batched form on a ragged batch of 64 synthetic pairs
- resolutions drawn from {640x480, 1280x720, 1920x1080, 3840x2160}, keypoint
caps from U[500, 8000], overlap U[20 %, 80 %], cross-checked Hamming matching,
RANSAC 4096, 6-band multiband blend. A step is that batch through S1-S6, with
the inputs resident in HBM before the timed region; `value` is the stream of
such batches (each batch's S1-S5 overlap the previous batch's S6 on a second
stream), and "batched" times one batch at a time. The same 64 pairs run
through the plain per-pair runner (the reference's execute_run flow: one
pair at a time, per-stage launches, two host synchronizations) for the ratio.
"c1_1080p" runs the 1080p load the same way: 64 pairs shaped like the
single-pair 1080p config (1920x1080, 30-50 % overlap, K 5000, cross-check,
RANSAC 4096, multiband_6), batched and streamed, against the plain runner on
the same pairs. The served form replays Poisson
arrival traces (adaptive legacy-v1 controller, 64 GB HBM budget, textured,
low-texture, low-overlap and repetitive scenes) and reports p50/p99.

"roofline": after the timed regions, K more batches run with direct launches
and HIP events around every S6 kernel on the render stream; the kernel with
the largest summed time is reported with its compulsory HBM bytes per launch
(computed from each canvas's pixel counts) over its mean launch duration,
against the 8 TB/s HBM peak. "traffic" is the measured HBM bytes per launch
of that kernel from the committed rocprofv3 PMC summary, when it covers it.
"""
from __future__ import annotations

import argparse
import gc
import json
import os
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
for p in (ROOT, ROOT / "crossstage-stitch_amd", ROOT / "tests"):
    sys.path.insert(0, str(p))

SIZES = [(640, 480), (1280, 720), (1920, 1080), (3840, 2160)]
R, BLEND = 4096, "multiband_6"
SEED0 = 20261022
POOL_BYTES = 48 << 30
HBM_PEAK_GBS = 8000.0  # MI355X HBM3E, published
# Measured HBM bytes per launch of the dominant kernel on this workload
# (rocprofv3 FETCH_SIZE + WRITE_SIZE, scaled by a copy-kernel calibration).
TRAFFIC_PROFILE = ROOT / "profiles" / "s6_warp_pull_mixed64_traffic.md"
COPY_CALIBRATION_BYTES = 1 << 30
SERVE_BUDGET = 64 << 30


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--pairs", type=int, default=64, help="pairs per batch")
    ap.add_argument("--serve-seconds", type=float, default=1.5, help="length of each arrival trace")
    ap.add_argument("--no-serve", action="store_true")
    ap.add_argument("--no-plain", action="store_true")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-subset", action="store_true", help="skip the 64-pair 1920x1080 batch (profiling runs)")
    ap.add_argument("--no-stream", action="store_true",
                    help="skip the stream of batches, so no two batches' kernels overlap (kernel-trace runs); "
                         "value is then the one-batch-at-a-time rate")
    ap.add_argument("--cpu-seconds", type=float, default=15.0, help="budget of the CPU oracle sample")
    ap.add_argument("--tune-cache", default=None)
    ap.add_argument("--no-tune", action="store_true", help="default launch shapes (profiling runs)")
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="first run one plain 1 GiB device copy (the reference kernel for PMC byte counters)")
    return ap.parse_args()


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def batch_workload(n):
    """(w, h, K, overlap) per pair, seeded."""
    rng = np.random.default_rng(SEED0)
    out = []
    for i in range(n):
        w, h = SIZES[int(rng.integers(len(SIZES)))]
        out.append((w, h, int(rng.integers(500, 8001)), float(rng.uniform(0.2, 0.8))))
    return out


def synth(seed, w, h, overlap, kind=0):
    from cs_stitch import pipeline as P
    return P.synth_pair(seed, w, h, overlap=(overlap, overlap), kind=kind)


def upload(a, b):
    import torch
    from cs_stitch import stages as st
    h, w = a.shape[:2]
    ta, tb = st.pitched(h, w, 3), st.pitched(h, w, 3)
    ta.copy_(torch.from_numpy(a).cuda())
    tb.copy_(torch.from_numpy(b).cuda())
    return ta, tb


def specs_for(dev, work):
    from cs_stitch import batch as B
    return [B.pair_spec(ta, tb, S=1.0, K=k, R=R, blend=BLEND, cross_check=True) for (ta, tb), (_, _, k, _) in
            zip(dev, work)]


def time_batched(br, specs, steps, warmup):
    import torch
    for _ in range(warmup):
        br.run(specs)
    torch.cuda.synchronize()
    times, last = [], None
    t0 = time.perf_counter()
    for _ in range(steps):
        s = time.perf_counter()
        last, tm = br.run(specs)
        times.append(time.perf_counter() - s)
    torch.cuda.synchronize()
    total = time.perf_counter() - t0
    return total, times, last, tm


def time_stream(br, specs, steps, warmup):
    """Stream of batches: each submit runs the batch's S1-S5 on the feature
    stream while the previous batch's S6 runs on the render stream. The timed
    region starts with nothing pending and ends once the last batch is done."""
    import torch
    for _ in range(warmup):
        br.submit(specs)
    br.finish()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        br.submit(specs)
    last = br.finish()
    torch.cuda.synchronize()
    return time.perf_counter() - t0, last


def probe_kernels(br, specs, steps):
    """Per-kernel S6 timing over `steps` batches: direct launches with HIP
    events around every kernel on the render stream, plus each launch's
    compulsory HBM bytes computed from the canvases and their pixel counts."""
    import torch
    br.set_probe(True)
    for _ in range(steps):
        br.run(specs)
    torch.cuda.synchronize()
    stats = br.probe_read()
    br.set_probe(False)
    return stats


def measured_traffic(kernel, algorithmic_bytes):
    """Bytes per launch from the committed PMC summary, or None. The summary
    names the kernel and the compulsory bytes per launch it was measured at;
    a different kernel version or workload gives other bytes, and then the
    measurement does not apply."""
    import re
    if not TRAFFIC_PROFILE.exists():
        return None
    text = TRAFFIC_PROFILE.read_text()
    entries = re.findall(r"^kernel:\s*(.+?)\s*\nalgorithmic_bytes_per_launch:\s*(\d+)\s*\n"
                         r"traffic_bytes_per_launch:\s*(\d+)\s*$", text, re.M)
    for name, alg, traffic in entries:
        if name == kernel and alg == str(int(algorithmic_bytes)):
            return int(traffic)
    return None


def roofline(stats):
    dom = max(stats, key=lambda e: e["total_ms"])
    launch_ms = dom["total_ms"] / dom["launches"]
    per_launch = dom["algorithmic_bytes"] / dom["launches"]
    achieved = per_launch / (launch_ms * 1e-3) / 1e9
    name = f"{dom['kernel']} level {dom['level']}"
    traffic = measured_traffic(name, per_launch)
    top = sorted(stats, key=lambda e: -e["total_ms"])[:10]
    return dict(bound="hbm", achieved=round(achieved, 1), peak=HBM_PEAK_GBS, unit="GB/s",
                frac=round(achieved / HBM_PEAK_GBS, 4), traffic=traffic, kernel=name,
                launch_ms=round(launch_ms, 4), algorithmic_bytes=int(per_launch),
                kernels=[dict(kernel=f"{e['kernel']} level {e['level']}", launches=e["launches"],
                              ms_per_launch=round(e["total_ms"] / e["launches"], 4),
                              gbs=round(e["algorithmic_bytes"] / (e["total_ms"] * 1e-3) / 1e9, 1)) for e in top])


def plain_runners(dev, work, outcomes):
    """One plain runner per canvas box, so every pair the batch stitched fits a
    runner (a box may hold at most 16,777,216 pixels, like the reference)."""
    from cs_stitch import pipeline as P
    boxes = {}
    assign = []
    for (ta, _), (w, h, k, _), o in zip(dev, work, outcomes):
        cw = max(o.geometry.width, w)
        ch = max(o.geometry.height, h)
        key = (w, h, (cw + 255) // 256 * 256, (ch + 127) // 128 * 128)
        if key[2] * key[3] > 16777216:
            key = (w, h, cw, ch)
        assign.append(key)
        boxes.setdefault(key, 0)
    runners = {}
    for key in boxes:
        w, h, cw, ch = key
        runners[key] = P.PipelineRunner(w, h, max_k=8000, max_r=R, canvas_width=cw, canvas_height=ch,
                                        workspace_mode=P.F.BLEND_NAMES[BLEND])
    return runners, assign


def time_plain(dev, work, runners, assign, passes=1):
    import torch
    from cs_stitch import pipeline as P
    lat, ok = [], 0
    t_total = 0.0
    for rep in range(passes + 1):  # first pass warms every runner
        t0 = time.perf_counter()
        for (ta, tb), (_, _, k, _), key in zip(dev, work, assign):
            r = runners[key]
            r.use_device_pair(ta, tb)
            rec = r.execute_run(P.harness_config(S=1.0, K=k, R=R, blend=BLEND, cross_check=True))
            if rep:
                lat.append(rec.t_e2e)
                ok += rec.status == b"OK"
        torch.cuda.synchronize()
        if rep:
            t_total += time.perf_counter() - t0
    n = passes * len(dev)
    return dict(pairs_per_s=n / t_total, ms_per_pair=1e3 * t_total / n, p50_ms=float(np.percentile(lat, 50)),
                p99_ms=float(np.percentile(lat, 99)), ok=int(ok), n=n)


def serve_pool(n):
    """Served scenes: sizes as the batch; half textured at 20-80 % overlap, the
    rest hard: low texture (blobs / 10), low overlap (10-25 %), repetitive."""
    rng = np.random.default_rng(SEED0 + 7)
    out = []
    for i in range(n):
        w, h = SIZES[int(rng.integers(len(SIZES)))]
        c = i % 6
        kind, ov = (0, float(rng.uniform(0.2, 0.8))) if c < 3 else (
            (1, float(rng.uniform(0.2, 0.8))) if c == 3 else ((0, float(rng.uniform(0.1, 0.25))) if c == 4 else
                                                               (2, float(rng.uniform(0.2, 0.8)))))
        out.append((w, h, kind, ov))
    return out


def run_serving(dev, rates, seconds):
    from cs_stitch import batch as B
    srv = B.Server(hbm_budget=SERVE_BUDGET, max_batch=64, max_live=256, adaptive=True)
    res = {}
    warm = [B.request(i, *dev[i % len(dev)], 0.0, seed=0) for i in range(min(8, len(dev)))]
    srv.run_trace(warm)
    for rate in rates:
        n = max(8, int(rate * seconds))
        arr = B.poisson_arrivals(rate, n, seed=int(rate))
        reqs = [B.request(i, *dev[i % len(dev)], t, seed=0) for i, t in enumerate(arr)]
        done, st = srv.run_trace(reqs)
        statuses = {}
        for c in done:
            statuses[c.status.decode()] = statuses.get(c.status.decode(), 0) + 1
        res[f"poisson_{int(rate)}"] = dict(requests=n, pairs_per_s=round(n / (st.makespan_ms / 1e3), 2),
                                           p50_ms=round(st.p50_ms, 3), p99_ms=round(st.p99_ms, 3),
                                           mean_feature_batch=round(st.mean_feature_batch, 2),
                                           mean_render_batch=round(st.mean_render_batch, 2), statuses=statuses,
                                           peak_pool_gb=round(st.peak_pool_bytes / 2 ** 30, 2))
        log("served", rate, res[f"poisson_{int(rate)}"])
    # Saturated: every request queued at t = 0 (capacity of the loop).
    n = 4 * len(dev)
    reqs = [B.request(i, *dev[i % len(dev)], 0.0, seed=0) for i in range(n)]
    done, st = srv.run_trace(reqs)
    res["saturated"] = dict(requests=n, pairs_per_s=round(n / (st.makespan_ms / 1e3), 2), p50_ms=round(st.p50_ms, 3),
                            p99_ms=round(st.p99_ms, 3), mean_feature_batch=round(st.mean_feature_batch, 2),
                            mean_render_batch=round(st.mean_render_batch, 2))
    log("served saturated", res["saturated"])
    srv.close()
    return res


def cpu_baseline(pairs, work, seconds):
    """Single-thread CPU oracle on the first pairs of the batch until the budget is spent."""
    import pipeline_oracle
    t0 = time.perf_counter()
    n = 0
    for (a, b, _), (w, h, k, _) in zip(pairs, work):
        pipeline_oracle.run(a, b, 1.0, k, R, BLEND, True)
        n += 1
        if time.perf_counter() - t0 > seconds:
            break
    dt = time.perf_counter() - t0
    sizes = ", ".join(f"{w}x{h}" for (w, h, _, _) in work[:n])
    return dict(value=round(n / dt, 4), unit="pairs/s", cores=1, kind="port",
                sample=f"first {n} pair(s) of the batch ({sizes}; K caps {[k for _, _, k, _ in work[:n]]}, R={R}, "
                       f"cross-check, {BLEND}) through the single-thread C oracle, {dt:.1f} s")


def cpu_threads():
    # The host share of one GPU: 16 threads at most, fewer where the process
    # may use fewer cores.
    try:
        avail = len(os.sched_getaffinity(0))
    except AttributeError:
        avail = os.cpu_count() or 1
    return max(1, min(16, avail))


def cpu_baseline_all_cores(pairs, work, seconds):
    """The same C oracle on the host's cores: pairs of the batch, in order, on a
    pool of threads (the oracle's C calls release the GIL), submitted while
    the budget lasts; the rate counts every submitted pair over the time
    until the last one finishes."""
    import concurrent.futures as cf
    import pipeline_oracle
    threads = cpu_threads()
    t0 = time.perf_counter()
    n = 0
    with cf.ThreadPoolExecutor(threads) as ex:
        pending = set()
        for (a, b, _), (w, h, k, _) in zip(pairs, work):
            if time.perf_counter() - t0 > seconds:
                break
            while len(pending) >= threads:
                _, pending = cf.wait(pending, return_when=cf.FIRST_COMPLETED)
            pending.add(ex.submit(pipeline_oracle.run, a, b, 1.0, k, R, BLEND, True))
            n += 1
        for f in cf.as_completed(pending):
            f.result()
    dt = time.perf_counter() - t0
    return dict(value=round(n / dt, 4), unit="pairs/s", cores=threads, kind="port",
                sample=f"first {n} pair(s) of the batch (R={R}, cross-check, {BLEND}) through the C oracle on "
                       f"{threads} threads (one pair per thread), {dt:.1f} s")


def main():
    args = parse()
    if args.gpus != 1:
        print(f"bench.py measures one GPU (--gpus {args.gpus} requested): pairs are independent, so the "
              "one-GPU figure is the measurement and the path shards by request across GPUs.", file=sys.stderr)
        sys.exit(2)
    import torch
    assert torch.cuda.is_available(), "bench.py needs a HIP device"
    from cs_stitch import _ffi
    from cs_stitch import batch as B
    _ffi.load()
    if args.calibrate_copy:
        src = torch.empty(COPY_CALIBRATION_BYTES, dtype=torch.uint8, device="cuda")
        dst = torch.empty_like(src)
        dst.copy_(src)
        torch.cuda.synchronize()
        del src, dst
    work = batch_workload(args.pairs)
    log("synthesizing", len(work), "pairs")
    pairs = [synth(SEED0 + i, w, h, ov) for i, (w, h, _, ov) in enumerate(work)]
    dev = [upload(a, b) for a, b, _ in pairs]
    torch.cuda.synchronize()
    specs = specs_for(dev, work)
    br = B.BatchRunner(max_pairs=args.pairs, pool_bytes=POOL_BYTES)
    if not args.no_tune:
        tune = br.autotune(specs, args.tune_cache)
        log("tuning", tune.s4_segment, tune.workgroups_per_cu, tune.graphs)
    # Python's cyclic collector can pause the host thread for tens of ms in
    # the middle of a timed step (many tracked objects once torch is loaded);
    # it stays off while batches are timed.
    gc.collect()
    gc.disable()
    total, times, outcomes, tm = time_batched(br, specs, args.steps, args.warmup)
    ok = sum(o.status == 0 for o in outcomes)
    batched = dict(pairs_per_s=args.pairs * args.steps / total, ms_per_step=1e3 * total / args.steps,
                   p50_ms=1e3 * float(np.percentile(times, 50)), p99_ms=1e3 * float(np.percentile(times, 99)),
                   feature_ms=tm.feature_ms, render_ms=tm.render_ms, host_ms=tm.host_ms, ok=ok,
                   steps_ms=[round(1e3 * t, 3) for t in times])
    log("batched", batched)
    stream = None
    if not args.no_stream:
        s_total, s_last = time_stream(br, specs, args.steps, args.warmup)
        assert [o.status for o in s_last] == [o.status for o in outcomes]
        stream = dict(pairs_per_s=args.pairs * args.steps / s_total, ms_per_step=1e3 * s_total / args.steps,
                      ok=sum(o.status == 0 for o in s_last))
        log("stream", stream)
    headline = stream if stream is not None else batched
    roof = roofline(probe_kernels(br, specs, args.steps))
    log("roofline", {k: v for k, v in roof.items() if k != "kernels"})
    # The 1080p load: a batch of 64 pairs shaped like the single-pair 1080p
    # config (1920x1080, 30-50 % overlap, K 5000, cross-check, RANSAC 4096,
    # multiband_6), against the plain runner on the same pairs.
    c1 = None
    if not args.no_subset:
        rng = np.random.default_rng(SEED0 + 2000)
        work1 = [(1920, 1080, 5000, float(rng.uniform(0.3, 0.5))) for _ in range(args.pairs)]
        pairs1 = [synth(SEED0 + 2000 + i, w, h, ov) for i, (w, h, _, ov) in enumerate(work1)]
        dev1 = [upload(a, b) for a, b, _ in pairs1]
        del pairs1
        torch.cuda.synchronize()
        specs1 = specs_for(dev1, work1)
        out1 = br.run(specs1)[0]
        t_b1, times_b1, last1, _ = time_batched(br, specs1, args.steps, args.warmup)
        t_s1 = time_stream(br, specs1, args.steps, args.warmup)[0] if not args.no_stream else t_b1
        c1 = dict(workload=f"{args.pairs} pairs of 1920x1080, overlap U[0.3,0.5], K 5000, cross-check, RANSAC {R}, "
                           f"{BLEND}", stream_pairs_per_s=args.pairs * args.steps / t_s1,
                  batched_pairs_per_s=args.pairs * args.steps / t_b1,
                  batched_p50_ms=1e3 * float(np.percentile(times_b1, 50)), ok=sum(o.status == 0 for o in last1))
        log("c1_1080p", c1)
    result = dict(stream=stream, batched=batched, c1_1080p=c1)
    if not args.no_plain:
        runners, assign = plain_runners(dev, work, outcomes)
        plain = time_plain(dev, work, runners, assign)
        log("plain", plain)
        for r in runners.values():
            r.close()
        if c1 is not None:
            runners1, assign1 = plain_runners(dev1, work1, out1)
            plain1 = time_plain(dev1, work1, runners1, assign1)
            log("plain 1080p", plain1)
            for r in runners1.values():
                r.close()
            c1.update(plain_pairs_per_s=plain1["pairs_per_s"], plain_p50_ms=plain1["p50_ms"],
                      plain_p99_ms=plain1["p99_ms"], speedup=c1["stream_pairs_per_s"] / plain1["pairs_per_s"])
        result["plain"] = plain
        result["speedup_vs_plain"] = headline["pairs_per_s"] / plain["pairs_per_s"]
        result["batched_speedup_vs_plain"] = batched["pairs_per_s"] / plain["pairs_per_s"]
    br.close()
    gc.enable()
    if not args.no_serve:
        pool = serve_pool(args.pairs)
        sdev = [upload(*synth(SEED0 + 1000 + i, w, h, ov, kind)[:2]) for i, (w, h, kind, ov) in enumerate(pool)]
        torch.cuda.synchronize()
        result["served"] = run_serving(sdev, [50, 100, 200, 400], args.serve_seconds)
    cpu = cpu_single = None
    if not args.no_cpu:
        cpu_single = cpu_baseline(pairs, work, args.cpu_seconds)
        log("cpu single thread", cpu_single)
        cpu = cpu_baseline_all_cores(pairs, work, args.cpu_seconds)
        log("cpu all cores", cpu)
    line = {
        "metric": "stitched pairs/s + p50/p99 latency, batched/served vs plain per-pair, 1 GPU",
        "value": round(headline["pairs_per_s"], 3),
        "unit": "pairs/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(headline["ms_per_step"], 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u8/int32/f32/f64",
        "data": "synthetic",
        "config": {"workload": f"batched form as {'a stream of batches' if stream else 'one batch at a time'}, "
                               f"{args.pairs} pairs per step: sizes from "
                               f"{{640x480,1280x720,1920x1080,3840x2160}}, K caps U[500,8000], overlap U[0.2,0.8], "
                               f"cross-checked Hamming, RANSAC {R}, {BLEND}", "pairs_per_step": args.pairs},
        **result,
        "roofline": roof,
        "cpu_baseline": cpu,
        "cpu_baseline_single_thread": cpu_single,
    }
    print(json.dumps(line))


if __name__ == "__main__":
    main()
