"""Stitched pairs/s and p50/p99 latency of the batched and served forms beside
the plain per-pair GPU path, on one MI355X. Prints one JSON line.

    python bench.py [--gpus 1] [--steps K] [--warmup W]
    python bench.py --stage s4|s1|strip  # one stage, or the strip, alone

Workload (value): the batched form on a ragged batch of 64 synthetic pairs
- resolutions drawn from {640x480, 1280x720, 1920x1080, 3840x2160}, keypoint
caps from U[500, 8000], overlap U[20 %, 80 %], cross-checked Hamming matching,
RANSAC 4096, 6-band multiband blend. A step is that batch through S1-S6, with
the inputs resident in HBM before the timed region; `value` counts the pairs
that stitched (status OK) per second of the stream of such batches (each
batch's S1-S5 overlap the previous batch's S6 on a second stream); pairs that
fail (a canvas over the 16,777,216-pixel budget) are listed, and the rate
over all pairs is reported beside it. "batched" times one batch at a time.
The same 64 pairs run through the plain per-pair runner (the reference's
execute_run flow: one pair at a time, per-stage launches, two host
synchronizations) for the ratio. "c1_1080p" runs the 1080p load the same
way: 64 pairs shaped like the single-pair 1080p config (1920x1080, 30-50 %
overlap, K 5000, cross-check, RANSAC 4096, multiband_6), batched and
streamed, against the plain runner on the same pairs. The served form
replays Poisson arrival traces with the same size mix and textured,
low-texture, low-overlap and repetitive scenes under a 64 GB HBM budget:
once with the adaptive legacy-v1 controller, once with fixed knobs (the
batch's), and saturated (every request at t = 0) so that the budget binds;
`pairs_per_s` there counts OK completions. "strip" stitches 8 frames of
3840x2160 in a row (30 % overlap, K 16000, RANSAC 8192, 7 bands) onto one
canvas of about 22000 x 2450: S1-S5 of the adjacent pairs in one batched
phase, one chain render.

"roofline": the S6 stage of one batch (one render-graph replay: every S6
kernel of every canvas) against the 8 TB/s HBM peak, with the compulsory
bytes 3(A + B) + 4P per stitched canvas (A, B input pixels, P canvas pixels)
over its event-timed duration on the render stream. "whole_path" adds the
other stages' compulsory bytes per pair - S1 4(A + B), S2/S3 (A + B) +
44(Ka + Kb), S4 32(Ka + Kb) + 12M - over ms_per_step. "s6_kernels" times
every S6 kernel with HIP events (direct launches) and lists each one's design
bytes (what it reads and writes by construction).
"""
from __future__ import annotations

import argparse
import gc
import json
import os
import platform
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
# Integer VALU: 256 CUs x 64 lanes x 2.4 GHz (one 32-bit op per lane per clock).
INT_PEAK_TOPS = 256 * 64 * 2.4e9 / 1e12
# Measured HBM bytes of the S6 stage of one batch (rocprofv3 FETCH_SIZE +
# WRITE_SIZE over the S6 kernels, scaled by a copy-kernel calibration).
TRAFFIC_PROFILE = ROOT / "profiles" / "s6_stage_mixed64_traffic.md"
COPY_CALIBRATION_BYTES = 1 << 30
SERVE_BUDGET = 64 << 30
# Poisson traces: (rate per second, requests).
SERVE_TRACES = [(400, 10000), (200, 2000), (100, 2000), (50, 2000)]
SERVE_FIXED_TRACES = [(400, 10000), (100, 2000)]


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--pairs", type=int, default=64, help="pairs per batch")
    ap.add_argument("--serve-scale", type=float, default=1.0, help="fraction of each trace's requests to replay")
    ap.add_argument("--no-serve", action="store_true")
    ap.add_argument("--no-plain", action="store_true")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-subset", action="store_true", help="skip the 64-pair 1920x1080 batch (profiling runs)")
    ap.add_argument("--no-stream", action="store_true",
                    help="skip the stream of batches, so no two batches' kernels overlap (kernel-trace runs); "
                         "value is then the one-batch-at-a-time rate")
    ap.add_argument("--no-probe", action="store_true", help="skip the per-kernel S6 probe")
    ap.add_argument("--cpu-seconds", type=float, default=15.0, help="budget of the CPU oracle sample")
    ap.add_argument("--tune-cache", default=None)
    ap.add_argument("--no-tune", action="store_true", help="default launch shapes (profiling runs)")
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="first run one plain 1 GiB device copy (the reference kernel for PMC byte counters)")
    ap.add_argument("--no-strip", action="store_true", help="skip the 8-frame 3840x2160 strip")
    ap.add_argument("--stage", choices=["s4", "s1", "strip"], default=None,
                    help="time one stage alone: s4 = 16000 x 16000 cross-checked Hamming match, "
                         "s1 = gray conversion of a 3840x2160 pair, strip = the 8-frame 3840x2160 strip")
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


def status_counts(outcomes):
    out = {}
    for o in outcomes:
        key = "OK" if o.status == 0 else f"{o.fail_stage.decode()}:{o.status}"
        out[key] = out.get(key, 0) + 1
    return out


def time_batched(br, specs, steps, warmup):
    import torch
    for _ in range(warmup):
        br.run(specs)
    torch.cuda.synchronize()
    times, render, feature, last = [], [], [], None
    t0 = time.perf_counter()
    for _ in range(steps):
        s = time.perf_counter()
        last, tm = br.run(specs)
        times.append(time.perf_counter() - s)
        render.append(tm.render_ms)
        feature.append(tm.feature_ms)
    torch.cuda.synchronize()
    total = time.perf_counter() - t0
    return total, times, last, tm, float(np.mean(render)), float(np.mean(feature))


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
    design bytes computed from the canvases and their pixel counts."""
    import torch
    br.set_probe(True)
    for _ in range(steps):
        br.run(specs)
    torch.cuda.synchronize()
    stats = br.probe_read()
    br.set_probe(False)
    return stats


def s6_compulsory(outcomes):
    """3(A + B) + 4P over the canvases that were rendered, with the pixel sums."""
    ab = p = n = 0
    for o in outcomes:
        if o.status != 0:
            continue
        g = o.geometry
        ab += g.width_a * g.height_a + g.width_b * g.height_b
        p += g.width * g.height
        n += 1
    return 3 * ab + 4 * p, ab, p, n


def path_compulsory(outcomes, specs):
    """Compulsory bytes of the whole path per batch: S1 4(A + B), S2/S3
    (A + B) + 44(Ka + Kb), S4 32(Ka + Kb) + 12M for every pair, S6 for the
    rendered canvases."""
    total = 0
    for o, s in zip(outcomes, specs):
        ab = s.a.width * s.a.height + s.b.width * s.b.height
        kk = o.num_kp_a + o.num_kp_b
        total += 4 * ab + ab + 44 * kk + 32 * kk + 12 * o.homography.num_matches
    return total + s6_compulsory(outcomes)[0]


def measured_traffic(algorithmic_bytes):
    """S6-stage HBM bytes per batch from the committed PMC summary, or None.
    The summary states the compulsory bytes of the batch it was measured on;
    another workload or kernel set reads other bytes, and then it does not
    apply."""
    import re
    if not TRAFFIC_PROFILE.exists():
        return None
    text = TRAFFIC_PROFILE.read_text()
    m = re.search(r"^algorithmic_bytes_per_launch:\s*(\d+)\s*\ntraffic_bytes_per_launch:\s*(\d+)\s*$", text, re.M)
    if m and m.group(1) == str(int(algorithmic_bytes)):
        return int(m.group(2))
    return None


def roofline(outcomes, render_ms, stats):
    nbytes, ab, p, n = s6_compulsory(outcomes)
    achieved = nbytes / (render_ms * 1e-3) / 1e9
    top = sorted(stats, key=lambda e: -e["total_ms"])[:12] if stats else []
    return dict(bound="hbm", achieved=round(achieved, 1), peak=HBM_PEAK_GBS, unit="GB/s",
                frac=round(achieved / HBM_PEAK_GBS, 4), traffic=measured_traffic(nbytes),
                kernel="s6 render stage (every S6 kernel of one batch, one graph replay)",
                launch_ms=round(render_ms, 4), algorithmic_bytes=int(nbytes),
                units=dict(canvases=n, input_pixels=int(ab), canvas_pixels=int(p),
                           bytes_per_canvas="3(A+B)+4P"),
                s6_kernels=[dict(kernel=f"{e['kernel']} level {e['level']}", launches=e["launches"],
                                 ms_per_launch=round(e["total_ms"] / e["launches"], 4),
                                 design_bytes_per_launch=int(e["algorithmic_bytes"] / e["launches"]),
                                 design_gbs=round(e["algorithmic_bytes"] / (e["total_ms"] * 1e-3) / 1e9, 1))
                            for e in top])


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
        if key[2] * key[3] > 16777216:  # the batch found no canvas that fits: any box rejects it
            key = (w, h, w, h)
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
    return dict(pairs_per_s=ok / t_total, all_pairs_per_s=n / t_total, ms_per_pair=1e3 * t_total / n,
                p50_ms=float(np.percentile(lat, 50)), p99_ms=float(np.percentile(lat, 99)), ok=int(ok), n=n)


def serve_pool(n):
    """Served scenes: sizes as the batch; half textured at 20-80 % overlap, the
    rest hard: low texture (blobs / 10), low overlap (10-25 %), repetitive.
    Each scene has its own keypoint cap for the fixed-knob pass."""
    rng = np.random.default_rng(SEED0 + 7)
    out = []
    for i in range(n):
        w, h = SIZES[int(rng.integers(len(SIZES)))]
        c = i % 6
        kind, ov = (0, float(rng.uniform(0.2, 0.8))) if c < 3 else (
            (1, float(rng.uniform(0.2, 0.8))) if c == 3 else ((0, float(rng.uniform(0.1, 0.25))) if c == 4 else
                                                               (2, float(rng.uniform(0.2, 0.8)))))
        out.append((w, h, kind, ov, int(rng.integers(500, 8001))))
    return out


def trace_summary(done, st, n):
    statuses = {}
    for c in done:
        statuses[c.status.decode()] = statuses.get(c.status.decode(), 0) + 1
    ok = statuses.get("OK", 0)
    secs = st.makespan_ms / 1e3
    return dict(requests=n, pairs_per_s=round(ok / secs, 2), requests_per_s=round(n / secs, 2),
                p50_ms=round(st.p50_ms, 3), p99_ms=round(st.p99_ms, 3), ok=ok, statuses=statuses,
                budget_exhausted_share=round(statuses.get("BUDGET_EXHAUSTED", 0) / n, 4),
                mean_feature_batch=round(st.mean_feature_batch, 2), mean_render_batch=round(st.mean_render_batch, 2),
                peak_pool_gb=round(st.peak_pool_bytes / 2 ** 30, 2))


def run_serving(dev, pool, scale):
    from cs_stitch import batch as B
    res = {}
    for adaptive in (True, False):
        srv = B.Server(hbm_budget=SERVE_BUDGET, max_batch=64, max_live=256, adaptive=adaptive)
        fixed = [B.pair_spec(ta, tb, S=1.0, K=k, R=R, blend=BLEND, cross_check=True)
                 for (ta, tb), (_, _, _, _, k) in zip(dev, pool)]

        def req(i, t):
            j = i % len(dev)
            return B.request(i, *dev[j], t, seed=0, fixed=None if adaptive else fixed[j])

        name = "legacy_v1" if adaptive else "fixed_knobs"
        srv.run_trace([req(i, 0.0) for i in range(min(16, len(dev)))])  # warm the graphs
        part = {}
        for rate, count in (SERVE_TRACES if adaptive else SERVE_FIXED_TRACES):
            n = max(16, int(count * scale))
            arr = B.poisson_arrivals(rate, n, seed=int(rate) + (0 if adaptive else 7))
            done, st = srv.run_trace([req(i, t) for i, t in enumerate(arr)])
            part[f"poisson_{int(rate)}"] = trace_summary(done, st, n)
            log("served", name, rate, part[f"poisson_{int(rate)}"])
        # Saturated: every request queued at t = 0 (capacity of the loop; the
        # HBM budget decides how many are in flight).
        n = 4 * len(dev)
        done, st = srv.run_trace([req(i, 0.0) for i in range(n)])
        part["saturated"] = trace_summary(done, st, n)
        log("served", name, "saturated", part["saturated"])
        srv.close()
        res[name] = part
    res["workload"] = (f"{len(dev)} resident scenes cycled by the requests: sizes from {{640x480,1280x720,"
                       f"1920x1080,3840x2160}}; half textured (20-80 % overlap), the rest low texture, low overlap "
                       f"(10-25 %) or repetitive; legacy_v1: the adaptive controller; fixed_knobs: K U[500,8000], "
                       f"R {R}, cross-check, {BLEND}; HBM budget {SERVE_BUDGET >> 30} GB")
    return res


def host_cpu():
    model = platform.processor() or ""
    try:
        for line in Path("/proc/cpuinfo").read_text().splitlines():
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    try:
        avail = len(os.sched_getaffinity(0))
    except AttributeError:
        avail = os.cpu_count() or 1
    return dict(cpu_model=model, host_cores=os.cpu_count(), cores_available=avail)


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
    return dict(value=round(n / dt, 4), unit="pairs/s", cores=1, kind="port", **host_cpu(),
                sample=f"first {n} pair(s) of the batch ({sizes}; K caps {[k for _, _, k, _ in work[:n]]}, R={R}, "
                       f"cross-check, {BLEND}) through the single-thread C oracle, {dt:.1f} s")


def cpu_threads():
    # The host share of one GPU: 16 threads at most, fewer where the process
    # may use fewer cores.
    return max(1, min(16, host_cpu()["cores_available"]))


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
    return dict(value=round(n / dt, 4), unit="pairs/s", cores=threads, kind="port", **host_cpu(),
                sample=f"first {n} pair(s) of the batch (R={R}, cross-check, {BLEND}) through the C oracle on "
                       f"{threads} threads (one pair per thread), {dt:.1f} s")


def stage_bench(kind, steps, warmup):
    """One stage alone on its largest configured shape, timed with HIP events
    on the stream it runs on (torch's current stream)."""
    import torch
    from cs_stitch import stages as st
    ev = [torch.cuda.Event(enable_timing=True) for _ in range(2)]
    if kind == "s4":
        n = 16000
        g = torch.Generator().manual_seed(SEED0)
        da = torch.randint(0, 256, (n, 32), dtype=torch.uint8, generator=g).cuda()
        db = torch.randint(0, 256, (n, 32), dtype=torch.uint8, generator=g).cuda()
        ca = torch.tensor([n], dtype=torch.int32, device="cuda")
        cb = torch.tensor([n], dtype=torch.int32, device="cuda")
        out, ws = st.s4_match(da, ca, db, cb, 0.8, True)
        for _ in range(warmup):
            st.s4_match(da, ca, db, cb, 0.8, True, out=out, ws=ws)
        torch.cuda.synchronize()
        ev[0].record()
        for _ in range(steps):
            st.s4_match(da, ca, db, cb, 0.8, True, out=out, ws=ws)
        ev[1].record()
        torch.cuda.synchronize()
        ms = ev[0].elapsed_time(ev[1]) / steps
        ops = 16 * n * n * 2
        return dict(stage="s4", workload=f"{n} x {n} random 256-bit descriptors, cross-check, ratio 0.8",
                    ms_per_launch=round(ms, 4), int_ops=ops, tops=round(ops / ms / 1e9, 2),
                    peak_tops=round(INT_PEAK_TOPS, 1), frac=round(ops / ms / 1e9 / INT_PEAK_TOPS, 4))
    w, h = 3840, 2160
    a, b, _ = synth(SEED0, w, h, 0.5)
    ta, tb = upload(a, b)
    for _ in range(warmup):
        st.s1_preprocess(ta, tb, 1.0)
    torch.cuda.synchronize()
    ev[0].record()
    for _ in range(steps):
        st.s1_preprocess(ta, tb, 1.0)
    ev[1].record()
    torch.cuda.synchronize()
    ms = ev[0].elapsed_time(ev[1]) / steps
    nbytes = 4 * 2 * w * h
    return dict(stage="s1", workload=f"{w}x{h} RGB8 pair to gray at S=1", ms_per_launch=round(ms, 4),
                bytes=nbytes, gbs=round(nbytes / ms / 1e6, 1), peak_gbs=HBM_PEAK_GBS,
                frac=round(nbytes / ms / 1e6 / HBM_PEAK_GBS, 4))


STRIP_FRAMES, STRIP_W, STRIP_H, STRIP_OVERLAP = 8, 3840, 2160, 0.3
STRIP_K, STRIP_R, STRIP_BLEND = 16000, 8192, "multiband_7"


def strip_bench(steps, warmup):
    """The large-frame strip: 8 frames of 3840x2160 in a row, 30 % of each
    frame under the next, K 16000, cross-checked Hamming, RANSAC 8192 and a
    7-band blend of every frame onto one canvas. S1-S5 of the seven adjacent
    pairs in one batched phase, one chain render; wall-clock per strip, with
    the event-timed phases beside it."""
    import torch
    from cs_stitch import pipeline as P
    from cs_stitch import stages as S
    n, w, h = STRIP_FRAMES, STRIP_W, STRIP_H
    frames, truth = P.synth_strip(SEED0, n, w, h, STRIP_OVERLAP)
    dev = []
    for f in frames:
        t = S.pitched(h, w, 3)
        t.copy_(torch.from_numpy(f).cuda())
        dev.append(t)
    del frames
    torch.cuda.synchronize()
    cfg = P.harness_config(S=1.0, K=STRIP_K, R=STRIP_R, blend=STRIP_BLEND, cross_check=True)
    strip = P.StripStitcher(n, w, h, max_k=STRIP_K, max_r=STRIP_R)
    for _ in range(max(1, warmup)):
        strip.run(dev, cfg)
    recs, status = [], []
    t0 = time.perf_counter()
    for _ in range(steps):
        st, rec, pairs, counts = strip.run(dev, cfg)
        recs.append(rec)
        status.append(st)
    total = time.perf_counter() - t0
    rec = recs[-1]
    canvas = rec.canvas_width * rec.canvas_height
    # Median distance of estimated and true images over each pair's overlap
    # (a 64-pixel grid of frame k + 1 that lands in frame k).
    err = []
    for k, p in enumerate(pairs):
        if p.status != 0:
            err.append(None)
            continue
        hm, d = np.array(p.H).reshape(3, 3), []
        for y in range(0, h, 64):
            for x in range(0, w, 64):
                q = truth[k] @ np.array([x, y, 1.0])
                if 0 <= q[0] / q[2] < w and 0 <= q[1] / q[2] < h:
                    e = hm @ np.array([x, y, 1.0])
                    d.append(float(np.hypot(*(e[:2] / e[2] - q[:2] / q[2]))))
        err.append(round(float(np.median(d)), 3))
    mean = lambda f: round(float(np.mean([getattr(r, f) for r in recs])), 3)  # noqa: E731
    render_ms = float(np.mean([r.t_render for r in recs]))
    render_bytes = 3 * n * w * h + 4 * canvas
    strip.close()
    del dev
    torch.cuda.empty_cache()
    return dict(
        workload=f"{n} frames of {w}x{h} in a strip, overlap {STRIP_OVERLAP}, K {STRIP_K}, cross-check, "
                 f"RANSAC {STRIP_R}, {STRIP_BLEND} onto one canvas", status=status[-1],
        fail_stage=rec.fail_stage.decode(), strips_per_s=round(steps / total, 3),
        ms_per_strip=round(1e3 * total / steps, 2), t_pairs_ms=mean("t_pairs"), t_geometry_ms=mean("t_geometry"),
        t_render_ms=round(render_ms, 3),
        canvas=f"{rec.canvas_width}x{rec.canvas_height}", canvas_mp=round(canvas / 1e6, 2),
        union_pixels=int(rec.union_pixels), overlap_pixels=int(rec.overlap_pixels),
        keypoints=counts, inliers=[p.num_inliers for p in pairs], overlap_error_px=err,
        render_algorithmic_bytes=render_bytes, render_bytes_formula="3 * sum(frame pixels) + 4 * canvas pixels",
        render_gbs=round(render_bytes / render_ms / 1e6, 1) if rec.status == 0 else None,
        render_frac=round(render_bytes / render_ms / 1e6 / HBM_PEAK_GBS, 4) if rec.status == 0 else None,
        steps=steps)


def children():
    """Processes whose parent is this one (left-behind workers show here)."""
    me, out = os.getpid(), []
    for d in Path("/proc").iterdir():
        if not d.name.isdigit():
            continue
        try:
            fields = (d / "stat").read_text().rsplit(")", 1)[1].split()
            if int(fields[1]) == me:
                out.append((int(d.name), (d / "cmdline").read_bytes().replace(b"\0", b" ").decode()[:80]))
        except (OSError, ValueError, IndexError):
            pass
    return out


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
    if args.stage == "strip":
        print(json.dumps(strip_bench(args.steps, args.warmup)))
        return
    if args.stage:
        r = stage_bench(args.stage, max(args.steps, 10), max(args.warmup, 3))
        print(json.dumps(r))
        return
    work = batch_workload(args.pairs)
    log("synthesizing", len(work), "pairs")
    pairs = [synth(SEED0 + i, w, h, ov) for i, (w, h, _, ov) in enumerate(work)]
    dev = [upload(a, b) for a, b, _ in pairs]
    torch.cuda.synchronize()
    specs = specs_for(dev, work)
    br = B.BatchRunner(max_pairs=args.pairs, pool_bytes=POOL_BYTES)
    tune = None
    if not args.no_tune:
        tune = br.autotune(specs, args.tune_cache)
        log("tuning", tune.s4_segment, tune.workgroups_per_cu, tune.graphs, list(tune.stage_workgroups_per_cu),
            tune.s1_tile_rows)
    # Python's cyclic collector can pause the host thread for tens of ms in
    # the middle of a timed step (many tracked objects once torch is loaded);
    # it stays off while batches are timed.
    gc.collect()
    gc.disable()
    total, times, outcomes, tm, render_ms, feature_ms = time_batched(br, specs, args.steps, args.warmup)
    ok = sum(o.status == 0 for o in outcomes)
    batched = dict(pairs_per_s=ok * args.steps / total, all_pairs_per_s=args.pairs * args.steps / total,
                   ms_per_step=1e3 * total / args.steps, p50_ms=1e3 * float(np.percentile(times, 50)),
                   p99_ms=1e3 * float(np.percentile(times, 99)), feature_ms=feature_ms, render_ms=render_ms,
                   host_ms=tm.host_ms, ok=ok, statuses=status_counts(outcomes),
                   steps_ms=[round(1e3 * t, 3) for t in times])
    log("batched", batched)
    stream = None
    if not args.no_stream:
        s_total, s_last = time_stream(br, specs, args.steps, args.warmup)
        assert [o.status for o in s_last] == [o.status for o in outcomes]
        s_ok = sum(o.status == 0 for o in s_last)
        stream = dict(pairs_per_s=s_ok * args.steps / s_total, all_pairs_per_s=args.pairs * args.steps / s_total,
                      ms_per_step=1e3 * s_total / args.steps, ok=s_ok)
        log("stream", stream)
    headline = stream if stream is not None else batched
    stats = None if args.no_probe else probe_kernels(br, specs, args.steps)
    roof = roofline(outcomes, render_ms, stats)
    whole = path_compulsory(outcomes, specs)
    whole_path = dict(algorithmic_bytes=int(whole), ms_per_step=round(headline["ms_per_step"], 4),
                      achieved=round(whole / (headline["ms_per_step"] * 1e-3) / 1e9, 1), unit="GB/s",
                      frac=round(whole / (headline["ms_per_step"] * 1e-3) / 1e9 / HBM_PEAK_GBS, 4),
                      bytes_per_pair="S1 4(A+B) + S2/S3 (A+B)+44(Ka+Kb) + S4 32(Ka+Kb)+12M + S6 3(A+B)+4P")
    log("roofline", {k: v for k, v in roof.items() if k != "s6_kernels"}, "whole path", whole_path)
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
        t_b1, times_b1, last1, _, _, _ = time_batched(br, specs1, args.steps, args.warmup)
        ok1 = sum(o.status == 0 for o in last1)
        t_s1 = time_stream(br, specs1, args.steps, args.warmup)[0] if not args.no_stream else t_b1
        c1 = dict(workload=f"{args.pairs} pairs of 1920x1080, overlap U[0.3,0.5], K 5000, cross-check, RANSAC {R}, "
                           f"{BLEND}", stream_pairs_per_s=ok1 * args.steps / t_s1,
                  batched_pairs_per_s=ok1 * args.steps / t_b1,
                  batched_p50_ms=1e3 * float(np.percentile(times_b1, 50)), ok=ok1)
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
        sdev = [upload(*synth(SEED0 + 1000 + i, w, h, ov, kind)[:2]) for i, (w, h, kind, ov, _) in enumerate(pool)]
        torch.cuda.synchronize()
        result["served"] = run_serving(sdev, pool, args.serve_scale)
    if not args.no_strip:
        result["strip"] = strip_bench(args.steps, 1)
        log("strip", result["strip"])
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
                               f"cross-checked Hamming, RANSAC {R}, {BLEND}; value counts the pairs that stitched "
                               f"(status OK)", "pairs_per_step": args.pairs,
                   "tuning": None if tune is None else dict(
                       s4_segment=tune.s4_segment, workgroups_per_cu=tune.workgroups_per_cu,
                       stage_workgroups_per_cu=dict(zip(("s1", "s2", "s3", "s4", "s5", "s6"),
                                                        list(tune.stage_workgroups_per_cu))),
                       s1_tile_rows=tune.s1_tile_rows)},
        **result,
        "roofline": roof,
        "whole_path": whole_path,
        "cpu_baseline": cpu,
        "cpu_baseline_single_thread": cpu_single,
    }
    print(json.dumps(line), flush=True)
    left = children()
    if left:
        log("child processes at exit:", left)


if __name__ == "__main__":
    main()
