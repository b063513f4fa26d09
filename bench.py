"""Benchmark of the MI355X imgproc hot path (one GPU).

Default workload: the single-image chain on a 3840x2160 3-channel u8 BGR
frame of uniform bytes: cvtColor BGR->GRAY, resize INTER_LINEAR to
1920x1080 (exact half), Sobel dx=1 dy=0 ksize=3 to int16, warpAffine
INTER_LINEAR BORDER_CONSTANT=0 with a seeded matrix (rotation U(-30, 30) deg,
scale U(0.8, 1.2), shift U(+-10 %)), and warpPerspective with the corners
jittered by +-5 %. One step = the whole chain on one frame, every output
written. Inputs are resident in HBM and rotate through 24 distinct frames
(597 MB, more than twice the 256 MB MALL). The chain is enqueued through the
C ABI and replayed from HIP graphs of 24 steps each; inside a graph, step j
runs on stream j % L (L frames in flight, each step writing its own output
set), so the short 1080p kernels of one frame overlap those of the next.
Every step still does all of its work; `--lanes 1` gives one frame at a time.

Prints one JSON line: value = input Gpix/s of the whole chain, a per-kernel
table (HIP-event durations, algorithmic bytes, GB/s), the roofline of the
dominant kernel, and the CPU oracle timed on a bounded sample of the same
workload. `--config 1` times the plain 1080p GaussianBlur 5x5 instead;
`--config 6` the tracking image of a 3840x2160 BGRA frame (grayDownsample,
BT.601 gray shrunk by 6 with box averaging, to 640x360); `--config 7` lens
undistortion of a 1280x720 BGR frame (the reference benchmark's camera and
5 distortion coefficients: f64 map, 1/32 px fixed point, 15-bit bilinear);
`--config 8` the reference benchmark's pyramid and box cases at 1920x1080
(pyrDown BGR, gray and f32 BGR, pyrUp gray, blur 3x3 and 15x15 BGR) and
adaptiveThreshold (block 11, mean and Gaussian) in one step.
"""
import argparse
import ctypes
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=240)
    ap.add_argument("--warmup", type=int, default=24)
    ap.add_argument("--config", type=int, default=2, choices=[1, 2, 3, 4, 5, 6, 7, 8])
    ap.add_argument("--streams", type=int, default=64)
    ap.add_argument("--lanes", type=int, default=0, choices=[0, 1, 2, 3, 4],
                    help="frames in flight in the graph-replayed workloads (step j runs on stream j %% lanes); "
                         "0: the measured best per workload (GaussianBlur 1, chain 3, pyramid/box/adaptive step 4, others 2)")
    ap.add_argument("--seconds", type=float, default=6.0, help="served workload: Poisson run length")
    ap.add_argument("--host-workers", type=int, default=8, help="served workload: connected-components threads")
    ap.add_argument("--morph-cases", type=int, default=63,
                    help="morphology workload: time the first N of the 63 cases (0: only the roofline kernel, "
                         "for counter passes)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="only run a plain 256 MiB device copy (scales the HBM byte counters of a profile)")
    return ap.parse_args()


def profiled_traffic(key):
    """HBM bytes per launch of `key` from the committed rocprofv3 counter
    passes (profiles/traffic.json), or None."""
    tf = os.path.join(ROOT, "profiles", "traffic.json")
    if not os.path.exists(tf):
        return None
    return json.load(open(tf)).get("traffic_bytes_per_launch", {}).get(key)


CPU_THREADS = 16  # the GPU host's CPU share for one GPU

# The camera and distortion of the reference's undistort benchmark
# (bench/bench_vs_opencv.py:240-241).
UD_K = np.array([[900.0, 0, 640], [0, 900, 360], [0, 0, 1]])
UD_DIST = np.array([-0.28, 0.09, 0.001, -0.0005, -0.01])


def cpu_lib():
    """The oracle module driving its -O3 -march=native build, made here on
    the host that times it."""
    import oracle
    from oracle import bands
    oracle.use_library(bands.build_baseline())
    return oracle


def cpu_threads():
    from oracle import bands
    return max(1, min(CPU_THREADS, bands.cpu_info()["usable_cpus"]))


def oc(oracle, name, *args):
    """A call of oracle C function `name` on pre-built arguments: images
    become visar_image-shaped descriptors once, so a band's call costs no
    Python work beyond the call itself."""
    f = getattr(oracle.lib(), name)
    conv = []
    keep = []
    for a in args:
        if isinstance(a, np.ndarray) and a.ndim in (2, 3) and a.dtype != np.float64:
            d = oracle.desc(a)
            keep.append((a, d))
            conv.append(ctypes.byref(d))
        elif isinstance(a, np.ndarray):
            arr = np.ascontiguousarray(a, np.float64).ravel()
            keep.append(arr)
            conv.append(arr.ctypes.data_as(ctypes.c_void_p))
        elif isinstance(a, float):
            conv.append(ctypes.c_double(a))
        else:
            conv.append(a)

    def call():
        rc = f(*conv)
        if rc != 0:
            raise RuntimeError(f"{name}: oracle status {rc}")
    call.keep = keep
    return call


def per_launch_ms(torch, fn, n, stream):
    """Average duration of one launch of `fn(i)`: n launches are queued
    behind a spin kernel (so host enqueue time never shows) and bracketed by
    one pair of events on the launch stream; the figure includes the
    dispatch gap between back-to-back kernels."""
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    torch.cuda._sleep(50_000_000)
    e0.record(stream)
    for i in range(n):
        fn(i)
    e1.record(stream)
    torch.cuda.synchronize()
    return e0.elapsed_time(e1) / n


def bench_batch(args, torch, va):
    """512 images of uneven sizes and mixed ops: one graph launch
    (graph replay, autotuned launch shape) against 512 plain C-ABI calls."""
    from visar_amd import autotune, batch
    stream = torch.cuda.current_stream()
    jobs = batch.workload(0, 512)
    specs = [j["spec"] for j in jobs]
    items = batch.items(jobs)
    nbytes = batch.algorithmic_bytes(jobs)
    out_px = int(sum(j["dst"].shape[0] * j["dst"].shape[1] for j in jobs))
    # Plain per-call path, two ways: each call returns when its kernels are
    # done (the C ABI's synchronous contract), and calls queued back to back
    # on the stream with one wait at the end.
    def plain_pass():
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for j in jobs:
            batch.run_plain(va, j)
        torch.cuda.synchronize()
        return time.perf_counter() - t0

    va.set_sync(True)
    plain_pass()
    plain_sync_s = min(plain_pass() for _ in range(3))
    va.set_sync(False)
    plain_pass()
    plain_async_s = min(plain_pass() for _ in range(3))
    b = batch.Batch()
    best, timings = autotune.tune_batch(b, items, specs)
    for _ in range(args.warmup):
        b.run()
    torch.cuda.synchronize()
    K = args.steps
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    t0 = time.perf_counter()
    e0.record(stream)
    for _ in range(K):
        b.run()
    e1.record(stream)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    ms = max(e0.elapsed_time(e1), wall * 1e3) / K
    tiles, wgs = b.info()
    gbs = nbytes / (ms * 1e-3) / 1e9
    line = {
        "metric": "images/s of the 512-image mixed batch (one graph launch)",
        "value": round(512 / (ms * 1e-3), 1),
        "unit": "images/s",
        "n_gpus": 1, "steps": K, "warmup": args.warmup, "ms_per_step": round(ms, 4),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "u8",
        "data": "synthetic (splitmix64 uniform bytes; sides U{64..2048}, 1 or 3 channels, seeded op per image)",
        "config": {"workload": "batch 512 mixed: gauss 3/5/7, area 1/2, linear 1.5x, warpAffine, dilate 5x5 rect",
                   "tiles": tiles, "workgroups": wgs, "tuning": best, "graph_replay": True},
        "gpix_per_s_out": round(out_px / (ms * 1e-3) / 1e9, 3),
        "plain_per_call": {"sync_ms_per_batch": round(plain_sync_s * 1e3, 3),
                           "sync_images_per_s": round(512 / plain_sync_s, 1),
                           "speedup_vs_sync": round(plain_sync_s * 1e3 / ms, 2),
                           "queued_ms_per_batch": round(plain_async_s * 1e3, 3),
                           "queued_images_per_s": round(512 / plain_async_s, 1),
                           "speedup_vs_queued": round(plain_async_s * 1e3 / ms, 2)},
        "roofline": {"kernel": "batch_kernel", "bound": "hbm", "achieved": round(gbs, 1), "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": round(gbs / HBM_PEAK_GBS, 4), "traffic": profiled_traffic("batch_graph"),
                     "algorithmic_bytes": nbytes},
        "autotune_ms": {k: round(v, 4) for k, v in timings.items()},
    }
    if not args.no_cpu:
        from oracle import bands, pipelines
        oracle = cpu_lib()
        # Every 16th image of the batch, whole calls shared by the threads
        # (the images are independent); host copies of the sources.
        sample = [dict(j, src=j["src"].cpu().numpy()) for j in jobs[::16]]
        line["cpu_baseline"] = bands.baseline(
            oracle.lib(), lambda i: [bands.Items([lambda j=j: pipelines.batch_item(oracle, j) for j in sample])],
            len(sample), "images/s", f"{len(sample)} images of the batch (every 16th), one whole call per image",
            cpu_threads())
    print(json.dumps(line))


MORPH_OPS = ("erode", "dilate", "open", "close", "gradient", "tophat", "blackhat")
MORPH_ELEMENTS = ((0, 3, "rect3"), (2, 5, "ellipse5"), (1, 7, "cross7"))


def bench_morph(args, torch, va):
    """The 63 morphology cases on a 4096x4096 {0,255} mask: 7 ops x {3x3
    RECT, 5x5 ELLIPSE, 7x7 CROSS} x iterations 1-3. Each case is timed with
    HIP events around 5 back-to-back C-ABI calls (device buffers, 16 rotating
    inputs of 16.7 MB so the 256 MB MALL holds few of them), best per-call
    average of 3 runs. The roofline kernel (dilate 3x3 rect, one pass, one
    launch per call) is timed first the same way over 20 calls."""
    from visar_amd import synth
    stream = torch.cuda.current_stream()
    S = 4096
    base = synth.morph_mask(0, S)
    R = 16
    ins = [torch.from_numpy(np.roll(base, 97 * i, axis=1)).cuda() for i in range(R)]
    out = torch.empty((S, S), dtype=torch.uint8, device="cuda")
    nbytes = 2 * S * S
    state = {"k": 0}

    def per_call_ms(op, shape, ks, it, calls):
        va.morphology(ins[state["k"] % R], op, shape, (ks, ks), it, dst=out)
        state["k"] += 1
        torch.cuda.synchronize()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record(stream)
        for _ in range(calls):
            va.morphology(ins[state["k"] % R], op, shape, (ks, ks), it, dst=out)
            state["k"] += 1
        e1.record(stream)
        torch.cuda.synchronize()
        return e0.elapsed_time(e1) / calls

    single = per_call_ms(1, 0, 3, 1, 20)
    cases = {}
    for op in range(7):
        for shape, ks, en in MORPH_ELEMENTS:
            for it in (1, 2, 3):
                if len(cases) < args.morph_cases:
                    cases[f"{MORPH_OPS[op]}/{en}/it{it}"] = min(per_call_ms(op, shape, ks, it, 5) for _ in range(3))
    if not cases:  # roofline kernel only
        print(json.dumps({"roofline_only": True, "ms_per_launch": round(single, 5)}))
        return
    total_ms = sum(cases.values())
    value = len(cases) * S * S / (total_ms * 1e-3) / 1e9
    gbs = nbytes / (single * 1e-3) / 1e9
    slow = max(cases, key=cases.get)
    line = {
        "metric": "Gpix/s over the 63 morphology cases on a 4096x4096 mask (output pixels / summed call time)",
        "value": round(value, 3), "unit": "Gpix/s", "n_gpus": 1, "steps": 3, "warmup": 1,
        "ms_per_step": round(total_ms, 4), "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "u8", "data": "synthetic 4096x4096 mask: Bernoulli(0.3)*255, 5x5 binomial blur, threshold 128",
        "config": {"workload": "morphologyEx 4096x4096 u8: 7 ops x {3x3 rect, 5x5 ellipse, 7x7 cross} x it 1-3",
                   "rotating_inputs": R, "calls_per_timing": 5},
        "algorithmic_gbs_over_cases": round(len(cases) * nbytes / (total_ms * 1e-3) / 1e9, 1),
        "cases_ms": {k2: round(v, 4) for k2, v in cases.items()},
        "slowest": {slow: round(cases[slow], 4)},
        "roofline": {"kernel": "morph_band_u8 (dilate 3x3 rect, one pass)", "bound": "hbm", "achieved": round(gbs, 1),
                     "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": round(gbs / HBM_PEAK_GBS, 4),
                     "traffic": profiled_traffic("morph_dilate_rect3"), "ms_per_launch": round(single, 5)},
    }
    if not args.no_cpu:
        line["cpu_baseline"] = morph_cpu_baseline(base)
    print(json.dumps(line))


def bench_serve(args, torch, va):
    """64 synthetic 1280x720 camera streams through the served front end:
    capacity (every stream submitting as fast as its queue allows), then a
    Poisson run at 30 frames/s per stream for latency, beside the plain
    per-call path (four synchronous C-ABI calls + host components per
    frame)."""
    import threading
    from visar_amd import serve
    W, H, N = 1280, 720, args.streams
    cams = serve.CameraSet(N, W, H, seed=0, pool=32)
    cycle = 30  # frames rendered per stream, replayed in order
    frames = [[cams.next_frame(s).copy() for _ in range(cycle)] for s in range(N)]
    srv = serve.Server(width=W, height=H, max_streams=N, max_batch=N, queue_depth=4, host_workers=args.host_workers,
                       max_inflight=2)
    ids = [srv.admit() for _ in range(N)]
    st0 = srv.stats()

    def collect(into):
        while True:
            got = srv.poll(1024)
            if not got:
                return
            into.extend(got)

    # Warm-up: every stream primes its model.
    for i, sid in enumerate(ids):
        srv.submit(sid, frames[i][0])
    srv.drain()
    collect([])

    # Capacity: one submitter thread per group of streams, as fast as the queues allow.
    n_cap = max(1, int(args.steps))
    res_cap = []
    groups = 4
    t0 = srv.now()

    def pump(g):
        for k in range(1, n_cap + 1):
            for i in range(g, N, groups):
                srv.submit(ids[i], frames[i][k % cycle], tag=k)

    th = [threading.Thread(target=pump, args=(g,)) for g in range(groups)]
    for t in th:
        t.start()
    for t in th:
        t.join()
    srv.drain()
    t1 = srv.now()
    collect(res_cap)
    cap_fps = N * n_cap / (t1 - t0)
    cap_seconds = t1 - t0
    cap_lat = np.array([r["t_done"] - r["t_submit"] for r in res_cap])

    # Poisson arrivals, 30 frames/s per stream.
    def poisson(server, sids, seconds, first, seed):
        """Frames arrive as Poisson processes at 30 fps per stream; returns
        (offered fps, results, latency from arrival, latency from submit,
        batch sizes, seconds run)."""
        rng = np.random.default_rng(seed)
        events = []
        for i in range(N):
            t = rng.exponential(1 / 30.0)
            while t < seconds:
                events.append((t, i))
                t += rng.exponential(1 / 30.0)
        events.sort()
        sched, res = {}, []
        base = server.now() + 0.05
        nxt = [first] * N
        for t, i in events:
            target = base + t
            while True:
                d = target - server.now()
                if d <= 0:
                    break
                time.sleep(min(d, 0.002) if d > 0.0005 else 0)
            k = nxt[i]
            nxt[i] += 1
            sched[(sids[i], k)] = target
            server.submit(sids[i], frames[i][k % cycle], tag=k)
            if len(res) % 64 == 0:
                while True:
                    got = server.poll(1024)
                    if not got:
                        break
                    res.extend(got)
        server.drain()
        t_end = server.now()
        while True:
            got = server.poll(1024)
            if not got:
                break
            res.extend(got)
        lat = np.array([r["t_done"] - sched[(r["stream_id"], r["tag"])] for r in res])
        lat_sub = np.array([r["t_done"] - r["t_submit"] for r in res])
        bsz = np.array([r["batch_size"] for r in res])
        return len(events) / seconds, res, lat, lat_sub, bsz, t_end - base

    offered, res_p, lat, lat_sub, bsz, p_seconds = poisson(srv, ids, args.seconds, 1 + n_cap, 1)
    st = srv.stats()
    srv.close()

    # The same Poisson load with a batching window: a launch waits up to
    # the window for more streams' frames (batch sizes against latency).
    sweep = []
    for window in (250, 500, 1000, 2000):
        sw = serve.Server(width=W, height=H, max_streams=N, max_batch=N, queue_depth=4,
                          host_workers=args.host_workers, max_inflight=2, batch_window_us=window)
        sids = [sw.admit() for _ in range(N)]
        for i, sid in enumerate(sids):
            sw.submit(sid, frames[i][0])
        sw.drain()
        while sw.poll(1024):
            pass
        off_w, res_w, lat_w, _, bsz_w, _ = poisson(sw, sids, 2.0, 1, 2)
        sw.close()
        hist = {int(b): int(c) for b, c in zip(*np.unique(bsz_w, return_counts=True))}
        sweep.append({"batch_window_us": window, "offered_fps": round(off_w, 1), "frames": len(res_w),
                      "latency_ms_p50": round(float(np.percentile(lat_w, 50)) * 1e3, 3),
                      "latency_ms_p99": round(float(np.percentile(lat_w, 99)) * 1e3, 3),
                      "mean_batch": round(float(bsz_w.mean()), 2), "batch_size_frames": hist})

    # The plain per-call path on one stream's frames.
    va.set_sync(True)  # each plain call returns with its output written, as the C ABI does by default
    plain = serve.PlainPipeline(va, W, H)
    plain.run(frames[0][0])
    n_plain = 60
    t0 = time.perf_counter()
    for k in range(n_plain):
        plain.run(frames[0][1 + k % (cycle - 1)])
    plain_fps = n_plain / (time.perf_counter() - t0)

    fused_bytes = W * H * (1 + 2 + 2 + 1)  # frame in, model read + write, mask out
    line = {
        "metric": "frames/s of 64 camera streams through the served front end (capacity), p50/p99 latency at 30 fps/stream",
        "value": round(cap_fps, 1),
        "unit": "frames/s",
        "n_gpus": 1, "steps": n_cap, "warmup": 1, "ms_per_step": round(1e3 / cap_fps * N, 4),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "u8",
        "data": f"synthetic 1280x720 streams (gradient 60->140 + N(0,4^2) from a pool of 32 noise fields, 5 moving "
                f"ellipses of 210); {cycle} frames per stream replayed in order",
        "config": {"workload": "served: blur 5x5 -> background (shift 4, thr 28) -> threshold -> open 3x3 -> host CC",
                   "streams": N, "width": W, "height": H, "max_batch": N, "queue_depth": 4,
                   "host_workers": args.host_workers},
        "capacity": {"frames": N * n_cap, "seconds": round(cap_seconds, 3), "fps": round(cap_fps, 1),
                     "latency_ms_p50": round(float(np.percentile(cap_lat, 50)) * 1e3, 3),
                     "latency_ms_p99": round(float(np.percentile(cap_lat, 99)) * 1e3, 3)},
        "poisson": {"offered_fps": round(offered, 1), "frames": len(res_p),
                    "achieved_fps": round(len(res_p) / p_seconds, 1),
                    "latency_ms_p50": round(float(np.percentile(lat, 50)) * 1e3, 3),
                    "latency_ms_p99": round(float(np.percentile(lat, 99)) * 1e3, 3),
                    "latency_from_submit_ms_p99": round(float(np.percentile(lat_sub, 99)) * 1e3, 3),
                    "mean_batch": round(float(bsz.mean()), 2), "max_batch": int(bsz.max()),
                    "batch_size_frames": {int(b): int(c) for b, c in zip(*np.unique(bsz, return_counts=True))}},
        "poisson_batch_window": sweep,
        "admission": {"admitted": st["admitted"], "rejected": st["rejected"],
                      "bytes_per_stream": st0["bytes_per_stream"], "device_bytes": st0["device_bytes"]},
        "plain_per_call": {"fps": round(plain_fps, 1), "speedup_served": round(cap_fps / plain_fps, 2)},
        "roofline": {"kernel": "serve_k", "bound": "hbm", "achieved": None, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": None, "traffic": None, "algorithmic_bytes_per_frame": fused_bytes},
    }
    # Fused kernel alone: per-frame duration over a full batch (HIP events).
    line["roofline"].update(serve_kernel_roofline(torch, va, W, H, N, fused_bytes))
    per_frame = profiled_traffic("serve_k_per_frame")
    if per_frame is not None:
        line["roofline"]["traffic"] = int(per_frame * line["roofline"]["frames_per_launch"])
    if not args.no_cpu:
        from oracle import bands
        oracle = cpu_lib()
        nt = cpu_threads()
        # One frame of each of nt streams per step, each stream's frame on
        # one thread (the streams are independent), through the detector
        # front end: blur 5x5 -> Background -> threshold -> open 3x3 -> CC.
        models = [oracle.Background(W, H, 4) for _ in range(nt)]
        bufs = [[np.empty((H, W), np.uint8) for _ in range(4)] for _ in range(nt)]

        def frame(k, i):
            b, m, t, o = bufs[k]
            f = frames[k % len(frames)][i % cycle]
            oracle.gaussian_blur(f, (5, 5), 0, dst=b)
            models[k].apply(b, m, 28)
            oracle.threshold(m, 127, 255, 0, dst=t)
            oracle.morphology(t, 2, 0, (3, 3), dst=o)
            oracle.connected_components_with_stats(o)

        line["cpu_baseline"] = bands.baseline(
            oracle.lib(), lambda i: [bands.Items([lambda k=k: frame(k, i) for k in range(nt)])], nt, "frames/s",
            f"one 1280x720 frame of each of {nt} streams per step, one stream per thread", nt)
    print(json.dumps(line))


def serve_kernel_roofline(torch, va, W, H, N, fused_bytes):
    """The fused per-frame kernel alone: a launch over N resident frames,
    timed with HIP events on the server's stream."""
    from visar_amd import serve
    with serve.Server(width=W, height=H, max_streams=N, max_batch=N, queue_depth=2, max_inflight=1) as srv:
        ids = [srv.admit() for _ in range(N)]
        f = np.zeros((H, W), np.uint8)
        for sid in ids:
            srv.submit(sid, f)
        srv.drain()
        srv.poll(1024)
        ms, n = srv.time_kernel(20)
    gbs = fused_bytes * n / (ms * 1e-3) / 1e9
    return {"achieved": round(gbs, 1), "frac": round(gbs / HBM_PEAK_GBS, 4), "ms_per_launch": round(ms, 4),
            "frames_per_launch": n}


def main():
    args = parse()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for --gpus {args.gpus}): images and camera streams shard "
              "across GPUs without exchange, so one GPU's throughput is the measured unit", file=sys.stderr)
        sys.exit(2)
    import torch

    import visar_amd as va
    from visar_amd import synth

    assert torch.cuda.is_available(), "bench.py needs a GPU"
    dev = torch.device("cuda:0")
    stream = torch.cuda.Stream()  # a dedicated stream (not the legacy default stream)
    torch.cuda.set_stream(stream)
    va.set_stream(stream)
    va.set_sync(False)

    if args.copy_calibration:
        n = 256 << 20
        a = torch.empty(n, dtype=torch.uint8, device=dev).fill_(1)
        b = torch.empty_like(a)
        for _ in range(4):
            b.copy_(a)
        torch.cuda.synchronize()
        print(json.dumps({"copy_bytes_per_launch": n, "launches": 4}))
        return
    if args.config == 3:
        return bench_morph(args, torch, va)
    if args.config == 4:
        return bench_batch(args, torch, va)
    if args.config == 5:
        return bench_serve(args, torch, va)

    R = 24
    if args.lanes == 0:
        args.lanes = {1: 1, 2: 3, 8: 4}.get(args.config, 2)
    NB = max(2, args.lanes)  # output buffer sets: step i writes set i % NB (one per frame in flight)
    if args.config == 2:
        W, H = 3840, 2160
        frames = [torch.from_numpy(synth.uniform_bytes(i, (H, W, 3))).to(dev) for i in range(R)]
        gray = [torch.empty((H, W), dtype=torch.uint8, device=dev) for _ in range(NB)]
        small = [torch.empty((H // 2, W // 2), dtype=torch.uint8, device=dev) for _ in range(NB)]
        sob = [torch.empty((H // 2, W // 2), dtype=torch.int16, device=dev) for _ in range(NB)]
        wa = [torch.empty((H // 2, W // 2), dtype=torch.uint8, device=dev) for _ in range(NB)]
        wp = [torch.empty((H // 2, W // 2), dtype=torch.uint8, device=dev) for _ in range(NB)]
        ma = synth.affine_matrix(0, W // 2, H // 2)
        mp = synth.perspective_matrix(0, W // 2, H // 2)
        sz = (W // 2, H // 2)
        ops = {
            "cvt_color_bgr2gray": (lambda i: va.cvt_color(frames[i % R], 6, dst=gray[i % NB]), W * H * 3 + W * H),
            "resize_linear_half": (lambda i: va.resize(gray[i % NB], sz, 1, dst=small[i % NB]), W * H + W * H // 4),
            "sobel_dx_i16": (lambda i: va.sobel(small[i % NB], np.int16, 1, 0, 3, dst=sob[i % NB]), W * H // 4 * 3),
            "warp_affine_linear": (lambda i: va.warp_affine(small[i % NB], ma, sz, 1, 0, 0, dst=wa[i % NB]), W * H // 2),
            "warp_perspective_linear": (lambda i: va.warp_perspective(small[i % NB], mp, sz, 1, 0, 0, dst=wp[i % NB]),
                                        W * H // 2),
        }
        pix_per_step = W * H
        workload = "chain 3840x2160x3 u8: gray, resize linear 1/2, sobel dx i16, warpAffine, warpPerspective"
    elif args.config == 6:
        W, H, K6 = 3840, 2160, 6
        frames = [torch.from_numpy(synth.uniform_bytes(i, (H, W, 4))).to(dev) for i in range(R)]
        outs = [torch.empty((H // K6, W // K6), dtype=torch.uint8, device=dev) for _ in range(NB)]
        ops = {"gray_downsample_k6": (lambda i: va.gray_downsample(frames[i % R], K6, dst=outs[i % NB]),
                                      W * H * 4 + (W // K6) * (H // K6))}
        pix_per_step = W * H
        workload = "grayDownsample 3840x2160 BGRA u8 by 6 to 640x360 (tracking image)"
    elif args.config == 7:
        W, H = 1280, 720
        frames = [torch.from_numpy(synth.uniform_bytes(i, (H, W, 3))).to(dev) for i in range(R)]
        outs = [torch.empty((H, W, 3), dtype=torch.uint8, device=dev) for _ in range(NB)]
        ops = {"undistort_bgr": (lambda i: va.undistort(frames[i % R], UD_K, UD_DIST, dst=outs[i % NB]), 2 * W * H * 3)}
        pix_per_step = W * H
        workload = "undistort 1280x720 BGR u8, K = [[900, 0, 640], [0, 900, 360], [0, 0, 1]], 5 coefficients"
    elif args.config == 8:
        W, H = 1920, 1080
        bgr = [torch.from_numpy(synth.uniform_bytes(i, (H, W, 3))).to(dev) for i in range(R)]
        gray = [torch.from_numpy(synth.uniform_bytes(100 + i, (H, W))).to(dev) for i in range(R)]
        fbgr = [(b.float() / 255.0).contiguous() for b in bgr[:4]]
        hd, wd = (H + 1) // 2, (W + 1) // 2
        def two(shape, dtype=torch.uint8):
            return [torch.empty(shape, dtype=dtype, device=dev) for _ in range(NB)]
        o_pd, o_pg, o_pu = two((hd, wd, 3)), two((hd, wd)), two((2 * H, 2 * W))
        o_pf, o_b3, o_b15 = two((hd, wd, 3), torch.float32), two((H, W, 3)), two((H, W, 3))
        o_am, o_ag = two((H, W)), two((H, W))
        ops = {
            "pyr_down_bgr": (lambda i: va.pyr_down(bgr[i % R], dst=o_pd[i % NB]), W * H * 3 + wd * hd * 3),
            "pyr_down_gray": (lambda i: va.pyr_down(gray[i % R], dst=o_pg[i % NB]), W * H + wd * hd),
            "pyr_up_gray": (lambda i: va.pyr_up(gray[i % R], dst=o_pu[i % NB]), W * H + 4 * W * H),
            "pyr_down_f32": (lambda i: va.pyr_down(fbgr[i % 4], dst=o_pf[i % NB]), 4 * (W * H * 3 + wd * hd * 3)),
            "blur_3x3_bgr": (lambda i: va.box_filter(bgr[i % R], (3, 3), dst=o_b3[i % NB]), 2 * W * H * 3),
            "blur_15x15_bgr": (lambda i: va.box_filter(bgr[i % R], (15, 15), dst=o_b15[i % NB]), 2 * W * H * 3),
            "adaptive_mean_11": (lambda i: va.adaptive_threshold(gray[i % R], 255, 0, 0, 11, 2, dst=o_am[i % NB]), 2 * W * H),
            "adaptive_gauss_11": (lambda i: va.adaptive_threshold(gray[i % R], 255, 1, 0, 11, 2, dst=o_ag[i % NB]), 2 * W * H),
        }
        pix_per_step = W * H
        workload = ("1920x1080: pyrDown BGR, gray, f32 BGR; pyrUp gray; blur 3x3 and 15x15 BGR; "
                    "adaptiveThreshold 11 mean and Gaussian (one frame each per step)")
    else:
        W, H = 1920, 1080
        frames = [torch.from_numpy(synth.uniform_bytes(i, (H, W))).to(dev) for i in range(R * 6)]
        outs = [torch.empty((H, W), dtype=torch.uint8, device=dev) for _ in range(NB)]
        ops = {"gaussian_blur_5x5": (lambda i: va.gaussian_blur(frames[i % len(frames)], (5, 5), 0, dst=outs[i % NB]),
                                     2 * W * H)}
        pix_per_step = W * H
        workload = "GaussianBlur 5x5 reflect-101 1920x1080 u8"
        R = len(frames)

    def step(i):
        for fn, _ in ops.values():
            fn(i)

    for i in range(args.warmup):
        step(i)
    torch.cuda.synchronize()

    # Graphs of G consecutive steps (input slots taken mod R, R a multiple of
    # G), so the GPU and not the host's launch rate sets the pace: K steps
    # replay K // G chunk graphs and then one graph of the K % G remaining
    # steps, captured the same way, so every timed step runs with the same
    # frames in flight whatever K is.
    # With L lanes, step j runs on lane j % L: lane 0 is the capture stream,
    # the others are streams forked from it inside the graph and joined at
    # its end, so L frames are in flight. Step j writes output set j % NB
    # (NB = L, at least 2), so the lanes never share a buffer.
    K = args.steps
    G = 24
    n_chunks = R // G
    cap_stream = torch.cuda.Stream()
    cap_stream.wait_stream(stream)
    side = [torch.cuda.Stream() for _ in range(args.lanes - 1)]

    def capture(indices):
        lanes = ([cap_stream] + side)[:max(1, min(len(indices), args.lanes))]
        if len(lanes) > 1:
            fork = torch.cuda.Event()
            fork.record(cap_stream)
            for st in lanes[1:]:
                st.wait_event(fork)
        for j in indices:
            st = lanes[j % len(lanes)]
            va.set_stream(st)
            with torch.cuda.stream(st):
                step(j)
        va.set_stream(cap_stream)
        for st in lanes[1:]:
            cap_stream.wait_stream(st)
        return len(lanes)

    # Chunk graph c holds steps c*G .. c*G + G - 1; the remainder graph
    # continues from step (K // G) * G, so consecutive graphs take
    # consecutive input slots.
    plan =[list(range(c * G, c * G + G)) for c in range(min(n_chunks, max(1, K // G)))] if K >= G else []
    rem = K % G
    rem_graph = None
    lanes_used = set()
    with torch.cuda.stream(cap_stream):
        va.set_stream(cap_stream)
        chunks = []
        for idx in plan:
            g = torch.cuda.CUDAGraph()
            with torch.cuda.graph(g, stream=cap_stream):
                lanes_used.add(capture(idx))
            chunks.append(g)
        if rem:
            first = ((K // G) * G) % R
            rem_graph = torch.cuda.CUDAGraph()
            with torch.cuda.graph(rem_graph, stream=cap_stream):
                lanes_used.add(capture(list(range(first, first + rem))))
    va.set_stream(stream)
    stream.wait_stream(cap_stream)
    torch.cuda.synchronize()
    for g in chunks + ([rem_graph] if rem_graph else []):  # graph warmup
        g.replay()
    torch.cuda.synchronize()

    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    e0.record(stream)
    for c in range(K // G):
        chunks[c % len(chunks)].replay()
    if rem_graph is not None:
        rem_graph.replay()
    e1.record(stream)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    gpu_ms = e0.elapsed_time(e1)
    ms_per_step = max(gpu_ms, wall * 1e3) / K
    graph_steps = [G] * (K // G) + ([rem] if rem else [])
    value = pix_per_step / (ms_per_step * 1e-3) / 1e9

    # Per-kernel durations, live, with HIP events on the launch stream.
    kernels = {}
    for name, (fn, nbytes) in ops.items():
        ms = per_launch_ms(torch, fn, 96, stream)
        kernels[name] = {"ms": round(ms, 5), "bytes": int(nbytes), "GB/s": round(nbytes / (ms * 1e-3) / 1e9, 1),
                         "frac_hbm_peak": round(nbytes / (ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 4)}
    dom = max(kernels, key=lambda k: kernels[k]["ms"])
    achieved = kernels[dom]["GB/s"]
    traffic = profiled_traffic(dom)

    cpu = None
    if not args.no_cpu:
        cpu = cpu_baseline(args, synth)

    line = {
        "metric": "Gpix/s (input pixels through the chain) + HBM GB/s per kernel" if args.config == 2 else
                  "Gpix/s (input pixels) + HBM GB/s per kernel",
        "value": round(value, 4),
        "unit": "Gpix/s",
        "n_gpus": 1,
        "steps": K,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 5),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u8",
        "data": "synthetic (splitmix64 uniform bytes, 24 rotating device frames)",
        "config": {"workload": workload, "width": W, "height": H, "rotating_inputs": R, "graph_replay": True,
                   "graph_steps": sorted(set(graph_steps)), "graphs_replayed": len(graph_steps),
                   "frames_in_flight": max(lanes_used)},
        "kernels": kernels,
        "roofline": {"kernel": dom, "bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic},
        "cpu_baseline": cpu,
    }
    print(json.dumps(line))


MORPH_SAMPLE = ((4, 0, 3, 1), (2, 2, 5, 2), (5, 1, 7, 3))  # (op, shape, size, iterations)


def morph_steps(oracle, src, case, out, tmp):
    """The ops of one morphologyEx case as the reference runs it: each
    erode or dilate pass one banded op over the rows (morph.zig:532-593),
    the saturating subtraction another. tmp: four scratch images."""
    from oracle import bands
    op, shape, ks, it = case
    mask = oracle.structuring_element(shape, ks, ks)
    S = src.shape[0]
    grain = max(16, 2 * ks)
    ops = []

    def passes(a, op_max, n, final, scratch):
        cur = a
        for k in range(n):
            dst = final if k == n - 1 else scratch[k % 2]
            ops.append(bands.Banded(oc(oracle, "orc_morph_pass", cur, dst, int(op_max), mask.ctypes.data_as(ctypes.c_void_p),
                                       ks, ks, ks // 2, ks // 2), S, grain))
            cur = dst
        ops[-1].call.keep.append(mask)

    t0, t1, t2, t3 = tmp
    if op in (0, 1):
        passes(src, op == 1, it, out, (t0, t1))
    elif op in (2, 3):
        passes(src, op == 3, it, t2, (t0, t1))
        passes(t2, op == 2, it, out, (t0, t1))
    elif op == 4:
        passes(src, True, it, t2, (t0, t1))
        passes(src, False, it, t3, (t0, t1))
        ops.append(bands.Banded(oc(oracle, "orc_sub_sat", t2, t3, out), S, grain))
    else:
        passes(src, op == 6, it, t2, (t0, t1))
        passes(t2, op == 5, it, t3, (t0, t1))
        a, b = (src, t3) if op == 5 else (t3, src)
        ops.append(bands.Banded(oc(oracle, "orc_sub_sat", a, b, out), S, grain))
    return ops


def morph_cpu_baseline(base):
    from oracle import bands
    oracle = cpu_lib()
    S = base.shape[0]
    out = np.empty_like(base)
    tmp = [np.empty_like(base) for _ in range(4)]
    steps = [op for case in MORPH_SAMPLE for op in morph_steps(oracle, base, case, out, tmp)]
    names = ", ".join(f"{MORPH_OPS[o]} {MORPH_ELEMENTS[[e[0] for e in MORPH_ELEMENTS].index(sh)][2]} it{it}"
                      for o, sh, _, it in MORPH_SAMPLE)
    return bands.baseline(oracle.lib(), lambda i: steps, len(MORPH_SAMPLE) * S * S / 1e9, "Gpix/s",
                          f"3 of the 63 cases on the 4096x4096 mask ({names}), every pass in row bands",
                          cpu_threads())


def cpu_baseline(args, synth):
    """The restatement, -O3 -march=native on this host, on the same
    workload: each step's ops one after the other, every op's destination
    rows split into one band per thread at the op's reference grain."""
    from oracle import bands
    oracle = cpu_lib()
    nt = cpu_threads()
    F = 4  # rotating host frames
    if args.config == 2:
        W, H = 3840, 2160
        ma = synth.affine_matrix(0, W // 2, H // 2)
        mp = synth.perspective_matrix(0, W // 2, H // 2)
        srcs = [synth.uniform_bytes(k, (H, W, 3)) for k in range(F)]
        g, sm = np.empty((H, W), np.uint8), np.empty((H // 2, W // 2), np.uint8)
        sob = np.empty((H // 2, W // 2), np.int16)
        wa, wp = np.empty_like(sm), np.empty_like(sm)
        h2 = H // 2
        steps = [[bands.Banded(oc(oracle, "orc_cvt_color", a, g, 6), H, 16),
                  bands.Banded(oc(oracle, "orc_resize", g, sm, 1), h2, 16),
                  bands.Banded(oc(oracle, "orc_sobel", sm, sob, 1, 0, 3, 1.0, 4), h2, max(8, 65536 // (W // 2))),
                  bands.Banded(oc(oracle, "orc_warp_affine", sm, wa, ma, 1, 0, 0.0), h2, 4),
                  bands.Banded(oc(oracle, "orc_warp_perspective", sm, wp, mp, 1, 0, 0.0), h2, 8)] for a in srcs]
        return bands.baseline(oracle.lib(), lambda i: steps[i % F], W * H / 1e9, "Gpix/s",
                              "whole chains (gray, resize, Sobel, warpAffine, warpPerspective) of 3840x2160x3", nt)
    if args.config == 6:
        W, H = 3840, 2160
        srcs = [synth.uniform_bytes(k, (H, W, 4)) for k in range(F)]
        o = np.empty((H // 6, W // 6), np.uint8)
        steps = [[bands.Banded(oc(oracle, "orc_gray_downsample", a, o, ctypes.c_uint32(6)), H // 6, 4)] for a in srcs]
        return bands.baseline(oracle.lib(), lambda i: steps[i % F], W * H / 1e9, "Gpix/s",
                              "grayDownsample of 3840x2160 BGRA by 6", nt)
    if args.config == 8:
        W, H = 1920, 1080
        hd, wd = (H + 1) // 2, (W + 1) // 2
        b = [synth.uniform_bytes(k, (H, W, 3)) for k in range(F)]
        gr = [synth.uniform_bytes(100 + k, (H, W)) for k in range(F)]
        f = [(x.astype(np.float32) / 255.0).astype(np.float32) for x in b]
        o_pd, o_pg, o_pu = np.empty((hd, wd, 3), np.uint8), np.empty((hd, wd), np.uint8), np.empty((2 * H, 2 * W), np.uint8)
        o_pf, o_b3, o_b15 = np.empty((hd, wd, 3), np.float32), np.empty((H, W, 3), np.uint8), np.empty((H, W, 3), np.uint8)
        o_am, o_ag = np.empty((H, W), np.uint8), np.empty((H, W), np.uint8)
        gbox = max(8, 65536 // (W * 3))

        def step(k):
            return [bands.Banded(oc(oracle, "orc_pyr_down", b[k], o_pd, 4), hd, 8),
                    bands.Banded(oc(oracle, "orc_pyr_down", gr[k], o_pg, 4), hd, 8),
                    bands.Banded(oc(oracle, "orc_pyr_up", gr[k], o_pu, 4), H, 8),
                    bands.Banded(oc(oracle, "orc_pyr_down", f[k], o_pf, 4), hd, 8),
                    bands.Banded(oc(oracle, "orc_box_filter", b[k], o_b3, 3, 3, 1, 4), H, gbox),
                    bands.Banded(oc(oracle, "orc_box_filter", b[k], o_b15, 15, 15, 1, 4), H, gbox),
                    # adaptiveThreshold builds a whole mean image first: one call, one thread.
                    bands.Items([oc(oracle, "orc_adaptive_threshold", gr[k], o_am, 255.0, 0, 0, 11, 2.0)]),
                    bands.Items([oc(oracle, "orc_adaptive_threshold", gr[k], o_ag, 255.0, 1, 0, 11, 2.0)])]
        steps = [step(k) for k in range(F)]
        return bands.baseline(oracle.lib(), lambda i: steps[i % F], W * H / 1e9, "Gpix/s",
                              "whole steps of the 8 pyramid / box / adaptive calls on 1920x1080 (the two "
                              "adaptiveThreshold calls on one thread each)", nt)
    if args.config == 7:
        W, H = 1280, 720
        srcs = [synth.uniform_bytes(k, (H, W, 3)) for k in range(F)]
        o = np.empty((H, W, 3), np.uint8)
        dist = np.asarray(UD_DIST, np.float64)
        steps = [[bands.Banded(oc(oracle, "orc_undistort", a, o, np.asarray(UD_K, np.float64), dist,
                                  ctypes.c_size_t(dist.size), None), H, 8)] for a in srcs]
        return bands.baseline(oracle.lib(), lambda i: steps[i % F], W * H / 1e9, "Gpix/s",
                              "undistort of 1280x720 BGR (the reference benchmark's camera)", nt)
    W, H = 1920, 1080
    srcs = [synth.uniform_bytes(k, (H, W)) for k in range(F)]
    o = np.empty((H, W), np.uint8)
    steps = [[bands.Banded(oc(oracle, "orc_gaussian_blur", a, o, 5, 5, 0.0, 0.0, 4), H, max(8, 65536 // W))]
             for a in srcs]
    return bands.baseline(oracle.lib(), lambda i: steps[i % F], W * H / 1e9, "Gpix/s",
                          "GaussianBlur 5x5 of 1920x1080", nt)


if __name__ == "__main__":
    main()
