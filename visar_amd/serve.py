"""Served tracking front end (include/visar_serve.h): many camera streams,
continuous batching on one GPU, per-stream background models resident on the
device, host connected components, per-frame latency.

Also the per-frame pipeline through the plain C-ABI calls (the path the
server is compared with) and the synthetic camera streams of the served
workload: 1280x720 gray, gradient 60 -> 140 plus N(0, 4^2) noise, five
ellipses of intensity 210 with radii U(10, 40) moving at constant seeded
velocities of up to 6 px/frame and bouncing off the edges.
"""
import ctypes

import numpy as np

from . import VisarError, _check, lib


class Config(ctypes.Structure):
    _fields_ = [
        ("width", ctypes.c_uint32), ("height", ctypes.c_uint32),
        ("shift", ctypes.c_int), ("threshold", ctypes.c_int), ("mask_threshold", ctypes.c_int),
        ("min_area", ctypes.c_uint32), ("hbm_budget", ctypes.c_uint64),
        ("max_streams", ctypes.c_uint32), ("max_batch", ctypes.c_uint32), ("queue_depth", ctypes.c_uint32),
        ("host_workers", ctypes.c_uint32), ("max_inflight", ctypes.c_uint32), ("batch_window_us", ctypes.c_uint32),
    ]


class Result(ctypes.Structure):
    _fields_ = [
        ("stream_id", ctypes.c_int32), ("status", ctypes.c_int32), ("tag", ctypes.c_uint64),
        ("frame_index", ctypes.c_uint64), ("components", ctypes.c_uint32), ("objects", ctypes.c_uint32),
        ("checksum", ctypes.c_uint64), ("t_submit", ctypes.c_double), ("t_done", ctypes.c_double),
        ("batch_size", ctypes.c_uint32), ("reserved", ctypes.c_uint32),
    ]


class Stats(ctypes.Structure):
    _fields_ = [
        ("frames_done", ctypes.c_uint64), ("batches", ctypes.c_uint64), ("admitted", ctypes.c_uint64),
        ("rejected", ctypes.c_uint64), ("device_bytes", ctypes.c_uint64), ("bytes_per_stream", ctypes.c_uint64),
        ("max_batch_seen", ctypes.c_uint32), ("active_streams", ctypes.c_uint32),
    ]


lib.visar_serve_now.restype = ctypes.c_double
lib.visar_serve_now.argtypes = [ctypes.c_void_p]
lib.visar_serve_free.argtypes = [ctypes.c_void_p]
lib.visar_serve_submit.argtypes = [ctypes.c_void_p, ctypes.c_int32, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_uint64,
                                   ctypes.c_void_p, ctypes.c_size_t]

RESULT_FIELDS = [f for f, _ in Result._fields_ if f != "reserved"]


def default_config(**overrides):
    c = Config()
    lib.visar_serve_config_default(ctypes.byref(c))
    for k, v in overrides.items():
        setattr(c, k, v)
    return c


class Server:
    def __init__(self, **config):
        self.config = default_config(**config)
        self._h = ctypes.c_void_p()
        _check(lib.visar_serve_create(ctypes.byref(self.config), ctypes.byref(self._h)))
        self._buf = (Result * 1024)()

    def admit(self):
        sid = ctypes.c_int32(-1)
        _check(lib.visar_serve_admit(self._h, ctypes.byref(sid)))
        return sid.value

    def try_admit(self):
        """The new stream id, or None when the HBM budget (or max_streams) is full."""
        try:
            return self.admit()
        except VisarError as e:
            if e.rc == -4:
                return None
            raise

    def release(self, sid):
        _check(lib.visar_serve_release(self._h, sid))

    def submit(self, sid, frame, tag=0, mask_out=None):
        """Copies a (height, width) u8 frame; mask_out (same shape) receives the frame's mask."""
        if frame.dtype != np.uint8 or frame.ndim != 2 or frame.strides[1] != 1:
            raise ValueError("frame must be a 2-D u8 array with packed rows")
        mo, ms = (None, 0) if mask_out is None else (mask_out.ctypes.data, mask_out.strides[0])
        _check(lib.visar_serve_submit(self._h, sid, frame.ctypes.data, frame.strides[0], tag, mo, ms))

    def poll(self, max_results=1024, wait_ms=0):
        n = ctypes.c_size_t(0)
        m = min(max_results, len(self._buf))
        _check(lib.visar_serve_poll(self._h, self._buf, m, ctypes.byref(n), wait_ms))
        return [{f: getattr(self._buf[i], f) for f in RESULT_FIELDS} for i in range(n.value)]

    def drain(self):
        _check(lib.visar_serve_drain(self._h))

    def stats(self):
        s = Stats()
        _check(lib.visar_serve_get_stats(self._h, ctypes.byref(s)))
        return {f: getattr(s, f) for f, _ in Stats._fields_}

    def now(self):
        return lib.visar_serve_now(self._h)

    def time_kernel(self, reps=20):
        """(ms per launch, frames per launch) of the fused kernel over every admitted stream."""
        ms, n = ctypes.c_double(0), ctypes.c_uint32(0)
        _check(lib.visar_serve_time_kernel(self._h, reps, ctypes.byref(ms), ctypes.byref(n)))
        return ms.value, n.value

    def close(self):
        if getattr(self, "_h", None) and self._h.value:
            lib.visar_serve_free(self._h)
            self._h = ctypes.c_void_p()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        self.close()


# ------------------------------------------------------ reference pipelines

class PlainPipeline:
    """The same front end through the plain C-ABI calls, one synchronous call
    after the other (each returns with its output written): the frame is
    uploaded once, the blur, background, threshold and open run on device
    buffers, the mask comes back and is labelled on the host. This is the
    per-call path the server is measured against."""

    def __init__(self, va, w=1280, h=720, shift=4, threshold=28, mask_threshold=127):
        import torch
        self.va = va
        self.bg = va.Background(w, h, shift)
        self.threshold, self.mask_threshold = threshold, mask_threshold
        dev = torch.device("cuda")
        self.f, self.b, self.m, self.t, self.o = (torch.empty((h, w), dtype=torch.uint8, device=dev) for _ in range(5))
        self.host = np.empty((h, w), np.uint8)

    def run(self, frame):
        import torch
        va = self.va
        self.f.copy_(torch.from_numpy(frame))
        torch.cuda.current_stream().synchronize()  # the upload is done before the library's stream reads it
        va.gaussian_blur(self.f, (5, 5), 0, dst=self.b)
        self.bg.apply(self.b, self.m, self.threshold)
        va.threshold(self.m, self.mask_threshold, 255, 0, dst=self.t)
        va.morphology(self.t, 2, 0, (3, 3), dst=self.o)
        self.host[:] = self.o.cpu().numpy()
        n = va.connected_components_with_stats(self.host)[0]
        return self.host, n


# ------------------------------------------------------ synthetic streams

class CameraSet:
    """Synthetic camera streams. The static part of each frame (gradient plus
    noise) comes from a pool of pre-drawn noise fields, one picked per frame
    by a seeded hash, so frames are produced at camera rate on one host
    thread; the ellipses are drawn per frame at their current positions."""

    def __init__(self, n_streams, w=1280, h=720, seed=0, pool=32):
        rng = np.random.default_rng(seed)
        grad = np.tile(np.linspace(60, 140, w), (h, 1))
        self.pool = [np.clip(np.rint(grad + rng.normal(0, 4, (h, w))), 0, 255).astype(np.uint8) for _ in range(pool)]
        self.w, self.h = w, h
        self.objects = []
        for s in range(n_streams):
            r = np.random.default_rng(seed * 1000 + s + 1)
            obj = []
            for _ in range(5):
                rx, ry = r.uniform(10, 40, 2)
                x = r.uniform(min(rx, w / 2), max(w - rx, w / 2))  # frames narrower than an ellipse: anywhere
                y = r.uniform(min(ry, h / 2), max(h - ry, h / 2))
                obj.append([x, y, r.uniform(-6, 6), r.uniform(-6, 6), rx, ry])
            self.objects.append(obj)
        self.count = [0] * n_streams
        self.buf = [np.empty((h, w), np.uint8) for _ in range(n_streams)]

    def next_frame(self, s):
        """Frame of stream s at its next time step (a view reused per stream)."""
        k = self.count[s]
        self.count[s] += 1
        img = self.buf[s]
        img[:] = self.pool[(s * 7919 + k * 104729) % len(self.pool)]
        for o in self.objects[s]:
            x, y, vx, vy, rx, ry = o
            x0, x1 = int(max(0, x - rx)), int(min(self.w, x + rx + 1))
            y0, y1 = int(max(0, y - ry)), int(min(self.h, y + ry + 1))
            yy, xx = np.ogrid[y0:y1, x0:x1]
            inside = ((xx - x) / rx) ** 2 + ((yy - y) / ry) ** 2 <= 1
            img[y0:y1, x0:x1][inside] = 210
            x, y = x + vx, y + vy
            if x < rx or x > self.w - rx:
                vx = -vx
            if y < ry or y > self.h - ry:
                vy = -vy
            o[:] = [x, y, vx, vy, rx, ry]
        return img
