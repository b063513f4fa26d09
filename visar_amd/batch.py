"""Batched execution (include/visar_batch.h): many images of uneven sizes and
mixed operations in one graph launch, plus the seeded mixed-op workload
(512 images, sides U{64..2048}, 1 or 3 channels, one op each)."""
import ctypes

import numpy as np

from . import Image, _check, _ref, desc, lib
from . import synth

GAUSSIAN_BLUR, RESIZE, WARP_AFFINE, MORPHOLOGY = 0, 1, 2, 3


class Item(ctypes.Structure):
    _fields_ = [
        ("op", ctypes.c_int),
        ("src", Image),
        ("dst", Image),
        ("ksize", ctypes.c_int),
        ("sigma", ctypes.c_double),
        ("border", ctypes.c_int),
        ("interpolation", ctypes.c_int),
        ("m", ctypes.c_double * 6),
        ("value", ctypes.c_double),
        ("morph_op", ctypes.c_int),
        ("shape", ctypes.c_int),
        ("iterations", ctypes.c_int),
    ]


def item(op, src, dst, ksize=0, sigma=0.0, border=4, interpolation=1, m=None, value=0.0, morph_op=1, shape=0,
         iterations=1):
    it = Item()
    it.op = op
    it.src = desc(src)
    it.dst = desc(dst)
    it.ksize = ksize
    it.sigma = sigma
    it.border = border
    it.interpolation = interpolation
    if m is not None:
        it.m = (ctypes.c_double * 6)(*np.asarray(m, np.float64).ravel()[:6])
    it.value = value
    it.morph_op = morph_op
    it.shape = shape
    it.iterations = iterations
    return it


class Batch:
    """A planned batch; run() launches it (replaying a HIP graph by default)."""

    def __init__(self):
        self._h = ctypes.c_void_p()
        _check(lib.visar_batch_create(ctypes.byref(self._h)))
        self._keep = None

    def plan(self, items):
        arr = (Item * len(items))(*items)
        bad = ctypes.c_size_t(0)
        rc = lib.visar_batch_plan(self._h, arr, ctypes.c_size_t(len(items)), ctypes.byref(bad))
        if rc != 0:
            raise RuntimeError(f"item {bad.value}: {rc} {lib.visar_last_error().decode()}")
        self._keep = (arr, items)
        return self

    def run(self, use_graph=True):
        _check(lib.visar_batch_run(self._h, int(bool(use_graph))))

    def info(self):
        t, w = ctypes.c_size_t(0), ctypes.c_size_t(0)
        _check(lib.visar_batch_info(self._h, ctypes.byref(t), ctypes.byref(w)))
        return t.value, w.value

    def __del__(self):
        # At interpreter exit the module globals may already be gone.
        if getattr(self, "_h", None) and self._h.value and lib is not None:
            lib.visar_batch_free(self._h)
            self._h = None


# ----------------------------------------------------- the mixed workload

def workload(seed=0, n=512, lo=64, hi=2048, device="cuda"):
    """Device inputs and outputs of the seeded mixed batch. Returns a list of
    dicts: spec, src, dst and the keyword arguments of the op."""
    import torch
    jobs = []
    for i, (w, h, cn, op, cseed, oseed) in enumerate(synth.batch_specs(seed, n, lo, hi)):
        shape = (h, w) if cn == 1 else (h, w, cn)
        src = torch.from_numpy(synth.uniform_bytes(cseed, shape)).to(device)
        if op == "area_half":
            dw, dh = w // 2, h // 2
        elif op == "linear_1p5":
            dw, dh = int(round(1.5 * w)), int(round(1.5 * h))
        else:
            dw, dh = w, h
        dst = torch.empty((dh, dw) if cn == 1 else (dh, dw, cn), dtype=torch.uint8, device=device)
        kw = {}
        if op.startswith("gauss"):
            kw = dict(op=GAUSSIAN_BLUR, ksize=int(op[-1]), sigma=0.0, border=4)
        elif op == "area_half":
            kw = dict(op=RESIZE, interpolation=3)
        elif op == "linear_1p5":
            kw = dict(op=RESIZE, interpolation=1)
        elif op == "warp_affine":
            kw = dict(op=WARP_AFFINE, m=synth.affine_matrix(oseed, w, h), interpolation=1, border=0, value=0.0)
        else:
            kw = dict(op=MORPHOLOGY, ksize=5, morph_op=1, shape=0, iterations=1)
        jobs.append({"spec": (w, h, cn, op), "src": src, "dst": dst, "kw": kw})
    return jobs


def items(jobs):
    return [item(src=j["src"], dst=j["dst"], **j["kw"]) for j in jobs]


def run_plain(va, j, dst=None):
    """The same work as one batch item, through the plain C-ABI call."""
    kw, src = j["kw"], j["src"]
    dst = j["dst"] if dst is None else dst
    op = kw["op"]
    if op == GAUSSIAN_BLUR:
        return va.gaussian_blur(src, (kw["ksize"], kw["ksize"]), kw["sigma"], border=kw["border"], dst=dst)
    if op == RESIZE:
        return va.resize(src, (dst.shape[1], dst.shape[0]), kw["interpolation"], dst=dst)
    if op == WARP_AFFINE:
        return va.warp_affine(src, kw["m"], (dst.shape[1], dst.shape[0]), kw["interpolation"], kw["border"],
                              kw["value"], dst=dst)
    return va.morphology(src, kw["morph_op"], kw["shape"], (kw["ksize"], kw["ksize"]), kw["iterations"], dst=dst)


def algorithmic_bytes(jobs):
    """Sum of every image's input and output bytes."""
    return int(sum(j["src"].numel() + j["dst"].numel() for j in jobs))
