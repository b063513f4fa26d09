"""Autotuner for the launch-shape knobs of libvisar_amd (visar_gpu_set_tuning).

A tuning problem is a workload function (launches work on the current
stream) and a grid of knob settings. Each setting is applied, the workload is
warmed up and timed with HIP events (median of a few repetitions), and the
fastest setting is kept. Results are cached as JSON keyed by the device name
and a workload signature, so a later process with the same device and
workload applies the cached setting without timing anything.
"""
import ctypes
import itertools
import json
import os
import time

from . import _check, device_info, lib

CACHE_ENV = "VISAR_AMD_TUNE_CACHE"


def cache_path():
    p = os.environ.get(CACHE_ENV)
    if p:
        return p
    base = os.environ.get("XDG_CACHE_HOME") or os.path.join(os.path.expanduser("~"), ".cache")
    return os.path.join(base, "visar_amd", "tuning.json")


def _load():
    try:
        with open(cache_path()) as f:
            return json.load(f)
    except (OSError, ValueError):
        return {}


def _store(db):
    path = cache_path()
    try:
        os.makedirs(os.path.dirname(path), exist_ok=True)
        tmp = path + f".{os.getpid()}"
        with open(tmp, "w") as f:
            json.dump(db, f, indent=1, sort_keys=True)
        os.replace(tmp, path)
    except OSError:
        pass  # a read-only home only loses the cache


def set_knob(key, value):
    _check(lib.visar_gpu_set_tuning(key.encode(), int(value)))


def get_knob(key):
    v = ctypes.c_int(0)
    _check(lib.visar_gpu_get_tuning(key.encode(), ctypes.byref(v)))
    return v.value


def apply(setting):
    for k, v in setting.items():
        set_knob(k, v)


def time_ms(run, reps=5, inner=3):
    """Median over `reps` of the average of `inner` back-to-back runs (HIP events)."""
    import torch
    stream = torch.cuda.current_stream()
    out = []
    for _ in range(reps):
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record(stream)
        for _ in range(inner):
            run()
        e1.record(stream)
        torch.cuda.synchronize()
        out.append(e0.elapsed_time(e1) / inner)
    out.sort()
    return out[len(out) // 2]


def tune(name, signature, space, run, prepare=None, force=False):
    """Picks the fastest setting of `space` ({knob: [values]}) for `run()`.
    `prepare(setting)` is called after applying a setting and before timing
    (e.g. to re-plan a batch). Returns (setting, {setting: ms})."""
    key = f"{device_info()}|{name}|{signature}"
    db = _load()
    if not force and key in db:
        best = db[key]["best"]
        apply(best)
        if prepare:
            prepare(best)
        return best, db[key]["timings"]
    knobs = sorted(space)
    timings = {}
    best, best_ms = None, float("inf")
    for values in itertools.product(*(space[k] for k in knobs)):
        setting = dict(zip(knobs, values))
        apply(setting)
        if prepare:
            prepare(setting)
        run()
        run()
        ms = time_ms(run)
        timings[json.dumps(setting, sort_keys=True)] = ms
        if ms < best_ms:
            best, best_ms = setting, ms
    apply(best)
    if prepare:
        prepare(best)
    db[key] = {"best": best, "timings": timings, "when": time.strftime("%Y-%m-%dT%H:%M:%S")}
    _store(db)
    return best, timings


def batch_signature(specs):
    """Workload signature of a batch: counts per (op, channels, size class)."""
    counts = {}
    for w, h, cn, op in specs:
        mp = w * h / 1e6
        size = "s" if mp < 0.25 else "m" if mp < 1.5 else "l"
        k = f"{op}/{cn}/{size}"
        counts[k] = counts.get(k, 0) + 1
    return ",".join(f"{k}:{counts[k]}" for k in sorted(counts))


BATCH_SPACE = {"batch_stencil_rows": [1, 2, 4], "batch_unit_kb": [4, 8, 16, 32, 64]}
STENCIL_SPACE = {"stencil_rows": [1, 2, 4]}


def tune_batch(batch_obj, items, specs, force=False):
    """Tunes the batch launch of `items` (visar_amd.batch.Item list)."""
    def prepare(_setting):
        batch_obj.plan(items)

    def run():
        batch_obj.run(use_graph=True)

    return tune("batch", batch_signature(specs), BATCH_SPACE, run, prepare, force)


def tune_stencil(run, signature, force=False):
    """Tunes rows per thread of the plain small-kernel stencils for `run()`."""
    return tune("stencil", signature, STENCIL_SPACE, run, None, force)
