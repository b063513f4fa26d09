"""Timing of the CPU restatement as a baseline: one image split across host
threads in row bands, as the reference's Exec.parallelFor does
(src/core/pool.zig:31-37, 177-187), on a -O3 -march=native build made on the
host that runs it.

TEST AND BASELINE INFRASTRUCTURE ONLY, like the rest of this package: the
CPU-baseline leg of bench.py and the tests use it; the product package never
imports it.

A step is a list of ops run in order, each finished on every thread before
the next starts (the reference's parallelFor joins). An op is either banded
(a call covering destination rows [lo, hi), split into one band per thread
at multiples of the op's grain: pool.zig claims chunks of at least `grain`
rows; one band per thread is the static form of that split for ops whose
rows cost the same) or a list of whole calls shared round-robin by the
threads (independent images or streams).
"""
import ctypes
import os
import statistics
import subprocess
import threading
import time

HERE = os.path.dirname(os.path.abspath(__file__))
BASELINE_PATH = os.path.join(HERE, "_build", "libbaseline.so")
BASELINE_FLAGS = "-O3 -march=native -ffp-contract=off -fno-fast-math"


def build_baseline():
    """Builds the restatement with BASELINE_FLAGS for this host's CPU (a few
    seconds) and returns the library path."""
    subprocess.run(["make", "-s", "-B", "-C", HERE, "baseline"], check=True)
    return BASELINE_PATH


def cpu_info():
    model = "unknown"
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    try:
        usable = len(os.sched_getaffinity(0))
    except AttributeError:
        usable = os.cpu_count() or 1
    return {"cpu_model": model, "nproc": os.cpu_count() or 1, "usable_cpus": usable}


class Banded:
    """call(): one oracle call whose destination rows are limited by the
    thread's band; rows: its destination row count; grain: rows per band
    at least (pool.zig's grain for the op)."""

    def __init__(self, call, rows, grain=8):
        self.call, self.rows, self.grain = call, int(rows), max(1, int(grain))


class Items:
    """Independent whole calls, shared round-robin by the threads."""

    def __init__(self, calls):
        self.calls = list(calls)


def bands(rows, grain, n):
    """[lo, hi) of each of up to n bands covering rows, band edges at
    multiples of grain (fewer bands when rows / grain < n)."""
    chunks = max(1, min(n, -(-rows // grain)))
    per = -(-rows // chunks)
    per = -(-per // grain) * grain
    out = []
    lo = 0
    while lo < rows:
        out.append((lo, min(rows, lo + per)))
        lo += per
    return out


class BandPool:
    """`threads` persistent host threads running steps; the oracle's C calls
    release Python's GIL, so the bands run in parallel."""

    def __init__(self, lib, threads):
        self.lib = lib
        self.n = max(1, int(threads))
        self._start = threading.Barrier(self.n + 1)
        self._done = threading.Barrier(self.n + 1)
        self._work = None
        self._stop = False
        self._errors = []
        self._threads = [threading.Thread(target=self._loop, args=(k,), daemon=True) for k in range(self.n)]
        for t in self._threads:
            t.start()

    def _loop(self, k):
        set_band = self.lib.orc_set_band
        set_band.argtypes = [ctypes.c_long, ctypes.c_long]
        while True:
            self._start.wait()
            if self._stop:
                return
            op = self._work
            try:
                if isinstance(op, Banded):
                    b = bands(op.rows, op.grain, self.n)
                    if k < len(b):
                        set_band(b[k][0], b[k][1])
                        try:
                            op.call()
                        finally:
                            set_band(0, -1)
                else:
                    for i in range(k, len(op.calls), self.n):
                        op.calls[i]()
            except Exception as e:  # reported by run()
                self._errors.append(e)
            self._done.wait()

    def run(self, step):
        for op in step:
            self._work = op
            self._start.wait()
            self._done.wait()
            if self._errors:
                raise self._errors[0]

    def close(self):
        self._stop = True
        self._start.wait()
        for t in self._threads:
            t.join()


def time_step(pool, make_step, runs=7, warmup=1):
    """Seconds of each of `runs` timed runs of one step (after `warmup`
    untimed ones). make_step(i) returns the ops of run i (rotating inputs)."""
    for i in range(warmup):
        pool.run(make_step(i))
    ts = []
    for i in range(runs):
        step = make_step(warmup + i)
        t0 = time.perf_counter()
        pool.run(step)
        ts.append(time.perf_counter() - t0)
    return ts


def baseline(lib, make_step, per_step, unit, what, threads, runs=7):
    """The cpu_baseline object: best-of and median rates over `runs` runs at
    one thread and at `threads` threads."""
    out = {}
    for t in sorted({1, threads}):
        pool = BandPool(lib, t)
        try:
            ts = time_step(pool, make_step, runs)
        finally:
            pool.close()
        out[t] = ts
    tn, t1 = out[threads], out[1]
    info = cpu_info()
    return {
        "value": round(per_step / min(tn), 6), "unit": unit, "cores": threads, "kind": "port",
        "value_median": round(per_step / statistics.median(tn), 6),
        "value_1_thread": round(per_step / min(t1), 6),
        "value_1_thread_median": round(per_step / statistics.median(t1), 6),
        "ms_per_step_best": round(min(tn) * 1e3, 3), "ms_per_step_1_thread_best": round(min(t1) * 1e3, 3),
        "runs": runs, "build": BASELINE_FLAGS, **info,
        "sample": f"{what}: best of {runs} runs (and median) on {threads} threads, one image in row bands per op "
                  f"(static split at each op's grain), and on one thread; the scalar C restatement of the "
                  f"reference arithmetic, built {BASELINE_FLAGS} on this host",
    }
