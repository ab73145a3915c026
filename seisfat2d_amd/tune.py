"""Schedule tuner for batched solves of the persistent tiled schedule.

The tile is fixed at 64 x 64: one node row per lane of a 64-wide wave (the
DPP shifts between neighbouring rows need the whole strip in one wave) and
64 steps, which is what a 16 KB LDS ring of old traveltimes holds.  What
varies per workload is how the tiles are launched and ordered, the options
of seisfat_sched_opts:

  pairs     both sweeps of a pair of sweeps in one launch (1) or one
            persistent launch per sweep (0)
  wpcu      persistent workgroups per CU (0 = as many as fit)
  skip      tiles whose every input is the not-reached value are not run
            where that is exact (1) or every tile runs (0)
  priority  wave priorities from each shot's scheduling slack (1) or equal (0)

Every combination gives bit-identical traveltimes, so tuning only times
them: coordinate descent from the defaults, each candidate a fresh batch
timed by the median device time of its sweep phase (HIP events inside the
graph) over `runs` runs after `warmup`; a candidate replaces the best only
when at least 0.5% faster.  The result is cached on disk per workload key:
device name, grid, spacings, solve type, shot count and the histogram of
window widths in 64-column bins (so two batches with the same geometry and
width mix share an entry), in the JSON file named by SEISFAT_TUNE_CACHE, or
~/.cache/seisfat2d_amd/schedule.json.
"""
from __future__ import annotations

import json
import os
import tempfile
import time

import numpy as np

from .capi import ISO

VERSION = 1
DEFAULT = {"pairs": 1, "wpcu": 0, "skip": 1, "priority": 1}
SPACE = {"pairs": (1, 0), "wpcu": (0, 8, 7, 6), "skip": (1, 0), "priority": (1, 0)}
ORDER = ("pairs", "wpcu", "skip", "priority")
MIN_GAIN = 0.005


def cache_path():
    p = os.environ.get("SEISFAT_TUNE_CACHE")
    if p:
        return p
    return os.path.join(os.path.expanduser("~"), ".cache", "seisfat2d_amd", "schedule.json")


def device_name():
    try:
        import torch
        if torch.cuda.is_available():
            p = torch.cuda.get_device_properties(0)
            return f"{p.name}/{getattr(p, 'gcnArchName', '')}/{p.multi_processor_count}CU"
    except Exception:
        pass
    return "unknown-device"


def width_histogram(shots):
    """Window widths in 64-column bins, as "bin:count" pairs of the occupied bins."""
    w = np.asarray([int(s[3]) for s in shots], np.int64)
    if w.size == 0:
        raise ValueError("no shots")
    h = np.bincount(w // 64)
    return ",".join(f"{i}:{c}" for i, c in enumerate(h) if c)


def workload_key(eik, shots, type=ISO, device=None):
    return (f"v{VERSION}|{device or device_name()}|{eik.nz}x{eik.nx}|dz={float(eik.dz)!r}|dx={float(eik.dx)!r}"
            f"|type={int(type)}|n={len(shots)}|w={width_histogram(shots)}")


def _load(path):
    try:
        with open(path) as f:
            d = json.load(f)
        return d if isinstance(d, dict) else {}
    except (OSError, ValueError):
        return {}


def _store(path, key, entry):
    d = _load(path)
    d[key] = entry
    os.makedirs(os.path.dirname(os.path.abspath(path)), exist_ok=True)
    fd, tmp = tempfile.mkstemp(dir=os.path.dirname(os.path.abspath(path)), suffix=".tmp")
    with os.fdopen(fd, "w") as f:
        json.dump(d, f, indent=1, sort_keys=True)
    os.replace(tmp, path)


def measure(eik, shots, rx, rz, type, opts, runs=5, warmup=2):
    """Median device time (ms) of the sweep phase of a batch with these
    options, and the options it ran with."""
    b = eik.batch(shots, rx, rz, type=type, opts=opts)
    try:
        for _ in range(warmup):
            b.run()
        b.status()
        t = []
        for _ in range(runs):
            b.run()
            t.append(b.sweep_ms())
        b.status()
        return float(np.median(t)), b.opts
    finally:
        b.close()


def tune(eik, shots, rx=None, rz=None, type=ISO, runs=5, warmup=2, cache=True, force=False, measure=measure,
         device=None, log=None):
    """Schedule options for this workload: from the cache, or measured (and
    cached) when it has none or force is set.  Returns a dict with "opts",
    "ms" (the best median sweep time), "key", "cached" and "trials" (each
    measured candidate: requested options, the options in effect, ms)."""
    key = workload_key(eik, shots, type, device)
    path = cache_path()
    if cache and not force:
        hit = _load(path).get(key)
        if hit and set(hit.get("opts", {})) == set(DEFAULT):
            return {"opts": dict(hit["opts"]), "ms": hit.get("ms"), "key": key, "cached": True, "trials": []}
    trials = []
    seen = {}

    def run(opts):
        k = tuple(opts[n] for n in ORDER)
        if k not in seen:
            ms, eff = measure(eik, shots, rx, rz, type, dict(opts), runs=runs, warmup=warmup)
            seen[k] = ms
            trials.append({"opts": dict(opts), "effective": dict(eff), "ms": ms})
            if log:
                log(f"tune {opts} -> {ms:.3f} ms")
        return seen[k]

    best = dict(DEFAULT)
    best_ms = run(best)
    for name in ORDER:
        for v in SPACE[name]:
            if v == best[name]:
                continue
            cand = dict(best)
            cand[name] = v
            ms = run(cand)
            if ms < best_ms * (1.0 - MIN_GAIN):
                best, best_ms = cand, ms
    entry = {"opts": best, "ms": best_ms, "trials": trials, "when": time.time()}
    if cache:
        _store(path, key, entry)
    return {"opts": dict(best), "ms": best_ms, "key": key, "cached": False, "trials": trials}
