"""Autotuner for the per-plan launch knobs (dsht_set_tuning).

For one plan and a representative column count it times the dynamics-step transform set
(3 inverse kinds over 4 fields, 1 forward over 6 tendencies) under each candidate
(column chunk x contraction column block) and keeps the fastest.  Results are cached per
(nmax, im, jm, precision, ncols bucket) in a JSON file so later runs reuse them.
"""
import json
import os
import time

from . import GRADLAT, GRADLON, TUNE_CHUNK_COLS, TUNE_LEG_CB, VAL

CACHE = os.environ.get("DSHT_TUNE_CACHE", os.path.join(os.path.expanduser("~"), ".cache", "dsht", "tuning.json"))


def _bucket(ncols):
    b = 16
    while b < ncols:
        b *= 2
    return b


def _key(plan, ncols):
    return f"{plan.nmax}x{plan.im}x{plan.jm}:p{plan.precision}:c{_bucket(ncols)}"


def candidates(plan, ncols):
    chunks = sorted({c for c in (32, 64, 128, 256, 512) if c < ncols} | {0})
    return [(c, cb) for c in chunks for cb in (16, 32)]


def _time_step(plan, bufs, stream, reps=3):
    import torch
    spec, grids, tend, spect = bufs
    for _ in range(2):
        for k, kind in enumerate((VAL, GRADLON, GRADLAT)):
            plan.inv_d(kind, spec, grids[k], stream=stream)
        plan.fwd_d(VAL, tend, spect, stream=stream)
    torch.cuda.synchronize()
    t = time.perf_counter()
    for _ in range(reps):
        for k, kind in enumerate((VAL, GRADLON, GRADLAT)):
            plan.inv_d(kind, spec, grids[k], stream=stream)
        plan.fwd_d(VAL, tend, spect, stream=stream)
    torch.cuda.synchronize()
    return (time.perf_counter() - t) / reps


def tune(plan, ncols, dev=None, reps=3, use_cache=True, verbose=False):
    """Pick (chunk, column block) for `plan` at `ncols` prognostic columns; applies and returns it."""
    import torch

    cache = {}
    if use_cache and os.path.exists(CACHE):
        try:
            cache = json.load(open(CACHE))
        except (OSError, ValueError):
            cache = {}
    key = _key(plan, ncols)
    if key in cache:
        chunk, cb = cache[key]["chunk"], cache[key]["cb"]
        plan.set_tuning(TUNE_CHUNK_COLS, chunk)
        plan.set_tuning(TUNE_LEG_CB, cb)
        return cache[key]
    dev = dev or torch.device("cuda:0")
    f64 = dict(dtype=torch.float64, device=dev)
    spec = torch.randn((plan.lm, ncols), **f64) * 1e-3
    grids = [torch.empty((ncols, plan.jm, plan.im), **f64) for _ in range(3)]
    nt = max(1, ncols * 3 // 2)
    tend = torch.randn((nt, plan.jm, plan.im), **f64)
    spect = torch.empty((plan.lm, nt), **f64)
    stream = torch.cuda.current_stream(dev).cuda_stream
    best = None
    results = []
    for chunk, cb in candidates(plan, ncols):
        plan.set_tuning(TUNE_CHUNK_COLS, chunk)
        plan.set_tuning(TUNE_LEG_CB, cb)
        dt = _time_step(plan, (spec, grids, tend, spect), stream, reps)
        results.append({"chunk": chunk, "cb": cb, "ms": 1e3 * dt})
        if verbose:
            print(f"  tune {key}: chunk={chunk} cb={cb} {1e3 * dt:.3f} ms", flush=True)
        if best is None or dt < best[0]:
            best = (dt, chunk, cb)
    _, chunk, cb = best
    plan.set_tuning(TUNE_CHUNK_COLS, chunk)
    plan.set_tuning(TUNE_LEG_CB, cb)
    entry = {"chunk": chunk, "cb": cb, "ms": 1e3 * best[0], "candidates": results}
    if use_cache:
        cache[key] = entry
        try:
            os.makedirs(os.path.dirname(CACHE), exist_ok=True)
            json.dump(cache, open(CACHE, "w"), indent=1)
        except OSError:
            pass
    return entry
