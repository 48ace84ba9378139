"""Autotuner for the launch-shape knobs of libclkmi.

`tune(run, key, candidates)` times a workload under each value of one knob
(HIP events on the current stream, median of repeats), keeps the fastest and
records it.  Results persist as JSON ({knob: value} plus the device and the
timings) so a later process can `apply()` them instead of re-tuning.  The
knobs only change launch shapes (segment lengths, tasks per wave, rounds of
blocks), never results.
"""
import json
import os
import time

from . import get_tuning, set_tuning

DEFAULT_PATH = os.environ.get("CLKMI_TUNING",
                              os.path.join(os.path.expanduser("~"), ".cache", "clkmi", "tuning.json"))

CANDIDATES = {
    "batch_seg": [16, 32, 64, 128, 256],
    "strided_rounds": [1, 2, 4],
    "x_sub": [1, 2],
}


def time_ms(run, repeats=3, warmup=1):
    """Median wall time of run() in ms on the current stream."""
    import torch
    for _ in range(warmup):
        run()
    torch.cuda.synchronize()
    times = []
    for _ in range(repeats):
        a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        a.record()
        run()
        b.record()
        torch.cuda.synchronize()
        times.append(a.elapsed_time(b))
    times.sort()
    return times[len(times) // 2]


def tune(run, key, candidates=None, repeats=3):
    """Try each candidate value of `key` on `run`; leave the fastest set and
    return {"key", "best", "ms": {value: ms}}."""
    cands = candidates if candidates is not None else CANDIDATES[key]
    old = get_tuning(key)
    res = {}
    try:
        for v in cands:
            set_tuning(key, v)
            res[v] = time_ms(run, repeats=repeats)
    except Exception:
        set_tuning(key, old)
        raise
    best = min(res, key=res.get)
    set_tuning(key, best)
    return {"key": key, "best": best, "ms": {str(k): round(v, 4) for k, v in res.items()}}


def save(results, path=DEFAULT_PATH):
    import torch
    doc = {"device": torch.cuda.get_device_name(0), "time": time.time(),
           "knobs": {r["key"]: r["best"] for r in results}, "trials": results}
    os.makedirs(os.path.dirname(path), exist_ok=True)
    tmp = path + ".tmp"
    with open(tmp, "w") as f:
        json.dump(doc, f, indent=1)
    os.replace(tmp, path)
    return doc


def apply(path=DEFAULT_PATH):
    """Set the knobs recorded in `path`; returns them ({} if none)."""
    try:
        with open(path) as f:
            doc = json.load(f)
    except (OSError, ValueError):
        return {}
    knobs = doc.get("knobs", {})
    for k, v in knobs.items():
        set_tuning(k, int(v))
    return knobs
