"""Launch-shape and layout autotuner.

Three knobs are tuned by measurement on the case itself:

* the step kernel's work-group width (block_x: 64, 128, 192 or 256 cells of
  one x-row per work-group), timed with the library's HIP events around every
  step launch on the solver's own stream, on a scratch solver of the same
  case so the run being tuned is not advanced;
* for batched sweeps, the number of cases per batched launch (all cases in
  one launch, or groups on separate streams that the GPU overlaps), timed
  as wall time of a few cadence chunks;
* the gap between the 19 population direction arrays (pop_stagger, 0-3.5 KiB
  in 512-byte steps): the step's 19 read and 19 write streams reach
  different HBM rates for different offsets between them, set by the
  address interleave. On the 1500x522x30 reach the best gap makes the step
  0.5 % (FP32) to 2 % (BF16) faster than packed arrays, gaps of an odd
  number of 512-byte units do best and some multiples of 1 KiB are slower
  than packed, and the ranking is not the same on every GPU, so it is timed
  on the GPU in use: a scratch solver per candidate, about 2 s per format.

Results are cached per (storage, closure, nx, ny, nz) in memory and in a JSON
file (WFS_AUTOTUNE_CACHE, default ~/.cache/water_flow_simulation_amd/
autotune.json) so later runs of the same shape reuse them.
"""
from __future__ import annotations

import json
import os
import time

WIDTHS = (64, 128, 192, 256)
_cache: dict | None = None


def _cache_path():
    return os.environ.get("WFS_AUTOTUNE_CACHE",
                          os.path.join(os.path.expanduser("~"), ".cache", "water_flow_simulation_amd", "autotune.json"))


def _load():
    global _cache
    if _cache is None:
        try:
            with open(_cache_path()) as f:
                _cache = json.load(f)
        except (OSError, ValueError):
            _cache = {}
    return _cache


def _store(key, value):
    c = _load()
    c[key] = value
    try:
        os.makedirs(os.path.dirname(_cache_path()), exist_ok=True)
        with open(_cache_path(), "w") as f:
            json.dump(c, f, indent=1, sort_keys=True)
    except OSError:
        pass   # the in-memory cache still serves this process


def _key(reach, storage, closure):
    return f"{storage}/{'mlp' if closure is not None else 'smag'}/{reach.nx}x{reach.ny}x{reach.nz}"


def time_widths(reach, storage="f32", closure=None, steps=40, warmup=6, widths=WIDTHS, rounds=2, settle=40):
    """Average step-kernel milliseconds at each width, on a scratch solver of
    the case: ``settle`` steps first (a compute-bound step such as the
    learned one runs about 12 % slow for its first 20-odd steps after the GPU
    sat idle during the set-up), then the widths in turn ``rounds`` times,
    the best round counting."""
    from .reach3d import Run3D
    run = Run3D(reach, storage=storage, closure=closure)
    lbm = run.lbm
    out = {bx: float("inf") for bx in widths}
    try:
        lbm.step(settle)
        for _ in range(rounds):
            for bx in widths:
                lbm.block_x = bx
                lbm.step(warmup)
                lbm.timing(True)
                lbm.step(steps)
                ms, n = lbm.timing_read()
                lbm.timing(False)
                out[bx] = min(out[bx], ms / max(n, 1))
        lbm.check_finite()
    finally:
        run.close()
    return out


def block_x_for(reach, storage="f32", closure=None, retune=False):
    """The tuned work-group width for this case shape (tunes on first use)."""
    key = _key(reach, storage, closure)
    c = _load()
    if not retune and key in c:
        return int(c[key]["block_x"])
    t = time_widths(reach, storage, closure)
    best = min(t, key=t.get)
    _store(key, {"block_x": int(best), "kernel_ms": {str(k): round(v, 5) for k, v in t.items()},
                 "tuned": time.strftime("%Y-%m-%d %H:%M:%S")})
    return int(best)


STAGGERS = (0, 512, 1024, 1536, 2048, 2560, 3072, 3584)


def time_staggers(reach, storage="f32", steps=60, warmup=10, staggers=STAGGERS, rounds=2):
    """Average step-kernel milliseconds at each population stagger (bytes),
    each on a scratch solver of the case, the candidates taken in turn
    ``rounds`` times (the best round counts; the plain step it times is
    bound by HBM, which needs no settling)."""
    from .reach3d import Run3D
    out = {s: float("inf") for s in staggers}
    for _ in range(rounds):
        for s in staggers:
            run = Run3D(reach, storage=storage, pop_stagger=s)
            try:
                lbm = run.lbm
                lbm.step(warmup)
                lbm.timing(True)
                lbm.step(steps)
                ms, n = lbm.timing_read()
                lbm.timing(False)
                out[s] = min(out[s], ms / max(n, 1))
            finally:
                run.close()
    return out


def pop_stagger_for(reach, storage="f32", retune=False):
    """The tuned population stagger for this case shape and storage format
    (tunes on first use; the closure does not change the access pattern)."""
    key = "stagger/" + _key(reach, storage, None)
    c = _load()
    if not retune and key in c:
        return int(c[key]["pop_stagger"])
    t = time_staggers(reach, storage)
    best = min(t, key=t.get)
    _store(key, {"pop_stagger": int(best), "kernel_ms": {str(k): round(v, 5) for k, v in t.items()},
                 "tuned": time.strftime("%Y-%m-%d %H:%M:%S")})
    return int(best)


def tune_group(batch, steps=40, candidates=None):
    """Times the batch with several cases-per-launch settings (each for
    ``steps`` steps of the real run, so tuning is part of the spin-up) and
    leaves the fastest set. Returns {group: seconds}."""
    n = len(batch.runs)
    if candidates is None:
        candidates = sorted({n, max(1, (n + 1) // 2), max(1, (n + 3) // 4), max(1, (n + 7) // 8)}, reverse=True)
    res = {}
    for k in candidates:
        batch.set_group(k)
        batch.advance(20)                     # capture the graph outside the timed part
        t0 = time.perf_counter()
        batch.advance(steps)
        res[k] = time.perf_counter() - t0
    best = min(res, key=res.get)
    batch.set_group(best)
    batch.tuning = {int(k): round(v, 5) for k, v in res.items()}
    return res


def summary():
    """The cached tuning results."""
    return dict(_load())


if __name__ == "__main__":
    from . import workloads
    from .closure import WEIGHTS, init_params, load_params
    r = workloads.trapezoid_tug()
    w = load_params()[0] if os.path.exists(WEIGHTS) else init_params(0)
    for st, cl in (("f32", None), ("f16dev", None), ("bf16dev", None), ("f32", w), ("f16dev", w)):
        t = time_widths(r, st, cl)
        print(st, "learned" if cl is not None else "smagorinsky", "block_x ms",
              {k: round(v, 4) for k, v in t.items()}, "best", min(t, key=t.get), flush=True)
    for st in ("f32", "f16dev", "bf16dev"):
        t = time_staggers(r, st)
        print(st, "pop_stagger ms", {k: round(v, 4) for k, v in t.items()}, "best", min(t, key=t.get), flush=True)
