"""Launch-shape autotuner for batched plans.

A plan's launch shape (``bldfm_tuning`` in include/bldfm_amd.h) changes how
the work is cut into workgroups, never a result:

* ``shoot_block``: modes per workgroup of the vertical solve (64, 128, 256).
  Small blocks waste fewer lanes on the partial last tile of each instance
  (batches of small grids); large blocks amortise the per-tile setup.
* ``xpass_layout``: how the x-pass reads the y-pass output -- 0 gathers each
  row at stride ny, 1 has the y-pass store its columns transposed, 2 runs a
  coalesced LDS-tiled transpose kernel between the passes.
* ``fft_threads``: threads per workgroup of the FFT passes (0: one thread per
  8 or 10 points of the longest transform, rounded up to a wave).  More
  threads leave some idle in the radix-8 stages but split the short-radix
  stages over more waves.

``tune`` times each candidate on the device with the batch's own instances
(coordinate search: block size, then layout, then FFT workgroup size, so
3 + 2 + up to 2 plans), keeps
the fastest (a candidate has to win by more than 1%, above run-to-run noise),
and caches the choice per (device, batch shape) in memory and,
when a path is given, in a JSON file, so a serving loop pays the search once
per shape.  The reference has no counterpart: its numba and FFTW back ends
size their own work (src/bldfm/solver.py:211-218, src/bldfm/fft_manager.py).
"""

import hashlib
import json
import os
import threading

import torch

from . import _lib

SHOOT_BLOCKS = (256, 128, 64)
XPASS_LAYOUTS = (0, 1, 2)
DEFAULT = {"shoot_block": 256, "xpass_layout": 0, "fft_threads": 0}

_FIELDS = ("nx", "ny", "nlx", "nly", "px", "py", "nz", "nslot", "footprint", "analytic", "single",
           "out_f32", "fast", "surrogate", "prof_from_met")
_memo = {}
_lock = threading.Lock()


def shape_key(instances, dev=None):
    """Cache key: device name and the geometry of every instance (the run
    time of a plan depends on these, not on the met values)."""
    dev = dev if dev is not None else _lib.device()
    name = torch.cuda.get_device_name(dev) if torch.cuda.is_available() else "cpu"
    rows = sorted(tuple(int(getattr(inst, f)) for f in _FIELDS) for inst in instances)
    h = hashlib.sha256(json.dumps([name, rows]).encode()).hexdigest()[:24]
    return f"{name}:{len(rows)}:{h}"


def fft_thread_candidates(instances):
    """FFT workgroup sizes worth timing for these instances: the default (0)
    and up to two larger multiples of 64 within the kernel family's bound
    (256 threads for transforms up to 2048 points, else 1024)."""
    n = max(max(int(i.nx) + 2 * int(i.px), int(i.ny) + 2 * int(i.py)) for i in instances)
    small = n <= 2048
    per = 8 if small else 10
    base = max(64, -(-(-(-n // per)) // 64) * 64)  # ceil(n/per) rounded up to a multiple of 64
    cap = 256 if small else 1024
    out = [0]
    for t in (base + 64, 2 * base):
        t = -(-t // 64) * 64
        if base < t <= cap and t not in out:
            out.append(t)
    return out


def _time_plan(instances, levels, dev, tuning, reps, warmup, prof, src):
    plan = _lib.Plan(instances, levels, dev, tuning=tuning)
    out_dtype = torch.float32 if all(int(i.out_f32) for i in instances) else torch.float64
    if prof is None:
        prof = torch.empty(max(plan.prof_doubles, 1), dtype=torch.float64, device=dev)
    if src is None and plan.src_doubles > 0:
        src = torch.zeros(plan.src_doubles, dtype=torch.float64, device=dev)
    conc = torch.empty(max(plan.out_elems, 1), dtype=out_dtype, device=dev)
    flx = torch.empty_like(conc)
    ws = torch.empty(max(plan.workspace_bytes, 1), dtype=torch.uint8, device=dev)
    for _ in range(warmup):
        plan.execute(prof, src, conc, flx, ws)
    ev = [torch.cuda.Event(enable_timing=True) for _ in range(2 * reps)]
    for r in range(reps):
        ev[2 * r].record()
        plan.execute(prof, src, conc, flx, ws)
        ev[2 * r + 1].record()
    torch.cuda.synchronize(dev)
    ms = sorted(ev[2 * r].elapsed_time(ev[2 * r + 1]) for r in range(reps))
    return ms[len(ms) // 2]


def _load(path):
    try:
        with open(path) as f:
            return json.load(f)
    except (OSError, ValueError):
        return {}


MARGIN = 0.99  # a candidate must beat the current choice by 1% (run-to-run noise is below that)


def tune(instances, levels, dev=None, reps=7, warmup=2, prof=None, src=None, cache_path=None, force=False):
    """Fastest launch shape for this batch; returns a dict with
    ``shoot_block``, ``xpass_layout``, the median milliseconds per execution
    of every candidate timed (``ms``) and of the choice (``best_ms``).  The
    device clocks are brought up by untimed executions first; a candidate
    replaces the current choice only when it is faster by more than 1%.
    ``prof`` / ``src``: the caller's profile and source buffers (needed only
    for instances that bring their own profiles; prof_from_met instances
    compute theirs).  ``cache_path`` (or $BLDFM_AMD_TUNE_CACHE): JSON file
    of earlier choices."""
    dev = dev if dev is not None else _lib.device()
    key = shape_key(instances, dev)
    cache_path = cache_path or os.environ.get("BLDFM_AMD_TUNE_CACHE")
    with _lock:
        if not force:
            if key in _memo:
                return dict(DEFAULT, **_memo[key], cached=True)
            if cache_path:
                hit = _load(cache_path).get(key)
                if hit:
                    _memo[key] = hit
                    return dict(DEFAULT, **hit, cached=True)  # entries may predate a coordinate
        timings = {}

        def run(cand):
            name = (f"shoot_block={cand['shoot_block']},xpass_layout={cand['xpass_layout']},"
                f"fft_threads={cand['fft_threads']}")
            if name not in timings:
                timings[name] = _time_plan(instances, levels, dev, cand, reps, warmup, prof, src)
            return timings[name]

        _time_plan(instances, levels, dev, DEFAULT, 1, 4 * warmup + reps, prof, src)  # clocks up
        best = dict(DEFAULT)
        best_ms = run(best)
        for b in SHOOT_BLOCKS:
            cand = dict(best, shoot_block=b)
            ms = run(cand)
            if ms < MARGIN * best_ms:
                best, best_ms = cand, ms
        for lay in XPASS_LAYOUTS:
            cand = dict(best, xpass_layout=lay)
            ms = run(cand)
            if ms < MARGIN * best_ms:
                best, best_ms = cand, ms
        for ft in fft_thread_candidates(instances):
            cand = dict(best, fft_threads=ft)
            ms = run(cand)
            if ms < MARGIN * best_ms:
                best, best_ms = cand, ms
        result = dict(best, ms=timings, best_ms=best_ms)
        _memo[key] = result
        if cache_path:
            data = _load(cache_path)
            data[key] = result
            tmp = cache_path + ".tmp"
            with open(tmp, "w") as f:
                json.dump(data, f, indent=1, sort_keys=True)
            os.replace(tmp, cache_path)
        return dict(result, cached=False)


def clear_memo():
    with _lock:
        _memo.clear()
