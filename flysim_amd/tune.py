"""Autotuner of the engine's launch schedule.

Times short runs of candidate schedules on the network at hand and keeps the
fastest.  The knobs: the step path (windowed: as many steps per launch as the
smallest synaptic delay allows; fused: one launch per step; two launches per
step: update, then delivery), the block size, and the NMDA gather lanes per
row (in the windowed path up to 16 of them take consecutive steps of one
term, the rest split the row).  The reference's only schedule knob is its
thread count (-t, code/NetSimGnl2.cpp:160-229).

Results are cached in a JSON file keyed by a signature of the network's
shape -- neurons, populations, copies, synapse counts by class, delay range,
precision, input mode -- and of the device, so that sweeps of many
same-shaped networks tune once.
"""
from __future__ import annotations

import hashlib
import itertools
import json
import math
import os
import statistics
import time
from typing import Dict, Iterable, List, Optional

from .engine import Engine, device_info

DEFAULT_CACHE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "_lib", "tune_cache.json")


def signature(eng: Engine, extra: str = "") -> str:
    st = eng.stats()
    dev, cu = device_info()
    key = json.dumps({"net": st, "device": dev, "cu": cu, "extra": extra}, sort_keys=True)
    return hashlib.sha1(key.encode()).hexdigest()[:16]


def candidates(stats: Dict[str, int]) -> List[Dict]:
    """Block sizes x gather lanes per row x the step paths the network
    admits: windowed (as many steps per launch as the smallest delay allows,
    one step per launch without delays), fused (needs a delay of at least one
    step on every synapse) and two launches per step.  Replicas in whole
    waves read one CSR row per wave, so one or two lanes per row suit them."""
    delayed = stats["min_delay"] >= 1
    has_nmda = stats["syn_nmda"] > 0
    reps = stats["replicas"]
    out = []
    lanes = [1, 2] if reps >= 64 else ([2, 4, 8, 16, 32, 64] if has_nmda else [8])
    if not delayed:  # one step per launch: a few shapes
        lanes = lanes[:2] if reps >= 64 else [4, 8]
    for b, L in itertools.product([128, 256, 512] if delayed else [256], lanes):
        if L * 2 <= b:
            out.append(dict(update_block=b, gather_lanes=L, fused=0, window=64))
    if delayed:
        for b, L in itertools.product([256, 512], [1, 2] if reps >= 64 else ([1, 4, 8] if has_nmda else [8])):
            out.append(dict(update_block=b, gather_lanes=L, fused=1, window=0))
    lanes = [1, 2] if reps >= 64 else ([1, 4, 8, 16] if has_nmda else [8])
    for b, L in itertools.product([256, 512, 1024], lanes):
        out.append(dict(update_block=b, gather_lanes=L, fused=0, window=0))
    return out


def refinements(sched: Dict, stats: Dict[str, int]) -> List[Dict]:
    """Variants of a windowed schedule on one network copy: blocks dealt to
    the XCDs in neighbouring runs (or not), smaller staged gather chunks
    (more blocks per compute unit) and several lanes per (row, step) pair
    of the gather."""
    if not sched.get("window") or stats["replicas"] != 1 or stats["syn_nmda"] == 0:
        return []
    x = 1 - int(sched.get("xcd_order", 0))
    out = [dict(sched, xcd_order=x)]
    for c in (128, 256):
        if sched.get("gather_chunk", 0) != c:
            out.append(dict(sched, gather_chunk=c))
            out.append(dict(sched, gather_chunk=c, xcd_order=x))
    # lanes per (row, step) pair of the gather: long rows want several
    for sp in (4, 16, 64):
        if sp <= sched.get("gather_lanes", 1) and sp != sched.get("gather_split", 0):
            out.append(dict(sched, gather_split=sp))
            out.append(dict(sched, gather_split=sp, xcd_order=x))
    return out


def time_schedule(eng: Engine, sched: Dict, steps: int, warmup: Optional[int] = None) -> float:
    """Seconds per step of `steps` steps on `sched`, after an untimed run of
    `warmup` steps (default: as many, so that the timed run replays the
    launch graph the warm-up captured)."""
    eng.set_schedule(**sched)
    eng.run(steps if warmup is None else warmup, record=False)
    t0 = time.perf_counter()
    eng.run(steps, record=False)
    return (time.perf_counter() - t0) / steps


def _load(cache: Optional[str]) -> Dict:
    if cache and os.path.exists(cache):
        try:
            return json.load(open(cache))
        except (OSError, ValueError):
            pass
    return {}


def tune(eng: Engine, steps: int = 200, trial_s: float = 0.03, cands: Optional[Iterable[Dict]] = None,
         cache: Optional[str] = DEFAULT_CACHE, key_extra: str = "", passes: int = 3, keep: int = 5,
         verbose: bool = False) -> Dict:
    """Pick the fastest schedule for `eng` and apply it.

    The trials run one after another on the live state, whose activity can
    drift (a network that ignites runs several times slower per step
    afterwards), so each candidate is timed between two runs of the engine's
    starting schedule and ranked by the ratio to them (a trial during which
    the starting schedule's time changed by more than 25 % is repeated once
    at the end): every candidate once, the `keep` best again in
    `passes - 1` more passes, by the median of their ratios; with the
    built-in candidates (cands None), the two best of the first pass are
    also tried with the windowed path's extra knobs (refinements).  A trial runs up to `steps` steps, fewer on
    networks whose steps are slow (about `trial_s` seconds per timed run).
    The trial runs advance the engine's state: reset it afterwards.
    steps_per_graph stays the engine's own unless a candidate sets it.
    Returns {"schedule", "s_per_step", "vs_start", "signature", "tune_s",
    "trials" (absent when the cache answered)}."""
    t_start = time.perf_counter()
    sig = signature(eng, key_extra)
    db = _load(cache)
    if sig in db:
        eng.set_schedule(**db[sig]["schedule"])
        return dict(db[sig], signature=sig, tune_s=time.perf_counter() - t_start)
    refine = cands is None  # a caller's own candidate list is tried as given
    if cands is None:
        cands = candidates(eng.stats())
    base = dict(eng.schedule)
    probe = time_schedule(eng, base, 10)
    n = int(min(steps, max(20, trial_s / max(probe, 1e-9))))
    t_base = [time_schedule(eng, base, n)]   # the starting schedule before each trial

    def trial(sched):
        """Candidate time over the geometric mean of the starting schedule's
        times just before and just after it; None when those two differ by
        more than 25 % (the activity changed during the trial)."""
        try:
            t1 = time_schedule(eng, sched, n)
            applied = dict(eng.schedule)
        finally:
            t_base.append(time_schedule(eng, base, n))
        a, b = t_base[-2], t_base[-1]
        if max(a, b) > 1.25 * min(a, b):
            return applied, None, t1
        return applied, t1 / math.sqrt(a * b), t1

    runs = []   # [applied schedule, [ratio, ...], [s_per_step, ...]]
    unsettled = []
    for sc in cands:
        try:
            applied, r, dt = trial(dict(base, **sc))
        except Exception as e:  # an invalid candidate is skipped, not fatal
            if verbose:
                print("schedule", sc, "failed:", e)
            continue
        if r is None:
            unsettled.append(applied)
            continue
        runs.append([applied, [r], [dt]])
        if verbose:
            print("%-100s %.2f us/step (%.3f x start)" % (applied, dt * 1e6, r), flush=True)
    for sc in unsettled:   # once more, after the activity changed
        applied, r, dt = trial(sc)
        if r is not None:
            runs.append([applied, [r], [dt]])
    if not runs:
        raise RuntimeError("no schedule candidate ran")
    # the two best first-pass schedules, varied in the knobs that only the
    # windowed path on one network copy has
    seen = {json.dumps(q[0], sort_keys=True) for q in runs}
    for q in (sorted(runs, key=lambda q: q[1][0])[:2] if refine else []):
        for sc in refinements(q[0], eng.stats()):
            if json.dumps(sc, sort_keys=True) in seen:
                continue
            seen.add(json.dumps(sc, sort_keys=True))
            try:
                applied, r, dt = trial(sc)
            except Exception as e:
                if verbose:
                    print("schedule", sc, "failed:", e)
                continue
            if r is not None:
                runs.append([applied, [r], [dt]])
                if verbose:
                    print("%-100s %.2f us/step (%.3f x start)" % (applied, dt * 1e6, r), flush=True)
    top = sorted(runs, key=lambda q: q[1][0])[:max(1, keep)]
    for p in range(1, passes if len(runs) > 1 else 1):
        for q in (top[::-1] if p % 2 else top):
            _, r, dt = trial(q[0])
            if r is not None:
                q[1].append(r)
                q[2].append(dt)
    rank = lambda q: float(statistics.median(q[1]))
    best = min(top, key=rank)
    eng.set_schedule(**best[0])
    trials = [{"schedule": q[0], "vs_start": rank(q), "s_per_step": float(statistics.median(q[2])),
               "runs": len(q[1])} for q in runs]
    out = {"schedule": best[0], "vs_start": rank(best), "s_per_step": float(statistics.median(best[2])),
           "trial_steps": n}
    if cache:
        db[sig] = out
        os.makedirs(os.path.dirname(cache), exist_ok=True)
        with open(cache, "w") as f:
            json.dump(db, f, indent=1)
    return dict(out, signature=sig, trials=trials, tune_s=time.perf_counter() - t_start)
