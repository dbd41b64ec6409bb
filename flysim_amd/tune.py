"""Autotuner of the engine's launch schedule.

Times short runs of candidate schedules (neuron-update block size, NMDA
gather lanes per row, delivery block size, steps per captured graph, and the
step path: windowed, fused one-launch or two-launch) on the network at hand
and keeps the fastest.  Results are cached in a JSON file keyed by a
signature of the network's shape (neuron and synapse counts by class,
delays, precision, device) so that sweeps of many same-shaped networks tune
once.
"""
from __future__ import annotations

import hashlib
import itertools
import json
import os
import time
from typing import Dict, Iterable, List, Optional

import numpy as np

from .engine import Engine

DEFAULT_CACHE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "_lib", "tune_cache.json")


def signature(eng: Engine, extra: str = "") -> str:
    key = "%d|%d|%d|%d|%s" % (eng.n, eng.n_pop, getattr(eng, "replicas", 1), eng.precision, extra)
    return hashlib.sha1(key.encode()).hexdigest()[:16]


def candidates(replicas: int = 1, has_nmda: bool = True, fused_ok: bool = True) -> List[Dict]:
    """Block sizes x NMDA lanes per row x step path (windowed: as many steps
    per launch as the smallest delay allows; fused one launch per step; two
    launches per step).  Replicas in whole waves read one CSR row per wave,
    so one lane per row is the natural choice there; single networks spread
    a row over 1-16 lanes."""
    blocks = [128, 256, 512, 1024]  # the windowed path takes at most 512
    if not has_nmda:
        lanes = [8]
    elif replicas >= 64:
        lanes = [1, 2]
    else:
        lanes = [1, 2, 4, 8, 16]
    out = []
    paths = [(0, 64)] + ([(1, 0)] if fused_ok else []) + [(0, 0)]
    for b, L in itertools.product(blocks, lanes):
        for f, w in paths:
            if w and b > 512:
                continue
            out.append(dict(update_block=b, gather_lanes=L, push_block=256, steps_per_graph=100,
                            use_graph=1, fused=f, window=w))
    return out


def time_schedule(eng: Engine, sched: Dict, steps: int, warmup: int) -> float:
    eng.set_schedule(**sched)
    eng.run(warmup, record=False)
    t0 = time.perf_counter()
    eng.run(steps, record=False)
    return (time.perf_counter() - t0) / steps


def tune(eng: Engine, steps: int = 400, warmup: int = 100, cands: Optional[Iterable[Dict]] = None,
         cache: Optional[str] = DEFAULT_CACHE, key_extra: str = "", verbose: bool = False) -> Dict:
    """Pick the fastest schedule for `eng` (state is advanced by the trial
    runs; reset the engine afterwards).  Returns {"schedule", "s_per_step",
    "trials"} and applies the schedule."""
    sig = signature(eng, key_extra)
    db = {}
    if cache and os.path.exists(cache):
        try:
            db = json.load(open(cache))
        except (OSError, ValueError):
            db = {}
    if sig in db:
        eng.set_schedule(**db[sig]["schedule"])
        return db[sig]
    if cands is None:
        cands = candidates(getattr(eng, "replicas", 1))
    trials = []
    best = None
    for sc in cands:
        try:
            dt = time_schedule(eng, sc, steps, warmup)
        except Exception as e:  # an invalid candidate is skipped, not fatal
            if verbose:
                print("schedule", sc, "failed:", e)
            continue
        trials.append({"schedule": sc, "s_per_step": dt})
        if verbose:
            print("%-90s %.2f us/step" % (sc, dt * 1e6))
        if best is None or dt < best["s_per_step"]:
            best = {"schedule": sc, "s_per_step": dt}
    if best is None:
        raise RuntimeError("no schedule candidate ran")
    eng.set_schedule(**best["schedule"])
    res = {"schedule": best["schedule"], "s_per_step": best["s_per_step"], "trials": trials}
    if cache:
        db[sig] = {"schedule": best["schedule"], "s_per_step": best["s_per_step"]}
        os.makedirs(os.path.dirname(cache), exist_ok=True)
        with open(cache, "w") as f:
            json.dump(db, f, indent=1)
    return res
