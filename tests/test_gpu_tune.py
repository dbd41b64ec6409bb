"""GPU test of the launch-schedule autotuner (flysim_amd/tune.py): it times
the candidates on the network at hand, applies the fastest, caches it by
network signature, and a tuned schedule changes no result (FP64 raster
identical to the default schedule's)."""
import json
import os

import numpy as np
import pytest

from flysim_amd import _capi as A
from flysim_amd import tune
from flysim_amd import workloads as W
from flysim_amd.engine import Engine, arrays_from_network

pytestmark = pytest.mark.gpu


def test_tune_picks_applies_and_caches(tmp_path):
    net = W.c2_network(seed=4, n=3000, n_mod=10, mod_edges=20, t_end_ms=100)
    arr = arrays_from_network(net)
    eng = Engine(arr, precision=A.FS_FP64)
    eng.reset(5)
    cands = [dict(update_block=b, gather_lanes=L, push_block=256, steps_per_graph=50, use_graph=1,
                  fused=f) for b, L, f in ((256, 8, 1), (128, 4, 1), (512, 2, 0))]
    cache = str(tmp_path / "tune.json")
    res = tune.tune(eng, steps=100, cands=cands, cache=cache)
    assert len(res["trials"]) == 3 and all(t["runs"] == 3 for t in res["trials"])
    best = min(res["trials"], key=lambda t: t["vs_start"])["schedule"]
    assert res["schedule"] == best
    got = eng.schedule
    for k in ("update_block", "gather_lanes", "steps_per_graph"):
        assert got[k] == best[k]
    db = json.load(open(cache))
    assert len(db) == 1
    # a second engine of the same shape takes the cached schedule without timing
    e2 = Engine(arr, precision=A.FS_FP64)
    r2 = tune.tune(e2, steps=100, cands=cands, cache=cache)
    assert "trials" not in r2 and r2["schedule"] == best
    # results do not depend on the schedule
    e2.reset(9)
    ra = e2.raster(600)
    e3 = Engine(arr, precision=A.FS_FP64)
    e3.reset(9)
    rb = e3.raster(600)
    np.testing.assert_array_equal(ra, rb)
