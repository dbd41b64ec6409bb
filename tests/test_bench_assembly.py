"""The bench line's assembly on the CPU: a run in which nothing fires still
prints a complete, valid JSON line (no division by a zero CPU value), the
requested steps are floors under the workload's simulated time, and runs on
more than one GPU are refused with a message."""
import json
import os
import subprocess
import sys
import types

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
import bench  # noqa: E402


def _args(steps=20, warmup=5, fp64=False):
    return types.SimpleNamespace(steps=steps, warmup=warmup, fp64=fp64)


def _line(ctr, kt, sched, steps=20000, wall=0.1):
    return bench.assemble(_args(), steps, 2000, wall, ctr, kt, sched, False, 20000, 1050577,
                          672000, 0.1, "fly20k")


ZERO = {"spikes": 0, "lin_pushes": 0, "nmda_out": 0, "syn_events": 0, "nmda_terms": 0}
KT = {"update_ms": 0.05, "deliver_ms": 0.0, "graph_step_ms": 0.0005, "n_timed": 20}


def test_zero_events_gives_a_valid_line():
    res = _line(ZERO, KT, {"window": 11})
    bench.add_cpu_leg(res, {"value": 0.0, "unit": "synaptic events/s", "cores": 1,
                            "kind": "reference", "sim_s_per_wall_s": 0.016, "sample": "x"})
    s = json.dumps(res)
    back = json.loads(s)
    assert back["value"] == 0.0
    assert back["speedup_vs_cpu"] is None
    assert back["sim_speedup_vs_cpu"] > 0
    assert back["roofline"]["kernel"] == "k_window"
    assert back["roofline"]["frac"] > 0  # the update bytes are there even without spikes
    for key in ("metric", "value", "unit", "n_gpus", "steps", "warmup", "ms_per_step",
                "higher_is_better", "scaling", "vs_baseline", "dtype", "data", "config",
                "roofline", "cpu_baseline"):
        assert key in back


def test_roofline_bytes_per_window_launch():
    ctr = {"spikes": 1000, "lin_pushes": 31000, "nmda_out": 21600, "syn_events": 52600,
           "nmda_terms": 432591 * 100}
    rf = bench.roofline_of({"update_ms": 0.0, "deliver_ms": 0.0, "graph_step_ms": 0.001,
                            "n_timed": 1}, ctr, {"window": 11}, False, 20000, 100, False)
    per_step = 84 * 20000 + (8 * 1000 + 17 * 31000 + 12 * 432591 * 100 + 8 * 20000 * 100) / 100
    assert rf["bytes_per_launch"] == pytest.approx(per_step * 11)
    assert rf["avg_launch_ms"] == pytest.approx(0.011)


def test_steps_are_floors():
    assert bench.FLOORS["fly20k"] == (20000, 2000)  # 2 s and 0.2 s at dt 0.1 ms
    res = _line(ZERO, KT, {"window": 0}, steps=max(20, bench.FLOORS["fly20k"][0]))
    assert res["steps"] == 20000 and res["steps_requested"] == 20
    assert res["config"]["sim_seconds"] == pytest.approx(2.0)


def test_more_than_one_gpu_is_refused():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 2
    assert "one GPU" in r.stderr
