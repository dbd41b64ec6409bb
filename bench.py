"""Benchmark of the Flysim core on one MI355X.

Workload (default): the synthetic 20,000-neuron fly-like connectome (50
lognormal modules, 70/30 E/I, intra p=0.05 / inter p=0.002 on a mean-degree-6
module graph, lognormal efficacies, 1-2 ms delays, Poisson drive 1-3 kHz per
module), FP32 state, dt = 0.1 ms.  One "step" is one 0.1 ms simulation step
of the whole network (LIF update + spike delivery + Poisson input).

Prints one JSON line: synaptic events/s (whole job), sim-s/wall-s, the
roofline of the dominant kernel (HIP events on the engine's stream over the
timed region) and the CPU baseline (the CPU restatement timed on this host's
cores over a bounded sample of the same network).
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "synaptic events/s + sim-s/wall-s, 20k-neuron fly net; sweep pts/s vs plain GPU"
HBM_PEAK_GBS = 8000.0
UPDATE_BYTES_FP32 = 84      # algorithmic bytes per neuron-step of the LIF update
UPDATE_BYTES_FP64 = 160


def parse():
    p = argparse.ArgumentParser()
    p.add_argument("--gpus", type=int, default=1)
    p.add_argument("--steps", type=int, default=20000)
    p.add_argument("--warmup", type=int, default=2000)
    p.add_argument("--workload", default="fly20k", choices=["fly20k", "ei1k", "sparse1m"])
    p.add_argument("--fp64", action="store_true")
    p.add_argument("--cpu-steps", type=int, default=-1,
                   help="steps of the CPU-baseline sample (-1: sized to ~15 s)")
    p.add_argument("--no-cpu", action="store_true")
    p.add_argument("--no-sweep", action="store_true", help="skip the sweep points/s figures")
    p.add_argument("--injection", action="store_true",
                   help="the seizure protocol's form: Gaussian current injection on every "
                        "population from step 0 (GaussMean ~ U(0.04, 0.16) nA, GaussSTD = 6 x mean)")
    p.add_argument("--schedule", default="", help="json dict of fs_schedule overrides")
    p.add_argument("--copy-calib", type=int, default=0, metavar="MIB",
                   help="only run 10 device-to-device copies of MIB MiB (the known-size "
                        "kernel that scales the FETCH_SIZE / WRITE_SIZE counters) and exit")
    return p.parse_args()


def copy_calib(mib):
    import torch
    a = torch.ones(mib << 18, dtype=torch.float32, device="cuda:0")
    b = torch.empty_like(a)
    for _ in range(10):
        b.copy_(a)
    torch.cuda.synchronize()
    print(json.dumps({"copy_calib_bytes_read": mib << 20, "copy_calib_bytes_written": mib << 20,
                      "launches": 10}))


def make_network(name):
    from flysim_amd import workloads as W
    if name == "fly20k":
        return W.c2_network(seed=2), ("fly20k: 20,000 LIF neurons in 50 lognormal modules "
                                      "(70/30 E/I), intra p=0.05 / inter p=0.002, "
                                      "delays 1-2 ms, Poisson drive 1-3 kHz, dt 0.1 ms")
    if name == "ei1k":
        return W.c1_network(seed=1), "ei1k: 800 E + 200 I, p=0.1, dt 0.1 ms"
    return W.c3_network(seed=3), ("sparse1m: 1,000,000 LIF neurons (80/20), in-degree 100, "
                                  "delays 1-5 ms, Poisson 2 kHz, dt 0.1 ms")


def cpu_baseline(net, arr, events_per_sim_s, n_steps):
    """The reference built from source (oracle/_ref/flysim_ref) timed on this
    host's cores on the module form of the same network; the CPU restatement
    when the reference binary is absent."""
    import re
    import subprocess
    import tempfile
    ref = os.path.join(ROOT, "oracle", "_ref", "flysim_ref")
    if not os.path.exists(ref):
        return cpu_baseline_port(net, arr, events_per_sim_s, n_steps)
    from flysim_amd import workloads as W
    cores = min(16, os.cpu_count() or 1)
    sim_ms = 200.0 if n_steps < 0 else n_steps * net.dt
    with tempfile.TemporaryDirectory() as d:
        info = W.write_module_conf(net, os.path.join(d, "network.conf"))
        W.write_pro(os.path.join(d, "network.pro"), sim_ms, seeds=[2], spike_file=None)
        best = None
        for threads in sorted({1, cores}):
            out = subprocess.run([ref, "-conf", "network.conf", "-pro", "network.pro", "-t",
                                  str(threads)], cwd=d, capture_output=True, text=True,
                                 timeout=900).stdout.replace("\r", "\n")
            m = re.findall(r"real time=([0-9.eE+-]+)s 100%", out)
            if not m:
                raise RuntimeError("reference run produced no timing line")
            wall = float(m[-1])
            if best is None or wall < best[1]:
                best = (threads, wall)
    threads, wall = best
    steps = int(np.floor(np.float32(sim_ms) * np.float32(1e-3) * np.float32(1e3) / np.float32(net.dt)))
    sim_per_wall = steps * net.dt * 1e-3 / wall
    return {"value": sim_per_wall * events_per_sim_s, "unit": "synaptic events/s",
            "cores": threads, "kind": "reference", "sim_s_per_wall_s": sim_per_wall,
            "sample": "reference flysim built from source, -t %d (best of 1 and %d threads), "
                      "%d steps (%.0f ms simulated) of the module form of the workload "
                      "(%d population links, %d synapses, SpikeDly %.2f ms); step-loop time "
                      "from its own 'real time=' line; events/s = GPU-measured events per "
                      "simulated second x CPU sim-s/wall-s"
                      % (threads, cores, steps, sim_ms, info["links"], info["approx_synapses"],
                         info["spike_delay_ms"])}


def cpu_baseline_port(net, arr, events_per_sim_s, n_steps):
    """Time the CPU restatement (the reference's O(S)-per-step pull algorithm
    with its std::thread strided partition) on this host's cores."""
    from oracle.oracle import Oracle
    from flysim_amd.engine import pop_conf_of
    order = np.lexsort((np.arange(net.n_syn), net.receptor, net.post))
    post, rec = net.post[order], net.receptor[order]
    key = post.astype(np.int64) * 8 + rec
    starts = np.concatenate([[0], np.nonzero(np.diff(key))[0] + 1, [len(key)]]).astype(np.uint32)
    layout = {
        "gate_pre": net.pre[order], "gate_rec": rec, "gate_g": net.w[order],
        "gate_post0": post, "row_post": post[starts[:-1]], "row_rec": rec[starts[:-1]],
        "row_beg": starts, "term": np.arange(len(order), dtype=np.uint32),
        "neuron_pop": arr.neuron_pop.astype(np.int32),
        "pop_conf": [pop_conf_of(p) for p in net.pops],
        "rfped": np.array([q.rfped for q in arr.pops], np.int32),
        "spidy": np.array([q.spidy for q in arr.pops], np.int32),
        "evt_time_s": np.zeros(0, np.float32), "evt_type": np.zeros(0, np.int32),
        "evt_pop": np.zeros(0, np.int32), "evt_sub": np.zeros(0, np.int32),
        "evt_var1": np.zeros(0, np.float32), "evt_var2": np.zeros(0, np.float32),
        "act_ctr": 1 | 2 | 4 | 8 | 32 | 64 | 256,
    }
    o = Oracle(layout, net.n, precision=0)
    o.reset(1)
    cores = min(16, os.cpu_count() or 1)
    if n_steps < 0:
        probe = 20
        t = o.time_run(probe, cores)
        n_steps = int(max(50, min(20000, 15.0 / max(t / probe, 1e-9))))
    wall = o.time_run(n_steps, cores)
    sim_per_wall = n_steps * net.dt * 1e-3 / wall
    return {"value": sim_per_wall * events_per_sim_s, "unit": "synaptic events/s",
            "cores": cores, "kind": "port",
            "sim_s_per_wall_s": sim_per_wall,
            "sample": "%d steps (%.3f s simulated) of the same network on the CPU restatement "
                      "(O(S) pull per step, %d std::threads, flush-to-zero on); per-synapse "
                      "delays not modelled; events/s = GPU-measured events per simulated "
                      "second x CPU sim-s/wall-s" % (n_steps, n_steps * net.dt * 1e-3, cores)}


def main():
    args = parse()
    if args.copy_calib:
        copy_calib(args.copy_calib)
        return
    if args.gpus != 1:
        print("bench.py: this build runs on one GPU; --gpus %d is not supported" % args.gpus,
              file=sys.stderr)
        sys.exit(2)
    from flysim_amd import _capi as A
    from flysim_amd.engine import Engine, arrays_from_network
    net, desc = make_network(args.workload)
    arr = arrays_from_network(net)
    prec = A.FS_FP64 if args.fp64 else A.FS_FP32
    eng = Engine(arr, precision=prec, ext_mode=A.FS_EXT_PHILOX, record=False)
    if args.schedule:
        eng.set_schedule(**json.loads(args.schedule))
    eng.reset(2024)
    if args.injection:
        means = np.random.default_rng(4).uniform(0.04, 0.16, size=len(arr.pops))
        for p, m in enumerate(means):
            eng.set_injection(p, float(m), float(6.0 * m))
        desc += "; Gaussian injection on every population (seizure protocol form)"
    if args.warmup:
        eng.run(args.warmup, record=False)
    eng.enable_timing(True)
    t0 = time.perf_counter()
    eng.run(args.steps, record=False)
    wall = time.perf_counter() - t0
    ctr = eng.counters()
    kt = eng.kernel_times()
    sim_s = args.steps * net.dt * 1e-3
    # synaptic events: every receptor-synapse a spike reaches (linear pushes
    # plus the NMDA synapses of the spiking neurons)
    ev_value = ctr["syn_events"] / wall
    # roofline of the dominant kernel (algorithmic bytes per launch): the LIF
    # update moves 84 B per neuron-step (160 B in FP64); delivery moves 8 B per
    # spike (row pointer) + 17 B per pushed synapse (col, weight, delay, atomic
    # read-modify-write) and the NMDA gather 12 B per term + 8 B per row
    n_launch = args.steps
    upd_bytes = (UPDATE_BYTES_FP64 if args.fp64 else UPDATE_BYTES_FP32) * net.n
    del_bytes = (8 * ctr["spikes"] + 17 * ctr["lin_pushes"] + 12 * ctr["nmda_terms"]
                 + 8 * net.n * n_launch) / n_launch
    upd_ms, del_ms = kt["update_ms"], kt["deliver_ms"]
    win = eng.schedule.get("window", 0)
    if win:
        # W steps per launch; graphs hold whole windows (steps_per_graph a
        # multiple of W), so a launch lasts W graph steps
        dom, b = "k_window", (upd_bytes + del_bytes) * win
        ms = kt["graph_step_ms"] * win if kt["graph_step_ms"] > 0 else upd_ms
    elif eng.fused:
        # one kernel per step: its launch time is the graph time per step
        # (the eager first step of a batch also carries its dispatch latency)
        dom, b = "k_step", upd_bytes + del_bytes
        ms = kt["graph_step_ms"] if kt["graph_step_ms"] > 0 else upd_ms
    elif del_ms >= upd_ms:
        dom, b, ms = "k_deliver", del_bytes, del_ms
    else:
        dom, b, ms = "k_update", upd_bytes, upd_ms
    achieved = b / (ms * 1e-3) / 1e9 if ms > 0 else 0.0
    res = {
        "metric": METRIC, "value": ev_value, "unit": "synaptic events/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": wall / args.steps * 1e3,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f64" if args.fp64 else "f32", "data": "synthetic",
        "config": {"workload": desc, "neurons": net.n, "receptor_synapses": net.n_syn,
                   "pairs": int(net.notes.get("pairs", net.n_syn)),
                   "sim_seconds": sim_s, "schedule": eng.schedule},
        "sim_s_per_wall_s": sim_s / wall,
        "spikes_per_s": ctr["spikes"] / wall,
        "mean_rate_hz": ctr["spikes"] / net.n / sim_s,
        "roofline": {"bound": "hbm", "kernel": dom, "achieved": achieved, "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS, "traffic": None,
                     "bytes_per_launch": b, "avg_launch_ms": ms,
                     "update_ms": upd_ms, "deliver_ms": del_ms,
                     "graph_step_ms": kt["graph_step_ms"],
                     "timed_launches": kt["n_timed"]},
    }
    traffic = None if args.injection else measured_traffic(args.workload, dom, res["dtype"])
    if traffic:
        res["roofline"]["traffic"] = traffic
    eng.close()
    if not args.no_sweep and args.workload == "fly20k":
        res["sweep"] = sweep_figures()
    if not args.no_cpu and not args.injection:  # the CPU legs run the workload without injection
        res["cpu_baseline"] = cpu_baseline(net, arr, ctr["syn_events"] / sim_s, args.cpu_steps)
        res["speedup_vs_cpu"] = res["value"] / res["cpu_baseline"]["value"]
    print(json.dumps(res))


def measured_traffic(workload, kernel, dtype):
    """HBM bytes per launch of the dominant kernel from the committed
    rocprofv3 FETCH_SIZE / WRITE_SIZE runs (profiles/traffic.json)."""
    path = os.path.join(ROOT, "profiles", "traffic.json")
    if not os.path.exists(path):
        return None
    rec = json.load(open(path)).get("%s/%s/%s" % (workload, kernel, dtype))
    return rec["bytes"] if rec else None


def sweep_figures(points_plain=16, replicas=128, mlp_points=1 << 20):
    """Sweep points/s on the fly20k topology (100 ms warm-up + 200 ms window
    per point): plain per-instance GPU simulation vs replicated launches of
    `replicas` points, and the surrogate's forward over 2^20 points (BF16
    matrix-core GEMMs; GPU time from HIP events) with its MFMA roofline."""
    from flysim_amd.surrogate import Mlp, SweepSpace
    from flysim_amd.sweep import make_sim
    space = SweepSpace()
    rng = np.random.default_rng(7)
    plain = make_sim(1)
    plain.run(space.sample(1, rng), seed=1)            # graph capture outside the timing
    plain.points, plain.wall_seconds = 0, 0.0
    plain.run(space.sample(points_plain, rng), seed=2)
    batched = make_sim(replicas)
    batched.run(space.sample(replicas, rng), seed=3)
    batched.points, batched.wall_seconds = 0, 0.0
    batched.run(space.sample(2 * replicas, rng), seed=4)
    sizes = [space.dim, 256, 256, 256, space.n_mod]
    mlp = Mlp(sizes, seed=0)
    x = space.sample(mlp_points, rng)
    mlp.predict(x[:4096])
    mlp.predict(x)
    flop = 2.0 * sum(sizes[i] * sizes[i + 1] for i in range(len(sizes) - 1)) * mlp_points
    tflops = flop / (mlp.last_ms * 1e-3) / 1e12
    return {"topology": "fly20k", "window_ms": 200, "warmup_ms": 100,
            "plain_points_per_s": plain.points_per_s,
            "batched_points_per_s": batched.points_per_s, "replicas": replicas,
            "batched_vs_plain": batched.points_per_s / plain.points_per_s,
            "surrogate_points_per_s": mlp_points / (mlp.last_ms * 1e-3),
            "surrogate_roofline": {"bound": "mfma", "kernel": "k_gemm (forward, 4 layers)",
                                   "achieved": tflops, "peak": 2500.0, "unit": "TFLOP/s",
                                   "frac": tflops / 2500.0, "points": mlp_points,
                                   "gpu_ms": mlp.last_ms}}


if __name__ == "__main__":
    main()
