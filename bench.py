"""Benchmark of the Flysim core on one MI355X.

Workload (default): the synthetic 20,000-neuron fly-like connectome (50
lognormal modules, 70/30 E/I, intra p=0.05 / inter p=0.002 on a mean-degree-6
module graph, lognormal efficacies, 1-2 ms delays, Poisson drive 1-3 kHz per
module), FP32 state, dt = 0.1 ms.  One "step" is one 0.1 ms simulation step
of the whole network (LIF update + spike delivery + Poisson input).

Prints one JSON line: synaptic events/s (whole job), sim-s/wall-s, the
roofline of the dominant kernel (HIP events on the engine's stream over the
timed region) and the CPU baseline (the reference built from source, timed on
this host's cores over a bounded sample of the same network).  --steps and
--warmup are floors: the timed run covers at least the simulated time the
workload's configuration states (2 s for the fly-like net).
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
                   help="steps of the CPU-baseline sample (-1: 1,000 steps, 100 ms simulated)")
    p.add_argument("--no-cpu", action="store_true")
    p.add_argument("--no-sweep", action="store_true", help="skip the sweep points/s figures")
    p.add_argument("--injection", action="store_true",
                   help="the seizure protocol's form: Gaussian current injection on every "
                        "population from step 0 (GaussMean ~ U(0.04, 0.16) nA, GaussSTD = 6 x mean)")
    p.add_argument("--schedule", default="",
                   help="json dict of fs_schedule overrides (default: autotuned on this device "
                        "before the timed region)")
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


def host_cpu():
    """Logical CPUs this process may use and the host's CPU model."""
    model = ""
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    try:
        usable = len(os.sched_getaffinity(0))
    except AttributeError:
        usable = os.cpu_count() or 1
    return {"nproc": usable, "cpu_model": model}


def cpu_baseline(net, arr, events_per_sim_s, n_steps):
    """The reference built from source (oracle/_ref/flysim_ref) timed on this
    host's cores on the same network written in its single-neuron-population
    form (every synapse its own single synapse with its own efficacy, the
    form of its fly-brain experiment); the CPU restatement when the reference
    binary is absent.  Step-loop time from the reference's own 'real time='
    line (code/NetSimGnl2.cpp:351-354), best of 1 thread and this process's
    CPU share (at most 16)."""
    import re
    import subprocess
    import tempfile
    ref = os.path.join(ROOT, "oracle", "_ref", "flysim_ref")
    if not os.path.exists(ref):
        return cpu_baseline_port(net, arr, events_per_sim_s, n_steps)
    from flysim_amd import workloads as W
    host = host_cpu()
    cores = max(1, min(16, host["nproc"]))
    sim_ms = 100.0 if n_steps < 0 else n_steps * net.dt
    single = W.single_neuron_form(net)
    runs = {}
    with tempfile.TemporaryDirectory() as d:
        W.write_conf(single, os.path.join(d, "network.conf"))
        W.write_pro(os.path.join(d, "network.pro"), sim_ms, seeds=[2], spike_file=None)
        for threads in sorted({1, cores}):
            t0 = time.perf_counter()
            out = subprocess.run([ref, "-conf", "network.conf", "-pro", "network.pro", "-t",
                                  str(threads)], cwd=d, capture_output=True, text=True,
                                 timeout=900).stdout.replace("\r", "\n")
            total = time.perf_counter() - t0
            m = re.findall(r"real time=([0-9.eE+-]+)s 100%", out)
            if not m:
                raise RuntimeError("reference run produced no timing line")
            runs[threads] = (float(m[-1]), total)
    threads = min(runs, key=lambda t: runs[t][0])
    wall = runs[threads][0]
    steps = int(np.floor(np.float32(sim_ms) * np.float32(1e-3) * np.float32(1e3) / np.float32(net.dt)))
    sim_per_wall = steps * net.dt * 1e-3 / wall
    return {"value": sim_per_wall * events_per_sim_s, "unit": "synaptic events/s",
            "cores": threads, "kind": "reference", "sim_s_per_wall_s": sim_per_wall,
            "host_nproc": host["nproc"], "cpu_model": host["cpu_model"],
            "step_loop_s": {str(t): runs[t][0] for t in runs},
            "process_s": {str(t): round(runs[t][1], 2) for t in runs},
            "sample": "reference flysim built from source, -t %d (best of %s threads), %d steps "
                      "(%.0f ms simulated) of the same network (%d neurons, %d receptor-synapses, "
                      "same efficacies and drives) as single-neuron populations; per-synapse "
                      "delays become SpikeDly %.2f ms (the format has no per-synapse delay; its "
                      "per-step cost does not depend on delays); step-loop time from its own "
                      "'real time=' line; events/s = GPU-measured events per simulated second x "
                      "CPU sim-s/wall-s"
                      % (threads, "/".join(str(t) for t in sorted(runs)), steps, sim_ms, net.n,
                         net.n_syn, single.notes["spike_delay_ms"])}


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


# Timed steps and warm-up steps never go below the simulated time the
# workload's configuration states (2 s for the fly-like net, 1 s for the
# others) and a tenth of it: shorter runs start from rest and fire nothing.
FLOORS = {"fly20k": (20000, 2000), "ei1k": (10000, 1000), "sparse1m": (10000, 1000)}


def roofline_of(kt, ctr, sched, fused, n, steps, fp64):
    """The dominant kernel, its algorithmic bytes per launch and its average
    launch duration.  The LIF update moves 84 B per neuron-step (160 B in
    FP64); delivery 8 B per spike (row pointer) + 17 B per pushed synapse
    (column, weight, delay, atomic read-modify-write); the NMDA gather 12 B
    per term + 8 B per row."""
    upd_bytes = (UPDATE_BYTES_FP64 if fp64 else UPDATE_BYTES_FP32) * n
    del_bytes = (8 * ctr["spikes"] + 17 * ctr["lin_pushes"] + 12 * ctr["nmda_terms"]
                 + 8 * n * steps) / max(steps, 1)
    upd_ms, del_ms = kt["update_ms"], kt["deliver_ms"]
    win = sched.get("window", 0)
    if win:
        # W steps per launch; graphs hold whole windows, so a launch lasts
        # W graph steps
        dom, b = "k_window", (upd_bytes + del_bytes) * win
        ms = kt["graph_step_ms"] * win if kt["graph_step_ms"] > 0 else upd_ms
    elif fused:
        dom, b = "k_step", upd_bytes + del_bytes
        ms = kt["graph_step_ms"] if kt["graph_step_ms"] > 0 else upd_ms
    elif del_ms >= upd_ms:
        dom, b, ms = "k_deliver", del_bytes, del_ms
    else:
        dom, b, ms = "k_update", upd_bytes, upd_ms
    achieved = b / (ms * 1e-3) / 1e9 if ms > 0 else 0.0
    return {"bound": "hbm", "kernel": dom, "achieved": achieved, "peak": HBM_PEAK_GBS,
            "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS, "traffic": None,
            "bytes_per_launch": b, "avg_launch_ms": ms, "update_ms": upd_ms,
            "deliver_ms": del_ms, "graph_step_ms": kt["graph_step_ms"],
            "timed_launches": kt["n_timed"]}


def assemble(args, steps, warmup, wall, ctr, kt, sched, fused, net_n, n_syn, pairs, dt,
             desc):
    """The bench line (without the CPU and sweep legs): whole-job synaptic
    events per second over the timed steps, with inputs resident in HBM."""
    sim_s = steps * dt * 1e-3
    wall = max(wall, 1e-12)
    return {
        "metric": METRIC, "value": ctr["syn_events"] / wall, "unit": "synaptic events/s",
        "n_gpus": 1, "steps": steps, "warmup": warmup,
        "steps_requested": args.steps, "warmup_requested": args.warmup,
        "ms_per_step": wall / max(steps, 1) * 1e3,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f64" if args.fp64 else "f32", "data": "synthetic",
        "config": {"workload": desc, "neurons": net_n, "receptor_synapses": n_syn,
                   "pairs": pairs, "sim_seconds": sim_s, "schedule": sched},
        "sim_s_per_wall_s": sim_s / wall,
        "spikes_per_s": ctr["spikes"] / wall,
        "mean_rate_hz": ctr["spikes"] / max(net_n, 1) / sim_s if sim_s > 0 else 0.0,
        "roofline": roofline_of(kt, ctr, sched, fused, net_n, steps, args.fp64),
    }


def add_cpu_leg(res, cpu):
    """Attach the CPU baseline; the GPU/CPU ratio only when both are positive."""
    res["cpu_baseline"] = cpu
    v = cpu.get("value") or 0.0
    res["speedup_vs_cpu"] = res["value"] / v if v > 0 and res["value"] > 0 else None
    s = cpu.get("sim_s_per_wall_s") or 0.0
    res["sim_speedup_vs_cpu"] = res["sim_s_per_wall_s"] / s if s > 0 else None
    return res


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
    steps = max(args.steps, FLOORS[args.workload][0])
    warmup = max(args.warmup, FLOORS[args.workload][1])
    net, desc = make_network(args.workload)
    arr = arrays_from_network(net)
    prec = A.FS_FP64 if args.fp64 else A.FS_FP32
    eng = Engine(arr, precision=prec, ext_mode=A.FS_EXT_PHILOX, record=False)
    tuned = None
    if args.schedule:
        eng.set_schedule(**json.loads(args.schedule))
    else:
        from flysim_amd.tune import tune
        eng.reset(2024)
        tuned = tune(eng, cache=None)   # timed live on this device, outside the timed region
    eng.reset(2024)
    if args.injection:
        means = np.random.default_rng(4).uniform(0.04, 0.16, size=len(arr.pops))
        for p, m in enumerate(means):
            eng.set_injection(p, float(m), float(6.0 * m))
        desc += "; Gaussian injection on every population (seizure protocol form)"
    eng.run(warmup, record=False)
    eng.enable_timing(True)
    t0 = time.perf_counter()
    eng.run(steps, record=False)
    wall = time.perf_counter() - t0
    ctr = eng.counters()
    res = assemble(args, steps, warmup, wall, ctr, eng.kernel_times(), eng.schedule, eng.fused,
                   net.n, net.n_syn, int(net.notes.get("pairs", net.n_syn)), net.dt, desc)
    traffic = None if args.injection else measured_traffic(args.workload, res["roofline"]["kernel"],
                                                          res["dtype"])
    if traffic:
        res["roofline"]["traffic"] = traffic
    res["config"]["schedule_source"] = "override" if args.schedule else "autotuned"
    if tuned:
        res["config"]["tuning"] = {"candidates": len(tuned["trials"]), "seconds": round(tuned["tune_s"], 2),
                                   "trial_steps": tuned["trial_steps"],
                                   "time_vs_default_schedule": round(tuned["vs_start"], 4)}
    eng.close()
    if not args.no_sweep and args.workload == "fly20k":
        res["sweep"] = sweep_figures()
    if not args.no_cpu and not args.injection:  # the CPU legs run the workload without injection
        sim_s = steps * net.dt * 1e-3
        add_cpu_leg(res, cpu_baseline(net, arr, ctr["syn_events"] / sim_s, args.cpu_steps))
    print(json.dumps(res))


def measured_traffic(workload, kernel, dtype):
    """HBM bytes per launch of the dominant kernel from the committed
    rocprofv3 FETCH_SIZE / WRITE_SIZE runs (profiles/traffic.json)."""
    path = os.path.join(ROOT, "profiles", "traffic.json")
    if not os.path.exists(path):
        return None
    rec = json.load(open(path)).get("%s/%s/%s" % (workload, kernel, dtype))
    return rec["bytes"] if rec else None


def sweep_figures(points_plain=32, replicas=128, batches=6, mlp_points=1 << 20):
    """Sweep points/s on the fly20k topology (100 ms warm-up + 200 ms window
    per point): plain per-instance GPU simulation vs replicated launches of
    `replicas` points (`batches` of them timed, after one untimed batch that
    tunes the schedule and captures the graphs), and the surrogate's forward over 2^20 points resident
    in device memory (one BF16 matrix-core kernel for all four layers; GPU
    time from HIP events) with its MFMA roofline."""
    from flysim_amd._capi import DeviceBuffer
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
    batched.run(space.sample(batches * replicas, rng), seed=4)
    sizes = [space.dim, 256, 256, 256, space.n_mod]
    mlp = Mlp(sizes, seed=0)
    x = np.ascontiguousarray(space.sample(mlp_points, rng), dtype=np.float32)
    xd = DeviceBuffer(x.nbytes)
    yd = DeviceBuffer(mlp_points * space.n_mod * 4)
    xd.upload(x)
    mlp.predict_device(xd.ptr, 4096, yd.ptr)
    gpu_ms = min(mlp.predict_device(xd.ptr, mlp_points, yd.ptr) for _ in range(3))
    flop = 2.0 * sum(sizes[i] * sizes[i + 1] for i in range(len(sizes) - 1)) * mlp_points
    tflops = flop / (gpu_ms * 1e-3) / 1e12
    return {"topology": "fly20k", "window_ms": 200, "warmup_ms": 100,
            "plain_points_per_s": plain.points_per_s,
            "batched_points_per_s": batched.points_per_s, "replicas": replicas,
            "batched_vs_plain": batched.points_per_s / plain.points_per_s,
            "surrogate_points_per_s": mlp_points / (gpu_ms * 1e-3),
            "surrogate_roofline": {"bound": "mfma", "kernel": "k_mlp_fused (forward, 4 layers)",
                                   "achieved": tflops, "peak": 2500.0, "unit": "TFLOP/s",
                                   "frac": tflops / 2500.0, "points": mlp_points,
                                   "gpu_ms": gpu_ms}}


if __name__ == "__main__":
    main()
