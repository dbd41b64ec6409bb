"""Batched parameter sweeps over many independent E/I networks of uneven size.

The 256-instance sweep: instances of the 1k E/I generator's kind (80/20 E/I,
Bernoulli(0.1) pairs without self-connections, g ~ U(0.5, 1.5) nS, E->X
pairs AMPA + NMDA (0.1 g), I->X pairs GABA, SpikeDly 1 ms, refractory 2 ms)
with sizes uniform in {1000..20000}, seeds 1000 + k, external AMPA rate
uniform in 0.5-4 kHz and NMDA / GABA gains uniform in 0.5-2, all in one
handle whose synapse tables the device generates (fs_create_ei_batch).

The sweep states its 1 ms delay as a synaptic delay of 10 steps (SpikeDly
0) instead of the reference form's SpikeDly of 1 ms.  The two are the same
network: SpikeDly only delays the spike flag (code/EqsGnl2.cpp:229-251, the
flag is set at Count == SpiDy, once per firing while SpiDy < RfPed - 1),
and the presynaptic gates are driven by that flag alone, so moving the flag
10 steps earlier shifts every gate trajectory 10 steps earlier, which the
10-step synaptic delay shifts back: each target receives the identical
conductance sequence.  Only the reported spike times differ (threshold
crossing instead of 1 ms later).  With every synapse delayed, one launch
runs 11 steps (the windowed path).

``ei_batch_network`` restates that generator in numpy (same hash, same float
rounding) for small instances: the tests build the same networks through the
host path (fs_create) and compare rasters.
"""
from __future__ import annotations

from typing import Optional

import numpy as np

from . import workloads as W
from .engine import EiBatchSpec, Engine

M64 = np.uint64(0xFFFFFFFFFFFFFFFF)


def mix64(x: np.ndarray) -> np.ndarray:
    """splitmix64 finaliser (Steele, Lea, Flood 2014), uint64 wrap-around."""
    with np.errstate(over="ignore"):
        z = (np.asarray(x, dtype=np.uint64) + np.uint64(0x9E3779B97F4A7C15))
        z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
        return z ^ (z >> np.uint64(31))


def c1_populations(f_ext: float = 2400.0, g_ext: float = W.C1_G_EXT, spike_delay_ms: float = 1.0):
    """E and I population templates of the 1k E/I network (sizes unused)."""
    def pop(name):
        return W.Population(name=name, n=1, cm=0.5, taum=20.0, vth=-50.0, rest=-70.0, reset=-55.0,
                            refractory_ms=2.0, spike_delay_ms=spike_delay_ms, self_connection=False,
                            receptors=[W.Receptor(W.AMPA, tau=2.0, rev=0.0, freq_ext=f_ext,
                                                  mean_ext_eff=g_ext, mean_ext_con=1.0),
                                       W.Receptor(W.NMDA, tau=100.0, tau_xt=2.0, rev=0.0),
                                       W.Receptor(W.GABA, tau=5.0, rev=-70.0)])
    return pop("E"), pop("I")


def c4_spec(n_inst: int = 256, seed: int = 4, n_lo: int = 1000, n_hi: int = 20000,
            f_lo: float = 500.0, f_hi: float = 4000.0, gain_lo: float = 0.5,
            gain_hi: float = 2.0, g_ext: float = W.C1_G_EXT, synaptic_delay: bool = True) -> EiBatchSpec:
    """The 256-instance sweep (sizes, rates and gains drawn with numpy's
    PCG64 from `seed`; instance k uses connectivity seed 1000 + k).  The
    1 ms delay is a 10-step synaptic delay (the same dynamics as SpikeDly
    1 ms, see the module docstring), or SpikeDly 1 ms with
    synaptic_delay=False."""
    rng = np.random.default_rng(seed)
    n = rng.integers(n_lo, n_hi + 1, n_inst)
    ext = rng.uniform(f_lo, f_hi, n_inst)
    gn = rng.uniform(gain_lo, gain_hi, n_inst)
    gg = rng.uniform(gain_lo, gain_hi, n_inst)
    e, i = c1_populations(g_ext=g_ext, spike_delay_ms=0.0 if synaptic_delay else 1.0)
    return EiBatchSpec(n, 1000 + np.arange(n_inst), ext, gn, gg, e, i,
                       delay_steps=10 if synaptic_delay else 0)


def ei_batch_network(spec: EiBatchSpec) -> W.Network:
    """Explicit synapse list of a (small) batch, identical to the device
    generator: pair (j, p) of instance k exists iff the high 32 bits of
    mix64(mix64(seed_k) ^ (j << 32 | p)) are below p_conn * 2^32; g from the
    low 24 bits.  Synapses are emitted in (pre, post) order."""
    thr = np.uint64(min(4294967295, int(np.floor(np.float64(np.float32(spec.p_conn)) * 4294967296.0))))
    g_lo = np.float32(spec.g_lo)
    g_span = np.float32(np.float32(spec.g_hi) - np.float32(spec.g_lo))
    ratio = np.float32(spec.nmda_ratio)
    n_exc = spec.n_exc()
    pops, pre_l, post_l, w_l, r_l = [], [], [], [], []
    first = 0
    for k, n in enumerate(spec.n_neuron):
        n = int(n)
        ne = int(n_exc[k])
        key = mix64(np.uint64(spec.seed[k]))
        j = np.repeat(np.arange(n, dtype=np.uint64), n)
        p = np.tile(np.arange(n, dtype=np.uint64), n)
        h = mix64(key ^ ((j << np.uint64(32)) | p))
        ok = ((h >> np.uint64(32)) < thr) & (j != p)
        u = (h & np.uint64(0xFFFFFF)).astype(np.float32) * np.float32(1.0 / 16777216.0)
        g = (g_lo + (g_span * u).astype(np.float32)).astype(np.float32)
        jj, pp, gg = j[ok].astype(np.int64), p[ok].astype(np.int64), g[ok]
        exc = jj < ne
        gn = np.float32(spec.gain_nmda[k])
        gb = np.float32(spec.gain_gaba[k])
        # per pair in (j, p) order: AMPA then NMDA for E, GABA for I
        for sel, w, rec in ((exc, gg, W.AMPA),
                            (exc, ((gg * ratio).astype(np.float32) * gn).astype(np.float32), W.NMDA),
                            (~exc, (gg * gb).astype(np.float32), W.GABA)):
            pre_l.append(jj[sel] + first)
            post_l.append(pp[sel] + first)
            w_l.append(w[sel])
            r_l.append(np.full(int(sel.sum()), rec, np.uint8))
        e, i = spec.exc, spec.inh
        for tmpl, size, name in ((e, ne, "E%d" % k), (i, n - ne, "I%d" % k)):
            q = W.Population(**{**tmpl.__dict__, "name": name, "n": size,
                                "receptors": [W.Receptor(**r.__dict__) for r in tmpl.receptors]})
            for r in q.receptors:
                if r.kind == W.AMPA:
                    r.freq_ext = float(spec.ext_hz[k])
            pops.append(q)
        first += n
    pre = np.concatenate(pre_l).astype(np.uint32)
    order = np.argsort(pre, kind="stable")
    return W.Network(name="ei_batch", pops=pops, pre=pre[order],
                     post=np.concatenate(post_l).astype(np.uint32)[order],
                     w=np.concatenate(w_l).astype(np.float32)[order],
                     receptor=np.concatenate(r_l)[order],
                     delay=np.full(pre.size, spec.delay_steps, np.uint8), dt=spec.dt)


def expected_pairs(spec: EiBatchSpec) -> float:
    n = spec.n_neuron.astype(np.float64)
    return float(np.sum(n * (n - 1)) * spec.p_conn)


def subset(spec: EiBatchSpec, idx) -> EiBatchSpec:
    """The instances `idx` of a batch as a batch of their own."""
    idx = np.atleast_1d(idx)
    return EiBatchSpec(spec.n_neuron[idx], spec.seed[idx], spec.ext_hz[idx], spec.gain_nmda[idx],
                       spec.gain_gaba[idx], spec.exc, spec.inh, spec.exc_frac, spec.p_conn,
                       spec.g_lo, spec.g_hi, spec.nmda_ratio, spec.dt, spec.delay_steps)


def run_batch(spec: EiBatchSpec, steps: int, seed: int = 1, schedule="auto", log=print):
    """Generate and run all instances in one handle; per-instance E/I rates.
    schedule: fs_schedule overrides, or "auto" to autotune the launch
    schedule first (cached by network signature and device)."""
    import time
    from . import _capi as A
    t0 = time.perf_counter()
    eng = Engine.ei_batch(spec, precision=A.FS_FP32, record=False)
    t_gen = time.perf_counter() - t0
    tuned = None
    if schedule == "auto":
        from .tune import tune
        eng.reset(seed)
        tuned = tune(eng)
    elif schedule:
        eng.set_schedule(**schedule)
    eng.reset(seed)
    eng.enable_timing(True)
    t1 = time.perf_counter()
    eng.run(steps, record=False)
    wall = time.perf_counter() - t1
    ctr = eng.counters()
    kt = eng.kernel_times()
    counts = eng.pop_counts().astype(np.float64).reshape(-1, 2)
    ne = spec.n_exc()
    sizes = np.stack([ne, spec.n_neuron - ne], axis=1).astype(np.float64)
    rates = counts / np.maximum(sizes, 1) / (steps * spec.dt * 1e-3)
    out = {"t_gen_s": t_gen, "wall_s": wall, "ms_per_step": wall / steps * 1e3,
           "counters": ctr, "kernel_times": kt, "rates": rates, "schedule": eng.schedule,
           "tune_s": tuned["tune_s"] if tuned else 0.0}
    eng.close()
    return out


def main(argv=None):
    """Batched vs plain per-instance simulation of the 256-instance sweep."""
    import argparse
    import json
    import time
    ap = argparse.ArgumentParser(prog="python -m flysim_amd.batch")
    ap.add_argument("--steps", type=int, default=10000)
    ap.add_argument("--instances", type=int, default=256)
    ap.add_argument("--plain", type=int, default=256,
                    help="instances also run one by one (0: skip)")
    ap.add_argument("--lanes", type=int, default=0,
                    help="NMDA gather lanes per row (0: autotuned schedule)")
    ap.add_argument("--schedule", default="",
                    help='fs_schedule fields as JSON, e.g. \'{"gather_lanes": 32}\' (instead of the autotuner)')
    ap.add_argument("--out", default="")
    a = ap.parse_args(argv)
    log = lambda m: print(m, flush=True)
    spec = c4_spec(n_inst=a.instances)
    sched = json.loads(a.schedule) if a.schedule else ({"gather_lanes": a.lanes} if a.lanes else "auto")
    n = spec.n_neuron.astype(np.float64)
    ne = spec.n_exc().astype(np.float64)
    pairs_expect = float(np.sum(n * (n - 1)) * spec.p_conn)
    log("batch: %d instances, %d neurons, ~%.3g pairs" % (n.size, n.sum(), pairs_expect))
    b = run_batch(spec, a.steps, schedule=sched, log=log)
    nmda = b["counters"]["nmda_terms"] / a.steps
    log("batched: gen %.1f s, %.3f ms/step, wall %.1f s" % (b["t_gen_s"], b["ms_per_step"], b["wall_s"]))
    res = {"instances": int(n.size), "neurons": int(n.sum()), "steps": a.steps,
           "sim_seconds_each": a.steps * spec.dt * 1e-3,
           "e_pairs_nmda": nmda, "pairs_expected": pairs_expect,
           "receptor_synapses": float(nmda * 2 + (pairs_expect - nmda)),
           "batched": {k: v for k, v in b.items() if k not in ("rates",)},
           "rates_quantiles_hz": np.quantile(b["rates"].mean(axis=1), [0, .25, .5, .75, 1]).tolist(),
           "frac_mean_rate_over_50hz": float((b["rates"].mean(axis=1) > 50).mean())}
    # gather bytes per step: 8 B per NMDA term (column+weight) + a 4 B gate read
    res["batched"]["gather_bytes_per_step"] = 12.0 * nmda + 8.0 * n.sum()
    if a.plain:
        plain_sched = sched
        if sched == "auto":
            # one schedule for every plain run: the one tuned on the largest
            # instance (tuning each of them would dominate their run time)
            big = int(np.argmax(spec.n_neuron))
            plain_sched = run_batch(subset(spec, big), 10, schedule="auto")["schedule"]
            res["plain_schedule"] = plain_sched
        t0 = time.perf_counter()
        walls = []
        for k in range(min(a.plain, n.size)):
            r = run_batch(subset(spec, k), a.steps, schedule=plain_sched)
            walls.append(r["t_gen_s"] + r["wall_s"])
            if (k + 1) % 16 == 0:
                log("plain: %d instances, %.1f s" % (k + 1, time.perf_counter() - t0))
        plain = float(np.sum(walls)) * n.size / len(walls)
        res["plain"] = {"instances_run": len(walls), "total_s": plain,
                        "instances_per_s": n.size / plain}
        res["batched_instances_per_s"] = n.size / (b["t_gen_s"] + b["wall_s"])
        res["speedup_batched_vs_plain"] = plain / (b["t_gen_s"] + b["wall_s"])
    txt = json.dumps(res, default=float)
    print(txt)
    if a.out:
        open(a.out, "w").write(txt)


if __name__ == "__main__":
    main()
