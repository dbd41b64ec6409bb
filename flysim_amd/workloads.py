"""Synthetic workloads for the Flysim simulation core.

Each generator returns a :class:`Network` (population parameter table plus an
explicit synapse list) built with numpy's PCG64 from a stated seed.  The same
network can be written in the reference's drop-in text format
(``network.conf`` / ``network.pro``, grammar of code/parser.cpp:297-907 and
code/parser.cpp:1032-1407) whenever the reference format can express it:

* the 1k E/I network is written as 1,000 single-neuron populations with one
  ``TargetPopulation`` block per synapse and ``Connectivity=1`` so that every
  link takes the reference's single-synapse path
  (code/NetSimGnl2_BulPp.cpp:571-695) with a per-synapse efficacy and no
  ``std::shuffle`` draw;
* the 20k fly-like net and the 1M net carry per-synapse delays, which the
  reference format cannot express; they are handed to the engine directly.

Units follow the reference (code/NetSimGnl2_info.cpp:119-127): nF, nS, ms, mV.
"""
from __future__ import annotations

import dataclasses
import math
import os
from typing import Dict, List, Optional, Sequence

import numpy as np

# Receptor ids used throughout the package (order of the reference's
# RVs/RPs members, code/neuron_affi.h:637-656).
AMPA, GABA, NMDA, ACH, GCL = 0, 1, 2, 3, 4
RECEPTOR_NAMES = {AMPA: "AMPA", GABA: "GABA", NMDA: "NMDA", ACH: "Ach", GCL: "GluCl"}
# network.conf keys of the short-term plasticity parameters (code/parser.cpp:368-383)
STP_KEYS = ("STP_pv", "STP_tD", "STP_aF", "STP_tF", "STP_aCap", "STP_tCap", "STP_aCar", "STP_tCar")


@dataclasses.dataclass
class Receptor:
    """One ``Receptor:`` block of a population (code/parser.cpp:731-801)."""
    kind: int
    tau: float = 5.0          # ms (default, code/parser.cpp:188)
    tau_xt: float = 2.0       # ms (default, code/parser.cpp:189)
    rev: float = 0.0          # mV (default 0 when declared, code/parser.cpp:190)
    freq_ext: float = 0.0     # Hz
    mean_ext_eff: float = 2.1  # nS (default, code/parser.cpp:192)
    mean_ext_con: float = 1.0  # (default, code/parser.cpp:193)


@dataclasses.dataclass
class Population:
    """One ``NeuralPopulation`` block (code/parser.cpp:346-467)."""
    name: str
    n: int
    cm: float = 1.0           # nF   (default, code/parser.cpp:212)
    taum: float = 2.0         # ms   (default, code/parser.cpp:208)
    vth: float = -50.0        # mV
    rest: float = -70.0       # mV
    reset: float = -50.0      # mV   (VRESET, code/Inc.h:26)
    refractory_ms: float = 1.8
    spike_delay_ms: float = 0.0
    self_connection: bool = False
    receptors: List[Receptor] = dataclasses.field(default_factory=list)
    # STP_pv, STP_tD, STP_aF, STP_tF, STP_aCap, STP_tCap, STP_aCar, STP_tCar
    # (None: the reference defaults, code/parser.cpp:229-243)
    stp: Optional[Sequence[float]] = None


@dataclasses.dataclass
class Network:
    """A network as population ranges plus an explicit synapse list.

    ``pre``/``post`` are global neuron ids; populations own contiguous id
    ranges in declaration order (code/NetSimGnl2_BulPp.cpp:40-47).
    ``delay`` is an extra per-synapse transmission delay in steps on top of the
    reference's per-population ``SpikeDly``; 0 reproduces the reference.
    """
    name: str
    pops: List[Population]
    pre: np.ndarray
    post: np.ndarray
    w: np.ndarray          # float32 efficacy in nS
    receptor: np.ndarray   # uint8 receptor id
    delay: np.ndarray      # uint8 extra delay in steps
    dt: float = 0.1
    t_end_ms: float = 1000.0
    seed: int = 1
    notes: Dict[str, object] = dataclasses.field(default_factory=dict)

    @property
    def n(self) -> int:
        return int(sum(p.n for p in self.pops))

    @property
    def n_syn(self) -> int:
        return int(self.pre.shape[0])

    def pop_offsets(self) -> np.ndarray:
        off = np.zeros(len(self.pops) + 1, dtype=np.int64)
        off[1:] = np.cumsum([p.n for p in self.pops])
        return off


def _f32(x) -> np.ndarray:
    return np.asarray(x, dtype=np.float32)


# ---------------------------------------------------------------------------
# 1k E/I network (800 E + 200 I), FP64 parity workload
# ---------------------------------------------------------------------------
C1_G_EXT = 1.6          # nS per external AMPA event (MeanExtEff, MeanExtCon=1)
C1_NMDA_RATIO = 0.1     # g_NMDA = ratio * g_AMPA of the same E->X pair


def c1_network(seed: int = 1, n_exc: int = 800, n_inh: int = 200, p: float = 0.1,
               g_lo: float = 0.5, g_hi: float = 1.5, f_ext: float = 2400.0,
               g_ext: float = C1_G_EXT, nmda_ratio: float = C1_NMDA_RATIO,
               t_end_ms: float = 1000.0) -> Network:
    """Plain E/I LIF network: 800 E + 200 I single-neuron populations.

    Cm 0.5 nF, gL 25 nS (Taum 20 ms), Vrest -70, Vth -50, Vreset -55 mV,
    refractory 2 ms, SpikeDly 1 ms.  AMPA tau 2 ms / Erev 0; NMDA tau 100 ms,
    tauXT 2 ms / Erev 0 (Mg 1); GABA tau 5 ms / Erev -70 mV.  Bernoulli(p)
    pairs without self-connections, g ~ U(g_lo, g_hi) nS per pair; E->X pairs
    carry AMPA (g) and NMDA (nmda_ratio*g); I->X pairs carry GABA (g).
    External Poisson AMPA at f_ext Hz with g_ext nS per event.
    """
    rng = np.random.default_rng(seed)
    n = n_exc + n_inh
    conn = rng.random((n, n)) < p
    np.fill_diagonal(conn, False)
    g = _f32(rng.uniform(g_lo, g_hi, size=(n, n)))
    pops = []
    for i in range(n):
        exc = i < n_exc
        pops.append(Population(
            name=("E%d" % i) if exc else ("I%d" % (i - n_exc)), n=1, cm=0.5,
            taum=20.0, vth=-50.0, rest=-70.0, reset=-55.0, refractory_ms=2.0,
            spike_delay_ms=1.0, self_connection=False,
            receptors=[Receptor(AMPA, tau=2.0, rev=0.0, freq_ext=f_ext,
                                mean_ext_eff=g_ext, mean_ext_con=1.0),
                       Receptor(NMDA, tau=100.0, tau_xt=2.0, rev=0.0),
                       Receptor(GABA, tau=5.0, rev=-70.0)]))
    pre_l, post_l, w_l, r_l = [], [], [], []
    for j in range(n):
        tg = np.nonzero(conn[j])[0]
        if j < n_exc:
            pre_l += [np.full(tg.size, j), np.full(tg.size, j)]
            post_l += [tg, tg]
            w_l += [g[j, tg], _f32(g[j, tg] * np.float32(nmda_ratio))]
            r_l += [np.full(tg.size, AMPA), np.full(tg.size, NMDA)]
        else:
            pre_l.append(np.full(tg.size, j))
            post_l.append(tg)
            w_l.append(g[j, tg])
            r_l.append(np.full(tg.size, GABA))
    pre = np.concatenate(pre_l).astype(np.uint32)
    post = np.concatenate(post_l).astype(np.uint32)
    w = _f32(np.concatenate(w_l))
    rec = np.concatenate(r_l).astype(np.uint8)
    return Network(name="ei1k", pops=pops, pre=pre, post=post, w=w, receptor=rec,
                   delay=np.zeros(pre.shape[0], np.uint8), dt=0.1, t_end_ms=t_end_ms,
                   seed=seed, notes={"g_ext_nS": g_ext, "nmda_ratio": nmda_ratio,
                                     "p": p, "f_ext_Hz": f_ext})


# ---------------------------------------------------------------------------
# Synthetic Drosophila-scale connectome (20k neurons, 50 modules)
# ---------------------------------------------------------------------------
C2_G_EXT = 1.1          # nS per external AMPA event (~12 Hz mean; >50 Hz from ~1.5 nS)
C2_NMDA_RATIO = 0.1


def _ei_receptors(f_ext: float, g_ext: float) -> List[Receptor]:
    return [Receptor(AMPA, tau=2.0, rev=0.0, freq_ext=f_ext, mean_ext_eff=g_ext, mean_ext_con=1.0),
            Receptor(NMDA, tau=100.0, tau_xt=2.0, rev=0.0),
            Receptor(GABA, tau=5.0, rev=-70.0)]


def _lif_pop(name: str, n: int, f_ext: float, g_ext: float) -> Population:
    return Population(name=name, n=int(n), cm=0.5, taum=20.0, vth=-50.0, rest=-70.0, reset=-55.0,
                      refractory_ms=2.0, spike_delay_ms=0.0, self_connection=False,
                      receptors=_ei_receptors(f_ext, g_ext))


def _emit(pre, post, g, exc_pre, nmda_ratio, delay):
    """E presynaptic pairs carry AMPA (g) and NMDA (ratio*g), I pairs GABA (g)."""
    e = exc_pre
    pre_o = np.concatenate([pre[e], pre[e], pre[~e]]).astype(np.uint32)
    post_o = np.concatenate([post[e], post[e], post[~e]]).astype(np.uint32)
    w_o = _f32(np.concatenate([g[e], g[e] * np.float32(nmda_ratio), g[~e]]))
    r_o = np.concatenate([np.full(e.sum(), AMPA), np.full(e.sum(), NMDA),
                          np.full((~e).sum(), GABA)]).astype(np.uint8)
    d_o = np.concatenate([delay[e], delay[e], delay[~e]]).astype(np.uint8)
    return pre_o, post_o, w_o, r_o, d_o


def _largest_remainder(x: np.ndarray, total: int) -> np.ndarray:
    x = x / x.sum() * total
    n = np.floor(x).astype(np.int64)
    rest = total - n.sum()
    n[np.argsort(-(x - n))[:rest]] += 1
    return n


def c2_network(seed: int = 2, n: int = 20000, n_mod: int = 50, mu: float = 5.8,
               sigma: float = 0.6, exc_frac: float = 0.7, p_in: float = 0.05,
               p_out: float = 0.002, mod_edges: int = 150, w_sigma: float = 0.8,
               d_lo: int = 10, d_hi: int = 20, f_lo: float = 1000.0, f_hi: float = 3000.0,
               g_ext: float = C2_G_EXT, nmda_ratio: float = C2_NMDA_RATIO,
               t_end_ms: float = 2000.0, drive: Optional[np.ndarray] = None) -> Network:
    """Synthetic fly-like connectome: n LIF neurons in n_mod modules with
    lognormal(mu, sigma) sizes renormalised to n; exc_frac excitatory
    (AMPA+NMDA) / rest inhibitory (GABA) in every module; Bernoulli pairs with
    p_in inside a module and p_out (both directions) along the edges of a
    G(n_mod, mod_edges) random module graph (mean degree 2*mod_edges/n_mod);
    efficacies lognormal(median 1 nS, w_sigma); delays uniform
    {d_lo..d_hi} steps (1-2 ms); Poisson AMPA drive U(f_lo, f_hi) Hz per
    module (or `drive`).  Populations: module m -> "M<m>E", "M<m>I"."""
    rng = np.random.default_rng(seed)
    sizes = _largest_remainder(rng.lognormal(mu, sigma, n_mod), n)
    pairs = [(a, b) for a in range(n_mod) for b in range(a + 1, n_mod)]
    edges = [pairs[i] for i in rng.choice(len(pairs), size=mod_edges, replace=False)]
    if drive is None:
        drive = rng.uniform(f_lo, f_hi, n_mod)
    pops, first, n_exc = [], np.zeros(n_mod + 1, np.int64), np.zeros(n_mod, np.int64)
    for m in range(n_mod):
        ne = int(round(sizes[m] * exc_frac))
        n_exc[m] = ne
        pops.append(_lif_pop("M%dE" % m, ne, float(drive[m]), g_ext))
        pops.append(_lif_pop("M%dI" % m, sizes[m] - ne, float(drive[m]), g_ext))
        first[m + 1] = first[m] + sizes[m]
    pre_l, post_l = [], []
    for m in range(n_mod):
        k = int(sizes[m])
        c = rng.random((k, k)) < p_in
        np.fill_diagonal(c, False)
        a, b = np.nonzero(c)
        pre_l.append(a + first[m]); post_l.append(b + first[m])
    for (a_m, b_m) in edges:
        for (x, y) in ((a_m, b_m), (b_m, a_m)):
            c = rng.random((int(sizes[x]), int(sizes[y]))) < p_out
            a, b = np.nonzero(c)
            pre_l.append(a + first[x]); post_l.append(b + first[y])
    pre = np.concatenate(pre_l)
    post = np.concatenate(post_l)
    mod_of = np.repeat(np.arange(n_mod), sizes)
    exc = (pre - first[mod_of[pre]]) < n_exc[mod_of[pre]]
    g = _f32(rng.lognormal(0.0, w_sigma, pre.shape[0]))
    d = rng.integers(d_lo, d_hi + 1, pre.shape[0])
    P, Q, Wt, R, Dl = _emit(pre, post, g, exc, nmda_ratio, d)
    return Network(name="fly20k", pops=pops, pre=P, post=Q, w=Wt, receptor=R, delay=Dl, dt=0.1,
                   t_end_ms=t_end_ms, seed=seed,
                   notes={"pairs": int(pre.shape[0]), "module_sizes": sizes.tolist(),
                          "module_edges": edges, "drive_Hz": [float(x) for x in drive],
                          "g_ext_nS": g_ext, "nmda_ratio": nmda_ratio})


# ---------------------------------------------------------------------------
# Scaled sparse network (1M neurons, fixed in-degree 100)
# ---------------------------------------------------------------------------
C3_G_EXT = 1.4   # the 1M net is bistable: quiet below ~1.3 nS, refractory-limited above ~1.5


def c3_network(seed: int = 3, n: int = 1_000_000, exc_frac: float = 0.8, k_in: int = 100,
               w_sigma: float = 0.8, d_lo: int = 10, d_hi: int = 50, f_ext: float = 2000.0,
               g_ext: float = C3_G_EXT, nmda_ratio: float = 0.1,
               t_end_ms: float = 1000.0) -> Network:
    """n LIF neurons (exc_frac E), each receiving exactly k_in synapses from
    uniformly drawn presynaptic neurons (self excluded); lognormal(median
    1 nS, w_sigma) efficacies; delays uniform {d_lo..d_hi} steps (1-5 ms);
    Poisson AMPA drive f_ext Hz.  E pairs carry AMPA+NMDA, I pairs GABA."""
    rng = np.random.default_rng(seed)
    ne = int(round(n * exc_frac))
    pops = [_lif_pop("E", ne, f_ext, g_ext), _lif_pop("I", n - ne, f_ext, g_ext)]
    post = np.repeat(np.arange(n, dtype=np.uint32), k_in)
    pre = rng.integers(0, n - 1, size=n * k_in, dtype=np.uint32)
    pre = pre + (pre >= post)  # skip self without bias
    g = _f32(rng.lognormal(0.0, w_sigma, pre.shape[0]))
    d = rng.integers(d_lo, d_hi + 1, pre.shape[0], dtype=np.uint8)
    exc = pre < ne
    P, Q, Wt, R, Dl = _emit(pre, post, g, exc, nmda_ratio, d)
    return Network(name="sparse1m", pops=pops, pre=P, post=Q, w=Wt, receptor=R, delay=Dl,
                   dt=0.1, t_end_ms=t_end_ms, seed=seed,
                   notes={"pairs": int(pre.shape[0]), "g_ext_nS": g_ext, "k_in": k_in})


# ---------------------------------------------------------------------------
# Reference text format writers (drop-in API)
# ---------------------------------------------------------------------------
def _fmt(x: float) -> str:
    """A decimal that the reference's ``stof`` reads back as the same float32."""
    return "%.9g" % float(np.float32(x))


def write_conf(net: Network, path: str) -> None:
    """Write ``network.conf`` in the reference grammar (code/parser.cpp:297-907).

    Only networks whose populations are single neurons (or whose synapse list
    is expressible as population links with Connectivity=1 and one target) can
    be written: every synapse becomes one ``TargetPopulation`` block.
    """
    if np.any(net.delay != 0):
        raise ValueError("per-synapse delays are not expressible in network.conf")
    off = net.pop_offsets()
    if not all(p.n == 1 for p in net.pops):
        raise ValueError("write_conf supports single-neuron populations only")
    order = np.lexsort((np.arange(net.n_syn), net.pre))
    by_pre: Dict[int, List[int]] = {}
    for s in order:
        by_pre.setdefault(int(net.pre[s]), []).append(int(s))
    out = []
    for i, pop in enumerate(net.pops):
        out.append("NeuralPopulation: %s\n" % pop.name)
        out.append("N=%d\nC=%s\nTaum=%s\nThreshold=%s\nRestPot=%s\nResetPot=%s\n"
                   "RefractoryPeriod=%s\nSpikeDly=%s\n" % (
                       pop.n, _fmt(pop.cm), _fmt(pop.taum), _fmt(pop.vth),
                       _fmt(pop.rest), _fmt(pop.reset), _fmt(pop.refractory_ms),
                       _fmt(pop.spike_delay_ms)))
        if pop.self_connection:
            out.append("SelfConnection=true\n")
        if pop.stp is not None:
            for key, v in zip(STP_KEYS, pop.stp):
                out.append("%s=%s\n" % (key, _fmt(v)))
        for r in pop.receptors:
            out.append("Receptor:%s\nTau=%s\nTauXT=%s\nRevPot=%s\nFreqExt=%s\n"
                       "MeanExtEff=%s\nMeanExtCon=%s\nEndReceptor\n" % (
                           RECEPTOR_NAMES[r.kind], _fmt(r.tau), _fmt(r.tau_xt),
                           _fmt(r.rev), _fmt(r.freq_ext), _fmt(r.mean_ext_eff),
                           _fmt(r.mean_ext_con)))
        for s in by_pre.get(int(off[i]), []):
            out.append("TargetPopulation: %s\nTargetReceptor=%s\nMeanEff=%s\n"
                       "EndTargetPopulation\n" % (
                           net.pops[int(net.post[s])].name,
                           RECEPTOR_NAMES[int(net.receptor[s])], _fmt(net.w[s])))
        out.append("EndNeuralPopulation\n\n")
    with open(path, "w") as f:
        f.write("".join(out))


def single_neuron_form(net: Network, spike_delay_ms: Optional[float] = None) -> Network:
    """The same neurons, synapses, efficacies and drives as `net`, with every
    neuron its own population ("N<id>", parameters of its population), the
    form the reference's fly-brain experiment uses (one population per
    neuron, experiment/seizure/network1_900_10nOvf.pro) and the one in which
    every link is a single synapse with its own efficacy
    (code/NetSimGnl2_BulPp.cpp:571-695).  Per-synapse delays have no
    counterpart there: they become one SpikeDly for every neuron (their mean
    when `spike_delay_ms` is None); the reference's cost per step does not
    depend on delays (it updates every synapse every step)."""
    off = net.pop_offsets()
    if spike_delay_ms is None:
        spike_delay_ms = float(np.mean(net.delay)) * net.dt if net.n_syn else 0.0
    pops = []
    for i, p in enumerate(net.pops):
        for k in range(p.n):
            pops.append(dataclasses.replace(p, name="N%d" % (off[i] + k), n=1,
                                            spike_delay_ms=float(spike_delay_ms)))
    return Network(name=net.name + "-single", pops=pops, pre=net.pre, post=net.post, w=net.w,
                   receptor=net.receptor, delay=np.zeros(net.n_syn, np.uint8), dt=net.dt,
                   t_end_ms=net.t_end_ms, seed=net.seed,
                   notes=dict(net.notes, spike_delay_ms=float(spike_delay_ms)))


def seizure_events(names: Sequence[str], seed: int, onset_ms: float = 1.0,
                   mean_lo: float = 0.04, mean_hi: float = 0.16,
                   std_ratio: float = 6.0) -> List[str]:
    """Events of the fly-brain seizure protocol's form
    (experiment/seizure/network1_900_10nOvf.pro): at ``onset_ms`` every
    population gets its own ChangeMembraneNoise with GaussMean ~ U(mean_lo,
    mean_hi) nA and GaussSTD = std_ratio * GaussMean (the protocol's ratio),
    one event block per population."""
    rng = np.random.default_rng(seed)
    means = rng.uniform(mean_lo, mean_hi, size=len(names))
    return ["EventTime %s\nType=ChangeMembraneNoise\nLabel=#1#\nPopulation:%s\n"
            "GaussMean=%s\nGaussSTD=%s\nEndEvent\n%%--------------------------------------\n\n"
            % (_fmt(onset_ms), n, _fmt(m), _fmt(std_ratio * m)) for n, m in zip(names, means)]


def write_pro(path: str, t_end_ms: float, seeds: Sequence[int] = (),
              spike_file: Optional[str] = "Spikes.dat",
              rate_file: Optional[str] = None, rate_window_ms: int = 50,
              rate_print_ms: int = 10, events: Sequence[str] = (),
              rate_sum: bool = False) -> None:
    """Write ``network.pro`` (code/parser.cpp:911-1407): seeds, events, outputs."""
    out = []
    if seeds:
        out.append("DefineMacro\nRandomSeeds: %s EndRandomSeeds\nEndDefineMacro\n\n"
                   % " ".join(str(int(s)) for s in seeds))
    out.extend(events)
    out.append("EventTime %s\nType=EndTrial\nLabel=End_of_the_trial\nEndEvent\n\n"
               % _fmt(t_end_ms))
    out.append("OutControl\n")
    if spike_file:
        out.append("FileName:%s\nType=Spike\npopulation:AllPopulation\nEndOutputFile\n"
                   % spike_file)
    if rate_file:
        out.append("FileName:%s\nType=FiringRate\nFiringRateWindow=%d\nPrintStep=%d\n"
                   "population:AllPopulation\n%sEndOutputFile\n"
                   % (rate_file, rate_window_ms, rate_print_ms,
                      "SumAllPopulation:true\n" if rate_sum else ""))
    out.append("EndOutControl\n")
    with open(path, "w") as f:
        f.write("".join(out))


def write_module_conf(net: Network, path: str, spike_delay_ms: Optional[float] = None) -> Dict:
    """Write a population-level ("module form") network.conf of `net`: one
    TargetPopulation link per (pre population, post population, receptor)
    with Connectivity = the pair density of that block and MeanEff = its mean
    efficacy.  The reference realises it with a fixed out-degree
    ceil(N_post * Connectivity) per presynaptic neuron (code/NetSimGnl2_BulPp.cpp:310)
    and one efficacy per link, so it is a network of the same size and class,
    not the same synapses; per-synapse delays become one SpikeDly per
    population (mean delay when `spike_delay_ms` is None).  Used to time the
    reference on the fly-like and 1M workloads."""
    off = net.pop_offsets()
    P = len(net.pops)
    pre_pop = np.searchsorted(off, net.pre, side="right") - 1
    post_pop = np.searchsorted(off, net.post, side="right") - 1
    key = (pre_pop.astype(np.int64) * P + post_pop) * 8 + net.receptor
    uniq, inv, cnt = np.unique(key, return_inverse=True, return_counts=True)
    wsum = np.bincount(inv, weights=net.w.astype(np.float64))
    links: Dict[int, list] = {}
    for k, c, ws in zip(uniq, cnt, wsum):
        rec = int(k % 8)
        pp, qq = divmod(int(k // 8), P)
        n_pre, n_post = net.pops[pp].n, net.pops[qq].n
        conn = c / float(n_pre) / float(n_post)
        links.setdefault(pp, []).append((qq, rec, conn, ws / c))
    if spike_delay_ms is None:
        spike_delay_ms = float(np.mean(net.delay)) * net.dt if net.n_syn else 0.0
    out = []
    for i, pop in enumerate(net.pops):
        out.append("NeuralPopulation: %s\nN=%d\nC=%s\nTaum=%s\nThreshold=%s\nRestPot=%s\n"
                   "ResetPot=%s\nRefractoryPeriod=%s\nSpikeDly=%s\n" % (
                       pop.name, pop.n, _fmt(pop.cm), _fmt(pop.taum), _fmt(pop.vth),
                       _fmt(pop.rest), _fmt(pop.reset), _fmt(pop.refractory_ms),
                       _fmt(spike_delay_ms)))
        for r in pop.receptors:
            out.append("Receptor:%s\nTau=%s\nTauXT=%s\nRevPot=%s\nFreqExt=%s\n"
                       "MeanExtEff=%s\nMeanExtCon=%s\nEndReceptor\n" % (
                           RECEPTOR_NAMES[r.kind], _fmt(r.tau), _fmt(r.tau_xt), _fmt(r.rev),
                           _fmt(r.freq_ext), _fmt(r.mean_ext_eff), _fmt(r.mean_ext_con)))
        for (qq, rec, conn, g) in links.get(i, []):
            out.append("TargetPopulation: %s\nTargetReceptor=%s\nMeanEff=%s\nConnectivity=%s\n"
                       "EndTargetPopulation\n" % (net.pops[qq].name, RECEPTOR_NAMES[rec],
                                                  _fmt(g), _fmt(conn)))
        out.append("EndNeuralPopulation\n\n")
    with open(path, "w") as f:
        f.write("".join(out))
    n_links = sum(len(v) for v in links.values())
    syn = sum(int(np.ceil(net.pops[qq].n * np.float32(conn))) * net.pops[pp].n
              for pp, v in links.items() for (qq, rec, conn, g) in v)
    return {"links": n_links, "approx_synapses": int(syn), "spike_delay_ms": spike_delay_ms}


def read_spikes(path: str) -> np.ndarray:
    """Read a reference ``Spike`` file ("%.7g time_s nid" lines) as (step?, nid).

    Returns an (n, 2) array of [time_s as float64, nid]."""
    if not os.path.exists(path) or os.path.getsize(path) == 0:
        return np.zeros((0, 2))
    return np.loadtxt(path, ndmin=2)
