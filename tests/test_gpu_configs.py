"""FP32 state against FP64 state at the full workload sizes, through the C
ABI, on the same network and the same counter-based input stream (the FP32
mode approximates the FP64 arithmetic of code/EqsGnl2.cpp:152-225, 284-337):

* the 20,000-neuron fly-like net for the 2 s its configuration states;
* the 1,000,000-neuron, in-degree-100 net for 1 s (it ignites after about
  0.4 s and then fires near its refractory limit);
* the 256 uneven E/I instances (1,000-20,000 neurons each, 2.76 M neurons,
  about 7.5e9 receptor-synapses) for 1,000 steps;

plus, on the full fly-like net in FP64, the windowed step path (11 steps per
launch) against the two-launch path (update, then delivery).  Each test
prints its measured errors.  Bars: per-population rates within 2 % (every
population with at least 500 FP64 spikes; smaller ones are inside Poisson
noise), every population's count difference within 3 Poisson standard
deviations, and per-neuron spike counts within a two-sample KS p > 0.05
(for the 1M net over its active half: it ignites a few steps earlier or
later in FP32, a shift a KS test on a million whole-run counts resolves)."""
import numpy as np
import pytest
from scipy import stats

from flysim_amd import _capi as A
from flysim_amd import workloads as W
from flysim_amd.batch import c4_spec
from flysim_amd.engine import Engine, arrays_from_network

pytestmark = pytest.mark.gpu

SEED = 77


def _counts(make, precision, steps, ks_from, chunk=1000):
    """Per-neuron spike counts of the whole run and of steps >= ks_from, and
    per-population counts of the whole run."""
    eng = make(precision)
    eng.reset(SEED)
    c = np.zeros(eng.n, np.int64)
    late = np.zeros(eng.n, np.int64)
    done = 0
    while done < steps:
        k = min(chunk, steps - done)
        off, ids = eng.run(k)
        c += np.bincount(ids, minlength=eng.n)
        if done + k > ks_from:
            late += np.bincount(ids[off[max(0, ks_from - done)]:], minlength=eng.n)
        done += k
    pops = eng.pop_counts().astype(np.int64)
    sched = eng.schedule
    eng.close()
    return c, late, pops, sched


def _compare(label, make, steps, pop_size, ks_from=0):
    """ks_from: first step of the window whose per-neuron counts the KS test
    compares (0: the whole run)."""
    c64, l64, p64, s64 = _counts(make, A.FS_FP64, steps, ks_from)
    c32, l32, p32, s32 = _counts(make, A.FS_FP32, steps, ks_from)
    assert s64 == s32
    sim_s = steps * 0.1e-3
    r64 = p64 / pop_size / sim_s
    r32 = p32 / pop_size / sim_s
    rel = np.abs(r32 - r64) / np.maximum(r64, 1e-12)
    z = (p32 - p64) / np.sqrt(np.maximum(p32 + p64, 1))
    big = p64 >= 500
    ks = stats.ks_2samp(l32, l64)
    ks_all = stats.ks_2samp(c32, c64)
    tot = abs(int(p32.sum()) - int(p64.sum())) / max(int(p64.sum()), 1)
    print("%s: %d steps, schedule %s; mean rate FP64 %.3f Hz, FP32 %.3f Hz (total %.2e rel.); "
          "per-population rate error max %.2e over the %d populations with >= 500 spikes "
          "(mean %.2e over all %d); Poisson z max %.3f; per-neuron count KS p = %.3f from step %d "
          "(%.3f over the whole run)"
          % (label, steps, s64, p64.sum() / pop_size.sum() / sim_s, p32.sum() / pop_size.sum() / sim_s,
             tot, rel[big].max() if big.any() else 0.0, int(big.sum()), rel.mean(), len(rel),
             np.abs(z).max(), ks.pvalue, ks_from, ks_all.pvalue), flush=True)
    assert big.any()
    assert rel[big].max() <= 0.02
    assert np.abs(z).max() <= 3.0
    assert ks.pvalue > 0.05
    return rel, z, ks


def test_c2_fly20k_fp32_tracks_fp64_2s():
    net = W.c2_network(seed=2)
    arr = arrays_from_network(net)
    print("case c2: 20,000 neurons, %d receptor-synapses, 20,000 steps" % net.n_syn, flush=True)
    size = np.array([p.n for p in net.pops], np.int64)
    _compare("c2 fly20k", lambda p: Engine(arr, precision=p), 20000, size)


def test_c2_fly20k_window_equals_two_launch_fp64():
    """11 steps per launch against one update + one delivery launch per step;
    the windowed gather with 8 lanes per row (8 time lanes, one entry
    sub-lane) sums each step's terms in the order of the two-launch gather
    with one lane per row."""
    net = W.c2_network(seed=2)
    arr = arrays_from_network(net)
    print("case c2 window vs two-launch: 20,000 neurons FP64, 3,000 steps", flush=True)
    out = {}
    for tag, sched in (("win", dict(window=64, gather_lanes=8)),
                       ("two", dict(window=0, fused=0, gather_lanes=1))):
        eng = Engine(arr, precision=A.FS_FP64)
        eng.set_schedule(**sched)
        eng.reset(5)
        out[tag] = (eng.raster(3000), eng.state("V"), eng.state("gN"), eng.schedule)
        eng.close()
    assert out["win"][3]["window"] == 11 and out["two"][3]["window"] == 0
    assert out["two"][3]["fused"] == 0
    rw, rt = out["win"][0], out["two"][0]
    print("window %d spikes, two-launch %d spikes, max |dV| %.3g mV"
          % (len(rw), len(rt), float(np.max(np.abs(out["win"][1] - out["two"][1])))), flush=True)
    assert len(rw) > 30000
    np.testing.assert_array_equal(rw, rt)
    np.testing.assert_allclose(out["win"][1], out["two"][1], rtol=0, atol=1e-9)
    np.testing.assert_allclose(out["win"][2], out["two"][2], rtol=1e-10, atol=1e-12)


def test_c3_sparse1m_fp32_tracks_fp64_1s():
    """The network is bistable: it ignites after about 0.4 s, and the two
    precisions cross over a few steps apart, a shift the KS test on a
    million whole-run counts resolves; the spike-count distributions are
    compared over the active half (steps 5,000-10,000), the rates over the
    whole second."""
    net = W.c3_network(seed=3)
    arr = arrays_from_network(net)
    size = np.array([p.n for p in net.pops], np.int64)
    print("case c3: 1,000,000 neurons, %d receptor-synapses, 10,000 steps" % net.n_syn, flush=True)
    del net
    _compare("c3 sparse1m", lambda p: Engine(arr, precision=p), 10000, size, ks_from=5000)


def test_c4_batch_fp32_tracks_fp64():
    spec = c4_spec()
    ne = spec.n_exc()
    size = np.stack([ne, spec.n_neuron - ne], axis=1).reshape(-1).astype(np.int64)
    print("case c4: %d instances, %d neurons, 1,000 steps" % (spec.n_neuron.size, int(spec.n_neuron.sum())),
          flush=True)
    _compare("c4 batch", lambda p: Engine.ei_batch(spec, precision=p), 1000, size)
