"""CPU checks of the batched E/I generator's numpy restatement
(flysim_amd.batch): splitmix64 known answers, no self-pairs, pair density and
efficacy range, determinism and the 256-instance sweep's parameter ranges."""
import numpy as np

from flysim_amd.batch import c1_populations, c4_spec, ei_batch_network, mix64
from flysim_amd.engine import EiBatchSpec
from flysim_amd import workloads as W


def test_splitmix64_known_answers():
    # splitmix64 state 0 -> first outputs (Vigna's reference generator:
    # x += golden; z = mix(x)); mix64(k * golden) is the k-th output
    g = np.uint64(0x9E3779B97F4A7C15)
    outs = mix64(np.arange(3, dtype=np.uint64) * g)
    assert int(outs[0]) == 0xE220A8397B1DCDAF
    assert int(outs[1]) == 0x6E789E6AA1B965F4
    assert int(outs[2]) == 0x06C45D188009454F


def test_restated_generator_properties():
    e, i = c1_populations()
    spec = EiBatchSpec(np.array([400, 250]), np.array([5, 6]), [2000.0, 3000.0], [1.5, 1.0],
                       [0.5, 2.0], e, i)
    net = ei_batch_network(spec)
    assert net.n == 650 and len(net.pops) == 4
    assert not np.any(net.pre == net.post)
    inst = lambda x: (x >= 400).astype(int)
    assert np.all(inst(net.pre) == inst(net.post))       # no cross-instance pairs
    pairs = np.sum(net.receptor != W.NMDA)
    n = np.array([400.0, 250.0])
    mean = 0.1 * np.sum(n * (n - 1))
    assert abs(pairs - mean) < 5 * np.sqrt(mean * 0.9)
    amp = net.w[net.receptor == W.AMPA]
    assert amp.min() >= 0.5 and amp.max() < 1.5 and abs(amp.mean() - 1.0) < 0.01
    # NMDA = g * 0.1 * gain_nmda of the same pair, GABA = g * gain_gaba
    a = net.receptor == W.AMPA
    m = net.receptor == W.NMDA
    np.testing.assert_array_equal(net.pre[a], net.pre[m])
    np.testing.assert_array_equal(net.post[a], net.post[m])
    gain = np.where(net.pre[m] < 400, np.float32(1.5), np.float32(1.0))
    np.testing.assert_array_equal(net.w[m], (net.w[a] * np.float32(0.1)).astype(np.float32) * gain)
    again = ei_batch_network(spec)
    np.testing.assert_array_equal(again.w, net.w)


def test_c4_spec_ranges():
    s = c4_spec()
    assert s.n_neuron.size == 256
    assert s.n_neuron.min() >= 1000 and s.n_neuron.max() <= 20000
    assert np.all(s.seed == 1000 + np.arange(256))
    assert s.ext_hz.min() >= 500 and s.ext_hz.max() <= 4000
    for g in (s.gain_nmda, s.gain_gaba):
        assert g.min() >= 0.5 and g.max() <= 2.0
    pairs = 0.1 * np.sum(s.n_neuron.astype(np.float64) * (s.n_neuron - 1))
    assert 3.0e9 < pairs < 4.2e9


def test_c4_states_its_delay_as_synaptic_delay():
    s = c4_spec(n_inst=3, n_lo=100, n_hi=200)
    assert s.delay_steps == 10 and s.exc.spike_delay_ms == 0.0 and s.inh.spike_delay_ms == 0.0
    net = ei_batch_network(s)
    assert np.all(net.delay == 10)
    r = c4_spec(n_inst=3, n_lo=100, n_hi=200, synaptic_delay=False)
    assert r.delay_steps == 0 and r.exc.spike_delay_ms == 1.0
    # same pairs and efficacies either way
    np.testing.assert_array_equal(ei_batch_network(r).w, net.w)
    e, i = c1_populations()
    for bad in (-1, 64):
        try:
            EiBatchSpec(np.array([10]), np.array([1]), [1.0], [1.0], [1.0], e, i, delay_steps=bad)
        except ValueError:
            continue
        raise AssertionError("delay_steps %d accepted" % bad)
