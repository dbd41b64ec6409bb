"""GPU tests of device-generated batches of uneven E/I instances
(fs_create_ei_batch).

Bars: the device generator builds exactly the networks of its numpy
restatement (flysim_amd.batch.ei_batch_network) -- same raster in FP64 with
replayed external trains through the host path (fs_create); an instance in a
batch runs spike for spike as the same instance alone (per-instance input
streams, FP64); the realised pair count matches Bernoulli(p_conn) within 5
standard deviations.
"""
import numpy as np
import pytest

from flysim_amd import _capi as A
from flysim_amd.batch import c1_populations, ei_batch_network
from flysim_amd.engine import EiBatchSpec, Engine, arrays_from_network

pytestmark = pytest.mark.gpu


def _spec(sizes, seeds, ext=(2400.0,), gn=(1.0,), gg=(1.0,), delay=0, spike_delay_ms=1.0):
    e, i = c1_populations(g_ext=1.3, spike_delay_ms=spike_delay_ms)
    k = len(sizes)
    return EiBatchSpec(np.array(sizes), np.array(seeds), np.resize(ext, k), np.resize(gn, k),
                       np.resize(gg, k), e, i, delay_steps=delay)


def _bits(n, steps, seed):
    rng = np.random.default_rng(seed)
    words = (n + 31) // 32
    x = np.zeros((steps, 5, words), np.uint32)
    a = rng.integers(0, 2**32, size=(steps, words), dtype=np.uint64).astype(np.uint32)
    b = rng.integers(0, 2**32, size=(steps, words), dtype=np.uint64).astype(np.uint32)
    x[:, 0, :] = a & b          # ~25 % AMPA events per step
    if n % 32:
        x[:, 0, -1] &= (1 << (n % 32)) - 1
    return x.reshape(-1)


@pytest.mark.parametrize("delay", [0, 10])
def test_device_generator_equals_host_restatement(delay):
    spec = _spec([150, 333, 271], [11, 12, 13], ext=(2400.0, 1800.0, 3000.0), gn=(1.0, 1.7, 0.6),
                 gg=(1.0, 0.5, 1.9), delay=delay, spike_delay_ms=0.0 if delay else 1.0)
    steps = 600
    print("case ei batch: device generator vs numpy restatement, 3 instances, delay %d, %d steps"
          % (delay, steps), flush=True)
    net = ei_batch_network(spec)
    bits = _bits(net.n, steps, 5)
    rasters, ctrs = [], []
    for make in (lambda: Engine.ei_batch(spec, precision=A.FS_FP64, ext_mode=A.FS_EXT_REPLAY),
                 lambda: Engine(arrays_from_network(net), precision=A.FS_FP64,
                                ext_mode=A.FS_EXT_REPLAY)):
        eng = make()
        eng.reset(3)
        eng.load_ext_trains(bits, 0, steps)
        rasters.append(eng.raster(steps))
        ctrs.append(eng.counters())
    assert len(rasters[0]) > 100
    np.testing.assert_array_equal(rasters[0], rasters[1])
    assert ctrs[0]["syn_events"] == ctrs[1]["syn_events"]
    assert ctrs[0]["nmda_terms"] == ctrs[1]["nmda_terms"]


def test_instance_in_batch_runs_as_alone():
    spec = _spec([400, 650, 300], [21, 22, 23], ext=(2600.0, 2200.0, 3100.0), gn=(1.2, 0.8, 1.5),
                 gg=(0.9, 1.4, 0.7))
    alone = _spec([650], [22], ext=(2200.0,), gn=(0.8,), gg=(1.4,))
    steps = 1500
    print("case ei batch: instance 1 of 3 vs alone, %d steps" % steps, flush=True)
    eb = Engine.ei_batch(spec, precision=A.FS_FP64)
    eb.reset(9)
    rb = eb.raster(steps)
    ea = Engine.ei_batch(alone, precision=A.FS_FP64)
    ea.reset(9)
    ra = ea.raster(steps)
    mine = rb[(rb[:, 1] >= 400) & (rb[:, 1] < 1050)] - np.array([0, 400])
    assert len(ra) > 50
    np.testing.assert_array_equal(mine, ra)
    np.testing.assert_array_equal(eb.pop_counts()[2:4], ea.pop_counts())


def test_generated_pair_count_is_bernoulli():
    spec = _spec([3000, 1200], [31, 32])
    eng = Engine.ei_batch(spec, precision=A.FS_FP32, record=False)
    eng.reset(1)
    eng.run(10, record=False)
    nmda = eng.counters()["nmda_terms"] / 10      # one NMDA synapse per E->X pair
    ne = spec.n_exc()
    n = spec.n_neuron.astype(np.float64)
    trials = float(np.sum(ne * (n - 1)))
    mean, sd = trials * 0.1, np.sqrt(trials * 0.1 * 0.9)
    print("E pairs", nmda, "expected", mean, "+-", sd)
    assert abs(nmda - mean) < 5 * sd


def test_synaptic_delay_form_equals_spike_delay_form():
    """SpikeDly 1 ms with no synaptic delay (the reference form: the spike
    flag follows the threshold crossing by 10 steps, code/EqsGnl2.cpp:229-251)
    and SpikeDly 0 with a 10-step synaptic delay (the sweep's form, run by
    the windowed path) give every neuron the same conductances: the
    membrane potentials agree to rounding at every step (pushes that reach
    one target in one step are added by float atomics in either order) and
    the rasters differ by the 10-step shift of the reported spike times
    (FP64, same input stream, the same gather_lanes on both paths)."""
    sizes, seeds = [420, 700, 310], [41, 42, 43]
    kw = dict(ext=(2600.0, 2300.0, 3300.0), gn=(1.3, 0.7, 1.6), gg=(0.8, 1.5, 1.1))
    ref = _spec(sizes, seeds, **kw)
    fold = _spec(sizes, seeds, delay=10, spike_delay_ms=0.0, **kw)
    steps, chunk = 2000, 250
    print("case ei batch: SpikeDly 1 ms vs 10-step synaptic delay, 3 instances, %d steps" % steps,
          flush=True)
    ea = Engine.ei_batch(ref, precision=A.FS_FP64)
    eb = Engine.ei_batch(fold, precision=A.FS_FP64)
    for e in (ea, eb):
        e.set_schedule(gather_lanes=8, update_block=256)
        e.reset(5)
    assert ea.schedule["window"] == 0 and eb.schedule["window"] == 11
    ra, rb, worst = [], [], 0.0
    for _ in range(steps // chunk):
        ra.append(ea.raster(chunk))
        rb.append(eb.raster(chunk))
        worst = max(worst, float(np.max(np.abs(ea.state("V") - eb.state("V")))))
    ra, rb = np.concatenate(ra), np.concatenate(rb)
    print("spikes %d / %d, max |dV| %.3g mV" % (len(ra), len(rb), worst))
    assert len(ra) > 500
    assert worst < 1e-9
    shifted = rb[rb[:, 0] < steps - 10] + np.array([10, 0])
    np.testing.assert_array_equal(ra[ra[:, 0] >= 10], shifted)
