"""GPU tests of replicated handles (parameter sweeps over one network).

A handle with B replicas holds B copies of one network that share its synapse
tables, with per-copy recurrent gains, external efficacy gain, external rates
and random-stream keys.  Bar: every copy reproduces, spike for spike, the
same network run alone with the same knobs (FP64, counter-based input keyed
per copy), on the one-launch (delayed synapses) and the two-launch
(zero-delay synapses) step paths.
"""
import numpy as np
import pytest

from flysim_amd import _capi as A
from flysim_amd import workloads as W
from flysim_amd.engine import Engine, arrays_from_network

pytestmark = pytest.mark.gpu

STEPS = 1500
KNOBS = [  # (gains by receptor id A, G, N, H, L), ext gain, stream key
    ((1.0, 1.0, 1.0, 1.0, 1.0), 1.0, 0),
    ((1.2, 0.8, 1.5, 1.0, 1.0), 1.05, 12345),
    ((0.7, 1.6, 0.5, 1.0, 1.0), 0.95, 0x9E3779B97F4A7C15),
    ((1.0, 1.0, 2.0, 1.0, 1.0), 1.1, 77),
]
EXT_CHANGE = (2, 0, W.AMPA, 3100.0)   # replica 2, population 0, AMPA -> 3.1 kHz


def _nets():
    delayed = W.c2_network(seed=8, n=1500, n_mod=6, mod_edges=10, t_end_ms=100)
    plain = W.c1_network(seed=9, n_exc=200, n_inh=50, g_ext=1.3, t_end_ms=100)
    return {"delayed": delayed, "zero_delay": plain}


# step paths: the default (windowed for delayed synapses, two launches for
# zero delays) and the alternative (fused one launch per step; windowed W = 1)
PATHS = {("delayed", "default"): {}, ("delayed", "alt"): {"window": 0, "fused": 1},
         ("zero_delay", "default"): {}, ("zero_delay", "alt"): {"window": 1}}


@pytest.mark.parametrize("kind,path", sorted(PATHS))
def test_replicas_equal_single_runs(kind, path):
    net = _nets()[kind]
    arr = arrays_from_network(net)
    B, P0, n0 = len(KNOBS), len(net.pops), net.n
    sched = PATHS[(kind, path)]
    print("case replicas %s/%s: %d copies of %d neurons, %d steps" % (kind, path, B, n0, STEPS),
          flush=True)
    singles = []
    for r, (g, eg, key) in enumerate(KNOBS):
        e = Engine(arr, precision=A.FS_FP64)
        e.set_replica(0, gain=g, ext_gain=eg, stream_key=key)
        if r == EXT_CHANGE[0]:
            e.set_ext_rate(EXT_CHANGE[1], EXT_CHANGE[2], EXT_CHANGE[3])
        e.reset(31)
        singles.append((e.raster(STEPS), e.pop_counts(), e.state("V")))
        e.close()
    eb = Engine(arr, precision=A.FS_FP64, replicas=B)
    if sched:
        eb.set_schedule(**sched)
    for r, (g, eg, key) in enumerate(KNOBS):
        eb.set_replica(r, gain=g, ext_gain=eg, stream_key=key)
    eb.set_ext_rate(EXT_CHANGE[0] * P0 + EXT_CHANGE[1], EXT_CHANGE[2], EXT_CHANGE[3])
    eb.reset(31)
    rb = eb.raster(STEPS)
    counts = eb.pop_counts().reshape(B, P0)
    vb = eb.state("V").reshape(n0, B)
    s = eb.schedule
    expect = {("delayed", "default"): (0, 11), ("delayed", "alt"): (1, 0),
              ("zero_delay", "default"): (0, 0), ("zero_delay", "alt"): (0, 1)}[(kind, path)]
    assert (s["fused"], s["window"]) == expect
    for r in range(B):
        mine = rb[rb[:, 1] % B == r]
        mine = np.stack([mine[:, 0], mine[:, 1] // B], axis=1)
        ref, ref_counts, ref_v = singles[r]
        assert len(ref) > 50
        np.testing.assert_array_equal(mine, ref)
        np.testing.assert_array_equal(counts[r], ref_counts)
        np.testing.assert_allclose(vb[:, r], ref_v, rtol=0, atol=1e-9)
    # the knobs matter: copies 0 and 1 differ
    assert len(singles[0][0]) != len(singles[1][0])


def test_replica_knobs_validate():
    net = _nets()["zero_delay"]
    eb = Engine(arrays_from_network(net), precision=A.FS_FP32, replicas=3)
    with pytest.raises(A.FlysimError):
        eb.set_replica(3)
    with pytest.raises(A.FlysimError):
        eb.set_ext_rate(3 * len(net.pops), W.AMPA, 100.0)


@pytest.mark.parametrize("kind", ["delayed", "zero_delay"])
def test_whole_wave_replicas_equal_single_runs(kind):
    """64 copies (one local neuron per wave: scalar-load update and gather,
    grouped push) equal the same network run alone, copy by copy (FP64)."""
    net = _nets()[kind]
    arr = arrays_from_network(net)
    B, n0 = 64, net.n
    knobs = [((0.8 + 0.01 * r, 1.3 - 0.01 * r, 0.7 + 0.02 * r, 1.0, 1.0), 0.9 + 0.004 * r, 7 * r + 1)
             for r in range(B)]
    print("case whole-wave replicas %s: %d copies of %d neurons, %d steps" % (kind, B, n0, STEPS),
          flush=True)
    eb = Engine(arr, precision=A.FS_FP64, replicas=B)
    eb.set_schedule(gather_lanes=1)
    for r, (g, eg, key) in enumerate(knobs):
        eb.set_replica(r, gain=g, ext_gain=eg, stream_key=key)
    eb.reset(5)
    rb = eb.raster(STEPS)
    total = 0
    for r in (0, 21, 42, 63):
        g, eg, key = knobs[r]
        e = Engine(arr, precision=A.FS_FP64)
        e.set_schedule(gather_lanes=1)
        e.set_replica(0, gain=g, ext_gain=eg, stream_key=key)
        e.reset(5)
        ref = e.raster(STEPS)
        e.close()
        mine = rb[rb[:, 1] % B == r]
        mine = np.stack([mine[:, 0], mine[:, 1] // B], axis=1)
        assert len(ref) > 20
        np.testing.assert_array_equal(mine, ref)
        total += len(ref)
    assert total > 200
