"""GPU parity tests of the windowed (temporal-blocking) step path, through the
C ABI: W consecutive steps per launch, W <= 1 + the smallest synaptic delay,
each block pushing its own spikes.  Bar: FP64 results identical to the fused
one-launch-per-step path (delayed synapses) and to the two-launch path
(zero delays, W = 1), including replicas, event counters and the golden
raster of the reference built from source."""
import numpy as np
import pytest

from flysim_amd import _capi as A
from flysim_amd import workloads as W
from flysim_amd.engine import Engine, arrays_from_network

pytestmark = pytest.mark.gpu


def _run(arr, sched, steps, seed=7, replicas=1, knobs=None):
    eng = Engine(arr, precision=A.FS_FP64, replicas=replicas)
    eng.set_schedule(**sched)
    for r, kn in enumerate(knobs or []):
        eng.set_replica(r, **kn)
    eng.reset(seed)
    r = eng.raster(steps)
    out = dict(raster=r, V=eng.state("V"), gN=eng.state("gN"), gA=eng.state("gA"),
               counts=eng.pop_counts(), ctr=eng.counters(), sched=eng.schedule)
    eng.close()
    return out


def _same(a, b, what):
    np.testing.assert_array_equal(a["raster"], b["raster"], err_msg=what)
    np.testing.assert_allclose(a["V"], b["V"], rtol=0, atol=1e-9, err_msg=what)
    np.testing.assert_allclose(a["gN"], b["gN"], rtol=1e-10, atol=1e-12, err_msg=what)
    np.testing.assert_allclose(a["gA"], b["gA"], rtol=1e-10, atol=1e-12, err_msg=what)
    np.testing.assert_array_equal(a["counts"], b["counts"], err_msg=what)
    assert a["ctr"]["syn_events"] == b["ctr"]["syn_events"], what
    assert a["ctr"]["nmda_terms"] == b["ctr"]["nmda_terms"], what
    assert a["ctr"]["spikes"] == b["ctr"]["spikes"], what


def test_window_equals_fused_path_delayed_synapses():
    net = W.c2_network(seed=4, n=3000, n_mod=10, mod_edges=20, t_end_ms=100)
    arr = arrays_from_network(net)
    dmin = int(net.delay.min())
    print("case window vs fused: 3000 neurons, delays %d..%d" % (dmin, int(net.delay.max())), flush=True)
    ref = _run(arr, dict(fused=1, window=0, steps_per_graph=50), 1200)
    assert ref["sched"]["fused"] == 1 and len(ref["raster"]) > 100
    for sched in [dict(window=64, steps_per_graph=50), dict(window=4, steps_per_graph=7),
                  dict(window=3, steps_per_graph=50, use_graph=0, gather_lanes=1, update_block=64),
                  dict(window=64, steps_per_graph=1000, gather_lanes=16, update_block=512),
                  dict(window=64, steps_per_graph=100, gather_lanes=64, update_block=128)]:
        got = _run(arr, sched, 1200)
        assert got["sched"]["window"] == min(sched["window"], dmin + 1)
        assert got["sched"]["fused"] == 0
        _same(got, ref, str(sched))


def test_window_rejects_blocks_over_512():
    net = W.c2_network(seed=4, n=600, n_mod=4, mod_edges=4, t_end_ms=100)
    eng = Engine(arrays_from_network(net), precision=A.FS_FP32)
    with pytest.raises(A.FlysimError):
        eng.set_schedule(window=64, update_block=1024)
    assert eng.schedule["update_block"] == 256  # the previous schedule stays


def test_window_one_step_equals_two_launch_path_zero_delay():
    net = W.c1_network(seed=9, n_exc=400, n_inh=100, g_ext=1.3, t_end_ms=100)
    arr = arrays_from_network(net)
    print("case window W=1 vs two-launch: 500 neurons, zero delays", flush=True)
    ref = _run(arr, dict(window=0, steps_per_graph=100), 1500)
    assert ref["sched"]["fused"] == 0 and len(ref["raster"]) > 100
    for sched in [dict(window=8, steps_per_graph=100), dict(window=1, steps_per_graph=9, gather_lanes=4)]:
        got = _run(arr, sched, 1500)
        assert got["sched"]["window"] == 1
        _same(got, ref, str(sched))


@pytest.mark.parametrize("B", [4, 64])
def test_window_replicas_equal_fused_path(B):
    net = W.c2_network(seed=8, n=1500, n_mod=6, mod_edges=10, t_end_ms=100)
    arr = arrays_from_network(net)
    knobs = [dict(gain=(0.8 + 0.01 * r, 1.3 - 0.01 * r, 0.7 + 0.02 * r, 1.0, 1.0),
                  ext_gain=0.9 + 0.004 * r, stream_key=7 * r + 1) for r in range(B)]
    print("case window replicas: %d copies of 1500 neurons" % B, flush=True)
    ref = _run(arr, dict(window=0, fused=1, gather_lanes=1, steps_per_graph=100), 1500,
               replicas=B, knobs=knobs)
    got = _run(arr, dict(window=64, gather_lanes=1, steps_per_graph=100), 1500, replicas=B, knobs=knobs)
    assert got["sched"]["window"] > 1
    assert len(ref["raster"]) > 25 * B
    _same(got, ref, "B=%d" % B)


def test_window_fp64_replay_reproduces_reference_raster(golden, tmp_path):
    """The 1k E/I network (zero delays: W = 1) with the reference's draws
    replayed reproduces the reference raster for 200 ms."""
    import os
    from flysim_amd.engine import Conf
    net = W.c1_network(seed=1, g_ext=1.15, t_end_ms=1000.0)
    W.write_conf(net, os.path.join(tmp_path, "network.conf"))
    W.write_pro(os.path.join(tmp_path, "network.pro"), 1000.0, seeds=[1])
    conf = Conf(os.path.join(tmp_path, "network.conf"), os.path.join(tmp_path, "network.pro"))
    steps = 2000
    bits = conf.ext_trains(steps)
    print("case window replay: c1 seed 1, %d steps" % steps, flush=True)
    eng = Engine(conf, precision=A.FS_FP64, ext_mode=A.FS_EXT_REPLAY)
    eng.set_schedule(window=16)
    eng.reset(conf.seed(0))
    eng.load_ext_trains(bits, 0, steps)
    r = eng.raster(steps)
    ref = golden["runs"]["c1_s1_raster"].astype(np.int64)
    np.testing.assert_array_equal(r, ref[ref[:, 0] < steps])


def test_window_blocks_of_several_populations():
    """600 populations of 1-22 neurons: the windowed blocks fall back to plain
    pieces of 32 neurons spanning several populations (600 population runs
    against 94 plain pieces), so each lane looks its population up instead
    of taking the block's from the block table.  FP64 results identical to
    the fused one-launch-per-step path."""
    net = W.c2_network(seed=6, n=3000, n_mod=300, mu=2.0, sigma=0.4, p_in=0.3, p_out=0.02,
                       mod_edges=900, t_end_ms=100)
    arr = arrays_from_network(net)
    print("case window, mixed-population blocks: %d populations, %d neurons" % (len(net.pops), net.n),
          flush=True)
    ref = _run(arr, dict(fused=1, window=0, steps_per_graph=100), 1500)
    got = _run(arr, dict(window=64, steps_per_graph=110), 1500)
    assert ref["sched"]["fused"] == 1 and got["sched"]["window"] == 11
    assert len(ref["raster"]) > 200
    _same(got, ref, "mixed-population blocks")
