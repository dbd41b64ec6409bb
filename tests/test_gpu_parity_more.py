"""More GPU parity cases through the C ABI, each against the reference built
from source (golden rasters) or the CPU restatement:

* the 1k E/I network with seeds 2 and 3 (seed 3 turns hyperactive) and the
  hyperactive drive 1.18 nS: FP64 replay of the reference's draws reproduces
  the reference raster exactly for 200 ms (250 ms for the hyperactive case),
  and population rates over the whole golden run within 1 %;
* FP32 state against the FP64 CPU restatement on the 1k network for 1 s
  (rates within 2 %, per-neuron spike counts KS p > 0.05);
* the refractory state machine and the dV clamp (code/EqsGnl2.cpp:176-192,
  229-251) step by step against the restatement: SPIKE_F at Count == SpiDy,
  sticky SPIKE_F when SpiDy == RfPed - 1, no spike when SpiDy >= RfPed,
  MEMDV_OVF under a huge excitatory drive and MEMDV_UDF under a huge
  inhibitory one;
* ACh and GluCl receptors (group and single links, external drive and
  ChangeExtFreq events on both): FP64 replay reproduces the reference raster
  exactly over 300 ms.
"""
import os
import tempfile

import numpy as np
import pytest

from flysim_amd import _capi as A
from flysim_amd import workloads as W
from flysim_amd.engine import Conf, Engine
from oracle.oracle import Oracle

pytestmark = pytest.mark.gpu


def _c1(d, seed, g_ext, t_ms=1000.0):
    net = W.c1_network(seed=seed, g_ext=g_ext, t_end_ms=t_ms)
    W.write_conf(net, os.path.join(d, "network.conf"))
    W.write_pro(os.path.join(d, "network.pro"), t_ms, seeds=[seed])
    return Conf(os.path.join(d, "network.conf"), os.path.join(d, "network.pro"))


def _replay(conf, steps, precision=A.FS_FP64):
    bits = conf.ext_trains(steps)
    eng = Engine(conf, precision=precision, ext_mode=A.FS_EXT_REPLAY)
    eng.reset(conf.seed(0))
    eng.load_ext_trains(bits, 0, steps)
    return eng, bits


def _rates(counts, n_exc=800, seconds=1.0):
    return np.array([counts[:n_exc].sum() / n_exc, counts[n_exc:].sum() / (len(counts) - n_exc)]) / seconds


@pytest.mark.parametrize("case,seed,g_ext,exact_steps,full_steps", [
    ("c1_s2", 2, 1.15, 2000, 10000), ("c1_s3", 3, 1.15, 2000, 10000), ("c1_hyper", 1, 1.18, 2500, 3000)])
def test_c1_cases_fp64_replay_exact_and_rates(golden, case, seed, g_ext, exact_steps, full_steps):
    meta = golden["meta"][case]
    ref = golden["runs"][case + "_raster"].astype(np.int64)
    ref_counts = golden["runs"][case + "_counts"].astype(np.int64)
    with tempfile.TemporaryDirectory() as d:
        conf = _c1(d, seed, g_ext, meta["t_ms"])
        print("case %s: seed %d, g_ext %.2f nS, %d reference spikes over %g ms"
              % (case, seed, g_ext, meta["spikes"], meta["t_ms"]), flush=True)
        eng, _ = _replay(conf, full_steps)
        r = eng.raster(exact_steps)
        want = ref[ref[:, 0] < exact_steps]
        assert len(want) > 100
        np.testing.assert_array_equal(r, want)
        rest = eng.raster(full_steps - exact_steps)
        counts = np.bincount(np.concatenate([r[:, 1], rest[:, 1]]), minlength=conf.n_neuron)
        assert int(ref_counts.sum()) == meta["spikes"]
        rg, rr = _rates(counts, seconds=full_steps * 1e-4), _rates(ref_counts, seconds=full_steps * 1e-4)
        print("rates (E, I) GPU FP64 %s, reference %s" % (rg, rr), flush=True)
        np.testing.assert_allclose(rg, rr, rtol=0.01)
        eng.close()


def test_c1_fp32_tracks_fp64_oracle_1s():
    from scipy.stats import ks_2samp
    steps = 10000
    with tempfile.TemporaryDirectory() as d:
        conf = _c1(d, 1, 1.15)
        eng, bits = _replay(conf, steps, A.FS_FP32)
        r32 = eng.raster(steps)
        eng.close()
        orc = Oracle(conf.ref_layout(), conf.n_neuron, precision=1)
        orc.reset(conf.seed(0))
        orc.replay(bits, steps)
        r64 = orc.raster(steps)
    c32 = np.bincount(r32[:, 1], minlength=1000)
    c64 = np.bincount(r64[:, 1], minlength=1000)
    p = ks_2samp(c32, c64).pvalue
    print("1 s: GPU FP32 %d spikes, FP64 oracle %d; rates (E, I) %s vs %s; KS p %.3f"
          % (len(r32), len(r64), _rates(c32), _rates(c64), p), flush=True)
    np.testing.assert_allclose(_rates(c32), _rates(c64), rtol=0.02)
    assert p > 0.05


SINGLE = """NeuralPopulation: A
N=3
C=0.5
Taum=20
Threshold=-50
RestPot=-60
ResetPot=-55
RefractoryPeriod=%s
SpikeDly=%s
Receptor:AMPA
Tau=2
RevPot=0
FreqExt=%s
MeanExtEff=%s
EndReceptor
%sEndNeuralPopulation
"""
INHIB = """Receptor:GABA
Tau=5
RevPot=-90
FreqExt=10000
MeanExtEff=2000
EndReceptor
"""
PRO = "DefineMacro\nRandomSeeds: 1 EndRandomSeeds\nEndDefineMacro\nEventTime 100\nType=EndTrial\nEndEvent\n"


@pytest.mark.parametrize("precision", [A.FS_FP64, A.FS_FP32])
@pytest.mark.parametrize("rfped_ms,spidy_ms,inhib", [
    ("2", "0", False), ("2", "1", False), ("1.9", "1.8", False), ("1", "2", False), ("2", "0", True)])
def test_refractory_machine_and_clamp_step_by_step(precision, rfped_ms, spidy_ms, inhib):
    """Three neurons under an external drive far above threshold (pb = 1,
    200 nS per event), or under a ten times larger inhibitory one: every step
    the GPU's V, refractory count and NState flags equal the restatement's of
    the same precision."""
    steps = 120
    with tempfile.TemporaryDirectory() as d:
        open(os.path.join(d, "network.conf"), "w").write(
            SINGLE % (rfped_ms, spidy_ms, "10000", "200", INHIB if inhib else ""))
        open(os.path.join(d, "network.pro"), "w").write(PRO)
        conf = Conf(os.path.join(d, "network.conf"), os.path.join(d, "network.pro"))
        L = conf.ref_layout()
        print("case RfPed %d SpiDy %d, %s, %s" % (int(L["rfped"][0]), int(L["spidy"][0]),
                                               "inhibited" if inhib else "driven",
                                               "fp64" if precision == A.FS_FP64 else "fp32"), flush=True)
        eng, bits = _replay(conf, steps, precision)
        orc = Oracle(L, conf.n_neuron, precision=1 if precision == A.FS_FP64 else 0)
        orc.reset(conf.seed(0))
        orc.replay(bits, steps)
        tol = 1e-9 if precision == A.FS_FP64 else 1e-3
        seen = 0
        for _ in range(steps):
            np.testing.assert_array_equal(eng.raster(1), orc.raster(1))
            ns = eng.state("nstate")
            np.testing.assert_array_equal(ns, orc.state("nstate"))
            np.testing.assert_array_equal(eng.state("count"), orc.state("count"))
            np.testing.assert_allclose(eng.state("V"), orc.state("V"), rtol=0, atol=tol)
            seen |= int(np.bitwise_or.reduce(ns.astype(np.int64)))
        eng.close()
    rf, sd = int(L["rfped"][0]), int(L["spidy"][0])
    assert seen & (32 if inhib else 16)           # MEMDV_UDF / MEMDV_OVF reached
    if not inhib:
        assert bool(seen & 2) == (sd < rf)        # a spike flag only when SpiDy < RfPed


def test_ach_gcl_fp64_replay_reproduces_reference(golden, tmp_path):
    """The drop-in command line (FP64 state, the reference's draws) on the
    ACh / GluCl network, its ChangeExtFreq events on both receptors
    included, writes the reference's Spike file."""
    import ctypes as C
    f = golden["files"]
    ref = golden["runs"]["achgcl_raster"].astype(np.int64)
    (tmp_path / "network.conf").write_text(f["achgcl_network.conf"])
    (tmp_path / "network.pro").write_text(f["achgcl_network.pro"])
    print("case ACh / GluCl: %d reference spikes in 300 ms" % len(ref), flush=True)
    cwd = os.getcwd()
    os.chdir(tmp_path)
    try:
        argv = [b"flysim_amd", b"-fp64", b"-ext", b"replay", b"-t", b"1"]
        arr = (C.c_char_p * len(argv))(*argv)
        A.check(A.lib().fs_cli_main(len(argv), arr))
    finally:
        os.chdir(cwd)
    s = W.read_spikes(str(tmp_path / "Spikes.dat"))
    r = np.stack([np.round(s[:, 0] / 1e-4).astype(np.int64), s[:, 1].astype(np.int64)], 1)
    np.testing.assert_array_equal(r, ref)
