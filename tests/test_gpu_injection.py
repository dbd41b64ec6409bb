"""GPU tests of Gaussian current injection (ChangeMembraneNoise,
code/NetSimGnl2.cpp:120-133).

Replay mode draws the reference's own stream: every Igauss call reseeds SHR3
from the next state of one LCG (code/StiExt.h:214-229), so the device jumps
the LCG to each neuron-step's call index and runs Marsaglia and Tsang's
ziggurat there.  Bars: the drop-in command line in FP64 with replayed draws
reproduces the reference rasters of two injection protocols exactly (the
second switches injection on and off for populations in and out of neuron
order); 4000 injected neurons keep the FP64 oracle's potentials to 1e-6 mV
over 450 ms (18 M draws).

The counter-based mode draws with Philox + Box-Muller instead, so its bar is
statistical: after 300 ms of noise-only drive the membrane potentials of 4000
non-spiking neurons have the same distribution as the oracle's (two-sample
KS p > 1e-3, means within 4 standard errors, standard deviations within 5 %).
"""
import ctypes as C
import os
import tempfile

import numpy as np
import pytest

from flysim_amd import _capi as A
from flysim_amd import workloads as W
from flysim_amd.engine import Conf, Engine
from oracle.oracle import Oracle

pytestmark = pytest.mark.gpu

CONF = """NeuralPopulation: Drive
N=4000
C=0.5
Taum=20
Threshold=1000
RestPot=-70
ResetPot=-55
RefractoryPeriod=2
Receptor:AMPA
Tau=2
RevPot=0
FreqExt=0
EndReceptor
EndNeuralPopulation
"""

PRO = """DefineMacro
RandomSeeds: 7 EndRandomSeeds
EndDefineMacro
EventTime 0.0
Type=ChangeMembraneNoise
Population: Drive
GaussMean=0.05
GaussSTD=0.4
EndEvent
EventTime 450.0
Type=EndTrial
EndEvent
OutControl
FileName:Spikes.dat
Type=Spike
population:AllPopulation
EndOutputFile
EndOutControl
"""


def test_injection_membrane_distribution_matches_oracle():
    from scipy.stats import ks_2samp
    steps = 3000
    with tempfile.TemporaryDirectory() as d:
        open(os.path.join(d, "network.conf"), "w").write(CONF)
        open(os.path.join(d, "network.pro"), "w").write(PRO)
        conf = Conf(os.path.join(d, "network.conf"), os.path.join(d, "network.pro"))
        print("case injection: 4000 neurons, %d steps, FP64" % steps, flush=True)
        orc = Oracle(conf.ref_layout(), conf.n_neuron, precision=1)
        orc.reset(conf.seed(0))
        orc.raster(steps)
        vo = orc.state("V")
        eng = Engine(conf, precision=A.FS_FP64)
        eng.set_injection(0, 0.05, 0.4)      # the t = 0 event, as DoEvt applies it
        eng.reset(11)
        r = eng.raster(steps)
        vg = eng.state("V")
    assert len(r) == 0
    se = np.hypot(vo.std(), vg.std()) / np.sqrt(len(vo))
    print("V mean gpu %.4f oracle %.4f (se %.4f), std gpu %.4f oracle %.4f"
          % (vg.mean(), vo.mean(), se, vg.std(), vo.std()))
    assert vo.std() > 0.5                     # the noise is visible
    assert abs(vg.mean() - vo.mean()) < 4 * se
    assert abs(vg.std() / vo.std() - 1) < 0.05
    assert ks_2samp(vg, vo).pvalue > 1e-3


def test_injection_replay_follows_the_reference_stream():
    # 4000 x 4500 = 18 M draws: call indices past 2^24 use every byte of the
    # LCG jump tables
    steps = 4500
    with tempfile.TemporaryDirectory() as d:
        open(os.path.join(d, "network.conf"), "w").write(CONF)
        open(os.path.join(d, "network.pro"), "w").write(PRO)
        conf = Conf(os.path.join(d, "network.conf"), os.path.join(d, "network.pro"))
        bits = conf.ext_trains(steps)
        print("case injection replay: 4000 neurons, %d steps, FP64" % steps, flush=True)
        orc = Oracle(conf.ref_layout(), conf.n_neuron, precision=1)
        orc.reset(conf.seed(0))
        orc.replay(bits, steps)
        eng = Engine(conf, precision=A.FS_FP64, ext_mode=A.FS_EXT_REPLAY)
        eng.reset(conf.seed(0))
        eng.load_ext_trains(bits, 0, steps)
        eng.set_injection(0, 0.05, 0.4)      # the t = 0 event, as DoEvt applies it
        for _ in range(3):
            assert len(eng.raster(steps // 3)) == 0
            orc.raster(steps // 3)
            vg, vo = eng.state("V"), orc.state("V")
            assert vo.std() > 0.5
            np.testing.assert_allclose(vg, vo, rtol=0, atol=1e-6)


@pytest.mark.parametrize("case,steps", [("noise", 1200), ("noise2", 3000)])
def test_cli_injection_fp64_replay_exact_raster(golden, case, steps, tmp_path):
    files = golden["files"]
    (tmp_path / "network.conf").write_text(files["module_network.conf"])
    (tmp_path / "network.pro").write_text(files[case + "_network.pro"])
    print("case injection %s: drop-in CLI, fp64 replay, %d steps" % (case, steps), flush=True)
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
    ref = golden["runs"][case + "_raster"].astype(np.int64)
    assert len(ref) > 300
    np.testing.assert_array_equal(r, ref)


def test_cli_seizure_protocol_form_1k(golden, tmp_path):
    """The fly-brain seizure protocol's form (one ChangeMembraneNoise per
    single-neuron population at 1 ms, GaussSTD = 6 x GaussMean; Spike and
    summed FiringRate outputs, window 50 ms printed every 1 ms) on the 1k E/I
    network, through the drop-in CLI in FP64 with replayed draws.  The
    network ignites at ~130 ms into ~300 Hz hyperactivity.  The raster equals
    the reference's for the first 200 ms and the FiringRate lines of those
    200 ms are identical; about 75 ms into the hyperactive phase FP64 and the
    reference's float arithmetic part ways (the FP64 oracle does the same), so
    over 300 ms the spike count is held to 1 %."""
    files = golden["files"]
    net = W.c1_network(seed=4, g_ext=0.8, t_end_ms=300.0)
    W.write_conf(net, str(tmp_path / "network.conf"))
    (tmp_path / "network.pro").write_text(files["seizure1k_network.pro"])
    print("case seizure form: 1000 populations, 1000 noise events, drop-in CLI, fp64 replay, 3000 steps",
          flush=True)
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
    ref = golden["runs"]["seizure1k_raster"].astype(np.int64)
    np.testing.assert_array_equal(r[r[:, 0] < 2000], ref[ref[:, 0] < 2000])
    assert (ref[:, 0] < 1200).sum() < 400 and (ref[:, 0] >= 2000).sum() > 25000  # quiet, then ignited
    print("spikes", len(r), "reference", len(ref))
    assert abs(len(r) / len(ref) - 1) < 0.01
    got = (tmp_path / "FRates.dat").read_text().split("\n")
    want = files["seizure1k_FRates.dat"].split("\n")
    assert len(got) == len(want)
    assert got[:200] == want[:200]
