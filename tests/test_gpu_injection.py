"""GPU test of Gaussian current injection (ChangeMembraneNoise,
code/NetSimGnl2.cpp:120-133).  The engine draws the Gaussian with Philox +
Box-Muller where the reference draws a sequential Marsaglia-Tsang ziggurat
stream (restated exactly by the oracle, pinned against the reference in
tests/test_oracle.py), so the bar is statistical: after 300 ms of noise-only
drive the membrane potentials of 4000 non-spiking neurons have the same
distribution as the oracle's (two-sample KS p > 1e-3, means within 4 standard
errors, standard deviations within 5 %)."""
import os
import tempfile

import numpy as np
import pytest

from flysim_amd import _capi as A
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
EventTime 300.0
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
