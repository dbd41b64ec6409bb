"""GPU parity tests of gap junctions (electrical synapses of the
single-synapse path, ThdGapSyn), through the C ABI.

Reference data: the golden raster of the reference built from source on the
module network with one-target gap links (tests/golden/make_golden.py,
gap_case) and the CPU restatement (oracle/).  Bars: the drop-in command line
in FP64 with the reference's external draws replayed reproduces the
reference raster exactly over 300 ms (ChangeExtFreq events included); the
membrane potentials follow the FP64 oracle to 1e-8 mV.
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


def test_cli_gap_junctions_fp64_replay_exact_raster(golden, tmp_path):
    files = golden["files"]
    (tmp_path / "network.conf").write_text(files["gap_network.conf"])
    (tmp_path / "network.pro").write_text(files["module_network.pro"])
    print("case gap junctions: drop-in CLI, fp64 replay, 3000 steps", flush=True)
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
    ref = golden["runs"]["gap_raster"].astype(np.int64)
    assert len(ref) > 1000
    np.testing.assert_array_equal(r, ref)


def test_gap_junction_state_tracks_fp64_oracle(golden):
    files = golden["files"]
    steps = 900
    with tempfile.TemporaryDirectory() as d:
        open(os.path.join(d, "network.conf"), "w").write(files["gap_network.conf"])
        open(os.path.join(d, "network.pro"), "w").write(files["module_network.pro"])
        conf = Conf(os.path.join(d, "network.conf"), os.path.join(d, "network.pro"))
        bits = conf.ext_trains(steps)
        print("case gap state: fp64 replay, %d steps" % steps, flush=True)
        eng = Engine(conf, precision=A.FS_FP64, ext_mode=A.FS_EXT_REPLAY)
        assert eng.schedule["window"] == 0 and eng.schedule["fused"] == 0
        eng.reset(conf.seed(0))
        eng.load_ext_trains(bits, 0, steps)
        orc = Oracle(conf.ref_layout(), conf.n_neuron, precision=1)
        orc.reset(conf.seed(0))
        orc.replay(bits, steps)
        for _ in range(steps // 100):
            rg, ro = eng.raster(100), orc.raster(100)
            np.testing.assert_array_equal(rg, ro)
            np.testing.assert_allclose(eng.state("V"), orc.state("V"), rtol=0, atol=1e-8)
        assert int(eng.pop_counts().sum()) > 100
