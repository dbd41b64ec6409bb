"""GPU parity tests of short-term plasticity (-STP, -STD, and -udfsed, which
switches STD on in the reference), through the C ABI.

Reference data: golden rasters of the reference simulator built from source
with those switches (tests/golden/make_golden.py) and the CPU restatement
(oracle/).  Bars: the drop-in command line in FP64 with the reference's
external trains replayed reproduces the reference raster exactly over the
300 ms module protocol (ChangeExtFreq events included); the presynaptic
depression / facilitation state follows the FP64 oracle to 1e-9.
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

CASES = {"stp": (["-STP"], dict(stp=True)), "std": (["-STD"], dict(std=True)),
         "udf": (["-udfsed", "4242"], dict(udf_seed=4242))}


@pytest.mark.parametrize("case", sorted(CASES))
def test_cli_plasticity_fp64_replay_exact_raster(golden, case, tmp_path):
    files = golden["files"]
    (tmp_path / "network.conf").write_text(files[case + "_network.conf"])
    (tmp_path / "network.pro").write_text(files["module_network.pro"])
    print("case plasticity %s: drop-in CLI, fp64 replay, 3000 steps" % case, flush=True)
    cwd = os.getcwd()
    os.chdir(tmp_path)
    try:
        argv = [b"flysim_amd", b"-fp64", b"-ext", b"replay", b"-t", b"1"]
        argv += [a.encode() for a in CASES[case][0]]
        arr = (C.c_char_p * len(argv))(*argv)
        A.check(A.lib().fs_cli_main(len(argv), arr))
    finally:
        os.chdir(cwd)
    s = W.read_spikes(str(tmp_path / "Spikes.dat"))
    r = np.stack([np.round(s[:, 0] / 1e-4).astype(np.int64), s[:, 1].astype(np.int64)], 1)
    ref = golden["runs"][case + "_raster"].astype(np.int64)
    assert len(ref) > 300
    np.testing.assert_array_equal(r, ref)


@pytest.mark.parametrize("case", ["stp", "std"])
def test_plasticity_state_tracks_fp64_oracle(golden, case):
    """D (and F under STP) of every neuron after 90 ms, before the first
    protocol event, against the FP64 oracle fed the same trains."""
    files = golden["files"]
    steps = 900
    with tempfile.TemporaryDirectory() as d:
        open(os.path.join(d, "network.conf"), "w").write(files[case + "_network.conf"])
        open(os.path.join(d, "network.pro"), "w").write(files["module_network.pro"])
        conf = Conf(os.path.join(d, "network.conf"), os.path.join(d, "network.pro"), **CASES[case][1])
        bits = conf.ext_trains(steps)
        print("case plasticity state %s: fp64 replay, %d steps" % (case, steps), flush=True)
        eng = Engine(conf, precision=A.FS_FP64, ext_mode=A.FS_EXT_REPLAY)
        eng.reset(conf.seed(0))
        eng.load_ext_trains(bits, 0, steps)
        orc = Oracle(conf.ref_layout(), conf.n_neuron, precision=1)
        orc.reset(conf.seed(0))
        orc.replay(bits, steps)
        rg, ro = eng.raster(steps), orc.raster(steps)
        np.testing.assert_array_equal(rg, ro)
        assert len(ro) > 15
        dg, do = eng.state("stp_D"), orc.state("stp_D")
        assert do.min() < 0.9
        np.testing.assert_allclose(dg, do, rtol=0, atol=1e-9)
        if case == "stp":
            np.testing.assert_allclose(eng.state("stp_F"), orc.state("stp_F"), rtol=0, atol=1e-9)
        np.testing.assert_allclose(eng.state("V"), orc.state("V"), rtol=0, atol=1e-8)


def test_cli_ltp_fp64_replay_exact_raster(golden, tmp_path):
    """-LTP: single-synapse weights learn from delayed LTP spikes and LP
    traces; the drop-in CLI in FP64 replay reproduces the reference raster."""
    files = golden["files"]
    (tmp_path / "network.conf").write_text(files["ltp_network.conf"])
    (tmp_path / "network.pro").write_text(files["module_network.pro"])
    print("case plasticity ltp: drop-in CLI, fp64 replay, 3000 steps", flush=True)
    cwd = os.getcwd()
    os.chdir(tmp_path)
    try:
        argv = [b"flysim_amd", b"-fp64", b"-ext", b"replay", b"-t", b"1", b"-LTP"]
        arr = (C.c_char_p * len(argv))(*argv)
        A.check(A.lib().fs_cli_main(len(argv), arr))
    finally:
        os.chdir(cwd)
    s = W.read_spikes(str(tmp_path / "Spikes.dat"))
    r = np.stack([np.round(s[:, 0] / 1e-4).astype(np.int64), s[:, 1].astype(np.int64)], 1)
    ref = golden["runs"]["ltp_raster"].astype(np.int64)
    assert len(ref) > 1000
    np.testing.assert_array_equal(r, ref)


def test_ltp_state_tracks_fp64_oracle(golden):
    files = golden["files"]
    steps = 900
    with tempfile.TemporaryDirectory() as d:
        open(os.path.join(d, "network.conf"), "w").write(files["ltp_network.conf"])
        open(os.path.join(d, "network.pro"), "w").write(files["module_network.pro"])
        conf = Conf(os.path.join(d, "network.conf"), os.path.join(d, "network.pro"), ltp=True)
        bits = conf.ext_trains(steps)
        print("case ltp state: fp64 replay, %d steps" % steps, flush=True)
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


def _c1_stp_conf(d):
    net = W.c1_network(seed=1, g_ext=1.25, t_end_ms=1000.0)
    W.write_conf(net, os.path.join(d, "network.conf"))
    W.write_pro(os.path.join(d, "network.pro"), 1000.0, seeds=[1])
    return Conf(os.path.join(d, "network.conf"), os.path.join(d, "network.pro"), stp=True)


def test_stp_1k_fp64_exact_200ms_and_fp32_rates(golden):
    """-STP on the 1k E/I network.  FP64 replay reproduces the FP64 oracle's
    raster exactly for 200 ms and the reference's up to one event: the
    reference's float arithmetic fires neuron 818 at step 905, FP64 at step
    906 (the FP32 oracle reproduces the reference's 200 ms exactly).  Rates
    over 1 s: FP64 within 1 % of the reference, FP32 within 2 %, and the
    per-neuron spike counts of FP32 within a KS p > 0.05."""
    from scipy.stats import ks_2samp
    ref = golden["runs"]["c1_stp_raster"].astype(np.int64)
    ref200 = ref[ref[:, 0] < 2000]
    with tempfile.TemporaryDirectory() as d:
        conf = _c1_stp_conf(d)
        bits = conf.ext_trains(10000)
        print("case c1 stp: fp64 replay 10000 steps, fp32 replay 10000 steps", flush=True)
        e64 = Engine(conf, precision=A.FS_FP64, ext_mode=A.FS_EXT_REPLAY)
        e64.reset(conf.seed(0))
        e64.load_ext_trains(bits, 0, 10000)
        r64 = e64.raster(10000)
        orc = Oracle(conf.ref_layout(), conf.n_neuron, precision=1)
        orc.reset(conf.seed(0))
        orc.replay(bits, 2000)
        g200 = r64[r64[:, 0] < 2000]
        np.testing.assert_array_equal(g200, orc.raster(2000))
        assert g200.shape == ref200.shape
        bad = np.nonzero((g200 != ref200).any(1))[0]
        assert bad.size <= 1
        for i in bad:
            assert tuple(ref200[i]) == (905, 818) and tuple(g200[i]) == (906, 818)
        e32 = Engine(conf, precision=A.FS_FP32, ext_mode=A.FS_EXT_REPLAY)
        e32.reset(conf.seed(0))
        e32.load_ext_trains(bits, 0, 10000)
        r32 = e32.raster(10000)
    rates = lambda r: np.array([(r[:, 1] < 800).sum() / 800.0, (r[:, 1] >= 800).sum() / 200.0])
    print("fp64 rates", rates(r64), "fp32 rates", rates(r32), "reference", rates(ref))
    np.testing.assert_allclose(rates(r64), rates(ref), rtol=0.01)
    np.testing.assert_allclose(rates(r32), rates(ref), rtol=0.02)
    p = ks_2samp(np.bincount(r32[:, 1], minlength=1000), np.bincount(ref[:, 1], minlength=1000)).pvalue
    print("KS p-value", p)
    assert p > 0.05


def test_plasticity_state_absent_without_switch(golden):
    files = golden["files"]
    with tempfile.TemporaryDirectory() as d:
        open(os.path.join(d, "network.conf"), "w").write(files["stp_network.conf"])
        open(os.path.join(d, "network.pro"), "w").write(files["module_network.pro"])
        conf = Conf(os.path.join(d, "network.conf"), os.path.join(d, "network.pro"))
        eng = Engine(conf, precision=A.FS_FP32)
        eng.reset(1)
        with pytest.raises(A.FlysimError):
            eng.state("stp_D")
