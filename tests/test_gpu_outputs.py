"""GPU test of the per-step output types of the drop-in command line:
DeltaMemPotLimit (plain and Align), IsynSeparatedSum, GatingVariable and
MemPot, against the files the reference built from source wrote for the same
network and protocol (tests/golden/make_golden.py, outputs_case).

Run in FP64 with the reference's external draws replayed.  Bar: the spike
and dV-limit files line for line (times and neuron ids are exact); the
printed currents, gates and potentials to 1e-4 relative (the reference
computes them in float and prints 6 significant digits).
"""
import ctypes as C
import os

import numpy as np
import pytest

from flysim_amd import _capi as A

pytestmark = pytest.mark.gpu

SAMPLE = {"Spikes.dat": 1, "dV.dat": 10, "dVal.dat": 10, "Isyn.dat": 30, "Gate.dat": 30,
          "Vm.dat": 30}


def _sampled(text, every):
    lines = text.split("\n")[:-1]
    return lines[::every]


def test_cli_per_step_outputs_match_reference(golden, tmp_path):
    files = golden["files"]
    (tmp_path / "network.conf").write_text(files["outputs_network.conf"])
    (tmp_path / "network.pro").write_text(files["outputs_network.pro"])
    print("case outputs: drop-in CLI, fp64 replay, 1500 steps, 6 output files", flush=True)
    cwd = os.getcwd()
    os.chdir(tmp_path)
    try:
        argv = [b"flysim_amd", b"-fp64", b"-ext", b"replay", b"-t", b"1"]
        arr = (C.c_char_p * len(argv))(*argv)
        A.check(A.lib().fs_cli_main(len(argv), arr))
    finally:
        os.chdir(cwd)
    meta = golden["meta"]["outputs"]
    for fn, every in SAMPLE.items():
        got_all = (tmp_path / fn).read_text()
        assert got_all.count("\n") == meta[fn]["lines"], fn
        got = _sampled(got_all, every)
        ref = _sampled(files["outputs_" + fn], 1)
        assert len(got) == len(ref), fn
        if fn in ("Spikes.dat", "dV.dat", "dVal.dat"):
            assert got == ref, fn
            continue
        for g, r in zip(got, ref):
            gv = np.array([float(x) for x in g.split()])
            rv = np.array([float(x) for x in r.split()])
            assert gv.shape == rv.shape, fn
            scale = np.max(np.abs(rv)) if rv.size else 0.0
            np.testing.assert_allclose(gv, rv, rtol=1e-4, atol=1e-5 * scale, err_msg=fn)
    # the drop-in CLI also writes the reference's ConfsInfo.log and network.log
    import hashlib
    for fn in ["ConfsInfo.log", "network.log"]:
        digest = hashlib.sha256((tmp_path / fn).read_bytes()).hexdigest()
        assert digest == golden["meta"]["logs"]["outputs"][fn]["sha256"], fn
    # the gates file carries values (neuron 3 spikes within the run)
    gates = _sampled(files["outputs_Gate.dat"], 1)
    assert any(float(x) != 0.0 for line in gates for x in line.split()[1:])
