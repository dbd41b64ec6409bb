"""Regenerate the golden fixtures from the reference simulator built from its
own sources (oracle/ref.mk -> oracle/_ref/flysim_ref).

Runs only where the reference sources exist (this build container).  For
each case it writes the network in the reference's text format, runs the
reference with `-t 1 -rp 1` and keeps only small outputs: spike rasters as
(step, neuron) pairs, synapse orders from network.log, the bytes of small
output files, and digests.  Usage (repo root):  python tests/golden/make_golden.py [plasticity|noise|achgcl|logs [dir]]
"""
from __future__ import annotations

import hashlib
import json
import os
import re
import shutil
import subprocess
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from flysim_amd import workloads as W  # noqa: E402

REF_CODE = "/root/reference/code"
REF_SEIZURE_PRO = "/root/reference/experiment/seizure/network1_900_10nOvf.pro"
REF_BIN = os.path.join(ROOT, "oracle", "_ref", "flysim_ref")


def build_ref():
    subprocess.run(["make", "-f", "oracle/ref.mk", "-j8"], cwd=ROOT, check=True,
                   stdout=subprocess.DEVNULL)


def run_ref(d, args=(), timeout=None, wait_for=None):
    cmd = [REF_BIN, "-conf", "network.conf", "-pro", "network.pro"] + list(args)
    if wait_for is None:
        r = subprocess.run(cmd, cwd=d, capture_output=True, text=True, timeout=timeout)
        return r.stdout
    # run until `wait_for` is complete (the file is written before the step loop)
    p = subprocess.Popen(cmd, cwd=d, stdout=subprocess.PIPE, stderr=subprocess.DEVNULL, text=True)
    out = []
    for line in p.stdout:
        out.append(line)
        if line.startswith("start processing"):
            break
    p.kill()
    p.wait()
    return "".join(out)


def raster(path, dt_s=1e-4):
    s = W.read_spikes(path)
    if len(s) == 0:
        return np.zeros((0, 2), np.int32)
    return np.stack([np.round(s[:, 0] / dt_s).astype(np.int32), s[:, 1].astype(np.int32)], 1)


def syn_sections(path):
    """Synapse lists of network.log in solver array order
    (code/NetSimGnl2_info.cpp:203-380): name -> int array of (pre, post)."""
    txt = open(path).read()
    out = {}
    for name in ["SynAMPA1", "SynGABA1", "SynACH1", "SynGCL1", "SynNMDA1"]:
        m = re.search(r"\n%s:(\d+)((?:;[^;\n]*\n?)*)" % name, txt)
        if m:
            items = [x for x in m.group(2).replace("\n", "").split(";") if x]
            out[name] = np.array([[int(float(v)) for v in it.split(",")[:2]] for it in items],
                                 np.int32).reshape(-1, 2)
    for name in ["PSynAMPA", "PSynGABA", "PSynNMDA", "PSynACH", "PSynGCL"]:
        m = re.search(r"\n%s:(\d+);([^\n]*(?:\n[^A-Za-z\n][^\n]*)*)" % name, txt)
        if m:
            items = [x for x in m.group(2).replace("\n", "").split(";") if x]
            out[name] = np.array([[int(float(v)) for v in it.split(",")[:2]] for it in items],
                                 np.int32).reshape(-1, 2)
    return out


def total_synapses(stdout):
    m = re.search(r"total synapses=(\d+)", stdout)
    return int(m.group(1)) if m else -1


def c1_case(name, seed, g_ext, t_ms, store, keep_syn=False, cut_steps=None, args=()):
    d = tempfile.mkdtemp(prefix="golden_")
    net = W.c1_network(seed=seed, g_ext=g_ext, t_end_ms=t_ms)
    W.write_conf(net, os.path.join(d, "network.conf"))
    W.write_pro(os.path.join(d, "network.pro"), t_ms, seeds=[seed], spike_file="Spikes.dat")
    out = run_ref(d, list(args))
    r = raster(os.path.join(d, "Spikes.dat"))
    store[name + "_counts"] = np.bincount(r[:, 1], minlength=net.n).astype(np.int32)
    store[name + "_raster"] = r if cut_steps is None else r[r[:, 0] < cut_steps]
    if keep_syn:
        for k, v in syn_sections(os.path.join(d, "network.log")).items():
            store[name + "_" + k] = v
    meta = {"seed": seed, "g_ext_nS": g_ext, "t_ms": t_ms, "spikes": int(len(r)),
            "raster_steps_kept": cut_steps, "total_synapses": total_synapses(out),
            "n_syn_generated": net.n_syn}
    shutil.rmtree(d)
    return meta


MODULE_CONF = """NeuralPopulation: ExcA
N=40
C=0.5
Taum=20
Threshold=-50
RestPot=-70
ResetPot=-55
RefractoryPeriod=2
SpikeDly=0.5
Receptor:AMPA
Tau=2
RevPot=0
FreqExt=2400
MeanExtEff=1.3
EndReceptor
Receptor:NMDA
Tau=100
TauXT=2
RevPot=0
EndReceptor
Receptor:GABA
Tau=5
RevPot=-70
EndReceptor
TargetPopulation: ExcA
TargetReceptor=AMPA
MeanEff=0.8
Connectivity=0.3
EndTargetPopulation
TargetPopulation: ExcA
TargetReceptor=NMDA
MeanEff=0.05
Connectivity=0.3
EndTargetPopulation
TargetPopulation: Inh
TargetReceptor=AMPA
MeanEff=0.9
Connectivity=0.4
EndTargetPopulation
TargetPopulation: Solo
TargetReceptor=AMPA
MeanEff=0.6
Connectivity=0.05
EndTargetPopulation
EndNeuralPopulation

NeuralPopulation: Inh
N=10
C=0.2
Taum=10
Threshold=-50
RestPot=-70
ResetPot=-55
RefractoryPeriod=1
Receptor:AMPA
Tau=2
RevPot=0
FreqExt=1800
MeanExtEff=1.1
EndReceptor
Receptor:GABA
Tau=5
RevPot=-70
EndReceptor
TargetPopulation: ExcA
TargetReceptor=GABA
MeanEff=1.5
Connectivity=0.5
EndTargetPopulation
TargetPopulation: Inh
TargetReceptor=GABA
MeanEff=1.0
EndTargetPopulation
EndNeuralPopulation

NeuralPopulation: Solo
N=20
C=0.5
Taum=20
Threshold=-50
RestPot=-70
ResetPot=-55
RefractoryPeriod=2
SelfConnection=true
Receptor:AMPA
Tau=2
RevPot=0
FreqExt=2600
MeanExtEff=1.2
EndReceptor
Receptor:GABA
Tau=5
RevPot=-70
EndReceptor
TargetPopulation: Inh
TargetReceptor=AMPA
MeanEff=0.7
Connectivity=0.1
EndTargetPopulation
TargetPopulation: Solo
TargetReceptor=AMPA
MeanEff=0.5
Connectivity=0.5
EndTargetPopulation
EndNeuralPopulation
"""

MODULE_PRO = """DefineMacro
GroupName:Drive
GroupMembers:ExcA,Solo
EndGroupMembers
RandomSeeds: 7 11 EndRandomSeeds
EndDefineMacro

EventTime 100.0
Type=ChangeExtFreq
Label=#1#
Population: Drive
Receptor: AMPA
FreqExt=3000.0
EndEvent

EventTime 180.0
Type=ChangeExtFreq
Label=#2#
Population: Inh
Receptor: AMPA
FreqExt=0
EndEvent

EventTime 240.0
Type=ChangeExtFreq
Label=#3#
Population: Inh
Receptor: AMPA
FreqExt=900.0
EndEvent

EventTime 300.0
Type=EndTrial
Label=End_of_the_trial
EndEvent

OutControl
FileName:Spikes.dat
Type=Spike
population:AllPopulation
EndOutputFile

FileName:FRates.dat
Type=FiringRate
FiringRateWindow=20
PrintStep=10
population:Drive
EndOutputFile

FileName:FRsum.dat
Type=FiringRate
FiringRateWindow=30
PrintStep=25
SumAllPopulation=true
population:AllPopulation
EndOutputFile

FileName:SpAlign.dat
Type=Spike
Align=true
population:Inh
EndOutputFile
EndOutControl
"""


# ACh and GluCl (code/EqsGnl2.cpp:307-325, 404-502, 937-1002): external drive
# on both receptors, group links (two or more targets) between ExcA and ExcB
# and one-target links through Solo1 / Solo2 (the single-synapse path).
# Every ACh group link has a GluCl twin with the same targets (Connectivity
# 1, so no shuffle): the reference builds its GluCl group rows' target map
# from the ACh rows (code/NetSimGnl2_BulPp.cpp:549-550), which is only
# defined, and then equal to the GluCl rows' own targets, when the two row
# structures coincide.  The linear receptors share one time constant (4 ms):
# every population projects onto several of them, and the engine keeps one
# presynaptic linear gate per neuron.
def _lif(name, n, receptors, targets, extra=""):
    return ("NeuralPopulation: %s\nN=%d\nC=0.5\nTaum=20\nThreshold=-50\nRestPot=-70\nResetPot=-55\n"
            "RefractoryPeriod=2\nSpikeDly=0.5\n%s%s%sEndNeuralPopulation\n\n"
            % (name, n, extra, receptors, targets))


def _rec(kind, tau, rev, fext=None, eff=None):
    s = "Receptor:%s\nTau=%s\nRevPot=%s\n" % (kind, tau, rev)
    if fext is not None:
        s += "FreqExt=%s\nMeanExtEff=%s\n" % (fext, eff)
    return s + "EndReceptor\n"


def _tgt(pop, rec, eff, conn=None):
    s = "TargetPopulation: %s\nTargetReceptor=%s\nMeanEff=%s\n" % (pop, rec, eff)
    if conn is not None:
        s += "Connectivity=%s\n" % conn
    return s + "EndTargetPopulation\n"


ACHGCL_CONF = (
    _lif("ExcA", 30, _rec("AMPA", 4, 0, 2400, 1.25) + _rec("Ach", 4, 0, 900, 0.7) + _rec("GluCl", 4, -70, 500, 0.4)
         + _rec("GABA", 5, -70),
         _tgt("ExcB", "Ach", 0.35, 1) + _tgt("ExcB", "GluCl", 0.2, 1) + _tgt("ExcA", "AMPA", 0.6, 0.3))
    + _lif("ExcB", 20, _rec("AMPA", 4, 0, 2200, 1.2) + _rec("Ach", 4, 0) + _rec("GluCl", 4, -70),
           _tgt("ExcA", "Ach", 0.3, 1) + _tgt("ExcA", "GluCl", 0.25, 1) + _tgt("Solo1", "Ach", 1.1))
    + _lif("Solo1", 1, _rec("AMPA", 4, 0, 2600, 1.4) + _rec("Ach", 4, 0, 700, 0.9) + _rec("GluCl", 4, -70),
           _tgt("Solo2", "GluCl", 1.5) + _tgt("ExcB", "Ach", 0.9, 0.05))
    + _lif("Solo2", 1, _rec("AMPA", 4, 0, 2500, 1.35) + _rec("Ach", 4, 0) + _rec("GluCl", 4, -70, 800, 0.5),
           _tgt("Solo1", "Ach", 1.2) + _tgt("ExcA", "GluCl", 0.8, 0.03)))

ACHGCL_PRO = """DefineMacro
RandomSeeds: 13 EndRandomSeeds
EndDefineMacro

EventTime 120.0
Type=ChangeExtFreq
Label=#1#
Population: ExcA
Receptor: Ach
FreqExt=1500.0
EndEvent

EventTime 200.0
Type=ChangeExtFreq
Label=#2#
Population: Solo2
Receptor: GluCl
FreqExt=0
EndEvent

EventTime 300.0
Type=EndTrial
Label=End_of_the_trial
EndEvent

OutControl
FileName:Spikes.dat
Type=Spike
population:AllPopulation
EndOutputFile
EndOutControl
"""


def achgcl_case(store, files):
    d = tempfile.mkdtemp(prefix="golden_")
    open(os.path.join(d, "network.conf"), "w").write(ACHGCL_CONF)
    open(os.path.join(d, "network.pro"), "w").write(ACHGCL_PRO)
    out = run_ref(d, ["-t", "1"])
    store["achgcl_raster"] = raster(os.path.join(d, "Spikes.dat"))
    for k, v in syn_sections(os.path.join(d, "network.log")).items():
        store["achgcl_" + k] = v
    files["achgcl_network.conf"] = ACHGCL_CONF
    files["achgcl_network.pro"] = ACHGCL_PRO
    meta = {"total_synapses": total_synapses(out), "spikes": int(len(store["achgcl_raster"])),
            "sections": sorted(k for k in store if k.startswith("achgcl_") and k != "achgcl_raster")}
    shutil.rmtree(d)
    return meta


def module_case(store, files):
    d = tempfile.mkdtemp(prefix="golden_")
    open(os.path.join(d, "network.conf"), "w").write(MODULE_CONF)
    open(os.path.join(d, "network.pro"), "w").write(MODULE_PRO)
    out = run_ref(d, ["-t", "1"])
    store["module_raster"] = raster(os.path.join(d, "Spikes.dat"))
    for k, v in syn_sections(os.path.join(d, "network.log")).items():
        store["module_" + k] = v
    for fn in ["Spikes.dat", "FRates.dat", "FRsum.dat", "SpAlign.dat"]:
        files["module_" + fn] = open(os.path.join(d, fn)).read()
    files["module_network.conf"] = MODULE_CONF
    files["module_network.pro"] = MODULE_PRO
    meta = {"total_synapses": total_synapses(out), "spikes": int(len(store["module_raster"]))}
    # the same network with -t 3 and -rp 2: seeds and the connectome change
    out3 = run_ref(d, ["-t", "3", "-rp", "2"])
    for k, v in syn_sections(os.path.join(d, "network.log")).items():
        store["module_t3_" + k] = v
    store["module_t3_raster_r2"] = raster(os.path.join(d, "Spikes.dat_2"))
    seeds = re.search(r"each repeat random seed=([ 0-9]+)", open(os.path.join(d, "network.log")).read())
    meta["t3_rp2_seeds"] = [int(x) for x in seeds.group(1).split()]
    meta["t3_total_synapses"] = total_synapses(out3)
    shutil.rmtree(d)
    return meta


NOISE_PRO = """DefineMacro
RandomSeeds: 5 EndRandomSeeds
EndDefineMacro
EventTime 20.0
Type=ChangeMembraneNoise
Population: Solo
GaussMean=0.3
GaussSTD=0.2
EndEvent
EventTime 60.0
Type=ChangeMembraneNoise
Population: ExcA
GaussMean=0.1
GaussSTD=0.4
EndEvent
EventTime 120.0
Type=EndTrial
EndEvent
OutControl
FileName:Spikes.dat
Type=Spike
population:AllPopulation
EndOutputFile
EndOutControl
"""


NOISE2_PRO = """DefineMacro
RandomSeeds: 9 EndRandomSeeds
EndDefineMacro
EventTime 10.0
Type=ChangeMembraneNoise
Population: Inh
GaussMean=0.1
GaussSTD=0.5
EndEvent
EventTime 30.0
Type=ChangeMembraneNoise
Population: Solo
GaussMean=0.2
GaussSTD=0.8
EndEvent
EventTime 80.0
Type=ChangeMembraneNoise
Population: ExcA
GaussMean=0.05
GaussSTD=1.0
EndEvent
EventTime 150.0
Type=ChangeMembraneNoise
Population: Inh
GaussMean=0
GaussSTD=0
EndEvent
EventTime 200.0
Type=ChangeMembraneNoise
Population: Solo
GaussMean=0.3
GaussSTD=0.3
EndEvent
EventTime 250.0
Type=ChangeMembraneNoise
Population: ExcA
GaussMean=0
GaussSTD=0
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


def noise2_case(store, files):
    """Injection switched on for three populations in turn (the middle one
    ahead of the others in neuron order), switched off for two, and changed
    in place for one: the stream's per-step draw count changes five times."""
    d = tempfile.mkdtemp(prefix="golden_")
    open(os.path.join(d, "network.conf"), "w").write(MODULE_CONF)
    open(os.path.join(d, "network.pro"), "w").write(NOISE2_PRO)
    run_ref(d)
    store["noise2_raster"] = raster(os.path.join(d, "Spikes.dat"))
    files["noise2_network.pro"] = NOISE2_PRO
    shutil.rmtree(d)
    return {"spikes": int(len(store["noise2_raster"]))}


def seizure_case_1k(store, files, seed=4, g_ext=0.8, t_ms=300.0):
    """The seizure protocol's form on the 1k E/I network: every single-neuron
    population gets its own ChangeMembraneNoise at 1 ms (GaussSTD = 6 x
    GaussMean), Spike and summed FiringRate outputs (window 50 ms, printed
    every 1 ms) as in experiment/seizure/network1_900_10nOvf.pro.  With
    g_ext = 0.8 nS the network stays nearly silent for ~130 ms, then ignites
    into ~300 Hz hyperactivity."""
    d = tempfile.mkdtemp(prefix="golden_")
    net = W.c1_network(seed=seed, g_ext=g_ext, t_end_ms=t_ms)
    W.write_conf(net, os.path.join(d, "network.conf"))
    ev = W.seizure_events([p.name for p in net.pops], seed)
    W.write_pro(os.path.join(d, "network.pro"), t_ms, seeds=[seed], spike_file="Spikes.dat",
                rate_file="FRates.dat", rate_window_ms=50, rate_print_ms=1, events=ev,
                rate_sum=True)
    run_ref(d)
    r = raster(os.path.join(d, "Spikes.dat"))
    store["seizure1k_raster"] = r
    files["seizure1k_FRates.dat"] = open(os.path.join(d, "FRates.dat")).read()
    files["seizure1k_network.pro"] = open(os.path.join(d, "network.pro")).read()
    shutil.rmtree(d)
    return {"seed": seed, "g_ext_nS": g_ext, "t_ms": t_ms, "spikes": int(len(r)),
            "events": len(ev)}


LOG_CASES = [("module", "module_network.conf", "module_network.pro", ["-t", "1"]),
             ("outputs", "outputs_network.conf", "outputs_network.pro", ["-t", "1"]),
             ("stp", "stp_network.conf", "module_network.pro", ["-t", "1", "-STP"]),
             ("udf", "udf_network.conf", "module_network.pro", ["-t", "1", "-udfsed", "4242"]),
             ("ltp", "ltp_network.conf", "module_network.pro", ["-t", "1", "-LTP"]),
             ("gap", "gap_network.conf", "module_network.pro", ["-t", "1"]),
             ("noise2", "module_network.conf", "noise2_network.pro", ["-t", "1", "-rp", "2"])]


def logs_case(files, keep_dir=None):
    """ConfsInfo.log and network.log of the reference built from source for
    stored confs and protocols: digest, size and line count of each."""
    out = {}
    for name, conf, pro, args in LOG_CASES:
        d = tempfile.mkdtemp(prefix="golden_")
        open(os.path.join(d, "network.conf"), "w").write(files[conf])
        open(os.path.join(d, "network.pro"), "w").write(files[pro])
        run_ref(d, args, wait_for="start processing")
        out[name] = {"args": args, "conf": conf, "pro": pro}
        for fn in ["ConfsInfo.log", "network.log"]:
            b = open(os.path.join(d, fn), "rb").read()
            out[name][fn] = {"sha256": hashlib.sha256(b).hexdigest(), "bytes": len(b),
                             "lines": b.count(b"\n")}
            if keep_dir:
                os.makedirs(keep_dir, exist_ok=True)
                open(os.path.join(keep_dir, name + "_" + fn), "wb").write(b)
        shutil.rmtree(d)
    return out


def update_logs(keep_dir=None):
    build_ref()
    meta = json.load(open(os.path.join(HERE, "reference_meta.json")))
    files = json.load(open(os.path.join(HERE, "reference_files.json")))
    meta["logs"] = logs_case(files, keep_dir)
    with open(os.path.join(HERE, "reference_meta.json"), "w") as f:
        json.dump(meta, f, indent=1)
    print(json.dumps(meta["logs"], indent=1)[:1500])


def noise_case(store, files):
    d = tempfile.mkdtemp(prefix="golden_")
    open(os.path.join(d, "network.conf"), "w").write(MODULE_CONF)
    open(os.path.join(d, "network.pro"), "w").write(NOISE_PRO)
    run_ref(d)
    store["noise_raster"] = raster(os.path.join(d, "Spikes.dat"))
    files["noise_network.pro"] = NOISE_PRO
    shutil.rmtree(d)
    return {"spikes": int(len(store["noise_raster"]))}


def example_case(store):
    d = tempfile.mkdtemp(prefix="golden_")
    for fn in ["network.conf", "network.pro"]:
        shutil.copy(os.path.join(REF_CODE, fn), d)
    out = run_ref(d, ["-STP"], wait_for="network.log")
    for k, v in syn_sections(os.path.join(d, "network.log")).items():
        store["example_" + k] = v
    meta = {"total_synapses": total_synapses(out),
            "seed": int(re.search(r"random seed= (\d+)", open(os.path.join(d, "network.log")).read()).group(1))}
    shutil.rmtree(d)
    return meta


# STP parameters on ExcA (the other populations keep the parser's defaults)
STP_KEYS = """STP_pv=0.45
STP_tD=180
STP_aF=0.004
STP_tF=5000
STP_aCap=200
STP_tCap=3
STP_aCar=0.2
STP_tCar=1500
"""


def stp_conf():
    return MODULE_CONF.replace("SpikeDly=0.5\n", "SpikeDly=0.5\n" + STP_KEYS, 1)


def std_conf():
    """Under -STD the reference's group NMDA gates add D*F with F never set
    (code/EqsGnl2.cpp:797-863 against :353-362), so the STD fixture keeps
    NMDA on single synapses (one target per presynaptic neuron)."""
    c = stp_conf()
    a = "TargetReceptor=NMDA\nMeanEff=0.05\nConnectivity=0.3\n"
    assert a in c
    # ceil(40 * 0.02f) = 1 target (0.025f would round up to 2)
    return c.replace(a, "TargetReceptor=NMDA\nMeanEff=0.3\nConnectivity=0.02\n")


GAP_LINKS = """TargetPopulation: Inh
TargetReceptor=GAP
MeanEff=2.0
Connectivity=0.09
EndTargetPopulation
"""


def gap_conf():
    """The module network with one-target gap links (the solver couples only
    single-synapse gap junctions; PreID is the population index): from Solo
    (population 2, so neuron 2) and Inh (population 1, neuron 1) onto Inh
    neurons."""
    c = MODULE_CONF
    a = "EndTargetPopulation\nEndNeuralPopulation\n\nNeuralPopulation: Solo"
    assert a in c
    c = c.replace(a, "EndTargetPopulation\n" + GAP_LINKS + "EndNeuralPopulation\n\nNeuralPopulation: Solo")
    a = "TargetPopulation: Inh\nTargetReceptor=AMPA\nMeanEff=0.7\nConnectivity=0.1\nEndTargetPopulation\n"
    assert a in c
    return c.replace(a, a + GAP_LINKS.replace("MeanEff=2.0", "MeanEff=3.5"))


LTP_KEYS = """LTP_tLP=30
LTP_SpiLPDy=0.2
LTP_PosISI=0.2
LTP_NegISI=-0.05
"""


def ltp_conf():
    """The dV-limit network (module + Blast, whose one-target links take the
    single-synapse path with PreID 3) plus one-target ExcA -> Solo AMPA links
    (PreID 0), with LTP parameters on ExcA (the populations of neurons 0 and
    3, whose parameters the single synapses use) and Inh."""
    c = MODULE_CONF + BLAST_POP
    c = c.replace("SpikeDly=0.5\n", "SpikeDly=0.5\n" + LTP_KEYS, 1)
    a = "TargetPopulation: Solo\nTargetReceptor=AMPA\nMeanEff=0.6\nConnectivity=0.05\nEndTargetPopulation\n"
    assert a in c
    c = c.replace(a, a + "TargetPopulation: Solo\nTargetReceptor=AMPA\nMeanEff=2.5\nConnectivity=0.045\n"
                  "EndTargetPopulation\n")
    a = "NeuralPopulation: Inh\nN=10\n"
    assert a in c
    return c.replace(a, a + LTP_KEYS.replace("PosISI=0.2", "PosISI=0.1"))


def ltp_case(store, files):
    d = tempfile.mkdtemp(prefix="golden_")
    conf = ltp_conf()
    open(os.path.join(d, "network.conf"), "w").write(conf)
    open(os.path.join(d, "network.pro"), "w").write(MODULE_PRO)
    out = run_ref(d, ["-t", "1", "-LTP"])
    store["ltp_raster"] = raster(os.path.join(d, "Spikes.dat"))
    files["ltp_network.conf"] = conf
    out0 = run_ref(d, ["-t", "1"])
    r0 = raster(os.path.join(d, "Spikes.dat"))
    meta = {"spikes": int(len(store["ltp_raster"])), "spikes_without_ltp": int(len(r0)),
            "total_synapses": total_synapses(out)}
    shutil.rmtree(d)
    return meta


def gap_case(store, files):
    d = tempfile.mkdtemp(prefix="golden_")
    conf = gap_conf()
    open(os.path.join(d, "network.conf"), "w").write(conf)
    open(os.path.join(d, "network.pro"), "w").write(MODULE_PRO)
    out = run_ref(d, ["-t", "1"])
    store["gap_raster"] = raster(os.path.join(d, "Spikes.dat"))
    files["gap_network.conf"] = conf
    meta = {"spikes": int(len(store["gap_raster"])), "total_synapses": total_synapses(out)}
    shutil.rmtree(d)
    return meta


def plasticity_case(store, files):
    """-STP, -STD (seed 0: -STD sets the user-seed bit) and -udfsed N (STD on)."""
    meta = {}
    for name, conf, args in [("stp", stp_conf(), ["-STP"]), ("std", std_conf(), ["-STD"]),
                             ("udf", std_conf(), ["-udfsed", "4242"])]:
        d = tempfile.mkdtemp(prefix="golden_")
        open(os.path.join(d, "network.conf"), "w").write(conf)
        open(os.path.join(d, "network.pro"), "w").write(MODULE_PRO)
        run_ref(d, ["-t", "1"] + args)
        store[name + "_raster"] = raster(os.path.join(d, "Spikes.dat"))
        files[name + "_network.conf"] = conf
        seed = re.search(r"random seed= (\d+)", open(os.path.join(d, "network.log")).read())
        meta[name] = {"args": args, "spikes": int(len(store[name + "_raster"])),
                      "seed": int(seed.group(1)) if seed else None}
        shutil.rmtree(d)
    return meta


BLAST_POP = """
NeuralPopulation: Blast
N=5
C=0.5
Taum=20
Threshold=-50
RestPot=-70
ResetPot=-55
RefractoryPeriod=2
Receptor:AMPA
Tau=2
RevPot=0
FreqExt=4000
MeanExtEff=300
EndReceptor
Receptor:GABA
Tau=5
RevPot=-70
EndReceptor
TargetPopulation: Inh
TargetReceptor=AMPA
MeanEff=0.7
Connectivity=0.09
EndTargetPopulation
TargetPopulation: ExcA
TargetReceptor=NMDA
MeanEff=0.2
Connectivity=0.02
EndTargetPopulation
EndNeuralPopulation
"""

OUTPUTS_PRO = """DefineMacro
RandomSeeds: 7 11 EndRandomSeeds
EndDefineMacro
EventTime 150.0
Type=EndTrial
EndEvent
OutControl
FileName:Spikes.dat
Type=Spike
population:AllPopulation
EndOutputFile
FileName:dV.dat
Type=DeltaMemPotLimit
population:AllPopulation
EndOutputFile
FileName:dVal.dat
Type=DeltaMemPotLimit
Align=true
population:Blast
EndOutputFile
FileName:Isyn.dat
Type=IsynSeparatedSum
population:Inh
EndOutputFile
FileName:Gate.dat
Type=GatingVariable
population:ExcA
EndOutputFile
FileName:Vm.dat
Type=MemPot
population:Inh
EndOutputFile
EndOutControl
"""


# lines kept of each output file: every n-th (1 = the whole file)
OUTPUT_SAMPLE = {"Spikes.dat": 1, "dV.dat": 10, "dVal.dat": 10, "Isyn.dat": 30, "Gate.dat": 30,
                 "Vm.dat": 30}


def outputs_case(store, files):
    """Every per-step output type on the module network plus a population
    driven hard enough to hit the dV limit, whose one-target links take the
    single-synapse path (PreID = its population index, 3), so that neuron 3
    has single-synapse gates to print (150 ms, -t 1).  Keeps the lines
    OUTPUT_SAMPLE names."""
    d = tempfile.mkdtemp(prefix="golden_")
    conf = MODULE_CONF + BLAST_POP
    open(os.path.join(d, "network.conf"), "w").write(conf)
    open(os.path.join(d, "network.pro"), "w").write(OUTPUTS_PRO)
    run_ref(d, ["-t", "1"])
    files["outputs_network.conf"] = conf
    files["outputs_network.pro"] = OUTPUTS_PRO
    meta = {}
    for fn in ["Spikes.dat", "dV.dat", "dVal.dat", "Isyn.dat", "Gate.dat", "Vm.dat"]:
        txt = open(os.path.join(d, fn)).read()
        lines = txt.split("\n")
        meta[fn] = {"lines": len(lines) - 1, "sha256": hashlib.sha256(txt.encode()).hexdigest()}
        every = OUTPUT_SAMPLE[fn]
        files["outputs_" + fn] = "\n".join(lines[i] for i in range(0, len(lines) - 1, every)) + "\n"
    shutil.rmtree(d)
    return meta


def event_digest_lines(log_text):
    i = log_text.index("time event information:")
    j = log_text.index("output files information:")
    return [l for l in log_text[i:j].split("\n")[1:] if l.startswith("Event Time")]


def seizure_case():
    """Parse the fly-brain seizure protocol through the reference and keep a
    digest of the expanded event list as network.log prints it."""
    if not os.path.exists(REF_SEIZURE_PRO):
        return None
    txt = open(REF_SEIZURE_PRO).read()
    names = []
    seen = set()
    for m in re.finditer(r"Population:\s*([^\s]+)", txt):
        n = m.group(1)
        if n not in seen and n not in ("AllPopulation",):
            seen.add(n)
            names.append(n)
    groups = set(re.findall(r"GroupName:\s*([^\s]+)", txt))
    names = [n for n in names if n not in groups]
    d = tempfile.mkdtemp(prefix="golden_")
    with open(os.path.join(d, "network.conf"), "w") as f:
        for n in names:
            f.write("NeuralPopulation: %s\nN=1\nEndNeuralPopulation\n" % n)
    shutil.copy(REF_SEIZURE_PRO, os.path.join(d, "network.pro"))
    run_ref(d, wait_for="network.log")
    lines = event_digest_lines(open(os.path.join(d, "network.log")).read())
    shutil.rmtree(d)
    h = hashlib.sha256("\n".join(lines).encode()).hexdigest()
    return {"n_pops": len(names), "n_events": len(lines), "sha256": h,
            "first": lines[:3], "last": lines[-3:], "pop_names_sha256":
            hashlib.sha256("\n".join(names).encode()).hexdigest()}


def update_noise():
    """Regenerate only the injection fixtures, keeping the others as stored."""
    build_ref()
    store = dict(np.load(os.path.join(HERE, "reference_runs.npz")))
    files = json.load(open(os.path.join(HERE, "reference_files.json")))
    meta = json.load(open(os.path.join(HERE, "reference_meta.json")))
    meta["noise"] = noise_case(store, files)
    meta["noise2"] = noise2_case(store, files)
    meta["seizure1k"] = seizure_case_1k(store, files)
    np.savez_compressed(os.path.join(HERE, "reference_runs.npz"), **store)
    with open(os.path.join(HERE, "reference_files.json"), "w") as f:
        json.dump(files, f, indent=0)
    with open(os.path.join(HERE, "reference_meta.json"), "w") as f:
        json.dump(meta, f, indent=1)
    print(meta["noise"], meta["noise2"], meta["seizure1k"])


def update_plasticity():
    """Regenerate only the plasticity fixtures, keeping the others as stored."""
    build_ref()
    store = dict(np.load(os.path.join(HERE, "reference_runs.npz")))
    files = json.load(open(os.path.join(HERE, "reference_files.json")))
    meta = json.load(open(os.path.join(HERE, "reference_meta.json")))
    meta["plasticity"] = plasticity_case(store, files)
    meta["outputs"] = outputs_case(store, files)
    meta["gap"] = gap_case(store, files)
    meta["c1_stp"] = c1_case("c1_stp", 1, 1.25, 1000.0, store, args=["-STP"])
    meta["ltp"] = ltp_case(store, files)
    np.savez_compressed(os.path.join(HERE, "reference_runs.npz"), **store)
    with open(os.path.join(HERE, "reference_files.json"), "w") as f:
        json.dump(files, f, indent=0)
    with open(os.path.join(HERE, "reference_meta.json"), "w") as f:
        json.dump(meta, f, indent=1)
    print(json.dumps(meta["plasticity"], indent=1))
    print(json.dumps(meta["outputs"], indent=1))
    print(json.dumps(meta["gap"], indent=1))
    print(json.dumps(meta["ltp"], indent=1))


def update_achgcl():
    """Regenerate only the ACh / GluCl fixture, keeping the others as stored."""
    build_ref()
    store = dict(np.load(os.path.join(HERE, "reference_runs.npz")))
    files = json.load(open(os.path.join(HERE, "reference_files.json")))
    meta = json.load(open(os.path.join(HERE, "reference_meta.json")))
    meta["achgcl"] = achgcl_case(store, files)
    np.savez_compressed(os.path.join(HERE, "reference_runs.npz"), **store)
    with open(os.path.join(HERE, "reference_files.json"), "w") as f:
        json.dump(files, f, indent=0)
    with open(os.path.join(HERE, "reference_meta.json"), "w") as f:
        json.dump(meta, f, indent=1)
    print(json.dumps(meta["achgcl"], indent=1))


def main():
    build_ref()
    store, files, meta = {}, {}, {}
    meta["c1_s1"] = c1_case("c1_s1", 1, 1.15, 1000.0, store, keep_syn=True)
    meta["c1_s2"] = c1_case("c1_s2", 2, 1.15, 1000.0, store)
    meta["c1_s3"] = c1_case("c1_s3", 3, 1.15, 1000.0, store, cut_steps=2000)
    meta["c1_hyper"] = c1_case("c1_hyper", 1, 1.18, 300.0, store, cut_steps=2500)
    meta["module"] = module_case(store, files)
    meta["noise"] = noise_case(store, files)
    meta["noise2"] = noise2_case(store, files)
    meta["seizure1k"] = seizure_case_1k(store, files)
    meta["example"] = example_case(store)
    meta["plasticity"] = plasticity_case(store, files)
    meta["outputs"] = outputs_case(store, files)
    meta["gap"] = gap_case(store, files)
    meta["c1_stp"] = c1_case("c1_stp", 1, 1.25, 1000.0, store, args=["-STP"])
    meta["ltp"] = ltp_case(store, files)
    meta["achgcl"] = achgcl_case(store, files)
    meta["logs"] = logs_case(files)
    sz = seizure_case()
    if sz:
        meta["seizure_pro"] = sz
    np.savez_compressed(os.path.join(HERE, "reference_runs.npz"), **store)
    with open(os.path.join(HERE, "reference_files.json"), "w") as f:
        json.dump(files, f, indent=0)
    with open(os.path.join(HERE, "reference_meta.json"), "w") as f:
        json.dump(meta, f, indent=1)
    print(json.dumps(meta, indent=1)[:3000])


if __name__ == "__main__":
    if sys.argv[1:] == ["plasticity"]:
        update_plasticity()
    elif sys.argv[1:] == ["noise"]:
        update_noise()
    elif sys.argv[1:] == ["achgcl"]:
        update_achgcl()
    elif sys.argv[1:2] == ["logs"]:
        update_logs(sys.argv[2] if len(sys.argv) > 2 else None)
    else:
        main()
