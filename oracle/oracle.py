"""ORACLE -- test infrastructure only.

ctypes binding of oracle/liboracle.so, the CPU restatement of the reference
GNL2 step (see oracle/flysim_oracle.h for the restated reference lines and
the parity pins).  Imported only by tests/, the package smoke check and the
CPU-baseline leg of bench.py.
"""
from __future__ import annotations

import ctypes as C
import os
from typing import Dict, Optional

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "liboracle.so")
REF_BIN = os.path.join(_HERE, "_ref", "flysim_ref")


class OrPop(C.Structure):
    _fields_ = [("cm", C.c_float), ("taum", C.c_float), ("vth", C.c_float), ("el", C.c_float),
                ("vreset", C.c_float), ("rfped", C.c_int32), ("spidy", C.c_int32),
                ("declared", C.c_int32 * 5), ("tau", C.c_float * 5), ("tau_xt", C.c_float * 5),
                ("vrev", C.c_float * 5), ("freq_ext", C.c_float * 5), ("ext_eff", C.c_float * 5),
                ("ext_con", C.c_float * 5), ("stp", C.c_float * 8), ("ltp", C.c_float * 4)]


class OrNet(C.Structure):
    _fields_ = [("n_neuron", C.c_int32), ("n_pop", C.c_int32),
                ("neuron_pop", C.POINTER(C.c_int32)), ("pops", C.POINTER(OrPop)),
                ("n_gate", C.c_int32), ("gate_pre", C.POINTER(C.c_uint32)),
                ("gate_rec", C.POINTER(C.c_uint8)), ("gate_g", C.POINTER(C.c_float)),
                ("gate_post0", C.POINTER(C.c_uint32)), ("n_row", C.c_int32),
                ("row_post", C.POINTER(C.c_uint32)), ("row_rec", C.POINTER(C.c_uint8)),
                ("row_beg", C.POINTER(C.c_uint32)), ("term_gate", C.POINTER(C.c_uint32)),
                ("act_ctr", C.c_uint32), ("dt", C.c_float), ("n_evt", C.c_int32),
                ("evt_time_s", C.POINTER(C.c_float)), ("evt_type", C.POINTER(C.c_int32)),
                ("evt_pop", C.POINTER(C.c_int32)), ("evt_sub", C.POINTER(C.c_int32)),
                ("evt_var1", C.POINTER(C.c_float)), ("evt_var2", C.POINTER(C.c_float)),
                ("gate_group", C.POINTER(C.c_uint8)), ("n_gap", C.c_int32), ("gap_pre", C.POINTER(C.c_uint32)),
                ("gap_post", C.POINTER(C.c_uint32)), ("gap_g", C.POINTER(C.c_float))]


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError("oracle library missing: %s (run make -C oracle)" % LIB_PATH)
        L = C.CDLL(LIB_PATH)
        L.or_create.restype = C.c_void_p
        L.or_create.argtypes = [C.POINTER(OrNet), C.c_int]
        L.or_destroy.argtypes = [C.c_void_p]
        L.or_reset.argtypes = [C.c_void_p, C.c_uint32]
        L.or_set_ext_bits.argtypes = [C.c_void_p, C.c_void_p, C.c_int64]
        L.or_record_ext_bits.argtypes = [C.c_void_p, C.c_void_p, C.c_int64]
        L.or_run.restype = C.c_int64
        L.or_run.argtypes = [C.c_void_p, C.c_int64, C.c_void_p, C.c_void_p, C.c_int64]
        L.or_step_index.restype = C.c_int64
        L.or_step_index.argtypes = [C.c_void_p]
        for f in ("or_get", "or_get_gates", "or_get_params"):
            getattr(L, f).argtypes = [C.c_void_p, C.c_int, C.c_void_p]
        L.or_mt19937_uniform.argtypes = [C.c_uint32, C.c_int64, C.c_void_p]
        L.or_time_run.restype = C.c_double
        L.or_time_run.argtypes = [C.c_void_p, C.c_int64, C.c_int]
        _lib = L
    return _lib


FIELDS = {"V": 0, "count": 1, "nstate": 2, "gA": 3, "gG": 4, "gN": 5, "gH": 6, "gL": 7,
          "Isyn": 8, "extA": 9, "extG": 10, "extH": 11, "extL": 12, "ext_xN": 14,
          "ext_sN": 15, "stp_D": 16, "stp_F": 17}
PARAMS = {"df": 0, "ef": 1, "gl": 2, "dvmax": 3, "dfst_A": 4, "dfst_G": 5, "dfst_N": 6,
          "dfxt_N": 7, "efst_N": 8, "pb_A": 9, "gext_A": 10}


def _p(a, t):
    return a.ctypes.data_as(C.POINTER(t))


class Oracle:
    """The reference step restated on the CPU, in float (precision=0) or
    double (precision=1), driven by a reference layout: gates, rows in the
    reference's summation order, conf-level population values and events."""

    def __init__(self, layout: Dict, n_neuron: int, dt: float = 0.1, precision: int = 0):
        self.n = n_neuron
        k = self._keep = {}
        pops = (OrPop * len(layout["pop_conf"]))()
        for i, pc in enumerate(layout["pop_conf"]):
            o = pops[i]
            o.cm, o.taum, o.vth, o.el, o.vreset = pc.cm, pc.taum, pc.vth, pc.el, pc.vreset
            # steps exactly as the reference parser derives them
            o.rfped = int(round(pc.refractory_ms / dt)) if "rfped" not in layout else int(layout["rfped"][i])
            o.spidy = int(round(pc.spike_delay_ms / dt)) if "spidy" not in layout else int(layout["spidy"][i])
            for r in range(8):
                o.stp[r] = pc.stp[r]
            for r in range(4):
                o.ltp[r] = pc.ltp[r]
            for r in range(5):
                o.declared[r] = pc.declared[r]
                o.tau[r], o.tau_xt[r], o.vrev[r] = pc.tau[r], pc.tau_xt[r], pc.vrev[r]
                o.freq_ext[r], o.ext_eff[r], o.ext_con[r] = pc.freq_ext[r], pc.ext_eff[r], pc.ext_con[r]
        k["pops"] = pops
        for name, dt_ in [("neuron_pop", np.int32), ("gate_pre", np.uint32), ("gate_rec", np.uint8),
                          ("gate_g", np.float32), ("gate_post0", np.uint32), ("row_post", np.uint32),
                          ("row_rec", np.uint8), ("row_beg", np.uint32), ("term", np.uint32),
                          ("evt_time_s", np.float32), ("evt_type", np.int32), ("evt_pop", np.int32),
                          ("evt_sub", np.int32), ("evt_var1", np.float32), ("evt_var2", np.float32)]:
            k[name] = np.ascontiguousarray(layout[name], dtype=dt_)
        net = OrNet()
        net.n_neuron = n_neuron
        net.n_pop = len(layout["pop_conf"])
        net.neuron_pop = _p(k["neuron_pop"], C.c_int32)
        net.pops = C.cast(pops, C.POINTER(OrPop))
        net.n_gate = int(k["gate_pre"].shape[0])
        net.gate_pre = _p(k["gate_pre"], C.c_uint32)
        net.gate_rec = _p(k["gate_rec"], C.c_uint8)
        net.gate_g = _p(k["gate_g"], C.c_float)
        net.gate_post0 = _p(k["gate_post0"], C.c_uint32)
        net.n_row = int(k["row_post"].shape[0])
        net.row_post = _p(k["row_post"], C.c_uint32)
        net.row_rec = _p(k["row_rec"], C.c_uint8)
        net.row_beg = _p(k["row_beg"], C.c_uint32)
        net.term_gate = _p(k["term"], C.c_uint32)
        net.act_ctr = int(layout["act_ctr"])
        net.dt = dt
        net.n_evt = int(k["evt_time_s"].shape[0])
        net.evt_time_s = _p(k["evt_time_s"], C.c_float)
        net.evt_type = _p(k["evt_type"], C.c_int32)
        net.evt_pop = _p(k["evt_pop"], C.c_int32)
        net.evt_sub = _p(k["evt_sub"], C.c_int32)
        net.evt_var1 = _p(k["evt_var1"], C.c_float)
        net.evt_var2 = _p(k["evt_var2"], C.c_float)
        k["gate_group"] = np.ascontiguousarray(
            layout.get("gate_group", np.ones(k["gate_pre"].shape[0])), dtype=np.uint8)
        net.gate_group = _p(k["gate_group"], C.c_uint8)
        for name, dt_ in [("gap_pre", np.uint32), ("gap_post", np.uint32), ("gap_g", np.float32)]:
            k[name] = np.ascontiguousarray(layout.get(name, np.zeros(0)), dtype=dt_)
        net.n_gap = int(k["gap_pre"].shape[0])
        net.gap_pre = _p(k["gap_pre"], C.c_uint32)
        net.gap_post = _p(k["gap_post"], C.c_uint32)
        net.gap_g = _p(k["gap_g"], C.c_float)
        k["net"] = net
        self._h = lib().or_create(C.byref(net), precision)

    def close(self):
        if getattr(self, "_h", None):
            lib().or_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def reset(self, seed: int) -> None:
        lib().or_reset(self._h, seed)

    def replay(self, bits: np.ndarray, n_steps: int) -> None:
        self._bits = np.ascontiguousarray(bits, dtype=np.uint32)
        lib().or_set_ext_bits(self._h, self._bits.ctypes.data, n_steps)

    def record(self, n_steps: int) -> np.ndarray:
        words = (self.n + 31) // 32
        self._rec = np.zeros(n_steps * 5 * words, dtype=np.uint32)
        lib().or_record_ext_bits(self._h, self._rec.ctypes.data, n_steps)
        return self._rec

    def raster(self, n_steps: int, cap: Optional[int] = None) -> np.ndarray:
        cap = cap or max(1 << 16, self.n * n_steps // 4)
        while True:
            ss = np.zeros(cap, np.int64)
            sn = np.zeros(cap, np.int32)
            s0 = self.step
            w = lib().or_run(self._h, n_steps, ss.ctypes.data, sn.ctypes.data, cap)
            if w >= 0:
                return np.stack([ss[:w], sn[:w].astype(np.int64)], axis=1)
            raise RuntimeError("oracle raster capacity exceeded (cap=%d)" % cap)

    @property
    def step(self) -> int:
        return int(lib().or_step_index(self._h))

    def state(self, field: str) -> np.ndarray:
        out = np.zeros(self.n, np.float64)
        lib().or_get(self._h, FIELDS[field], out.ctypes.data)
        return out

    def params(self, name: str) -> np.ndarray:
        out = np.zeros(self.n, np.float64)
        lib().or_get_params(self._h, PARAMS[name], out.ctypes.data)
        return out

    def time_run(self, n_steps: int, threads: int) -> float:
        return float(lib().or_time_run(self._h, n_steps, threads))


def mt19937_uniform(seed: int, n: int) -> np.ndarray:
    out = np.zeros(n, np.float32)
    lib().or_mt19937_uniform(seed, n, out.ctypes.data)
    return out
