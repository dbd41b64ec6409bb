"""ctypes binding of include/flysim.h (the engine and the drop-in host side).

The library is the in-tree ``flysim_amd/_lib/libflysim.so`` built by
``make -C flysim_amd``.  Loading fails loudly when it is missing: there is no
fallback implementation.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "_lib", "libflysim.so")

FS_OK = 0
FS_AMPA, FS_GABA, FS_NMDA, FS_ACH, FS_GCL = 0, 1, 2, 3, 4
FS_GAP = 5
FS_ACT_AMPA, FS_ACT_GABA, FS_ACT_NMDA, FS_ACT_GAP = 1, 2, 4, 8
FS_ACT_ACH, FS_ACT_GCL, FS_ACT_INJ, FS_ACT_LIF = 32, 64, 128, 256
FS_ACT_LTP, FS_ACT_STD, FS_ACT_STP = 8192, 16384, 32768
FS_FP32, FS_FP64 = 0, 1
FS_EXT_PHILOX, FS_EXT_REPLAY = 0, 1
DEFAULT_ACT = (FS_ACT_AMPA | FS_ACT_GABA | FS_ACT_NMDA | FS_ACT_ACH | FS_ACT_GCL
               | FS_ACT_GAP | FS_ACT_LIF)


class PopParams(C.Structure):
    _fields_ = [("vth", C.c_float), ("el", C.c_float), ("gl", C.c_float), ("cm", C.c_float),
                ("vreset", C.c_float), ("df", C.c_float), ("ef", C.c_float),
                ("dvmax", C.c_float), ("rfped", C.c_int32), ("spidy", C.c_int32),
                ("vrev", C.c_float * 5), ("dfst", C.c_float * 5), ("dfxt_n", C.c_float),
                ("efst_n", C.c_float), ("as_n", C.c_float), ("mg_n", C.c_float),
                ("pb", C.c_float * 5), ("gext", C.c_float * 5), ("act", C.c_uint32),
                ("inj_mean", C.c_float), ("inj_std", C.c_float), ("taum", C.c_float),
                ("stp_pv", C.c_float), ("stp_dfD", C.c_float), ("stp_efD", C.c_float),
                ("stp_aF", C.c_float), ("stp_dfF", C.c_float), ("stp_efF", C.c_float),
                ("stp_aCap", C.c_float), ("stp_dfCap", C.c_float), ("stp_aCar", C.c_float),
                ("stp_dfCar", C.c_float), ("stp_efCar", C.c_float),
                ("ltp_dfLP", C.c_float), ("ltp_pos", C.c_float), ("ltp_neg", C.c_float),
                ("ltp_spidy", C.c_int32)]


class PopConf(C.Structure):
    _fields_ = [("cm", C.c_float), ("taum", C.c_float), ("vth", C.c_float), ("el", C.c_float),
                ("vreset", C.c_float), ("refractory_ms", C.c_float),
                ("spike_delay_ms", C.c_float), ("declared", C.c_int32 * 5),
                ("tau", C.c_float * 5), ("tau_xt", C.c_float * 5), ("vrev", C.c_float * 5),
                ("freq_ext", C.c_float * 5), ("ext_eff", C.c_float * 5),
                ("ext_con", C.c_float * 5), ("stp", C.c_float * 8), ("ltp", C.c_float * 4)]

# STP_pv, STP_tD, STP_aF, STP_tF, STP_aCap, STP_tCap, STP_aCar, STP_tCar defaults
# of the reference parser (code/parser.cpp:229-243)
STP_DEFAULTS = (0.6, 300.0, 0.0035, 7000.0, 250.0, 2.0, 0.1, 2000.0)
# LTP_tLP (ms), LTP_SpiLPDy (steps), LTP_PosISI, LTP_NegISI defaults (code/parser.cpp:246-253)
LTP_DEFAULTS = (50.0, 0.0, 1e-5, 1e-5)


class NetDesc(C.Structure):
    _fields_ = [("n_neuron", C.c_int32), ("n_pop", C.c_int32),
                ("neuron_pop", C.POINTER(C.c_uint32)), ("pops", C.POINTER(PopParams)),
                ("n_syn", C.c_int64), ("pre", C.POINTER(C.c_uint32)),
                ("post", C.POINTER(C.c_uint32)), ("w", C.POINTER(C.c_float)),
                ("receptor", C.POINTER(C.c_uint8)), ("delay", C.POINTER(C.c_uint8)),
                ("plastic", C.POINTER(C.c_uint8)),
                ("dt", C.c_float), ("precision", C.c_int32), ("ext_mode", C.c_int32),
                ("record_spikes", C.c_int32)]


class Schedule(C.Structure):
    _fields_ = [("update_block", C.c_int32), ("gather_lanes", C.c_int32),
                ("push_block", C.c_int32), ("steps_per_graph", C.c_int32),
                ("use_graph", C.c_int32), ("fused", C.c_int32), ("window", C.c_int32),
                ("xcd_order", C.c_int32), ("gather_chunk", C.c_int32),
                ("gather_split", C.c_int32)]


class EiBatchDesc(C.Structure):
    _fields_ = [("n_inst", C.c_int32), ("n_neuron", C.POINTER(C.c_int32)),
                ("seed", C.POINTER(C.c_uint64)), ("ext_hz", C.POINTER(C.c_float)),
                ("gain_nmda", C.POINTER(C.c_float)), ("gain_gaba", C.POINTER(C.c_float)),
                ("exc_frac", C.c_float), ("p_conn", C.c_float), ("g_lo", C.c_float),
                ("g_hi", C.c_float), ("nmda_ratio", C.c_float), ("exc", PopParams),
                ("inh", PopParams), ("dt", C.c_float), ("precision", C.c_int32),
                ("ext_mode", C.c_int32), ("record_spikes", C.c_int32),
                ("delay_steps", C.c_int32)]


class ReplicaParams(C.Structure):
    _fields_ = [("gain", C.c_float * 5), ("ext_gain", C.c_float), ("stream_key", C.c_uint64)]


class Adam(C.Structure):
    _fields_ = [("lr", C.c_float), ("beta1", C.c_float), ("beta2", C.c_float), ("eps", C.c_float),
                ("weight_decay", C.c_float)]


class CliOpts(C.Structure):
    _fields_ = [("dt", C.c_float), ("threads", C.c_int32), ("repeats", C.c_int32),
                ("stp", C.c_int32), ("std_", C.c_int32), ("ltp", C.c_int32),
                ("udf_seed", C.c_int32), ("udf_seed_value", C.c_uint32),
                ("out_mem", C.c_char_p), ("out_spi", C.c_char_p), ("out_fr", C.c_char_p)]


class RefLayout(C.Structure):
    _fields_ = [("n_gate", C.c_int64), ("n_row", C.c_int64),
                ("gate_pre", C.POINTER(C.c_uint32)), ("gate_post0", C.POINTER(C.c_uint32)),
                ("gate_rec", C.POINTER(C.c_uint8)), ("gate_group", C.POINTER(C.c_uint8)),
                ("gate_g", C.POINTER(C.c_float)), ("row_post", C.POINTER(C.c_uint32)),
                ("row_beg", C.POINTER(C.c_uint32)), ("term", C.POINTER(C.c_uint32)),
                ("row_rec", C.POINTER(C.c_uint8)), ("pop_conf", C.POINTER(PopConf)),
                ("pop_rfped", C.POINTER(C.c_int32)), ("pop_spidy", C.POINTER(C.c_int32)),
                ("neuron_pop", C.POINTER(C.c_uint32)), ("n_evt", C.c_int64),
                ("evt_time_s", C.POINTER(C.c_float)), ("evt_var1", C.POINTER(C.c_float)),
                ("evt_var2", C.POINTER(C.c_float)), ("evt_type", C.POINTER(C.c_int32)),
                ("evt_pop", C.POINTER(C.c_int32)), ("evt_sub", C.POINTER(C.c_int32)),
                ("act_ctr", C.c_uint32), ("n_gap", C.c_int64),
                ("gap_pre", C.POINTER(C.c_uint32)), ("gap_post", C.POINTER(C.c_uint32)),
                ("gap_g", C.POINTER(C.c_float))]


EXPORTS = [
    "fs_derive_params", "fs_create", "fs_reset", "fs_set_ext_rate", "fs_set_injection",
    "fs_set_pop_params", "fs_load_ext_trains", "fs_run", "fs_step_index", "fs_get_state",
    "fs_pop_spike_counts", "fs_get_schedule", "fs_set_schedule", "fs_enable_timing",
    "fs_record_currents",
    "fs_kernel_times", "fs_run_counters", "fs_destroy", "fs_last_error", "fs_conf_load",
    "fs_conf_free", "fs_conf_desc", "fs_conf_sizes", "fs_conf_seed", "fs_conf_ext_trains",
    "fs_conf_ref_layout", "fs_conf_write_outputs", "fs_conf_write_logs", "fs_cli_main", "fs_create_replicated",
    "fs_set_replica", "fs_create_ei_batch", "fs_set_ext_rates", "fs_mlp_create", "fs_mlp_destroy", "fs_mlp_n_params", "fs_mlp_get_params",
    "fs_mlp_set_params", "fs_mlp_loss_grad", "fs_mlp_predict", "fs_mlp_predict_noalloc",
    "fs_mlp_set_data", "fs_mlp_train",
    "fs_device_alloc", "fs_device_free", "fs_device_copy", "fs_network_stats", "fs_device_info",
]

_lib = None


class FlysimError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__("flysim error %d: %s" % (code, msg))
        self.code = code


def lib() -> C.CDLL:
    """The loaded C library (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError("flysim_amd native library missing: %s (run make -C flysim_amd)"
                              % LIB_PATH)
        L = C.CDLL(LIB_PATH)
        P = C.POINTER
        L.fs_last_error.restype = C.c_char_p
        sig = {
            "fs_derive_params": [P(PopConf), C.c_float, C.c_uint32, P(PopParams)],
            "fs_create": [P(NetDesc), P(C.c_void_p)],
            "fs_create_replicated": [P(NetDesc), C.c_int32, P(C.c_void_p)],
            "fs_set_replica": [C.c_void_p, C.c_int32, P(ReplicaParams)],
            "fs_create_ei_batch": [P(EiBatchDesc), P(C.c_void_p)],
            "fs_set_ext_rates": [C.c_void_p, C.c_int32, C.c_void_p, C.c_int32],
            "fs_reset": [C.c_void_p, C.c_uint64],
            "fs_set_ext_rate": [C.c_void_p, C.c_int32, C.c_int32, C.c_float],
            "fs_set_injection": [C.c_void_p, C.c_int32, C.c_float, C.c_float],
            "fs_set_pop_params": [C.c_void_p, C.c_int32, P(PopParams)],
            "fs_load_ext_trains": [C.c_void_p, C.c_void_p, C.c_int64, C.c_int64],
            "fs_run": [C.c_void_p, C.c_int64, C.c_void_p, C.c_void_p, C.c_int64],
            "fs_step_index": [C.c_void_p, P(C.c_int64)],
            "fs_get_state": [C.c_void_p, C.c_int32, C.c_void_p],
            "fs_pop_spike_counts": [C.c_void_p, C.c_void_p],
            "fs_get_schedule": [C.c_void_p, P(Schedule)],
            "fs_set_schedule": [C.c_void_p, P(Schedule)],
            "fs_enable_timing": [C.c_void_p, C.c_int32],
            "fs_record_currents": [C.c_void_p, C.c_int32],
            "fs_kernel_times": [C.c_void_p, P(C.c_double), P(C.c_double), P(C.c_double),
                                P(C.c_int64)],
            "fs_run_counters": [C.c_void_p, C.c_void_p],
            "fs_network_stats": [C.c_void_p, C.c_void_p],
            "fs_device_info": [C.c_char_p, C.c_int32, P(C.c_int32)],
            "fs_destroy": [C.c_void_p],
            "fs_conf_load": [C.c_char_p, C.c_char_p, P(CliOpts), P(C.c_void_p)],
            "fs_conf_free": [C.c_void_p],
            "fs_conf_desc": [C.c_void_p, P(NetDesc)],
            "fs_conf_sizes": [C.c_void_p, C.c_void_p],
            "fs_conf_seed": [C.c_void_p, C.c_int32, P(C.c_uint32)],
            "fs_conf_ext_trains": [C.c_void_p, C.c_int32, C.c_int64, C.c_void_p],
            "fs_conf_ref_layout": [C.c_void_p, P(RefLayout)],
            "fs_conf_write_outputs": [C.c_void_p, C.c_int32, C.c_int64, C.c_void_p, C.c_void_p,
                                      C.c_char_p],
            "fs_conf_write_logs": [C.c_void_p, C.c_char_p],
            "fs_cli_main": [C.c_int, P(C.c_char_p)],
            "fs_mlp_create": [C.c_void_p, C.c_int32, C.c_uint64, P(C.c_void_p)],
            "fs_mlp_destroy": [C.c_void_p],
            "fs_mlp_n_params": [C.c_void_p, P(C.c_int64)],
            "fs_mlp_get_params": [C.c_void_p, C.c_void_p],
            "fs_mlp_set_params": [C.c_void_p, C.c_void_p],
            "fs_mlp_loss_grad": [C.c_void_p, C.c_void_p, C.c_void_p, C.c_int32, P(C.c_float),
                                 C.c_void_p],
            "fs_mlp_predict": [C.c_void_p, C.c_void_p, C.c_int64, C.c_void_p, C.c_int32,
                               P(C.c_double)],
            "fs_mlp_predict_noalloc": [C.c_void_p, C.c_void_p, C.c_int64, C.c_void_p],
            "fs_mlp_set_data": [C.c_void_p, C.c_void_p, C.c_void_p, C.c_int64],
            "fs_mlp_train": [C.c_void_p, C.c_int32, C.c_int32, P(Adam), C.c_uint64, C.c_int32,
                             C.c_void_p],
            "fs_device_alloc": [C.c_int64, P(C.c_void_p)],
            "fs_device_free": [C.c_void_p],
            "fs_device_copy": [C.c_void_p, C.c_void_p, C.c_int64, C.c_int32],
        }
        for name, args in sig.items():
            f = getattr(L, name)
            f.argtypes = args
            f.restype = C.c_int
        _lib = L
    return _lib


class DeviceBuffer:
    """Device memory owned by the library's HIP runtime (the process must
    not also initialise another HIP runtime, e.g. PyTorch's bundled one)."""

    def __init__(self, nbytes: int):
        p = C.c_void_p()
        check(lib().fs_device_alloc(int(nbytes), C.byref(p)))
        self.ptr = int(p.value or 0)
        self.nbytes = int(nbytes)

    def upload(self, a) -> None:
        import numpy as np
        a = np.ascontiguousarray(a)
        assert a.nbytes <= self.nbytes
        check(lib().fs_device_copy(C.c_void_p(self.ptr), C.c_void_p(a.ctypes.data), a.nbytes, 0))

    def download(self, a) -> None:
        assert a.flags["C_CONTIGUOUS"] and a.nbytes <= self.nbytes
        check(lib().fs_device_copy(C.c_void_p(a.ctypes.data), C.c_void_p(self.ptr), a.nbytes, 1))

    def free(self) -> None:
        if self.ptr:
            lib().fs_device_free(C.c_void_p(self.ptr))
            self.ptr = 0

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


def check(rc: int) -> None:
    if rc != FS_OK:
        raise FlysimError(rc, lib().fs_last_error().decode(errors="replace"))


def ptr(a: np.ndarray, ctype):
    return a.ctypes.data_as(C.POINTER(ctype))
