"""Python front end of the Flysim core: the drop-in configuration loader and
the GPU engine, both through the C ABI of include/flysim.h."""
from __future__ import annotations

import ctypes as C
from typing import Dict, Optional, Sequence, Tuple

import numpy as np

from . import _capi as A
from .workloads import Network, Population

STATE_FIELDS = {"V": 0, "count": 1, "nstate": 2, "gA": 3, "gG": 4, "gN": 5, "gH": 6, "gL": 7,
                "s_lin": 8, "x_nmda": 9, "s_nmda": 10, "extA": 11, "extG": 12, "extH": 13,
                "extL": 14, "ext_xN": 15, "ext_sN": 16, "stp_D": 17, "stp_F": 18,
                "IrepA": 19, "IrepG": 20, "IrepN": 21, "IrepH": 22, "IrepL": 23}


def pop_conf_of(p: Population) -> A.PopConf:
    """Conf-level values of one population (what network.conf states)."""
    c = A.PopConf()
    c.cm, c.taum, c.vth, c.el, c.vreset = p.cm, p.taum, p.vth, p.rest, p.reset
    c.refractory_ms, c.spike_delay_ms = p.refractory_ms, p.spike_delay_ms
    for k, v in enumerate(getattr(p, "stp", None) or A.STP_DEFAULTS):
        c.stp[k] = v
    for k, v in enumerate(A.LTP_DEFAULTS):
        c.ltp[k] = v
    for r in p.receptors:
        k = r.kind
        c.declared[k] = 1
        c.tau[k], c.tau_xt[k], c.vrev[k] = r.tau, r.tau_xt, r.rev
        c.freq_ext[k], c.ext_eff[k], c.ext_con[k] = r.freq_ext, r.mean_ext_eff, r.mean_ext_con
    return c


def derive_params(p: Population, dt: float = 0.1, act: int = A.DEFAULT_ACT) -> A.PopParams:
    """Derived per-population parameters, computed by the library exactly as
    the reference's SetPp computes them (float arithmetic)."""
    out = A.PopParams()
    A.check(A.lib().fs_derive_params(C.byref(pop_conf_of(p)), C.c_float(dt), act, C.byref(out)))
    return out


class NetArrays:
    """Owns the numpy arrays an fs_net_desc points into."""

    def __init__(self, neuron_pop, pops, pre, post, w, receptor, delay, dt):
        self.neuron_pop = np.ascontiguousarray(neuron_pop, dtype=np.uint32)
        self.pops = (A.PopParams * len(pops))(*pops)
        self.pre = np.ascontiguousarray(pre, dtype=np.uint32)
        self.post = np.ascontiguousarray(post, dtype=np.uint32)
        self.w = np.ascontiguousarray(w, dtype=np.float32)
        self.receptor = np.ascontiguousarray(receptor, dtype=np.uint8)
        self.delay = None if delay is None else np.ascontiguousarray(delay, dtype=np.uint8)
        self.dt = float(dt)

    @property
    def n(self) -> int:
        return int(self.neuron_pop.shape[0])

    def desc(self, precision: int, ext_mode: int, record: bool = True) -> A.NetDesc:
        d = A.NetDesc()
        d.n_neuron = self.n
        d.n_pop = len(self.pops)
        d.neuron_pop = A.ptr(self.neuron_pop, C.c_uint32)
        d.pops = C.cast(self.pops, C.POINTER(A.PopParams))
        d.n_syn = int(self.pre.shape[0])
        d.pre = A.ptr(self.pre, C.c_uint32)
        d.post = A.ptr(self.post, C.c_uint32)
        d.w = A.ptr(self.w, C.c_float)
        d.receptor = A.ptr(self.receptor, C.c_uint8)
        d.delay = A.ptr(self.delay, C.c_uint8) if self.delay is not None else None
        d.dt = self.dt
        d.precision = precision
        d.ext_mode = ext_mode
        d.record_spikes = 1 if record else 0
        return d


class EiBatchSpec:
    """Instances of the batched E/I generator (fs_ei_batch_desc)."""

    def __init__(self, n_neuron, seed, ext_hz, gain_nmda, gain_gaba, exc: Population,
                 inh: Population, exc_frac: float = 0.8, p_conn: float = 0.1, g_lo: float = 0.5,
                 g_hi: float = 1.5, nmda_ratio: float = 0.1, dt: float = 0.1, delay_steps: int = 0):
        self.n_neuron = np.ascontiguousarray(n_neuron, dtype=np.int32)
        k = self.n_neuron.size
        self.seed = np.ascontiguousarray(np.broadcast_to(seed, k), dtype=np.uint64)
        self.ext_hz = np.ascontiguousarray(np.broadcast_to(ext_hz, k), dtype=np.float32)
        self.gain_nmda = np.ascontiguousarray(np.broadcast_to(gain_nmda, k), dtype=np.float32)
        self.gain_gaba = np.ascontiguousarray(np.broadcast_to(gain_gaba, k), dtype=np.float32)
        self.exc, self.inh = exc, inh
        self.exc_frac, self.p_conn, self.g_lo, self.g_hi = exc_frac, p_conn, g_lo, g_hi
        self.nmda_ratio, self.dt = nmda_ratio, dt
        if not 0 <= int(delay_steps) <= 63:
            raise ValueError("delay_steps must be 0..63")
        self.delay_steps = int(delay_steps)

    def n_exc(self) -> np.ndarray:
        return np.floor(np.float64(np.float32(self.exc_frac)) * self.n_neuron + 0.5).astype(np.int64)

    def desc(self, precision: int, ext_mode: int, record: bool):
        d = A.EiBatchDesc()
        d.n_inst = int(self.n_neuron.size)
        d.n_neuron = A.ptr(self.n_neuron, C.c_int32)
        d.seed = A.ptr(self.seed, C.c_uint64)
        d.ext_hz = A.ptr(self.ext_hz, C.c_float)
        d.gain_nmda = A.ptr(self.gain_nmda, C.c_float)
        d.gain_gaba = A.ptr(self.gain_gaba, C.c_float)
        d.exc_frac, d.p_conn, d.g_lo, d.g_hi = self.exc_frac, self.p_conn, self.g_lo, self.g_hi
        d.nmda_ratio = self.nmda_ratio
        d.exc = derive_params(self.exc, self.dt)
        d.inh = derive_params(self.inh, self.dt)
        d.dt = self.dt
        d.precision, d.ext_mode, d.record_spikes = precision, ext_mode, int(record)
        d.delay_steps = self.delay_steps
        return d, self


def arrays_from_network(net: Network, act: int = A.DEFAULT_ACT) -> NetArrays:
    pops = [derive_params(p, net.dt, act) for p in net.pops]
    off = net.pop_offsets()
    npop = np.repeat(np.arange(len(net.pops), dtype=np.uint32), np.diff(off))
    return NetArrays(npop, pops, net.pre, net.post, net.w, net.receptor, net.delay, net.dt)


class Conf:
    """A network.conf / network.pro pair loaded and built by the library
    (parser and builder of the drop-in API)."""

    def __init__(self, conf: str, pro: str, dt: float = 0.1, threads: int = 1, repeats: int = 1,
                 stp: bool = False, std: bool = False, udf_seed: Optional[int] = None,
                 ltp: bool = False):
        """stp / std / ltp / udf_seed: the -STP, -STD, -LTP and -udfsed switches
        (code/main.cpp:90-97); -udfsed also switches STD on, and -STD alone
        seeds with 0, as the reference does."""
        o = A.CliOpts()
        o.dt, o.threads, o.repeats = dt, threads, repeats
        o.stp, o.std_, o.ltp = int(stp), int(std), int(ltp)
        if udf_seed is not None:
            o.udf_seed, o.udf_seed_value = 1, int(udf_seed)
        h = C.c_void_p()
        A.check(A.lib().fs_conf_load(conf.encode(), pro.encode(), C.byref(o), C.byref(h)))
        self._h = h
        s = np.zeros(7, dtype=np.int64)
        A.check(A.lib().fs_conf_sizes(self._h, s.ctypes.data))
        (self.n_neuron, self.n_pop, self.n_syn, self.total_synapses, self.n_events,
         self.n_outputs, self.n_repeats) = [int(x) for x in s]

    def close(self):
        if self._h:
            A.lib().fs_conf_free(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def seed(self, repeat: int = 0) -> int:
        v = C.c_uint32()
        A.check(A.lib().fs_conf_seed(self._h, repeat, C.byref(v)))
        return int(v.value)

    def desc(self) -> A.NetDesc:
        d = A.NetDesc()
        A.check(A.lib().fs_conf_desc(self._h, C.byref(d)))
        return d

    def ext_trains(self, n_steps: int, repeat: int = 0) -> np.ndarray:
        words = (self.n_neuron + 31) // 32
        bits = np.zeros(n_steps * 5 * words, dtype=np.uint32)
        A.check(A.lib().fs_conf_ext_trains(self._h, repeat, n_steps, bits.ctypes.data))
        return bits

    def write_outputs(self, offsets: np.ndarray, ids: np.ndarray, directory: str,
                      repeat: int = 1) -> None:
        """Write the protocol's Spike / FiringRate files for a raster given in
        the fs_run layout (offsets[n_steps+1], ids)."""
        off = np.ascontiguousarray(offsets, dtype=np.int64)
        idv = np.ascontiguousarray(ids, dtype=np.uint32)
        A.check(A.lib().fs_conf_write_outputs(self._h, repeat, len(off) - 1, off.ctypes.data,
                                              idv.ctypes.data, directory.encode()))

    def write_logs(self, directory: str) -> None:
        """Write ConfsInfo.log and network.log into ``directory``."""
        A.check(A.lib().fs_conf_write_logs(self._h, directory.encode()))

    def ref_layout(self) -> Dict[str, np.ndarray]:
        L = A.RefLayout()
        A.check(A.lib().fs_conf_ref_layout(self._h, C.byref(L)))

        def arr(p, n, dt):
            if n == 0:
                return np.zeros(0, dtype=dt)
            return np.ctypeslib.as_array(p, shape=(n,)).astype(dt, copy=True)

        n_term = int(L.row_beg[L.n_row]) if L.n_row else 0
        pc = [L.pop_conf[i] for i in range(self.n_pop)]
        return {
            "gate_pre": arr(L.gate_pre, L.n_gate, np.uint32),
            "gate_post0": arr(L.gate_post0, L.n_gate, np.uint32),
            "gate_rec": arr(L.gate_rec, L.n_gate, np.uint8),
            "gate_group": arr(L.gate_group, L.n_gate, np.uint8),
            "gate_g": arr(L.gate_g, L.n_gate, np.float32),
            "row_post": arr(L.row_post, L.n_row, np.uint32),
            "row_rec": arr(L.row_rec, L.n_row, np.uint8),
            "row_beg": arr(L.row_beg, L.n_row + 1, np.uint32),
            "term": arr(L.term, n_term, np.uint32),
            "neuron_pop": arr(L.neuron_pop, self.n_neuron, np.uint32),
            "pop_conf": pc,
            "rfped": arr(L.pop_rfped, self.n_pop, np.int32),
            "spidy": arr(L.pop_spidy, self.n_pop, np.int32),
            "evt_time_s": arr(L.evt_time_s, L.n_evt, np.float32),
            "evt_var1": arr(L.evt_var1, L.n_evt, np.float32),
            "evt_var2": arr(L.evt_var2, L.n_evt, np.float32),
            "evt_type": arr(L.evt_type, L.n_evt, np.int32),
            "evt_pop": arr(L.evt_pop, L.n_evt, np.int32),
            "evt_sub": arr(L.evt_sub, L.n_evt, np.int32),
            "act_ctr": int(L.act_ctr),
            "gap_pre": arr(L.gap_pre, L.n_gap, np.uint32),
            "gap_post": arr(L.gap_post, L.n_gap, np.uint32),
            "gap_g": arr(L.gap_g, L.n_gap, np.float32),
        }


class Engine:
    """One network (or a batch of independent instances) resident on the GPU.

    With ``replicas = B > 1`` the handle holds B copies of the network that
    share its synapse tables (parameter sweeps): neuron ids are interleaved
    (global = local * B + replica) and population ids replica-major
    (global = replica * n_pop_base + pop)."""

    def __init__(self, net, precision: int = A.FS_FP32, ext_mode: int = A.FS_EXT_PHILOX,
                 record: bool = True, replicas: int = 1):
        self._keep = net
        if isinstance(net, Conf):
            d = net.desc()
            d.precision, d.ext_mode, d.record_spikes = precision, ext_mode, int(record)
            n0, p0 = net.n_neuron, net.n_pop
        else:
            d = net.desc(precision, ext_mode, record)
            n0, p0 = net.n, len(net.pops)
        self.replicas = int(replicas)
        self.n_base, self.n_pop_base = n0, p0
        self.n = n0 * self.replicas
        self.n_pop = p0 * self.replicas
        self.dt = float(d.dt)
        h = C.c_void_p()
        if self.replicas == 1:
            A.check(A.lib().fs_create(C.byref(d), C.byref(h)))
        else:
            A.check(A.lib().fs_create_replicated(C.byref(d), self.replicas, C.byref(h)))
        self._h = h
        self.precision = precision

    @classmethod
    def ei_batch(cls, spec: "EiBatchSpec", precision: int = A.FS_FP32,
                 ext_mode: int = A.FS_EXT_PHILOX, record: bool = True) -> "Engine":
        """Batched random E/I instances generated on the device
        (fs_create_ei_batch); populations 2k / 2k+1 are the E / I
        populations of instance k."""
        self = cls.__new__(cls)
        d, keep = spec.desc(precision, ext_mode, record)
        self._keep = keep
        self.replicas = 1
        self.n = self.n_base = int(np.sum(spec.n_neuron))
        self.n_pop = self.n_pop_base = 2 * len(spec.n_neuron)
        self.dt = float(spec.dt)
        h = C.c_void_p()
        A.check(A.lib().fs_create_ei_batch(C.byref(d), C.byref(h)))
        self._h = h
        self.precision = precision
        return self

    def set_replica(self, replica: int, gain=(1.0, 1.0, 1.0, 1.0, 1.0), ext_gain: float = 1.0,
                    stream_key: Optional[int] = None) -> None:
        """Per-copy recurrent gains by receptor id, external efficacy gain and
        random-stream key (default: replica * 0x9E3779B97F4A7C15)."""
        p = A.ReplicaParams()
        for i, g in enumerate(gain):
            p.gain[i] = float(g)
        p.ext_gain = float(ext_gain)
        if stream_key is None:
            stream_key = (replica * 0x9E3779B97F4A7C15) & 0xFFFFFFFFFFFFFFFF
        p.stream_key = int(stream_key) & 0xFFFFFFFFFFFFFFFF
        A.check(A.lib().fs_set_replica(self._h, int(replica), C.byref(p)))

    def close(self):
        if getattr(self, "_h", None):
            A.lib().fs_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def reset(self, seed: int = 0) -> None:
        A.check(A.lib().fs_reset(self._h, int(seed)))

    def load_ext_trains(self, bits: np.ndarray, first_step: int, n_steps: int) -> None:
        self._bits = np.ascontiguousarray(bits, dtype=np.uint32)
        A.check(A.lib().fs_load_ext_trains(self._h, self._bits.ctypes.data, first_step, n_steps))

    # id buffer of a recorded run: the run goes in chunks whose worst case
    # (every neuron spiking every step) fits it, so fs_run never overflows
    RECORD_IDS = 1 << 26

    def run(self, n_steps: int, record: bool = True) -> Tuple[np.ndarray, np.ndarray]:
        """Run n steps; returns (offsets[n+1], ids) of the per-step spike lists."""
        if not record:
            A.check(A.lib().fs_run(self._h, n_steps, None, None, 0))
            return np.zeros(0, np.int64), np.zeros(0, np.uint32)
        chunk = max(1, min(n_steps, self.RECORD_IDS // max(self.n, 1)))
        ids = np.zeros(max(1024, self.n * min(chunk, n_steps)), dtype=np.uint32)
        off = np.zeros(n_steps + 1, dtype=np.int64)
        parts = []
        done = 0
        while done < n_steps:
            k = min(chunk, n_steps - done)
            o = np.zeros(k + 1, dtype=np.int64)
            A.check(A.lib().fs_run(self._h, k, o.ctypes.data, ids.ctypes.data, ids.size))
            off[done + 1: done + k + 1] = o[1:] + off[done]
            parts.append(ids[: o[-1]].copy())
            done += k
        return off, (np.concatenate(parts) if parts else np.zeros(0, np.uint32))

    def raster(self, n_steps: int) -> np.ndarray:
        """Run n steps and return an (k, 2) int64 array of (step, nid)."""
        s0 = self.step
        off, ids = self.run(n_steps)
        steps = np.repeat(np.arange(s0, s0 + n_steps, dtype=np.int64), np.diff(off))
        return np.stack([steps, ids.astype(np.int64)], axis=1)

    @property
    def step(self) -> int:
        v = C.c_int64()
        A.check(A.lib().fs_step_index(self._h, C.byref(v)))
        return int(v.value)

    def state(self, field: str) -> np.ndarray:
        out = np.zeros(self.n, dtype=np.float64)
        A.check(A.lib().fs_get_state(self._h, STATE_FIELDS[field], out.ctypes.data))
        return out

    def pop_counts(self) -> np.ndarray:
        out = np.zeros(self.n_pop, dtype=np.uint64)
        A.check(A.lib().fs_pop_spike_counts(self._h, out.ctypes.data))
        return out

    def set_ext_rate(self, pop: int, receptor: int, hz: float) -> None:
        A.check(A.lib().fs_set_ext_rate(self._h, pop, receptor, C.c_float(hz)))

    def set_ext_rates(self, receptor: int, hz) -> None:
        """External rate of every population (global ids) at once."""
        self._hz = np.ascontiguousarray(hz, dtype=np.float32)
        A.check(A.lib().fs_set_ext_rates(self._h, receptor, self._hz.ctypes.data, self._hz.size))

    def set_injection(self, pop: int, mean: float, std: float) -> None:
        A.check(A.lib().fs_set_injection(self._h, pop, C.c_float(mean), C.c_float(std)))

    def set_pop_params(self, pop: int, p: A.PopParams) -> None:
        A.check(A.lib().fs_set_pop_params(self._h, pop, C.byref(p)))

    @property
    def schedule(self) -> Dict[str, int]:
        s = A.Schedule()
        A.check(A.lib().fs_get_schedule(self._h, C.byref(s)))
        return {k: getattr(s, k) for k, _ in A.Schedule._fields_}

    @property
    def fused(self) -> bool:
        """True while the one-launch-per-step path is in effect."""
        return bool(self.schedule["fused"])

    def set_schedule(self, **kw) -> None:
        cur = self.schedule
        cur.update(kw)
        s = A.Schedule(**cur)
        A.check(A.lib().fs_set_schedule(self._h, C.byref(s)))

    def record_currents(self, on: bool = True) -> None:
        """Keep each neuron's per-receptor synaptic currents (state fields
        IrepA..IrepL, pA) of its last non-refractory step."""
        A.check(A.lib().fs_record_currents(self._h, int(on)))

    def enable_timing(self, on: bool = True) -> None:
        A.check(A.lib().fs_enable_timing(self._h, int(on)))

    def kernel_times(self) -> Dict[str, float]:
        u, p, g, n = C.c_double(), C.c_double(), C.c_double(), C.c_int64()
        A.check(A.lib().fs_kernel_times(self._h, C.byref(u), C.byref(p), C.byref(g), C.byref(n)))
        return {"update_ms": u.value, "deliver_ms": p.value, "graph_step_ms": g.value,
                "n_timed": n.value}

    def stats(self) -> Dict[str, int]:
        """Shape of the network (fs_network_stats): neurons and populations
        of one copy, copies, linear and NMDA receptor-synapses, smallest
        delay, largest linear and NMDA delays (steps), precision, input mode."""
        out = np.zeros(10, np.int64)
        A.check(A.lib().fs_network_stats(self._h, out.ctypes.data))
        keys = ("neurons", "pops", "replicas", "syn_lin", "syn_nmda", "min_delay", "max_delay_lin",
                "max_delay_nmda", "precision", "ext_mode")
        return {k: int(v) for k, v in zip(keys, out)}

    def counters(self) -> Dict[str, int]:
        """Counters of the last run(): spikes, linear receptor-synapses
        pushed, NMDA receptor-synapses of the spiking neurons, their sum
        (synaptic events) and the NMDA terms gathered."""
        out = np.zeros(4, np.int64)
        A.check(A.lib().fs_run_counters(self._h, out.ctypes.data))
        return {"spikes": int(out[0]), "lin_pushes": int(out[1]), "nmda_out": int(out[2]),
                "syn_events": int(out[1] + out[2]), "nmda_terms": int(out[3])}


def device_info() -> Tuple[str, int]:
    """Name / architecture of the current device and its compute units."""
    buf = C.create_string_buffer(256)
    cu = C.c_int32()
    A.check(A.lib().fs_device_info(buf, 256, C.byref(cu)))
    return buf.value.decode(errors="replace"), int(cu.value)
