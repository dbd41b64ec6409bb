"""Sweep driver and learned rate surrogate.

* :class:`Mlp` -- the surrogate network of include/flysim_mlp.h (BF16
  matrix-core GEMMs, FP32 master weights, Adam, all on the GPU).
* :class:`SweepSimulator` -- runs sweep points of one network (per-module
  external rates and receptor gains) as replicas of a replicated handle, B
  points per launch, and returns per-module rates; ``replicas=1`` is the plain
  per-instance GPU simulation it is compared with.
* :class:`RateSurrogate` -- input/output scaling around :class:`Mlp`.
* :func:`screen_sweep` -- the surrogate screens sweep points for the
  hyperactivity regime (mean rate above a threshold); the simulator re-runs
  every flagged point plus a random audit and is the judge of the
  surrogate's error.

The reference has no sweep driver (one network per process,
code/main.cpp:67-264); a sweep point is what one reference run with edited
``FreqExt`` and efficacies would compute.
"""
from __future__ import annotations

import ctypes as C
import time
from dataclasses import dataclass, field
from typing import Dict, List, Optional, Sequence, Tuple

import numpy as np

from . import _capi as A
from .engine import Engine, arrays_from_network
from .workloads import AMPA, Network

HIDDEN = (256, 256, 256)
Y_SCALE = 100.0   # Hz per unit of the network's output


class Mlp:
    """GPU MLP (GELU hidden layers) through include/flysim_mlp.h."""

    def __init__(self, sizes: Sequence[int], seed: int = 0):
        self.sizes = [int(s) for s in sizes]
        arr = (C.c_int32 * len(self.sizes))(*self.sizes)
        h = C.c_void_p()
        A.check(A.lib().fs_mlp_create(arr, len(self.sizes) - 1, int(seed), C.byref(h)))
        self._h = h
        n = C.c_int64()
        A.check(A.lib().fs_mlp_n_params(self._h, C.byref(n)))
        self.n_params = int(n.value)

    def close(self):
        if getattr(self, "_h", None):
            A.lib().fs_mlp_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def params(self) -> np.ndarray:
        out = np.zeros(self.n_params, np.float32)
        A.check(A.lib().fs_mlp_get_params(self._h, out.ctypes.data))
        return out

    @params.setter
    def params(self, p: np.ndarray) -> None:
        p = np.ascontiguousarray(p, dtype=np.float32)
        assert p.size == self.n_params
        A.check(A.lib().fs_mlp_set_params(self._h, p.ctypes.data))

    def loss_grad(self, x: np.ndarray, y: np.ndarray) -> Tuple[float, np.ndarray]:
        x = np.ascontiguousarray(x, dtype=np.float32)
        y = np.ascontiguousarray(y, dtype=np.float32)
        g = np.zeros(self.n_params, np.float32)
        loss = C.c_float()
        A.check(A.lib().fs_mlp_loss_grad(self._h, x.ctypes.data, y.ctypes.data, x.shape[0],
                                         C.byref(loss), g.ctypes.data))
        return float(loss.value), g

    def predict(self, x: np.ndarray) -> np.ndarray:
        x = np.ascontiguousarray(x, dtype=np.float32)
        y = np.zeros((x.shape[0], self.sizes[-1]), np.float32)
        ms = C.c_double()
        A.check(A.lib().fs_mlp_predict(self._h, x.ctypes.data, x.shape[0], y.ctypes.data, 0,
                                       C.byref(ms)))
        self.last_ms = float(ms.value)
        return y

    def predict_device(self, x_ptr: int, n: int, y_ptr: int) -> float:
        """Forward of n rows already in device memory (f32 [n][in] -> f32
        [n][out]); returns the GPU time in ms."""
        ms = C.c_double()
        A.check(A.lib().fs_mlp_predict(self._h, C.c_void_p(x_ptr), int(n), C.c_void_p(y_ptr), 1,
                                       C.byref(ms)))
        return float(ms.value)

    def set_data(self, x: np.ndarray, y: np.ndarray) -> None:
        self._x = np.ascontiguousarray(x, dtype=np.float32)
        self._y = np.ascontiguousarray(y, dtype=np.float32)
        A.check(A.lib().fs_mlp_set_data(self._h, self._x.ctypes.data, self._y.ctypes.data,
                                        self._x.shape[0]))

    def train(self, epochs: int, batch: int, lr: float = 1e-3, beta1: float = 0.9,
              beta2: float = 0.999, eps: float = 1e-8, weight_decay: float = 0.0,
              seed: int = 0, shuffle: bool = True) -> np.ndarray:
        opt = A.Adam(lr, beta1, beta2, eps, weight_decay)
        out = np.zeros(max(epochs, 1), np.float32)
        A.check(A.lib().fs_mlp_train(self._h, int(epochs), int(batch), C.byref(opt), int(seed),
                                     int(shuffle), out.ctypes.data))
        return out[:epochs]


# ----------------------------------------------------------------- sweep space
@dataclass
class SweepSpace:
    """Sweep point = per-module external AMPA rates (kHz) followed by the
    recurrent AMPA, GABA and NMDA gains and the external efficacy gain."""
    n_mod: int = 50
    rate_khz: Tuple[float, float] = (0.5, 4.0)
    gain: Tuple[float, float] = (0.5, 2.0)
    ext_gain: Tuple[float, float] = (0.5, 1.5)

    @property
    def dim(self) -> int:
        return self.n_mod + 4

    def sample(self, n: int, rng: np.random.Generator) -> np.ndarray:
        x = np.empty((n, self.dim), np.float32)
        x[:, :self.n_mod] = rng.uniform(*self.rate_khz, size=(n, self.n_mod))
        x[:, self.n_mod:self.n_mod + 3] = rng.uniform(*self.gain, size=(n, 3))
        x[:, self.n_mod + 3] = rng.uniform(*self.ext_gain, size=n)
        return x


def module_pops(net: Network) -> List[List[int]]:
    """Population indices of each module ("M<m>E", "M<m>I")."""
    mods: Dict[int, List[int]] = {}
    for i, p in enumerate(net.pops):
        m = int(p.name[1:-1])
        mods.setdefault(m, []).append(i)
    return [mods[m] for m in sorted(mods)]


class SweepSimulator:
    """Per-module rates of sweep points on one network, B points per launch
    (replicated handle); replicas=1 is the plain per-instance simulation.
    schedule: fs_schedule overrides, or "auto" to autotune the launch
    schedule on this network (flysim_amd/tune.py; the result is cached by
    network signature and device)."""

    def __init__(self, net: Network, replicas: int = 128, warm_steps: int = 1000,
                 meas_steps: int = 2000, precision: int = A.FS_FP32,
                 schedule=None):
        self.net = net
        self.B = int(replicas)
        self.warm, self.meas = int(warm_steps), int(meas_steps)
        self.mods = module_pops(net)
        self.n_mod = len(self.mods)
        self.P0 = len(net.pops)
        self.mod_size = np.array([sum(net.pops[i].n for i in m) for m in self.mods], np.float64)
        self.eng = Engine(arrays_from_network(net), precision=precision, record=False,
                          replicas=self.B)
        self.tuned = None
        if schedule == "auto":
            from .tune import tune
            self.eng.reset(0)
            self.tuned = tune(self.eng)
        elif schedule:
            self.eng.set_schedule(**schedule)
        self.sim_seconds = 0.0
        self.wall_seconds = 0.0
        self.points = 0

    def _apply(self, r: int, x: np.ndarray, key: int) -> None:
        nm = self.n_mod
        gA, gG, gN, gE = (float(v) for v in x[nm:nm + 4])
        self.eng.set_replica(r, gain=(gA, gG, gN, 1.0, 1.0), ext_gain=gE, stream_key=key)
        for m, pops in enumerate(self.mods):
            for pi in pops:
                self._hz[r * self.P0 + pi] = float(x[m]) * 1000.0

    def run(self, X: np.ndarray, seed: int = 0) -> np.ndarray:
        """Rates (Hz) [n, n_mod] over the measurement window after warm-up."""
        X = np.asarray(X, np.float32)
        n = X.shape[0]
        out = np.zeros((n, self.n_mod), np.float32)
        t0 = time.perf_counter()
        for b0 in range(0, n, self.B):
            k = min(self.B, n - b0)
            self._hz = np.zeros(self.B * self.P0, np.float32)
            for r in range(self.B):
                # unused replicas of a partial batch repeat the last point
                i = b0 + min(r, k - 1)
                self._apply(r, X[i], key=(seed * 1_000_003 + i) * 0x9E3779B97F4A7C15 % (1 << 64))
            self.eng.set_ext_rates(AMPA, self._hz)
            self.eng.reset(seed)
            if self.warm:
                self.eng.run(self.warm, record=False)
            c0 = self.eng.pop_counts().astype(np.int64).reshape(self.B, self.P0)
            self.eng.run(self.meas, record=False)
            c1 = self.eng.pop_counts().astype(np.int64).reshape(self.B, self.P0)
            d = c1 - c0
            per_mod = np.stack([d[:, pops].sum(axis=1) for pops in self.mods], axis=1)
            secs = self.meas * self.net.dt * 1e-3
            out[b0:b0 + k] = (per_mod[:k] / self.mod_size / secs).astype(np.float32)
        self.wall_seconds += time.perf_counter() - t0
        self.points += n
        self.sim_seconds += n * (self.warm + self.meas) * self.net.dt * 1e-3
        return out

    @property
    def points_per_s(self) -> float:
        return self.points / self.wall_seconds if self.wall_seconds else 0.0


class RateSurrogate:
    """Sweep point -> per-module rates (Hz), standardised inputs and
    outputs scaled by 1/100 Hz."""

    def __init__(self, n_in: int, n_out: int, hidden: Sequence[int] = HIDDEN, seed: int = 0):
        self.mlp = Mlp([n_in, *hidden, n_out], seed=seed)
        self.mu = np.zeros(n_in, np.float32)
        self.sd = np.ones(n_in, np.float32)

    def _x(self, X):
        return ((np.asarray(X, np.float32) - self.mu) / self.sd).astype(np.float32)

    def fit(self, X: np.ndarray, Y: np.ndarray, epochs: int = 60, batch: int = 512,
            lr: float = 1e-3, seed: int = 0, decay: Sequence[float] = (1.0, 0.3, 0.1)) -> np.ndarray:
        """Adam on the MSE of rates / 100 Hz; the learning rate steps down
        through the factors of `decay`, equal shares of the epochs each."""
        X = np.asarray(X, np.float32)
        self.mu = X.mean(axis=0)
        self.sd = X.std(axis=0) + 1e-6
        self.mlp.set_data(self._x(X), np.asarray(Y, np.float32) / Y_SCALE)
        losses = []
        for i, f in enumerate(decay):
            losses.append(self.mlp.train(max(1, epochs // len(decay)), batch, lr=lr * f, seed=seed + i))
        return np.concatenate(losses)

    def predict(self, X: np.ndarray) -> np.ndarray:
        return self.mlp.predict(self._x(X)) * Y_SCALE


@dataclass
class ScreenResult:
    flagged: np.ndarray            # indices the surrogate flagged
    audited: np.ndarray            # random audit indices (not flagged)
    sim_rates: Dict[int, np.ndarray] = field(default_factory=dict)
    pred: Optional[np.ndarray] = None
    t_predict: float = 0.0
    t_simulate: float = 0.0


def screen_sweep(sur: RateSurrogate, sim: SweepSimulator, X: np.ndarray, threshold_hz: float = 50.0,
                 margin_hz: float = 10.0, audit_frac: float = 0.01, seed: int = 0) -> ScreenResult:
    """Flag points whose predicted mean rate exceeds threshold - margin,
    re-run all flagged points plus a random audit of audit_frac of the rest
    on the simulator."""
    t0 = time.perf_counter()
    pred = sur.predict(X)
    t1 = time.perf_counter()
    mean = pred.mean(axis=1)
    flagged = np.nonzero(mean > threshold_hz - margin_hz)[0]
    rest = np.setdiff1d(np.arange(X.shape[0]), flagged)
    rng = np.random.default_rng(seed)
    n_aud = int(round(audit_frac * rest.size))
    audited = np.sort(rng.choice(rest, size=n_aud, replace=False)) if n_aud else rest[:0]
    todo = np.concatenate([flagged, audited])
    rates = sim.run(X[todo], seed=seed) if todo.size else np.zeros((0, sim.n_mod), np.float32)
    t2 = time.perf_counter()
    res = ScreenResult(flagged=flagged, audited=audited, pred=pred, t_predict=t1 - t0,
                       t_simulate=t2 - t1)
    res.sim_rates = {int(i): rates[k] for k, i in enumerate(todo)}
    return res
