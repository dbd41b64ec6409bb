"""Surrogate-screened parameter sweeps of the 20k-neuron fly-like network.

Stages (each a command; files are numpy .npz):

  gen     simulate N random sweep points (per-module external rates + 4 gains)
          on the GPU, B points per launch -> training set
  train   fit the rate surrogate on the GPU (BF16 MLP, Adam) -> model
  screen  predict M sweep points, flag predicted mean rate > threshold-margin,
          re-simulate every flagged point plus a random audit of the rest
          (optionally a slice of that list per call) -> verified rates
  band    the threshold-locating screen end to end: simulate only the points
          predicted near the threshold plus a 1 % audit, classify the rest by
          the surrogate
  moderr  per-module error of the surrogate against repeated simulator runs
  predict GPU time of the surrogate's forward over device-resident points
  report  merge the screen slices: hyperactive points found, surrogate error
          against the simulator, points/s of plain per-instance GPU simulation,
          batched simulation and batched + surrogate

    python -m flysim_amd.sweep gen --points 100000 --out d/train.npz
    python -m flysim_amd.sweep train --data d/train.npz --out d/model.npz
    python -m flysim_amd.sweep screen --model d/model.npz --points 1000000 --out d/screen.npz
    python -m flysim_amd.sweep report --train d/train.npz --model d/model.npz --screen d/screen*.npz

A sweep point's rates are what one run of the network with those external
rates and gains gives over a 200 ms window after 100 ms of warm-up (dt 0.1 ms).
"""
from __future__ import annotations

import argparse
import glob
import json
import sys
import time
from typing import Dict

import numpy as np

from . import workloads as W
from .surrogate import RateSurrogate, SweepSimulator, SweepSpace, screen_sweep

WARM_STEPS = 1000
MEAS_STEPS = 2000
THRESHOLD_HZ = 50.0


def network():
    return W.c2_network(seed=2)


def make_sim(replicas: int, warm: int = WARM_STEPS, meas: int = MEAS_STEPS,
             schedule="auto") -> SweepSimulator:
    """The sweep's simulator; its launch schedule autotuned unless given."""
    return SweepSimulator(network(), replicas=replicas, warm_steps=warm, meas_steps=meas,
                          schedule=schedule)


def gen(points: int, seed: int, replicas: int, warm: int = WARM_STEPS, meas: int = MEAS_STEPS,
        log=print) -> Dict[str, np.ndarray]:
    space = SweepSpace()
    X = space.sample(points, np.random.default_rng(seed))
    sim = make_sim(replicas, warm, meas)
    Y = np.zeros((points, sim.n_mod), np.float32)
    chunk = replicas * 16
    t0 = time.perf_counter()
    for a in range(0, points, chunk):
        b = min(points, a + chunk)
        Y[a:b] = sim.run(X[a:b], seed=seed * 7919 + a)
        log("gen %d/%d points, %.1f points/s" % (b, points, b / (time.perf_counter() - t0)))
    return {"X": X, "Y": Y, "wall_s": np.float64(time.perf_counter() - t0),
            "points_per_s": np.float64(sim.points_per_s), "replicas": np.int64(replicas),
            "warm": np.int64(warm), "meas": np.int64(meas)}


def train(X: np.ndarray, Y: np.ndarray, epochs: int = 60, batch: int = 512, lr: float = 1e-3,
          holdout: float = 0.1, seed: int = 0, decay=(1.0, 0.3, 0.1)) -> Dict[str, np.ndarray]:
    n = X.shape[0]
    rng = np.random.default_rng(seed)
    idx = rng.permutation(n)
    n_val = int(round(holdout * n))
    val, tr = idx[:n_val], idx[n_val:]
    sur = RateSurrogate(X.shape[1], Y.shape[1], seed=seed)
    t0 = time.perf_counter()
    losses = sur.fit(X[tr], Y[tr], epochs=epochs, batch=min(batch, len(tr)), lr=lr, seed=seed, decay=decay)
    t_fit = time.perf_counter() - t0
    out = {"params": sur.mlp.params, "mu": sur.mu, "sd": sur.sd, "losses": losses,
           "fit_s": np.float64(t_fit), "val_idx": val}
    if n_val:
        P = sur.predict(X[val])
        err = np.abs(P - Y[val])
        mt, mp = Y[val].mean(axis=1) > THRESHOLD_HZ, P.mean(axis=1) > THRESHOLD_HZ
        out.update(val_mae_hz=np.float64(err.mean()),
                   val_mae_sub_hz=np.float64(err[~mt].mean() if (~mt).any() else 0.0),
                   val_acc=np.float64((mt == mp).mean()))
    return out


def load_surrogate(model: Dict[str, np.ndarray], n_in: int, n_out: int) -> RateSurrogate:
    sur = RateSurrogate(n_in, n_out)
    sur.mlp.params = model["params"]
    sur.mu, sur.sd = model["mu"].astype(np.float32), model["sd"].astype(np.float32)
    return sur


def screen(model: Dict[str, np.ndarray], points: int, seed: int, replicas: int,
           margin_hz: float = 10.0, audit_frac: float = 0.01, part: int = 0, parts: int = 1,
           log=print) -> Dict[str, np.ndarray]:
    """Screen `points` sweep points; simulate part `part` of `parts` of the
    flagged + audit list."""
    space = SweepSpace()
    X = space.sample(points, np.random.default_rng(seed))
    sur = load_surrogate(model, X.shape[1], space.n_mod)
    t0 = time.perf_counter()
    pred = sur.predict(X)
    t_pred = time.perf_counter() - t0
    mean = pred.mean(axis=1)
    flagged = np.nonzero(mean > THRESHOLD_HZ - margin_hz)[0]
    rest = np.setdiff1d(np.arange(points), flagged)
    rng = np.random.default_rng(seed + 1)
    audited = np.sort(rng.choice(rest, size=int(round(audit_frac * rest.size)), replace=False))
    todo = np.concatenate([flagged, audited])
    mine = np.array_split(todo, parts)[part]
    sim = make_sim(replicas)
    rates = np.zeros((mine.size, space.n_mod), np.float32)
    chunk = replicas * 16
    t1 = time.perf_counter()
    for a in range(0, mine.size, chunk):
        b = min(mine.size, a + chunk)
        rates[a:b] = sim.run(X[mine[a:b]], seed=seed * 7919 + int(mine[a]))
        log("screen part %d/%d: verified %d/%d points, %.1f points/s"
            % (part + 1, parts, b, mine.size, b / (time.perf_counter() - t1)))
    t_sim = time.perf_counter() - t1
    # rates and predictions stored as float16 (11-bit mantissa: ~0.05 %)
    return {"points": np.int64(points), "seed": np.int64(seed), "flagged": flagged,
            "audited": audited, "idx": mine, "rates": rates.astype(np.float16),
            "pred": pred[mine].astype(np.float16),
            "pred_mean_all": (mean if part == 0 else mean[:0]).astype(np.float32),
            "t_predict_s": np.float64(t_pred),
            "predict_gpu_ms": np.float64(sur.mlp.last_ms), "t_sim_s": np.float64(t_sim),
            "sim_points_per_s": np.float64(mine.size / t_sim if t_sim > 0 else 0.0)}


def band(model: Dict[str, np.ndarray], points: int, seed: int, replicas: int, lo: float = 45.0,
         hi: float = 60.0, audit_frac: float = 0.01, log=print) -> Dict[str, np.ndarray]:
    """The threshold-locating screen, run end to end: predict every point,
    simulate the points whose predicted mean rate lies in the uncertain band
    (lo, hi] around the threshold plus a random audit of `audit_frac` of the
    others, and classify every point -- by its simulated rate where it was
    simulated, by the surrogate elsewhere (hyperactive above the band, quiet
    below).  The audit measures how often the surrogate's call outside the
    band is wrong."""
    space = SweepSpace()
    X = space.sample(points, np.random.default_rng(seed))
    sur = load_surrogate(model, X.shape[1], space.n_mod)
    t0 = time.perf_counter()
    pred = sur.predict(X)
    t_pred = time.perf_counter() - t0
    mean = pred.mean(axis=1)
    in_band = (mean > lo) & (mean <= hi)
    bidx = np.nonzero(in_band)[0]
    rest = np.nonzero(~in_band)[0]
    rng = np.random.default_rng(seed + 1)
    aidx = np.sort(rng.choice(rest, size=int(round(audit_frac * rest.size)), replace=False))
    todo = np.concatenate([bidx, aidx])
    sim = make_sim(replicas)
    rates = np.zeros((todo.size, space.n_mod), np.float32)
    chunk = replicas * 16
    t1 = time.perf_counter()
    for a in range(0, todo.size, chunk):
        b = min(todo.size, a + chunk)
        rates[a:b] = sim.run(X[todo[a:b]], seed=seed * 7919 + int(todo[a]))
        log("band: simulated %d/%d points, %.1f points/s" % (b, todo.size, b / (time.perf_counter() - t1)))
    t_sim = time.perf_counter() - t1
    call = mean > THRESHOLD_HZ                        # the surrogate's call
    sim_mean = rates.mean(axis=1)
    call[todo] = sim_mean > THRESHOLD_HZ              # the simulator's, where it ran
    return {"points": np.int64(points), "seed": np.int64(seed), "band_hz": np.array([lo, hi]),
            "band_idx": bidx, "audit_idx": aidx, "rates": rates.astype(np.float16),
            "pred": pred[todo].astype(np.float16), "pred_mean_all": mean.astype(np.float32),
            "hyperactive": call, "t_predict_s": np.float64(t_pred), "t_sim_s": np.float64(t_sim),
            "predict_gpu_ms": np.float64(sur.mlp.last_ms),
            "sim_points_per_s": np.float64(todo.size / t_sim if t_sim > 0 else 0.0)}


def band_report(train_meta: Dict, model: Dict[str, np.ndarray], bd: Dict[str, np.ndarray],
                plain_pts_s: float) -> Dict:
    """Wall time of the end-to-end band screen (training data, fit,
    prediction, simulation of the band and the audit) against simulating
    every point plainly, and the surrogate's wrong calls outside the band as
    the audit measures them."""
    points = int(bd["points"])
    lo, hi = (float(v) for v in bd["band_hz"])
    nb, na = bd["band_idx"].size, bd["audit_idx"].size
    pm = bd["pred_mean_all"].astype(np.float64)
    aud_sim = bd["rates"][nb:].astype(np.float64).mean(axis=1)
    aud_pm = pm[bd["audit_idx"]]
    above, below = aud_pm > hi, aud_pm <= lo
    false_hyper = int((aud_sim[above] <= THRESHOLD_HZ).sum())
    missed = int((aud_sim[below] > THRESHOLD_HZ).sum())
    n_above, n_below = int((pm > hi).sum()), int((pm <= lo).sum())
    t_gen, t_fit = float(train_meta["wall_s"]), float(model["fit_s"])
    t_pred, t_sim = float(bd["t_predict_s"]), float(bd["t_sim_s"])
    t_total = t_gen + t_fit + t_pred + t_sim
    t_plain = points / plain_pts_s
    return {"points": points, "threshold_hz": THRESHOLD_HZ, "band_hz": [lo, hi],
            "band_points": nb, "band_frac": nb / points, "audit_points": na,
            "hyperactive_points": int(np.asarray(bd["hyperactive"]).sum()),
            "above_band": n_above, "below_band": n_below,
            "audit_above": int(above.sum()), "audit_false_hyperactive": false_hyper,
            "audit_below": int(below.sum()), "audit_missed_hyperactive": missed,
            "est_wrong_calls_above": false_hyper / max(1, int(above.sum())) * n_above,
            "est_wrong_calls_below": missed / max(1, int(below.sum())) * n_below,
            "train_points": int(train_meta["points"]), "train_gen_s": t_gen, "fit_s": t_fit,
            "predict_s": t_pred, "predict_gpu_ms": float(bd["predict_gpu_ms"]),
            "verify_s": t_sim, "verify_points_per_s": float(bd["sim_points_per_s"]),
            "total_s": t_total, "plain_points_per_s": plain_pts_s, "plain_total_s": t_plain,
            "speedup_vs_plain": t_plain / t_total,
            "speedup_vs_batched_all": points / float(bd["sim_points_per_s"]) / t_total}


def module_error(model: Dict[str, np.ndarray], points: int, seed: int, replicas: int, repeats: int = 8,
                 log=print) -> Dict:
    """Per-module error of the surrogate against the simulator on held-out
    points: each point is simulated `repeats` times with independent input
    streams, and the surrogate's prediction is compared with the mean of the
    runs (the quantity it estimates); one run's deviation from that mean is
    the simulator's own noise at this protocol."""
    space = SweepSpace()
    X = space.sample(points, np.random.default_rng(seed))
    sur = load_surrogate(model, X.shape[1], space.n_mod)
    pred = sur.predict(X).astype(np.float64)
    sim = make_sim(replicas)
    runs = np.zeros((repeats, points, space.n_mod))
    for k in range(repeats):
        runs[k] = sim.run(X, seed=seed * 1000 + 17 * k + 1)
        log("module error: repeat %d/%d" % (k + 1, repeats))
    ref = runs.mean(axis=0)
    err = np.abs(pred - ref)
    noise = np.abs(runs[0] - ref) * np.sqrt(repeats / (repeats - 1.0))
    sub = ref.mean(axis=1) <= THRESHOLD_HZ
    out = {"points": points, "repeats": repeats,
           "module_mae_hz": err.mean(axis=0).tolist(),
           "module_mae_hz_max": float(err.mean(axis=0).max()),
           "mae_hz": float(err.mean()),
           "subthreshold_points": int(sub.sum()),
           "module_mae_hz_max_subthreshold": float(err[sub].mean(axis=0).max()) if sub.any() else None,
           "mae_hz_subthreshold": float(err[sub].mean()) if sub.any() else None,
           "one_run_noise_mae_hz": float(noise.mean()),
           "one_run_noise_module_max_hz": float(noise.mean(axis=0).max()),
           "modules_within_2hz": int((err.mean(axis=0) <= 2.0).sum())}
    return out


def predict_rate(points: int, repeats: int = 5) -> Dict:
    """GPU time of the surrogate's forward over `points` sweep points already
    in device memory (an untrained network of the sweep's shape): best of
    `repeats` runs, and the matrix-core rate it reaches."""
    from ._capi import DeviceBuffer
    from .surrogate import Mlp
    space = SweepSpace()
    sizes = [space.dim, 256, 256, 256, space.n_mod]
    mlp = Mlp(sizes, seed=0)
    x = np.ascontiguousarray(space.sample(points, np.random.default_rng(7)), dtype=np.float32)
    xd, yd = DeviceBuffer(x.nbytes), DeviceBuffer(points * space.n_mod * 4)
    xd.upload(x)
    mlp.predict_device(xd.ptr, min(points, 4096), yd.ptr)
    ms = min(mlp.predict_device(xd.ptr, points, yd.ptr) for _ in range(repeats))
    flop = 2.0 * sum(sizes[i] * sizes[i + 1] for i in range(len(sizes) - 1)) * points
    return {"points": points, "gpu_ms": ms, "points_per_s": points / (ms * 1e-3),
            "tflops": flop / (ms * 1e-3) / 1e12}


def plain_rate(points: int, seed: int) -> float:
    """Points/s of plain per-instance GPU simulation (one network per run)."""
    X = SweepSpace().sample(points, np.random.default_rng(seed))
    sim = make_sim(1)
    sim.run(X[:1], seed=seed)     # warm-up (graph capture)
    sim.points, sim.wall_seconds = 0, 0.0
    sim.run(X, seed=seed)
    return sim.points_per_s


def band_design(screens, t_fixed_s: float, plain_pts_s: float, lo: float = 45.0, hi: float = 60.0,
                audit_frac: float = 0.01) -> Dict:
    """The threshold-locating variant of the screen, evaluated on the data of
    the conservative one: simulate only the points whose predicted mean rate
    lies in the uncertain band (lo, hi] around the 50 Hz threshold, plus an
    audit of `audit_frac` of the others, and let the surrogate classify the
    rest (hyperactive above the band, quiet below).  Every point the
    conservative screen flagged (predicted above the threshold minus its
    margin) was simulated, so the surrogate's misclassifications above the
    band are counted exactly; below it they are counted where simulated and
    estimated from the audit elsewhere.  Cost: the fixed part (training data,
    fit, prediction) plus the band and its audit at the measured
    verification rate."""
    points = int(screens[0]["points"])
    pm = screens[0]["pred_mean_all"].astype(np.float64)
    if pm.size != points:
        raise ValueError("the band design needs part 0 of the screen (all predictions)")
    flag_lo = float(np.min(pm[screens[0]["flagged"]])) if screens[0]["flagged"].size else np.inf
    idx = np.concatenate([s["idx"] for s in screens])
    sim = np.full(points, np.nan)
    sim[idx] = np.concatenate([s["rates"] for s in screens]).astype(np.float64).mean(axis=1)
    t_sim = sum(float(s["t_sim_s"]) for s in screens)
    rate = idx.size / t_sim
    band = (pm > lo) & (pm <= hi)
    n_band = int(band.sum())
    n_audit = int(round(audit_frac * (points - n_band)))
    t_verify = (n_band + n_audit) / rate
    above = pm > hi
    seen_above = above & ~np.isnan(sim)
    fp = int((sim[seen_above] <= THRESHOLD_HZ).sum())
    below = pm <= lo
    seen_below = below & ~np.isnan(sim)
    below_flagged = below & (pm >= flag_lo)            # simulated in full by the conservative screen
    below_audit = below & ~below_flagged & ~np.isnan(sim)
    fn_seen = int((sim[below_flagged & seen_below] > THRESHOLD_HZ).sum())
    aud_rate = float((sim[below_audit] > THRESHOLD_HZ).mean()) if below_audit.any() else 0.0
    n_below_unflagged = int((below & ~below_flagged).sum())
    t_total = t_fixed_s + t_verify
    return {"band_hz": [lo, hi], "band_points": n_band, "band_frac": n_band / points,
            "audit_points": n_audit, "verify_s": t_verify, "total_s": t_total,
            "speedup_vs_plain": (points / plain_pts_s) / t_total,
            "above_band": int(above.sum()), "above_band_simulated": int(seen_above.sum()),
            "false_hyperactive_above": fp,
            "below_band": int(below.sum()), "below_band_simulated": int(seen_below.sum()),
            "missed_hyperactive_below_simulated": fn_seen,
            "audit_miss_rate_below": aud_rate,
            "est_missed_below_total": fn_seen + aud_rate * n_below_unflagged}


def report(train_d, model, screens, plain_pts_s: float, batched_pts_s: float) -> Dict:
    points = int(screens[0]["points"])
    flagged = screens[0]["flagged"]
    audited = screens[0]["audited"]
    idx = np.concatenate([s["idx"] for s in screens])
    rates = np.concatenate([s["rates"] for s in screens]).astype(np.float32)
    pred = np.concatenate([s["pred"] for s in screens]).astype(np.float32)
    is_f = np.isin(idx, flagged)
    sim_mean = rates.mean(axis=1)
    hyper = sim_mean > THRESHOLD_HZ
    err = np.abs(pred - rates)
    n_todo = flagged.size + audited.size
    t_sim = sum(float(s["t_sim_s"]) for s in screens)
    verified = idx.size
    # time of the full screen: prediction + simulation of every flagged and
    # audited point (measured when all parts ran, else at the measured rate)
    t_verify = t_sim if verified == n_todo else n_todo / (verified / t_sim)
    t_gen = float(train_d["wall_s"])
    t_fit = float(model["fit_s"])
    t_pred = float(screens[0]["t_predict_s"])
    t_total = t_gen + t_fit + t_pred + t_verify
    t_plain = points / plain_pts_s
    t_batched = points / batched_pts_s
    audit_hyper = hyper[~is_f]
    rep = {
        "points": points, "threshold_hz": THRESHOLD_HZ,
        "flagged": int(flagged.size), "audited": int(audited.size),
        "verified_this_run": int(verified), "to_verify": int(n_todo),
        "hyperactive_among_flagged": int(hyper[is_f].sum()),
        "precision_flagged": float(hyper[is_f].mean()) if is_f.any() else None,
        "audit_hyperactive_missed": int(audit_hyper.sum()),
        "audit_miss_rate": float(audit_hyper.mean()) if audit_hyper.size else None,
        "est_missed_total": float(audit_hyper.mean() * (points - flagged.size)) if audit_hyper.size else None,
        "surrogate_mae_hz_all": float(err.mean()),
        "surrogate_mae_hz_subthreshold": float(err[~hyper].mean()) if (~hyper).any() else None,
        "surrogate_module_mae_hz_subthreshold_max": float(err[~hyper].mean(axis=0).max()) if (~hyper).any() else None,
        "train_points": int(train_d["X"].shape[0]), "train_gen_s": t_gen, "fit_s": t_fit,
        "val_mae_hz": float(model.get("val_mae_hz", np.nan)),
        "predict_s": t_pred, "predict_gpu_ms": float(screens[0]["predict_gpu_ms"]),
        "verify_s": t_verify, "total_s": t_total,
        "plain_points_per_s": plain_pts_s, "batched_points_per_s": batched_pts_s,
        "plain_total_s": t_plain, "batched_total_s": t_batched,
        "speedup_vs_plain": t_plain / t_total, "speedup_vs_plain_excl_training": t_plain / (t_total - t_gen - t_fit),
        "speedup_vs_batched": t_batched / t_total,
        "effective_points_per_s": points / t_total,
    }
    if screens[0]["pred_mean_all"].size == points:
        rep["band_design"] = band_design(screens, t_gen + t_fit + t_pred, plain_pts_s)
    return rep


def main(argv=None):
    ap = argparse.ArgumentParser(prog="python -m flysim_amd.sweep")
    sub = ap.add_subparsers(dest="cmd", required=True)
    g = sub.add_parser("gen")
    g.add_argument("--points", type=int, default=100000)
    g.add_argument("--seed", type=int, default=1)
    g.add_argument("--replicas", type=int, default=128)
    g.add_argument("--out", required=True)
    t = sub.add_parser("train")
    t.add_argument("--data", required=True)
    t.add_argument("--epochs", type=int, default=60)
    t.add_argument("--batch", type=int, default=512)
    t.add_argument("--lr", type=float, default=1e-3)
    t.add_argument("--decay", default="1,0.3,0.1",
                   help="learning-rate factors, equal shares of the epochs each")
    t.add_argument("--out", required=True)
    s = sub.add_parser("screen")
    s.add_argument("--model", required=True)
    s.add_argument("--points", type=int, default=1000000)
    s.add_argument("--seed", type=int, default=2)
    s.add_argument("--replicas", type=int, default=128)
    s.add_argument("--part", type=int, default=0)
    s.add_argument("--parts", type=int, default=1)
    s.add_argument("--out", required=True)
    bd = sub.add_parser("band")
    bd.add_argument("--model", required=True)
    bd.add_argument("--points", type=int, default=1000000)
    bd.add_argument("--seed", type=int, default=2)
    bd.add_argument("--replicas", type=int, default=128)
    bd.add_argument("--lo", type=float, default=45.0)
    bd.add_argument("--hi", type=float, default=60.0)
    bd.add_argument("--out", required=True)
    br = sub.add_parser("bandreport")
    br.add_argument("--train", required=True, help="json written by the gen command")
    br.add_argument("--model", required=True)
    br.add_argument("--band", required=True)
    br.add_argument("--plain", required=True)
    br.add_argument("--moderr", default="", help="json written by the moderr command")
    br.add_argument("--out", default="")
    me = sub.add_parser("moderr")
    me.add_argument("--model", required=True)
    me.add_argument("--points", type=int, default=1024)
    me.add_argument("--seed", type=int, default=5)
    me.add_argument("--replicas", type=int, default=128)
    me.add_argument("--repeats", type=int, default=8)
    me.add_argument("--out", required=True)
    pr = sub.add_parser("predict")
    pr.add_argument("--points", type=int, default=1 << 20)
    pr.add_argument("--repeats", type=int, default=5)
    pl = sub.add_parser("plain")
    pl.add_argument("--points", type=int, default=64)
    pl.add_argument("--out", required=True)
    r = sub.add_parser("report")
    r.add_argument("--train", required=True)
    r.add_argument("--model", required=True)
    r.add_argument("--screen", required=True, nargs="+")
    r.add_argument("--plain", required=True, help="json written by the plain command")
    r.add_argument("--out", default="")
    a = ap.parse_args(argv)
    log = lambda m: print(m, flush=True)
    if a.cmd == "gen":
        d = gen(a.points, a.seed, a.replicas, log=log)
        np.savez(a.out, **d)
        meta = {k: float(v) for k, v in d.items() if np.ndim(v) == 0}
        meta["points"] = int(d["X"].shape[0])
        json.dump(meta, open(a.out[:-4] + ".json" if a.out.endswith(".npz") else a.out + ".json", "w"))
    elif a.cmd == "train":
        d = np.load(a.data)
        m = train(d["X"], d["Y"], epochs=a.epochs, batch=a.batch, lr=a.lr,
                  decay=tuple(float(v) for v in a.decay.split(",")))
        log("trained: final loss %.4g, holdout MAE %.2f Hz, fit %.1f s"
            % (m["losses"][-1], m.get("val_mae_hz", np.nan), m["fit_s"]))
        np.savez(a.out, **m)
    elif a.cmd == "screen":
        m = dict(np.load(a.model))
        np.savez(a.out, **screen(m, a.points, a.seed, a.replicas, part=a.part, parts=a.parts, log=log))
    elif a.cmd == "band":
        m = dict(np.load(a.model))
        np.savez(a.out, **band(m, a.points, a.seed, a.replicas, a.lo, a.hi, log=log))
    elif a.cmd == "moderr":
        m = dict(np.load(a.model))
        json.dump(module_error(m, a.points, a.seed, a.replicas, a.repeats, log=log), open(a.out, "w"), indent=1)
    elif a.cmd == "bandreport":
        rep = band_report(json.load(open(a.train)), dict(np.load(a.model)), dict(np.load(a.band)),
                          float(json.load(open(a.plain))["points_per_s"]))
        if a.moderr:
            rep["module_error"] = {k: v for k, v in json.load(open(a.moderr)).items() if k != "module_mae_hz"}
        txt = json.dumps(rep, indent=1)
        print(txt)
        if a.out:
            open(a.out, "w").write(txt)
    elif a.cmd == "predict":
        print(json.dumps(predict_rate(a.points, a.repeats)))
    elif a.cmd == "plain":
        rate = plain_rate(a.points, 99)
        log("plain per-instance simulation: %.2f points/s" % rate)
        json.dump({"points": a.points, "points_per_s": rate}, open(a.out, "w"))
    else:
        if a.train.endswith(".json"):
            meta = json.load(open(a.train))
            tr = {"X": np.zeros((int(meta["points"]), 0)), "wall_s": meta["wall_s"]}
        else:
            tr = np.load(a.train)
        m = dict(np.load(a.model))
        files = sorted(set(f for p in a.screen for f in glob.glob(p)))
        scr = [dict(np.load(f)) for f in files]
        plain = float(json.load(open(a.plain))["points_per_s"])
        batched = float(np.mean([float(s["sim_points_per_s"]) for s in scr]))
        rep = report(tr, m, scr, plain, batched)
        txt = json.dumps(rep, indent=1)
        print(txt)
        if a.out:
            open(a.out, "w").write(txt)


if __name__ == "__main__":
    main()
