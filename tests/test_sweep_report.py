"""CPU test of the sweep report's bookkeeping (flysim_amd.sweep.report) on
synthetic screen slices: counts, precision, audit miss rate and the time
accounting of plain vs batched vs surrogate-screened sweeps."""
import numpy as np

from flysim_amd.sweep import band_design, report


def test_report_accounting():
    rng = np.random.default_rng(0)
    points = 1000
    flagged = np.arange(0, 400)
    audited = np.arange(900, 906)
    todo = np.concatenate([flagged, audited])
    parts = np.array_split(todo, 2)
    rates = {int(i): np.full(50, 80.0 if i < 350 else 20.0, np.float32) for i in todo}
    rates[903] = np.full(50, 70.0, np.float32)          # one audited miss
    screens = []
    for k, idx in enumerate(parts):
        r = np.stack([rates[int(i)] for i in idx])
        screens.append({"points": np.int64(points), "flagged": flagged, "audited": audited,
                        "idx": idx, "rates": r.astype(np.float16),
                        "pred": (r + 1.0).astype(np.float16), "t_predict_s": np.float64(0.5),
                        "predict_gpu_ms": np.float64(3.0), "t_sim_s": np.float64(10.0),
                        "sim_points_per_s": np.float64(idx.size / 10.0),
                        "pred_mean_all": np.zeros(0, np.float32)})
    train_d = {"X": np.zeros((100, 54)), "wall_s": np.float64(30.0)}
    model = {"fit_s": np.float64(5.0), "val_mae_hz": np.float64(1.5)}
    rep = report(train_d, model, screens, plain_pts_s=10.0, batched_pts_s=20.0)
    assert rep["flagged"] == 400 and rep["audited"] == 6 and rep["verified_this_run"] == 406
    assert rep["hyperactive_among_flagged"] == 350
    assert abs(rep["precision_flagged"] - 350 / 400) < 1e-12
    assert rep["audit_hyperactive_missed"] == 1
    assert abs(rep["surrogate_mae_hz_all"] - 1.0) < 1e-3
    assert abs(rep["verify_s"] - 20.0) < 1e-9
    assert abs(rep["total_s"] - (30.0 + 5.0 + 0.5 + 20.0)) < 1e-9
    assert abs(rep["speedup_vs_plain"] - 100.0 / 55.5) < 1e-9
    assert "band_design" not in rep  # needs part 0's predictions of every point


def test_band_design_accounting():
    """Threshold-band screen: band (45, 60] simulated, the rest classified by
    the surrogate; misclassifications counted where simulated, estimated
    from the audit below the conservative screen's flag threshold."""
    points = 100
    pm = np.zeros(points, np.float32)
    pm[:10] = 100.0        # above the band
    pm[10:30] = 50.0       # in the band
    pm[30:40] = 42.0       # below the band, flagged by the conservative screen (> 40)
    pm[40:] = 10.0         # below, not flagged
    flagged = np.arange(0, 40)
    audited = np.array([50, 60, 70, 80])
    idx = np.concatenate([flagged, audited])
    sim = np.full(points, 20.0, np.float32)
    sim[:10] = 90.0
    sim[3] = 30.0          # one false hyperactive above the band
    sim[31] = 55.0         # one missed hyperactive among the simulated below-band points
    sim[60] = 70.0         # one audited miss: 1 of 4
    rates = np.repeat(sim[idx][:, None], 50, axis=1)
    scr = [{"points": np.int64(points), "flagged": flagged, "audited": audited, "idx": idx,
            "rates": rates.astype(np.float16), "pred_mean_all": pm, "t_sim_s": np.float64(4.4)}]
    b = band_design(scr, t_fixed_s=10.0, plain_pts_s=1.0, lo=45.0, hi=60.0, audit_frac=0.05)
    assert b["band_points"] == 20 and b["audit_points"] == 4
    assert abs(b["verify_s"] - 24 / 10.0) < 1e-9          # 44 points in 4.4 s
    assert abs(b["speedup_vs_plain"] - 100.0 / 12.4) < 1e-9
    assert b["false_hyperactive_above"] == 1 and b["above_band_simulated"] == 10
    assert b["missed_hyperactive_below_simulated"] == 1
    assert abs(b["audit_miss_rate_below"] - 0.25) < 1e-12
    assert abs(b["est_missed_below_total"] - (1 + 0.25 * 60)) < 1e-9


def test_band_report_end_to_end_accounting():
    """End-to-end band screen: band and audit simulated, wrong calls outside
    the band estimated from the audit on each side, the time of every stage
    (training data, fit, prediction, simulation) against plain simulation."""
    from flysim_amd.sweep import band_report
    points = 1000
    pm = np.full(points, 10.0, np.float32)
    pm[:300] = 90.0                       # above the band (45, 60]
    pm[300:350] = 50.0                    # in the band
    bidx = np.arange(300, 350)
    aidx = np.array([0, 1, 2, 3, 500, 501, 502, 503, 504])   # 4 above, 5 below
    sim_b = np.full(bidx.size, 52.0)
    sim_a = np.array([80.0, 80.0, 40.0, 80.0, 10.0, 10.0, 70.0, 10.0, 10.0])
    rates = np.repeat(np.concatenate([sim_b, sim_a])[:, None], 50, axis=1)
    hyper = pm > 50.0
    hyper[bidx] = True
    bd = {"points": np.int64(points), "band_hz": np.array([45.0, 60.0]), "band_idx": bidx,
          "audit_idx": aidx, "rates": rates.astype(np.float16), "pred_mean_all": pm,
          "hyperactive": hyper, "t_predict_s": np.float64(1.0), "t_sim_s": np.float64(5.9),
          "predict_gpu_ms": np.float64(2.0), "sim_points_per_s": np.float64(59 / 5.9)}
    rep = band_report({"wall_s": 20.0, "points": 100}, {"fit_s": np.float64(3.0)}, bd, plain_pts_s=2.0)
    assert rep["band_points"] == 50 and rep["audit_points"] == 9
    assert rep["audit_above"] == 4 and rep["audit_false_hyperactive"] == 1
    assert rep["audit_below"] == 5 and rep["audit_missed_hyperactive"] == 1
    assert abs(rep["est_wrong_calls_above"] - 300 / 4) < 1e-9
    assert abs(rep["est_wrong_calls_below"] - 650 / 5) < 1e-9
    assert abs(rep["total_s"] - (20.0 + 3.0 + 1.0 + 5.9)) < 1e-9
    assert abs(rep["speedup_vs_plain"] - 500.0 / 29.9) < 1e-9
    assert rep["hyperactive_points"] == 350
