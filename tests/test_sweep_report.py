"""CPU test of the sweep report's bookkeeping (flysim_amd.sweep.report) on
synthetic screen slices: counts, precision, audit miss rate and the time
accounting of plain vs batched vs surrogate-screened sweeps."""
import numpy as np

from flysim_amd.sweep import report


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
                        "sim_points_per_s": np.float64(idx.size / 10.0)})
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
