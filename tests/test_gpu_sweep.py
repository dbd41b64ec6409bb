"""GPU tests of surrogate-screened sweeps on the 20k-neuron fly-like network.

The surrogate (BF16 MLP trained on the GPU on rates the simulator itself
produced) screens sweep points for the hyperactive regime (mean rate > 50
Hz); the simulator is the judge.  Ablation: against simulating every point
(the method without the network) screening must keep recall >= 0.9 of the
hyperactive points while simulating well under all of them; the trained
network must beat the same network untrained and a least-squares linear
model on the same training set.  Short windows (30 + 50 ms) keep the test
fast; the full protocol is flysim_amd.sweep.
"""
import numpy as np
import pytest

from flysim_amd.surrogate import RateSurrogate, SweepSpace
from flysim_amd.sweep import THRESHOLD_HZ, make_sim

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def data():
    print("case sweep: 8192 training + 1024 test points, 128 replicas per launch", flush=True)
    sim = make_sim(128, warm=300, meas=500)
    space = SweepSpace()
    Xtr = space.sample(8192, np.random.default_rng(1))
    Ytr = sim.run(Xtr, seed=1)
    Xte = space.sample(1024, np.random.default_rng(2))
    Yte = sim.run(Xte, seed=2)
    return Xtr, Ytr, Xte, Yte


def test_screening_recall_and_savings(data):
    Xtr, Ytr, Xte, Yte = data
    sur = RateSurrogate(Xtr.shape[1], Ytr.shape[1], seed=0)
    losses = sur.fit(Xtr, Ytr, epochs=300, batch=128, lr=2e-3)
    assert losses[-1] < 0.5 * losses[0]
    hyper = Yte.mean(axis=1) > THRESHOLD_HZ
    assert 0.05 < hyper.mean() < 0.8
    pred = sur.predict(Xte)
    flag = pred.mean(axis=1) > THRESHOLD_HZ - 10.0
    recall = (flag & hyper).sum() / hyper.sum()
    frac = flag.mean()
    mae = np.abs(pred - Yte).mean()
    # untrained network, same scaling
    raw = RateSurrogate(Xtr.shape[1], Ytr.shape[1], seed=0)
    raw.mu, raw.sd = sur.mu, sur.sd
    pu = raw.predict(Xte)
    mae_u = np.abs(pu - Yte).mean()
    fu = pu.mean(axis=1) > THRESHOLD_HZ - 10.0
    recall_u = (fu & hyper).sum() / hyper.sum()
    # least-squares linear model on the same training set
    A = np.c_[Xtr, np.ones(len(Xtr))]
    coef, *_ = np.linalg.lstsq(A, Ytr, rcond=None)
    pl = np.c_[Xte, np.ones(len(Xte))] @ coef
    mae_l = np.abs(pl - Yte).mean()
    fl = pl.mean(axis=1) > THRESHOLD_HZ - 10.0
    recall_l = (fl & hyper).sum() / hyper.sum()
    print("hyperactive %.3f | trained: recall %.3f, simulated %.3f, MAE %.2f Hz | untrained: recall "
          "%.3f, MAE %.2f | linear: recall %.3f, simulated %.3f, MAE %.2f"
          % (hyper.mean(), recall, frac, mae, recall_u, mae_u, recall_l, fl.mean(), mae_l))
    assert recall >= 0.9
    assert frac < 0.75                      # vs 1.0 for simulating every point
    assert frac < hyper.mean() + 0.2        # few flags beyond the true positives
    assert mae < 0.25 * mae_u
    assert mae < 0.5 * mae_l
