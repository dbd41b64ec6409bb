"""The autotuner's host logic on a simulated engine (no GPU): candidates
follow the network's statistics, the cache key changes with anything that
can change the best schedule, and ranking by the ratio to the starting
schedule picks the truly fastest candidate although the network's activity
(and so every step's cost) triples halfway through the trials."""
import pytest

from flysim_amd import tune

BASE = dict(update_block=256, gather_lanes=8, push_block=256, steps_per_graph=100, use_graph=1,
            fused=0, window=11)
STATS = dict(neurons=20000, pops=100, replicas=1, syn_lin=400000, syn_nmda=430000, min_delay=10,
             max_delay_lin=20, max_delay_nmda=20, precision=0, ext_mode=1)


class FakeEngine:
    """Per-step cost = cost(schedule) x activity, where activity steps from 1
    to 3 once `ignite_at` steps have run; time is a virtual clock."""

    def __init__(self, cost, ignite_at, stats=STATS, base=BASE):
        self.cost, self.ignite_at, self._stats = cost, ignite_at, dict(stats)
        self._sched = dict(base)
        self.clock, self.steps = 0.0, 0

    @property
    def schedule(self):
        return dict(self._sched)

    def stats(self):
        return dict(self._stats)

    def set_schedule(self, **kw):
        bad = set(kw) - set(self._sched)
        if bad:
            raise ValueError("unknown schedule keys %s" % bad)
        if kw.get("gather_lanes", 1) > 32:
            raise ValueError("gather_lanes out of range")
        self._sched.update(kw)

    def run(self, n, record=False):
        for _ in range(n):
            act = 3.0 if self.steps >= self.ignite_at else 1.0
            self.clock += self.cost(self._sched) * act
            self.steps += 1


@pytest.fixture
def clocked(monkeypatch):
    holder = {}
    monkeypatch.setattr(tune.time, "perf_counter", lambda: holder["eng"].clock)
    monkeypatch.setattr(tune, "device_info", lambda: ("AMD Instinct MI355X", 256))
    return holder


def test_ratio_ranking_survives_an_activity_jump(clocked):
    # candidate i costs 1.0 - 0.01 i: the last is the fastest, but the
    # network ignites before it runs
    cands = [dict(gather_lanes=2 ** i) for i in range(6)]
    cost = lambda s: 1e-6 * (1.0 - 0.01 * s["gather_lanes"].bit_length())
    eng = FakeEngine(cost, ignite_at=1500)
    clocked["eng"] = eng
    res = tune.tune(eng, steps=100, cands=cands, cache=None, keep=3)
    assert res["schedule"]["gather_lanes"] == 32
    assert eng.schedule["gather_lanes"] == 32
    assert res["vs_start"] == pytest.approx(0.94 / 0.96)
    assert len(res["trials"]) == 6
    for t in res["trials"]:   # the trial across the jump was repeated, not kept
        assert t["vs_start"] == pytest.approx(cost(t["schedule"]) / cost(BASE), rel=0.1)
    # ranked by absolute first-pass times the early, pre-ignition candidates
    # would have won
    assert min(res["trials"], key=lambda t: t["s_per_step"])["schedule"]["gather_lanes"] < 32


def test_invalid_candidates_are_skipped(clocked):
    eng = FakeEngine(lambda s: 1e-6 / s["gather_lanes"], ignite_at=10 ** 9)
    clocked["eng"] = eng
    res = tune.tune(eng, cands=[dict(gather_lanes=64), dict(gather_lanes=16), dict(bogus=1)], cache=None)
    assert res["schedule"]["gather_lanes"] == 16
    assert [t["schedule"]["gather_lanes"] for t in res["trials"]] == [16]


def test_cache_hit_applies_without_timing(clocked, tmp_path):
    cache = str(tmp_path / "c.json")
    eng = FakeEngine(lambda s: 1e-6 / s["gather_lanes"], ignite_at=10 ** 9)
    clocked["eng"] = eng
    tune.tune(eng, cands=[dict(gather_lanes=4), dict(gather_lanes=16)], cache=cache)
    e2 = FakeEngine(lambda s: 1e-6, ignite_at=10 ** 9)
    clocked["eng"] = e2
    r2 = tune.tune(e2, cands=[dict(gather_lanes=4)], cache=cache)
    assert "trials" not in r2 and e2.steps == 0
    assert e2.schedule["gather_lanes"] == 16


def test_signature_separates_what_can_change_the_schedule(clocked):
    clocked["eng"] = FakeEngine(lambda s: 1e-6, 0)
    sigs = set()
    for key, val in [(None, None), ("min_delay", 5), ("max_delay_lin", 40), ("max_delay_nmda", 40),
                     ("syn_nmda", 1), ("syn_lin", 1), ("replicas", 128), ("precision", 1), ("ext_mode", 0),
                     ("neurons", 20001), ("pops", 7)]:
        st = dict(STATS)
        if key:
            st[key] = val
        sigs.add(tune.signature(FakeEngine(lambda s: 1e-6, 0, st)))
    assert len(sigs) == 11


def test_candidates_follow_the_network():
    delayed = tune.candidates(STATS)
    assert any(c["window"] for c in delayed) and any(c["fused"] for c in delayed)
    undelayed = tune.candidates(dict(STATS, min_delay=0))
    assert not any(c["fused"] for c in undelayed)
    assert any(c["window"] for c in undelayed)  # one step per launch
    assert any(not c["window"] and not c["fused"] for c in undelayed)
    reps = tune.candidates(dict(STATS, replicas=128))
    assert {c["gather_lanes"] for c in reps} <= {1, 2}
    for c in delayed:
        assert c["gather_lanes"] * 2 <= c["update_block"] or not c["window"]


def test_windowed_refinements_are_tried(clocked):
    # the knobs only the windowed path on one network copy has are varied
    # around the best first-pass schedules
    base = dict(BASE, xcd_order=0, gather_chunk=0)
    fast = lambda s: s.get("gather_chunk") == 128 and s.get("xcd_order") == 1
    cost = lambda s: 1e-6 * (0.5 if fast(s) else 1.0) * (1.0 - 0.01 * s["gather_lanes"].bit_length())
    eng = FakeEngine(cost, 10 ** 9, base=base)
    clocked["eng"] = eng
    cands = [dict(gather_lanes=4), dict(gather_lanes=16)]
    monkey = tune.candidates
    tune.candidates = lambda st: cands
    try:
        res = tune.tune(eng, cache=None)
    finally:
        tune.candidates = monkey
    assert res["schedule"]["gather_chunk"] == 128 and res["schedule"]["xcd_order"] == 1
    assert res["schedule"]["gather_lanes"] == 16
    # replicas have no such variants
    assert tune.refinements(base, dict(STATS, replicas=128)) == []
    assert tune.refinements(dict(base, window=0), STATS) == []
