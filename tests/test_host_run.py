"""Host logic of a recorded run (no GPU): Engine.run splits a run into
chunks whose worst case (every neuron spiking every step) fits its id
buffer, hands fs_run that buffer's true capacity and stitches the per-step
offsets and ids of the chunks together."""
import ctypes as C

import numpy as np

from flysim_amd import _capi as A
from flysim_amd.engine import Engine


class FakeLib:
    """fs_run of a network in which neuron (step % n) spikes every step."""

    def __init__(self, n):
        self.n, self.step, self.calls = n, 0, []

    def fs_run(self, h, k, off_p, ids_p, cap):
        self.calls.append((k, cap))
        assert cap >= self.n * k              # the chunk's worst case fits
        off = np.ctypeslib.as_array(C.cast(off_p, C.POINTER(C.c_int64)), shape=(k + 1,))
        ids = np.ctypeslib.as_array(C.cast(ids_p, C.POINTER(C.c_uint32)), shape=(cap,))
        for j in range(k):
            ids[j] = (self.step + j) % self.n
            off[j + 1] = j + 1
        self.step += k
        return 0


def _engine(n, record_ids):
    eng = Engine.__new__(Engine)
    eng._h = C.c_void_p(1)
    eng.n = n
    eng.RECORD_IDS = record_ids
    return eng


def test_recorded_run_is_chunked_to_the_worst_case(monkeypatch):
    fake = FakeLib(n=1000)
    monkeypatch.setattr(A, "lib", lambda: fake)
    eng = _engine(1000, record_ids=1000 * 7)
    off, ids = eng.run(50)
    assert [k for k, _ in fake.calls] == [7] * 7 + [1]
    assert off.shape == (51,) and off[0] == 0
    np.testing.assert_array_equal(np.diff(off), 1)
    np.testing.assert_array_equal(ids, np.arange(50) % 1000)
    eng._h = None


def test_short_run_is_one_call(monkeypatch):
    fake = FakeLib(n=10)
    monkeypatch.setattr(A, "lib", lambda: fake)
    eng = _engine(10, record_ids=1 << 26)
    off, ids = eng.run(30)
    assert len(fake.calls) == 1 and fake.calls[0][1] >= 300
    np.testing.assert_array_equal(ids, np.arange(30) % 10)
    eng._h = None
