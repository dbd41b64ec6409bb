"""The allocation-free surrogate forward (fs_mlp_predict_noalloc) through the
C ABI: on two calls in a row, the second with fewer rows, its output equals
fs_mlp_predict's bit for bit, and device memory in use does not change
across its calls -- for the sweep's network (the fused kernel, no
intermediate memory) and for a wider one whose layer-by-layer GEMMs keep
their intermediates in the library's fixed device scratch array."""
import ctypes as C

import numpy as np
import pytest

from flysim_amd import _capi as A
from flysim_amd.surrogate import Mlp

pytestmark = pytest.mark.gpu

SIZES = [54, 256, 256, 256, 50]


def _noalloc():
    return A.lib().fs_mlp_predict_noalloc


def _hip_runtime():
    """The HIP runtime the library itself uses (the file mapped for it)."""
    A.lib()
    with open("/proc/self/maps") as f:
        paths = {line.split()[-1] for line in f if "libamdhip64.so" in line}
    assert paths, "HIP runtime not mapped"
    return C.CDLL(sorted(paths)[0])


def _device_free_bytes():
    hip = _hip_runtime()
    free, total = C.c_size_t(), C.c_size_t()
    assert hip.hipMemGetInfo(C.byref(free), C.byref(total)) == 0
    return free.value


@pytest.mark.parametrize("sizes", [SIZES, [54, 320, 320, 50]])
def test_noalloc_forward_equals_predict_twice(sizes):
    f = _noalloc()
    print("case noalloc forward sizes=%s" % sizes, flush=True)
    mlp = Mlp(sizes, seed=11)
    rng = np.random.default_rng(5)
    params = mlp.params.copy()
    params += rng.normal(0, 0.05, size=params.shape).astype(np.float32)  # non-zero biases too
    mlp.params = params
    n1, n2 = 20000, 777  # 20000 rows span several scratch chunks; 777 a partial tile
    x = rng.normal(size=(n1, sizes[0])).astype(np.float32)
    xd = A.DeviceBuffer(x.nbytes)
    yd_ref = A.DeviceBuffer(n1 * sizes[-1] * 4)
    yd = A.DeviceBuffer(n1 * sizes[-1] * 4)
    xd.upload(x)
    for n in (n1, n2):
        mlp.predict_device(xd.ptr, n, yd_ref.ptr)
        ref = np.zeros((n, sizes[-1]), np.float32)
        yd_ref.download(ref)
        before = _device_free_bytes()
        A.check(f(mlp._h, C.c_void_p(xd.ptr), n, C.c_void_p(yd.ptr)))
        after = _device_free_bytes()
        got = np.zeros((n, sizes[-1]), np.float32)
        yd.download(got)
        print("rows %d: max |noalloc - predict| = %g, device free bytes %d -> %d"
              % (n, float(np.max(np.abs(got - ref))), before, after))
        assert np.array_equal(got, ref)
        assert after == before
    # rows past the second call's end keep the first call's output (no stray writes)
    tail = np.zeros((n1, sizes[-1]), np.float32)
    yd.download(tail)
    first = np.zeros((n1, sizes[-1]), np.float32)
    mlp.predict_device(xd.ptr, n1, yd_ref.ptr)
    yd_ref.download(first)
    assert np.array_equal(tail[n2:], first[n2:])


def test_noalloc_arguments():
    f = _noalloc()
    mlp = Mlp([8, 64, 3], seed=1)
    assert f(mlp._h, None, 10, None) == -1       # FS_ERR_ARG
    xd = A.DeviceBuffer(64)
    assert f(mlp._h, C.c_void_p(xd.ptr), 0, C.c_void_p(xd.ptr)) == 0  # nothing to do
