"""The allocation-free unpack (atkgpu_unpack_dc_noalloc): raw bytes, samples
and mean in device memory, the mean's scratch a static device array. On the
same bytes its samples equal atkdsp_unpack_dc's bit for bit and its mean is
the same double, for every format, on calls in a row whose second input is
smaller than the first (the scratch still holds the first call's pass sums).

Device memory comes from the HIP runtime the library itself is linked with
(opened by its soname once the library is loaded), and the work goes on the
null stream: a torch wheel bundles a HIP runtime of its own, and a pointer or
stream of one runtime means nothing to another."""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

FORMATS = [("cu8", 0), ("ci8", 1), ("ci16", 2), ("ci16q11", 3), ("cf32", 4)]
BPS = {0: 2, 1: 2, 2: 4, 3: 4, 4: 8}
MAX_SAMPLES = 512 << 19         # ATKGPU_NOALLOC_MAX_SAMPLES
E_CAP = -3


class _Hip:
    """hipMalloc / hipMemcpy / hipFree of the library's own HIP runtime."""
    H2D, D2H = 1, 2

    def __init__(self):
        self.h = C.CDLL("libamdhip64.so.7")
        self.live = []

    def alloc(self, nbytes):
        p = C.c_void_p()
        assert self.h.hipMalloc(C.byref(p), C.c_size_t(max(nbytes, 256))) == 0
        self.live.append(p)
        return p.value

    def put(self, dst, a: np.ndarray):
        assert self.h.hipMemcpy(C.c_void_p(dst), a.ctypes.data_as(C.c_void_p), C.c_size_t(a.nbytes), self.H2D) == 0

    def get(self, src, a: np.ndarray):
        assert self.h.hipMemcpy(a.ctypes.data_as(C.c_void_p), C.c_void_p(src), C.c_size_t(a.nbytes), self.D2H) == 0
        return a

    def sync(self):
        assert self.h.hipDeviceSynchronize() == 0

    def free(self):
        for p in self.live:
            self.h.hipFree(p)
        self.live = []


def _entry(lib):
    f = lib.atkgpu_unpack_dc_noalloc
    f.restype = C.c_ssize_t
    f.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int, C.c_void_p, C.c_size_t, C.c_float, C.c_float,
                  C.c_void_p]
    return f


def _raw(code, n, seed):
    rng = np.random.default_rng(seed)
    if code == 4:
        return (rng.standard_normal(2 * n) * 0.3).astype(np.float32).view(np.uint8)
    return rng.integers(0, 256, n * BPS[code], dtype=np.uint8)


@pytest.mark.parametrize("fmt,code", FORMATS)
def test_noalloc_unpack_equals_pooled_entry(gpu, fmt, code):
    f = _entry(gpu.load())
    hip = _Hip()
    first, second = 5_000_003, 70_001          # ten passes of 2^19 samples (across two pool chunks), then one
    raw_d = hip.alloc(first * BPS[code])
    out_d = hip.alloc(first * 8)
    mean_d = hip.alloc(16)
    hip.put(mean_d, np.full(2, np.nan))
    off = complex(0.0123, -0.0456)
    try:
        for i, n in enumerate((first, second)):
            raw = _raw(code, n, 10 * code + i)
            ref, ref_mean = gpu.unpack_dc(raw, fmt, off)
            hip.put(raw_d, raw)
            got = f(None, raw_d, raw.nbytes, code, out_d, n, off.real, off.imag, mean_d)
            assert got == n
            hip.sync()
            y = hip.get(out_d, np.empty(2 * n, np.float32))
            assert np.array_equal(y.view(np.uint32), ref.view(np.uint32)), (fmt, n)
            m = hip.get(mean_d, np.empty(2))
            assert m[0] == ref_mean.real and m[1] == ref_mean.imag, (fmt, n, m, ref_mean)
    finally:
        hip.free()


def test_noalloc_unpack_empty_and_refusal(gpu):
    f = _entry(gpu.load())
    hip = _Hip()
    raw_d, out_d, mean_d = hip.alloc(64), hip.alloc(256), hip.alloc(16)
    hip.put(mean_d, np.full(2, np.nan))
    try:
        # nothing whole: no samples, the mean is the offset (as atkdsp_unpack_dc)
        assert f(None, raw_d, 1, 0, out_d, 32, 0.25, -0.5, mean_d) == 0
        hip.sync()
        assert hip.get(mean_d, np.empty(2)).tolist() == [0.25, -0.5]
        # more samples than the scratch covers: refused on the host before any launch
        rc = f(None, raw_d, (MAX_SAMPLES + 1) * 2, 0, out_d, MAX_SAMPLES + 1, 0.0, 0.0, mean_d)
        assert rc == E_CAP
        hip.sync()
        assert hip.get(mean_d, np.empty(2)).tolist() == [0.25, -0.5]
        # no mean asked: samples only
        raw = _raw(3, 8, 5)
        ref, _ = gpu.unpack_dc(raw, "ci16q11", 0j)
        hip.put(raw_d, raw)
        assert f(None, raw_d, raw.nbytes, 3, out_d, 8, 0.0, 0.0, None) == 8
        hip.sync()
        assert np.array_equal(hip.get(out_d, np.empty(16, np.float32)).view(np.uint32), ref.view(np.uint32))
    finally:
        hip.free()
