"""The filterbank in matrix-core form (include/atkgpu.h atkgpu_pfbx_*) from Python.

    px = PfbX(form="bf16x2", fmt="cu8", fs=2.4e6, nchan=64, oversample=1, taps_per_poly=8,
              streams=128, block_max=1 << 20)
    px.upload(s, raw_bytes)          # per stream, host -> HBM
    m = px.run(n)                    # one block of n samples for every stream
    y = px.read(s)                   # [nchan][m] complex64

Forms: "fp32" (unpack, then the fused FP32 filterbank), "fp32raw" (the
fused FP32 filterbank reading the raw rows and converting each sample as it
loads it: bit-identical to "fp32", and the like-for-like baseline of the
matrix-core forms, which read the same raw rows), "fp32mm" (the FP32 fold on
the matrix cores: bit-identical to "fp32" for finite input), "bf16x2" / "bf16" (8-bit samples exact in BF16, taps
as hi + lo or one BF16 term), "fp16x2" / "fp16" (8- or 12-bit samples exact
in FP16, taps scaled by a power of two and split likewise). Every form
accumulates in FP32.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import AtkDspError, _Cf32, _check, _fmt, load

FORMS = {"fp32": 0, "bf16x2": 1, "bf16": 2, "fp16x2": 3, "fp16": 4, "fp32mm": 5, "fp32raw": 6}


class _Cfg(C.Structure):
    _fields_ = [("fmt", C.c_int), ("fs", C.c_double), ("nchan", C.c_uint), ("oversample", C.c_uint),
                ("taps_per_poly", C.c_uint), ("atten_db", C.c_double), ("form", C.c_int),
                ("block_max", C.c_size_t), ("off_re", C.c_float), ("off_im", C.c_float), ("coherent", C.c_int)]


def _bind(lib) -> None:
    if getattr(lib, "_pfbx_bound", False):
        return
    P, vp, sz = C.POINTER, C.c_void_p, C.c_size_t
    sig = {
        "atkgpu_pfbx_create": (vp, [P(_Cfg), C.c_uint]),
        "atkgpu_pfbx_destroy": (None, [vp]),
        "atkgpu_pfbx_reset": (C.c_int, [vp]),
        "atkgpu_pfbx_input": (vp, [vp, C.c_uint]),
        "atkgpu_pfbx_upload": (C.c_int, [vp, C.c_uint, vp, sz]),
        "atkgpu_pfbx_run": (C.c_ssize_t, [vp, sz]),
        "atkgpu_pfbx_sync": (C.c_int, [vp]),
        "atkgpu_pfbx_timing": (C.c_int, [vp, P(C.c_float), P(C.c_float), P(C.c_float), P(C.c_float)]),
        "atkgpu_pfbx_read_cov": (C.c_int, [vp, P(_Cf32), P(C.c_uint32)]),
        "atkgpu_pfbx_read_power": (C.c_int, [vp, P(C.c_float), P(C.c_float)]),
        "atkgpu_pfbx_cov_reset": (C.c_int, [vp]),
        "atkgpu_pfbx_read": (C.c_ssize_t, [vp, C.c_uint, P(_Cf32), sz]),
    }
    for name, (res, args) in sig.items():
        f = getattr(lib, name)
        f.restype, f.argtypes = res, args
    lib._pfbx_bound = True


class PfbX:
    def __init__(self, form="bf16x2", fmt="cu8", fs=2.4e6, nchan=64, oversample=1, taps_per_poly=8,
                 atten_db=80.0, streams=1, block_max=1 << 20, offset: complex = 0j, coherent=False):
        self._lib = load()
        _bind(self._lib)
        if form not in FORMS:
            raise AtkDspError(f"unknown form {form!r}")
        cfg = _Cfg(_fmt(fmt), float(fs), int(nchan), int(oversample), int(taps_per_poly), float(atten_db),
                   FORMS[form], int(block_max), float(offset.real), float(offset.imag), int(bool(coherent)))
        self._h = self._lib.atkgpu_pfbx_create(C.byref(cfg), int(streams))
        if not self._h:
            raise AtkDspError(f"atkgpu_pfbx_create failed (no GPU, or form {form} does not support this "
                              f"shape/format)")
        self.nchan, self.streams, self.form = int(nchan), int(streams), form
        self._m_cap = (int(block_max) + nchan) // (nchan // oversample) + 2

    def close(self):
        h, self._h = getattr(self, "_h", None), None
        if h:
            self._lib.atkgpu_pfbx_destroy(h)

    def __del__(self):
        self.close()

    def reset(self) -> None:
        _check(self._lib.atkgpu_pfbx_reset(self._h), "pfbx_reset")

    def upload(self, s: int, raw) -> None:
        b = np.ascontiguousarray(np.frombuffer(raw, np.uint8) if not isinstance(raw, np.ndarray)
                                 else raw.view(np.uint8).ravel())
        _check(self._lib.atkgpu_pfbx_upload(self._h, int(s), b.ctypes.data, b.size), "pfbx_upload")

    def run(self, n: int) -> int:
        return _check(self._lib.atkgpu_pfbx_run(self._h, int(n)), "pfbx_run")

    def sync(self) -> None:
        _check(self._lib.atkgpu_pfbx_sync(self._h), "pfbx_sync")

    def timing(self):
        """(ms of the whole run, of the unpack, of the filterbank kernel, of W1 + W2) of the last run."""
        a, b, c, d = C.c_float(), C.c_float(), C.c_float(), C.c_float()
        _check(self._lib.atkgpu_pfbx_timing(self._h, C.byref(a), C.byref(b), C.byref(c), C.byref(d)),
               "pfbx_timing")
        return a.value, b.value, c.value, d.value

    def read_cov(self):
        """W1 accumulated so far: R [nchan][nel][nel] complex64, counts [nchan] uint32."""
        R = np.empty((self.nchan, self.streams, self.streams), np.complex64)
        cnt = np.empty(self.nchan, np.uint32)
        _check(self._lib.atkgpu_pfbx_read_cov(self._h, R.ctypes.data_as(C.POINTER(_Cf32)),
                                              cnt.ctypes.data_as(C.POINTER(C.c_uint32))), "pfbx_read_cov")
        return R, cnt

    def read_power(self):
        """W2 of the last run: pwr [nel][nchan], mean_pwr [nchan] float32."""
        p = np.empty((self.streams, self.nchan), np.float32)
        mp = np.empty(self.nchan, np.float32)
        _check(self._lib.atkgpu_pfbx_read_power(self._h, p.ctypes.data_as(C.POINTER(C.c_float)),
                                                mp.ctypes.data_as(C.POINTER(C.c_float))), "pfbx_read_power")
        return p, mp

    def cov_reset(self) -> None:
        _check(self._lib.atkgpu_pfbx_cov_reset(self._h), "pfbx_cov_reset")

    def read(self, s: int) -> np.ndarray:
        out = np.empty(self.nchan * self._m_cap, np.complex64)
        m = _check(self._lib.atkgpu_pfbx_read(self._h, int(s), out.ctypes.data_as(C.POINTER(_Cf32)), out.size),
                   "pfbx_read")
        return out[: self.nchan * m].reshape(self.nchan, m)

    def process(self, raws, bps: int) -> list[np.ndarray]:
        """One block per stream (equal lengths): upload, run, read."""
        n = len(raws[0]) // bps
        for s, r in enumerate(raws):
            self.upload(s, r)
        self.run(n)
        return [self.read(s) for s in range(len(raws))]
