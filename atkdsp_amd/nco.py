"""The device-resident NCO (include/atkgpu.h atkgpu_nco_*) from Python.

    run = NcoRun(freq_hz=40e6 * 1234 / 10007, fs=40e6, chunk_max=65536, log_max=8192)
    run.upload(x)                 # device input buffer (kept across runs)
    for _ in range(k): run.run(65536)
    phases = run.log()            # the FP64 phase after each run, from device memory
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import AtkDspError, _Cf32, _cf32, _check, load


def _bind(lib) -> None:
    if getattr(lib, "_nco_bound", False):
        return
    vp, sz = C.c_void_p, C.c_size_t
    sig = {
        "atkgpu_nco_create": (vp, [C.c_double, C.c_double, sz, sz]),
        "atkgpu_nco_destroy": (None, [vp]),
        "atkgpu_nco_upload": (C.c_int, [vp, C.POINTER(_Cf32), sz]),
        "atkgpu_nco_run": (C.c_int, [vp, sz]),
        "atkgpu_nco_sync": (C.c_int, [vp]),
        "atkgpu_nco_read": (C.c_ssize_t, [vp, C.POINTER(_Cf32), sz]),
        "atkgpu_nco_phase": (C.c_int, [vp, C.POINTER(C.c_double), C.POINTER(C.c_longlong)]),
        "atkgpu_nco_log": (C.c_ssize_t, [vp, C.POINTER(C.c_double), sz]),
    }
    for name, (res, args) in sig.items():
        f = getattr(lib, name)
        f.restype, f.argtypes = res, args
    lib._nco_bound = True


class NcoRun:
    def __init__(self, freq_hz: float, fs: float, chunk_max: int = 65536, log_max: int = 0):
        self._lib = load()
        _bind(self._lib)
        self._h = self._lib.atkgpu_nco_create(float(freq_hz), float(fs), int(chunk_max), int(log_max))
        if not self._h:
            raise AtkDspError("atkgpu_nco_create failed (no GPU?)")
        self.chunk_max, self.log_max = int(chunk_max), int(log_max)

    def close(self):
        h, self._h = getattr(self, "_h", None), None
        if h:
            self._lib.atkgpu_nco_destroy(h)

    def __del__(self):
        self.close()

    def upload(self, x) -> None:
        x = _cf32(x, "x")
        _check(self._lib.atkgpu_nco_upload(self._h, x.ctypes.data_as(C.POINTER(_Cf32)), x.size), "nco_upload")

    def run(self, n: int) -> None:
        _check(self._lib.atkgpu_nco_run(self._h, int(n)), "nco_run")

    def sync(self) -> None:
        _check(self._lib.atkgpu_nco_sync(self._h), "nco_sync")

    def read(self, n: int) -> np.ndarray:
        out = np.empty(int(n), np.complex64)
        _check(self._lib.atkgpu_nco_read(self._h, out.ctypes.data_as(C.POINTER(_Cf32)), out.size), "nco_read")
        return out

    def phase(self):
        ph, ns = C.c_double(), C.c_longlong()
        _check(self._lib.atkgpu_nco_phase(self._h, C.byref(ph), C.byref(ns)), "nco_phase")
        return ph.value, ns.value

    def log(self) -> np.ndarray:
        out = np.empty(max(self.log_max, 1), np.float64)
        k = _check(self._lib.atkgpu_nco_log(self._h, out.ctypes.data_as(C.POINTER(C.c_double)), out.size),
                   "nco_log")
        return out[:k]
