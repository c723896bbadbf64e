"""The device-resident multi-stream chain (include/atkgpu.h) from Python.

    ch = Chain(fmt="ci16q11", fs=40e6, nchan=1024, oversample=2, taps_per_poly=16,
               nfft=4096, fft_hop=2048, line_cols=1024, trace_bins=8192, streams=128,
               block_max=2_000_000)
    ch.upload(s, raw_bytes)          # per stream, host -> HBM
    m = ch.run(n)                    # one block of n samples for every stream
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import AtkDspError, _Cf32, _check, _fmt, load, WIN, DET


class _Cfg(C.Structure):
    _fields_ = [("fmt", C.c_int), ("fs", C.c_double), ("nchan", C.c_uint), ("oversample", C.c_uint),
                ("taps_per_poly", C.c_uint), ("atten_db", C.c_double), ("nfft", C.c_uint),
                ("fft_hop", C.c_uint), ("window", C.c_int), ("detector", C.c_int),
                ("line_cols", C.c_uint), ("trace_bins", C.c_uint), ("trace_lo_hz", C.c_double),
                ("trace_hz_per_bin", C.c_double), ("fc0_hz", C.c_double), ("fc_step_hz", C.c_double),
                ("dc_mode", C.c_int), ("block_max", C.c_size_t)]


def _bind_chain(lib) -> None:
    if getattr(lib, "_chain_bound", False):
        return
    P, vp, sz = C.POINTER, C.c_void_p, C.c_size_t
    sig = {
        "atkgpu_chain_create": (vp, [P(_Cfg), C.c_uint]),
        "atkgpu_chain_destroy": (None, [vp]),
        "atkgpu_chain_reset": (None, [vp]),
        "atkgpu_chain_input": (vp, [vp, C.c_uint]),
        "atkgpu_chain_upload": (C.c_int, [vp, C.c_uint, vp, sz]),
        "atkgpu_chain_run": (C.c_ssize_t, [vp, sz]),
        "atkgpu_chain_sync": (C.c_int, [vp]),
        "atkgpu_chain_timing": (C.c_int, [vp, P(C.c_float), P(C.c_float), P(C.c_float)]),
        "atkgpu_chain_set_timing": (C.c_int, [vp, C.c_int]),
        "atkgpu_chain_read_pfb": (C.c_ssize_t, [vp, C.c_uint, P(_Cf32), sz]),
        "atkgpu_chain_read_line": (C.c_int, [vp, C.c_uint, P(C.c_float), P(C.c_float), P(C.c_float)]),
        "atkgpu_chain_read_mean": (C.c_int, [vp, C.c_uint, P(C.c_double)]),
        "atkgpu_chain_frames": (C.c_longlong, [vp]),
    }
    for name, (res, args) in sig.items():
        f = getattr(lib, name)
        f.restype, f.argtypes = res, args
    lib._chain_bound = True


class Chain:
    def __init__(self, fmt="ci16q11", fs=40e6, nchan=1024, oversample=2, taps_per_poly=16, atten_db=80.0,
                 nfft=4096, fft_hop=2048, window="hann", detector="max", line_cols=1024, trace_bins=8192,
                 trace_lo_hz=None, trace_hz_per_bin=None, fc0_hz=0.0, fc_step_hz=0.0, dc_mode=1,
                 streams=1, block_max=2_000_000):
        self._lib = load()
        _bind_chain(self._lib)
        if trace_hz_per_bin is None:
            trace_hz_per_bin = 8 * fs / max(trace_bins, 1)
        if trace_lo_hz is None:
            trace_lo_hz = fc0_hz - trace_hz_per_bin * trace_bins / 2
        cfg = _Cfg(_fmt(fmt), float(fs), int(nchan), int(oversample), int(taps_per_poly), float(atten_db),
                   int(nfft), int(fft_hop), WIN[window] if isinstance(window, str) else int(window),
                   DET[detector] if isinstance(detector, str) else int(detector), int(line_cols),
                   int(trace_bins), float(trace_lo_hz), float(trace_hz_per_bin), float(fc0_hz),
                   float(fc_step_hz), int(dc_mode), int(block_max))
        self.cfg = cfg
        self._h = self._lib.atkgpu_chain_create(C.byref(cfg), int(streams))
        if not self._h:
            raise AtkDspError("atkgpu_chain_create failed (shape, memory or no GPU)")
        self.streams, self.nchan, self.nfft = int(streams), int(nchan), int(nfft)
        self.line_cols, self.trace_bins = int(line_cols), int(trace_bins)
        self.last_m = 0

    def close(self):
        h, self._h = getattr(self, "_h", None), None
        if h:
            self._lib.atkgpu_chain_destroy(h)

    def __del__(self):
        self.close()

    def reset(self):
        self._lib.atkgpu_chain_reset(self._h)

    def upload(self, s: int, raw) -> None:
        b = np.frombuffer(raw, np.uint8) if not isinstance(raw, np.ndarray) else raw.view(np.uint8).ravel()
        _check(self._lib.atkgpu_chain_upload(self._h, int(s), b.ctypes.data, b.size), "chain_upload")

    def run(self, n: int) -> int:
        self.last_m = _check(self._lib.atkgpu_chain_run(self._h, int(n)), "chain_run")
        return self.last_m

    def sync(self) -> None:
        _check(self._lib.atkgpu_chain_sync(self._h), "chain_sync")

    def set_timing(self, on: bool) -> None:
        """Record (default) or skip the per-run timing events timing() reads."""
        _check(self._lib.atkgpu_chain_set_timing(self._h, int(bool(on))), "chain_set_timing")

    def timing(self):
        a, b, d = C.c_float(), C.c_float(), C.c_float()
        _check(self._lib.atkgpu_chain_timing(self._h, C.byref(a), C.byref(b), C.byref(d)), "chain_timing")
        return a.value, b.value, d.value

    def frames(self) -> int:
        return int(self._lib.atkgpu_chain_frames(self._h))

    def read_pfb(self, s: int) -> np.ndarray:
        out = np.empty(self.nchan * max(self.last_m, 1), np.complex64)
        m = _check(self._lib.atkgpu_chain_read_pfb(self._h, int(s), out.ctypes.data_as(C.POINTER(_Cf32)),
                                                   out.size), "chain_read_pfb")
        return out[: self.nchan * m].reshape(self.nchan, m)

    def read_line(self, s: int):
        line = np.empty(self.nfft, np.float32)
        cols = np.empty(self.line_cols, np.float32)
        trace = np.empty(max(self.trace_bins, 1), np.float32)
        fp = lambda a: a.ctypes.data_as(C.POINTER(C.c_float))
        _check(self._lib.atkgpu_chain_read_line(self._h, int(s), fp(line), fp(cols),
                                                fp(trace) if self.trace_bins else None), "chain_read_line")
        return line, cols, (trace if self.trace_bins else None)

    def read_mean(self, s: int) -> complex:
        m = (C.c_double * 2)()
        _check(self._lib.atkgpu_chain_read_mean(self._h, int(s), m), "chain_read_mean")
        return complex(m[0], m[1])
