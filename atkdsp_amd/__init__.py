"""atkdsp_amd — the wideband front end of atkdsp on one MI355X.

A ctypes binding over ``atkdsp_amd/atkdsp.so`` with the names, argument
meaning and error behaviour of the reference binding's front-end subset
(reference python/atkdsp/__init__.py: ``load`` :238-268, ``unpack`` :337,
``unpack_dc`` :361, ``Fft`` :531-599, ``window`` :602, ``db_to_pixels`` :673,
``decimate_max`` :686, ``stitch_max`` :719, ``Pfb`` :4159-4219,
``pfb_cov_accumulate`` :4222, ``pfb_channel_power`` :4252), so code written
against ``atkdsp`` runs against this package unchanged. The library is also a
drop-in for the reference binding itself: point ``ATKDSP_LIB`` at
``atkdsp_amd/atkdsp.so``.

There is no CPU fallback: without the built library or without a GPU every
call raises :class:`AtkDspError`.
"""
from __future__ import annotations

import ctypes as C
import os
from pathlib import Path

import numpy as np

ABI_VERSION = 26

FMT = {"cu8": 0, "ci8": 1, "ci16": 2, "ci16_le": 2, "cs16": 2,
       "ci16q11": 3, "cf32": 4, "cf32_le": 4}
DET = {"max": 0, "min": 1, "avg": 2}
WIN = {"rectangular": 0, "hann": 1, "hamming": 2, "blackmanharris": 3}

LIB_NAME = "atkdsp.so"
_HERE = Path(__file__).resolve().parent


class AtkDspError(RuntimeError):
    """A kernel returned an error code, or the library could not be loaded."""


class _Cf32(C.Structure):
    _fields_ = [("re", C.c_float), ("im", C.c_float)]


class _ChannelC(C.Structure):
    _fields_ = [("start_bin", C.c_int), ("end_bin", C.c_int), ("centroid_bin", C.c_float),
                ("peak_db", C.c_float), ("snr_db", C.c_float)]


class Channel:
    """One detected channel, in bins [start_bin, end_bin) (the caller knows the bin width)."""
    __slots__ = ("start_bin", "end_bin", "centroid_bin", "peak_db", "snr_db")

    def __init__(self, start_bin, end_bin, centroid_bin, peak_db, snr_db):
        self.start_bin, self.end_bin = int(start_bin), int(end_bin)
        self.centroid_bin, self.peak_db, self.snr_db = float(centroid_bin), float(peak_db), float(snr_db)

    def as_tuple(self):
        return (self.start_bin, self.end_bin, self.centroid_bin, self.peak_db, self.snr_db)

    def __repr__(self) -> str:
        return (f"Channel(bins {self.start_bin}..{self.end_bin}, centroid {self.centroid_bin:.1f}, "
                f"peak {self.peak_db:.1f} dB, snr {self.snr_db:.1f} dB)")


class _CfarConfig(C.Structure):
    _fields_ = [("mode", C.c_int32), ("n_ref", C.c_int32), ("n_guard", C.c_int32), ("k_frames", C.c_int32),
                ("pfa", C.c_double), ("looks", C.c_double), ("os_k", C.c_int32), ("open_m", C.c_int32),
                ("open_n", C.c_int32), ("close_k", C.c_int32), ("merge_gap", C.c_int32), ("min_bins", C.c_int32),
                ("max_tracks", C.c_int32), ("reserved", C.c_int32), ("rho2", C.c_double * 4)]


CFAR_MODES = {"ca": 0, "os": 1, "go": 2, "so": 3}
CFAR_EVENT = {1: "open", 2: "update", 3: "close"}
# the 64-byte event record (atkdsp_cfar_event) as a numpy record type
CFAR_EVENT_DTYPE = np.dtype([("t0", "<f8"), ("t1", "<f8"), ("fc_bin", "<f8"), ("type", "<i4"), ("id", "<u4"),
                             ("flags", "<u4"), ("lo_bin", "<i4"), ("hi_bin", "<i4"), ("n_hits", "<i4"),
                             ("level_db", "<f4"), ("noise_db", "<f4"), ("snr_db", "<f4"),
                             ("band_snr_db", "<f4")])
assert CFAR_EVENT_DTYPE.itemsize == 64


class _CfarCounters(C.Structure):
    _fields_ = [(k, C.c_uint64) for k in ("frames", "warmup", "cells_tested", "cell_hits", "untestable",
                                           "opened", "closed", "updates", "dropped", "resets", "tracks")]


class _RingMeta(C.Structure):
    _fields_ = [("seq", C.c_uint64), ("first_sample", C.c_uint64), ("t_utc_ns", C.c_int64),
                ("fc_hz", C.c_double), ("fs_hz", C.c_double), ("gain_db", C.c_float), ("format", C.c_int32),
                ("flags", C.c_uint32), ("nbytes", C.c_uint32)]


class _RingCounters(C.Structure):
    _fields_ = [("committed", C.c_uint64), ("oldest", C.c_uint64), ("block_bytes", C.c_uint64),
                ("nblocks", C.c_uint64), ("total_bytes", C.c_uint64), ("torn", C.c_uint64),
                ("max_consumers", C.c_uint32), ("active", C.c_uint32), ("active_mask", C.c_uint32),
                ("reserved", C.c_uint32), ("lost", C.c_uint64 * 16), ("read", C.c_uint64 * 16),
                ("lag", C.c_uint64 * 16)]


class _Nco(C.Structure):
    _fields_ = [("phase", C.c_double), ("step", C.c_double)]


_ERRORS = {-1: "bad argument", -2: "out of memory", -3: "output buffer too small",
           -4: "FFT backend or device failure", -7: "nothing to return yet"}

_lib = None
_lib_path: Path | None = None


def _check(rc: int, what: str) -> int:
    if rc < 0:
        raise AtkDspError(f"{what}: {_ERRORS.get(rc, f'error {rc}')}")
    return rc


def _bind(lib) -> None:
    P = C.POINTER
    cf, fp, vp, sz, u, d = P(_Cf32), P(C.c_float), C.c_void_p, C.c_size_t, C.c_uint, C.c_double
    sig = {
        "atkdsp_abi_version": (C.c_int, []),
        "atkdsp_version_string": (C.c_char_p, []),
        "atkdsp_build_info": (C.c_char_p, []),
        "atkdsp_bytes_per_sample": (C.c_int, [C.c_int]),
        "atkdsp_set_threads": (C.c_int, [C.c_int]),
        "atkdsp_get_threads": (C.c_int, []),
        "atkdsp_unpack": (C.c_ssize_t, [vp, sz, C.c_int, cf, sz, C.c_float, cf]),
        "atkdsp_unpack_dc": (C.c_ssize_t, [vp, sz, C.c_int, cf, sz, C.c_float, C.c_float,
                                           P(d), P(d)]),
        "atkdsp_fft_create": (vp, [sz]),
        "atkdsp_fft_destroy": (None, [vp]),
        "atkdsp_fft_length": (sz, [vp]),
        "atkdsp_fft_exec": (C.c_int, [vp, cf, cf, C.c_int]),
        "atkdsp_window": (C.c_int, [C.c_int, sz, fp]),
        "atkdsp_power_db": (C.c_int, [vp, cf, fp, fp]),
        "atkdsp_spectrum_reduce": (C.c_ssize_t, [vp, cf, sz, sz, fp, C.c_int, fp]),
        "atkdsp_spectrum_stats": (C.c_ssize_t, [vp, cf, sz, sz, fp, fp, fp, fp, fp]),
        "atkdsp_window_stats": (C.c_int, [fp, sz, P(d), P(d)]),
        "atkdsp_db_to_pixels": (None, [fp, sz, C.c_float, C.c_float, P(C.c_uint32), P(C.c_uint32)]),
        "atkdsp_decimate_max": (None, [fp, sz, fp, sz]),
        "atkdsp_stitch_max": (None, [fp, sz, d, d, d, d, d, d, fp, sz]),
        "atkdsp_median": (C.c_float, [fp, sz, fp]),
        "atkdsp_cfar_defaults": (None, [C.POINTER(_CfarConfig)]),
        "atkdsp_cfar_alpha": (C.c_int, [C.c_int, C.c_int, d, d, C.c_int, C.POINTER(d), C.POINTER(d), C.POINTER(d)]),
        "atkdsp_cfar_n_eff": (d, [C.c_int, C.POINTER(d)]),
        "atkdsp_cfar_window_rho2": (C.c_int, [fp, sz, C.POINTER(d)]),
        "atkdsp_cfar_create": (vp, [sz, C.POINTER(_CfarConfig)]),
        "atkdsp_cfar_destroy": (None, [vp]),
        "atkdsp_cfar_reset": (None, [vp]),
        "atkdsp_cfar_set_mask": (C.c_int, [vp, vp]),
        "atkdsp_cfar_events_max": (sz, [vp, sz]),
        "atkdsp_cfar_process": (C.c_ssize_t, [vp, fp, sz, d, d, vp, sz]),
        "atkdsp_cfar_flush": (C.c_ssize_t, [vp, vp, sz]),
        "atkdsp_cfar_last": (C.c_int, [vp, fp, fp, vp]),
        "atkdsp_cfar_info": (C.c_int, [vp, C.POINTER(d), C.POINTER(d), C.POINTER(d), C.POINTER(_CfarConfig)]),
        "atkdsp_cfar_describe": (C.c_int, [vp, C.c_char_p, sz]),
        "atkdsp_cfar_stats": (C.c_int, [vp, C.POINTER(_CfarCounters)]),
        "atkdsp_detect_channels": (C.c_ssize_t, [fp, sz, C.c_float, C.c_float, C.c_int, C.c_int, fp,
                                                 C.POINTER(_ChannelC), sz]),
        "atkdsp_pfb_create": (vp, [d, u, u, u, d, u]),
        "atkdsp_pfb_destroy": (None, [vp]),
        "atkdsp_pfb_reset": (None, [vp]),
        "atkdsp_pfb_nchan": (u, [vp]),
        "atkdsp_pfb_out_rate": (d, [vp]),
        "atkdsp_pfb_out_hops": (sz, [vp, sz]),
        "atkdsp_pfb_process": (C.c_ssize_t, [vp, cf, sz, cf, sz]),
        "atkdsp_pfb_process_n": (C.c_ssize_t, [vp, P(cf), u, sz, cf, sz]),
        "atkdsp_pfb_block_origin": (C.c_int64, [vp, C.c_int64]),
        "atkdsp_pfb_group_delay": (d, [vp]),
        "atkdsp_pfb_set_cal": (C.c_int, [vp, cf, u]),
        "atkdsp_pfb_describe": (C.c_int, [vp, C.c_char_p, sz]),
        "atkdsp_pfb_cov_accumulate": (None, [cf, u, u, sz, P(C.c_uint8), cf, P(C.c_uint32)]),
        "atkdsp_pfb_cov_reset": (None, [cf, P(C.c_uint32), u, u, P(C.c_uint8)]),
        "atkdsp_pfb_channel_power": (None, [cf, u, u, sz, fp, fp]),
        "atkdsp_nco_init": (None, [P(_Nco), d, d]),
        "atkdsp_nco_set_freq": (None, [P(_Nco), d, d]),
        "atkdsp_nco_mix": (None, [P(_Nco), cf, cf, sz]),
        "atkdsp_design_lowpass": (C.c_ssize_t, [d, d, d, d, fp, sz]),
        "atkdsp_arb_resampler_create": (vp, [d, d, d, u]),
        "atkdsp_arb_resampler_destroy": (None, [vp]),
        "atkdsp_arb_resampler_reset": (None, [vp]),
        "atkdsp_arb_resampler_set_ratio": (None, [vp, d]),
        "atkdsp_arb_resampler_ratio": (d, [vp]),
        "atkdsp_arb_resampler_out_max": (sz, [vp, sz]),
        "atkdsp_arb_resampler_process": (C.c_ssize_t, [vp, cf, sz, cf, sz]),
        "atkdsp_osx_create": (vp, [d, sz, d, u]),
        "atkdsp_osx_destroy": (None, [vp]),
        "atkdsp_osx_reset": (None, [vp]),
        "atkdsp_osx_add": (C.c_int, [vp, d, d, d]),
        "atkdsp_osx_retune": (C.c_int, [vp, C.c_int, d]),
        "atkdsp_osx_remove": (C.c_int, [vp, C.c_int]),
        "atkdsp_osx_channels": (C.c_int, [vp]),
        "atkdsp_osx_chan_rate": (d, [vp, C.c_int]),
        "atkdsp_osx_out_max": (sz, [vp, C.c_int, sz]),
        "atkdsp_osx_process": (C.c_ssize_t, [vp, cf, sz, P(cf), P(sz), sz]),
        "atkdsp_osx_process_n": (C.c_ssize_t, [vp, P(cf), u, sz, P(cf), P(sz), sz]),
        "atkdsp_osx_describe": (C.c_int, [vp, C.c_char_p, sz]),
        "atkdsp_ring_bytes": (sz, [sz, sz, C.c_int]),
        "atkdsp_ring_init": (vp, [vp, sz, sz, sz, C.c_int]),
        "atkdsp_ring_attach": (vp, [vp, sz]),
        "atkdsp_ring_claim": (vp, [vp]),
        "atkdsp_ring_commit": (C.c_int, [vp, P(_RingMeta)]),
        "atkdsp_ring_open": (C.c_int, [vp, C.c_int]),
        "atkdsp_ring_close": (C.c_int, [vp, C.c_int]),
        "atkdsp_ring_read": (C.c_ssize_t, [vp, C.c_int, vp, sz, P(_RingMeta), P(C.c_uint64)]),
        "atkdsp_ring_peek": (vp, [vp, C.c_int, P(_RingMeta), P(C.c_uint64)]),
        "atkdsp_ring_release": (C.c_int, [vp, C.c_int]),
        "atkdsp_ring_copy_range": (C.c_ssize_t, [vp, C.c_uint64, C.c_uint64, vp, sz, P(_RingMeta), sz]),
        "atkdsp_ring_stats": (C.c_int, [vp, P(_RingCounters)]),
        "atkdsp_ring_describe": (C.c_int, [vp, C.c_char_p, sz]),
    }
    for name, (res, args) in sig.items():
        f = getattr(lib, name)
        f.restype = res
        f.argtypes = args


def _candidates() -> list[Path]:
    out = []
    env = os.environ.get("ATKDSP_AMD_LIB")
    if env:
        out.append(Path(env))
    out.append(_HERE / LIB_NAME)
    return out


def load(path: str | os.PathLike | None = None):
    """Load the library once and return the ctypes handle; raises AtkDspError."""
    global _lib, _lib_path
    if _lib is not None and path is None:
        return _lib
    errors = []
    for cand in ([Path(path)] if path else _candidates()):
        if not cand.exists():
            errors.append(f"{cand}: not found (run python -m atkdsp_amd.build)")
            continue
        try:
            lib = C.CDLL(str(cand))
        except OSError as e:
            errors.append(f"{cand}: {e}")
            continue
        lib.atkdsp_abi_version.restype = C.c_int
        abi = int(lib.atkdsp_abi_version())
        if abi != ABI_VERSION:
            errors.append(f"{cand}: ABI {abi}, this binding needs {ABI_VERSION}")
            continue
        _bind(lib)
        _lib, _lib_path = lib, cand
        return lib
    raise AtkDspError("could not load atkdsp_amd: " + "; ".join(errors))


def available() -> bool:
    """True when the library loads (it may still find no GPU). Never raises."""
    try:
        load()
        return True
    except AtkDspError:
        return False


def lib_path() -> Path | None:
    return _lib_path


def version() -> str:
    return load().atkdsp_version_string().decode()


def build_info() -> str:
    return load().atkdsp_build_info().decode()


def set_threads(n: int) -> int:
    return int(load().atkdsp_set_threads(int(n)))


def get_threads() -> int:
    return int(load().atkdsp_get_threads())


def bytes_per_sample(fmt) -> int:
    return int(load().atkdsp_bytes_per_sample(_fmt(fmt)))


def _fmt(fmt) -> int:
    if isinstance(fmt, str):
        if fmt not in FMT:
            raise AtkDspError(f"unknown sample format {fmt!r}")
        return FMT[fmt]
    return int(fmt)


def _cf32(a, name="array") -> np.ndarray:
    a = np.ascontiguousarray(a, dtype=np.complex64)
    if a.ndim != 1:
        raise AtkDspError(f"{name} must be one-dimensional")
    return a


def _f32(a, name="array") -> np.ndarray:
    a = np.ascontiguousarray(a, dtype=np.float32)
    if a.ndim != 1:
        raise AtkDspError(f"{name} must be one-dimensional")
    return a


def _cfp(a: np.ndarray):
    return a.ctypes.data_as(C.POINTER(_Cf32))


def _fp(a: np.ndarray):
    return a.ctypes.data_as(C.POINTER(C.c_float))


def _raw_bytes(raw) -> np.ndarray:
    if isinstance(raw, np.ndarray):
        return np.ascontiguousarray(raw).view(np.uint8).reshape(-1)
    return np.frombuffer(raw, dtype=np.uint8)


# ---- unpack ------------------------------------------------------------------
def unpack(raw, fmt, dc_alpha: float = 0.0, dc_state: np.ndarray | None = None,
           out: np.ndarray | None = None) -> np.ndarray:
    """Raw bytes -> complex64. With ``dc_alpha > 0`` the one-pole DC block runs
    with its state in ``dc_state`` (a 1-element complex64 array kept by the caller)."""
    lib = load()
    src = _raw_bytes(raw)
    n = src.size // bytes_per_sample(fmt) if bytes_per_sample(fmt) else 0
    if out is None:
        out = np.empty(n, dtype=np.complex64)
    elif out.dtype != np.complex64 or not out.flags.c_contiguous:
        raise AtkDspError("out must be a contiguous complex64 array")
    st = None
    if dc_alpha > 0:
        if dc_state is None or dc_state.dtype != np.complex64 or dc_state.size != 1:
            raise AtkDspError("dc_state must be a 1-element complex64 array")
        st = _cfp(dc_state)
    got = _check(lib.atkdsp_unpack(src.ctypes.data, src.size, _fmt(fmt), _cfp(out), out.size,
                                   float(dc_alpha), st), "unpack")
    return out[:got]


def unpack_dc(raw, fmt, offset: complex = 0j, out: np.ndarray | None = None,
              want_mean: bool = True):
    """Raw bytes -> complex64 minus ``offset``; returns ``(x, mean)`` where
    ``mean`` is the mean of the samples before the offset was removed."""
    lib = load()
    src = _raw_bytes(raw)
    bps = bytes_per_sample(fmt)
    n = src.size // bps if bps else 0
    if out is None:
        out = np.empty(n, dtype=np.complex64)
    elif out.dtype != np.complex64 or not out.flags.c_contiguous:
        raise AtkDspError("out must be a contiguous complex64 array")
    mr, mi = C.c_double(0.0), C.c_double(0.0)
    got = _check(lib.atkdsp_unpack_dc(src.ctypes.data, src.size, _fmt(fmt), _cfp(out), out.size,
                                      float(offset.real), float(offset.imag),
                                      C.byref(mr) if want_mean else None,
                                      C.byref(mi) if want_mean else None), "unpack_dc")
    return out[:got], complex(mr.value, mi.value)


# ---- FFT ------------------------------------------------------------------------
class Fft:
    """A plan for one length (any n >= 2). Forward unscaled, inverse scaled 1/n."""

    def __init__(self, n: int):
        self._lib = load()
        self._h = self._lib.atkdsp_fft_create(int(n))
        if not self._h:
            raise AtkDspError(f"fft_create({n}) failed (no GPU or out of device memory)")
        self.n = int(n)

    def close(self) -> None:
        h, self._h = getattr(self, "_h", None), None
        if h:
            self._lib.atkdsp_fft_destroy(h)

    def __del__(self):
        self.close()

    def exec(self, x, inverse: bool = False, out: np.ndarray | None = None) -> np.ndarray:
        x = _cf32(x, "x")
        if x.size != self.n:
            raise AtkDspError(f"fft: expected {self.n} samples, got {x.size}")
        if out is None:
            out = np.empty(self.n, dtype=np.complex64)
        _check(self._lib.atkdsp_fft_exec(self._h, _cfp(x), _cfp(out), 1 if inverse else 0), "fft_exec")
        return out

    def power_db(self, x, win=None, out: np.ndarray | None = None) -> np.ndarray:
        x = _cf32(x, "x")
        if x.size != self.n:
            raise AtkDspError(f"power_db: expected {self.n} samples, got {x.size}")
        w = None if win is None else _f32(win, "win")
        if out is None:
            out = np.empty(self.n, dtype=np.float32)
        _check(self._lib.atkdsp_power_db(self._h, _cfp(x), None if w is None else _fp(w), _fp(out)),
               "power_db")
        return out

    def spectrum_reduce(self, x, hop: int | None = None, win=None, detector="max",
                        out: np.ndarray | None = None):
        """Every frame of x reduced per bin into one dB line. Returns (line, frames)."""
        x = _cf32(x, "x")
        w = None if win is None else _f32(win, "win")
        if out is None:
            out = np.empty(self.n, dtype=np.float32)
        det = DET[detector] if isinstance(detector, str) else int(detector)
        frames = _check(self._lib.atkdsp_spectrum_reduce(
            self._h, _cfp(x), x.size, int(hop or self.n), None if w is None else _fp(w), det,
            _fp(out)), "spectrum_reduce")
        return out, int(frames)

    def spectrum_stats(self, x, hop: int | None = None, win=None,
                       want=("max", "avg", "min", "sk")):
        """Per-bin max/avg/min (dB) and spectral kurtosis in one pass; returns a
        dict of the requested arrays plus ``frames``."""
        x = _cf32(x, "x")
        w = None if win is None else _f32(win, "win")
        want = set(want)
        outs = {k: (np.empty(self.n, dtype=np.float32) if k in want else None)
                for k in ("max", "avg", "min", "sk")}
        ptr = {k: (_fp(v) if v is not None else None) for k, v in outs.items()}
        frames = _check(self._lib.atkdsp_spectrum_stats(
            self._h, _cfp(x), x.size, int(hop or self.n), None if w is None else _fp(w),
            ptr["max"], ptr["avg"], ptr["min"], ptr["sk"]), "spectrum_stats")
        res = {k: v for k, v in outs.items() if v is not None}
        res["frames"] = int(frames)
        return res


def window(kind, n: int) -> np.ndarray:
    k = WIN[kind] if isinstance(kind, str) else int(kind)
    out = np.empty(int(n), dtype=np.float32)
    _check(load().atkdsp_window(k, int(n), _fp(out)), "window")
    return out


def power_db(x, win=None) -> np.ndarray:
    x = _cf32(x, "x")
    return Fft(x.size).power_db(x, win)


def spectrum_reduce(x, n: int, hop: int | None = None, win=None, detector="max"):
    return Fft(n).spectrum_reduce(x, hop, win, detector)


def spectrum_stats(x, n: int, hop: int | None = None, win=None, want=("max", "avg", "min", "sk")):
    return Fft(n).spectrum_stats(x, hop, win, want)


def window_stats(win) -> tuple:
    """(coherent_gain, enbw_bins) of a window array."""
    w = _f32(win, "win")
    cg, enbw = C.c_double(0.0), C.c_double(0.0)
    _check(load().atkdsp_window_stats(_fp(w), w.size, C.byref(cg), C.byref(enbw)), "window_stats")
    return cg.value, enbw.value


# ---- display ----------------------------------------------------------------------
def db_to_pixels(db, lo: float, hi: float, lut256, out: np.ndarray | None = None) -> np.ndarray:
    db = _f32(db, "db")
    lut = np.ascontiguousarray(lut256, dtype=np.uint32)
    if lut.size != 256:
        raise AtkDspError("lut256 must have 256 entries")
    if out is None:
        out = np.empty(db.size, dtype=np.uint32)
    load().atkdsp_db_to_pixels(_fp(db), db.size, float(lo), float(hi),
                               lut.ctypes.data_as(C.POINTER(C.c_uint32)),
                               out.ctypes.data_as(C.POINTER(C.c_uint32)))
    return out


def decimate_max(x, m: int, out: np.ndarray | None = None) -> np.ndarray:
    """Peak-preserving reduction of x to m columns (m > len(x) pads with the last value)."""
    x = _f32(x, "x")
    m = int(m)
    if out is None:
        out = np.empty(m, dtype=np.float32)
    load().atkdsp_decimate_max(_fp(x), x.size, _fp(out), m)
    return out


def median(x, scratch: np.ndarray | None = None) -> float:
    """The median of x (a robust noise floor; numpy's rule for even n)."""
    x = _f32(x, "x")
    if scratch is None or scratch.size < x.size:
        scratch = np.empty(max(x.size, 1), dtype=np.float32)
    return float(load().atkdsp_median(_fp(x), x.size, _fp(scratch)))


def _channels(buf, n) -> list:
    return [Channel(c.start_bin, c.end_bin, c.centroid_bin, c.peak_db, c.snr_db) for c in buf[:n]]


def detect_channels(line, floor_db: float, threshold_db: float, gap_bins: int, min_run: int,
                    max_channels: int = 256, scratch: np.ndarray | None = None) -> list:
    """Runs of the 5-bin-smoothed line above floor + threshold (gaps under
    gap_bins bridged, runs under min_run dropped), strongest first."""
    line = _f32(line, "line")
    if scratch is None or scratch.size < line.size:
        scratch = np.empty(max(line.size, 1), dtype=np.float32)
    buf = (_ChannelC * max(int(max_channels), 1))()
    got = _check(int(load().atkdsp_detect_channels(_fp(line), line.size, float(floor_db), float(threshold_db),
                                                   int(gap_bins), int(min_run), _fp(scratch), buf,
                                                   int(max_channels))), "detect_channels")
    return _channels(buf, got)


def _rho2_arg(rho2):
    if rho2 is None:
        return None
    vals = [float(v) for v in rho2]
    if len(vals) != 4:
        raise AtkDspError("rho2 needs four correlations (bins 1..4 apart)")
    return (C.c_double * 4)(*vals)


def cfar_defaults() -> dict:
    """The default CFAR configuration as a dict."""
    c = _CfarConfig()
    load().atkdsp_cfar_defaults(C.byref(c))
    d = {k: getattr(c, k) for k, _ in _CfarConfig._fields_ if k not in ("reserved", "rho2")}
    d["rho2"] = tuple(c.rho2)
    return d


def cfar_alpha(mode, n_ref: int, pfa: float, looks: float = 1.0, os_k: int = 0, rho2=None) -> tuple:
    """(alpha, E[T]): the CFAR threshold factor for a false-alarm rate pfa per
    cell, with the window's bin correlation rho2 (four values, or None)."""
    mode = CFAR_MODES[mode] if isinstance(mode, str) else int(mode)
    a, et = C.c_double(), C.c_double()
    _check(load().atkdsp_cfar_alpha(mode, int(n_ref), float(pfa), float(looks), int(os_k), _rho2_arg(rho2),
                                    C.byref(a), C.byref(et)), "cfar_alpha")
    return a.value, et.value


def cfar_n_eff(n_ref: int, rho2=None) -> float:
    """Equivalent independent cells among 2 n_ref window-correlated ones."""
    return float(load().atkdsp_cfar_n_eff(int(n_ref), _rho2_arg(rho2)))


def cfar_window_rho2(window) -> tuple:
    """The power correlation of periodogram bins 1..4 apart through window."""
    w = _f32(window, "window")
    out = (C.c_double * 4)()
    _check(load().atkdsp_cfar_window_rho2(_fp(w), w.size, out), "cfar_window_rho2")
    return tuple(float(v) for v in out)


class Cfar:
    """CFAR detection over spectrum frames and the burst tracker (atkdsp_cfar,
    reference include/atkdsp.h:1733-1925, src/cfar.c:270-819). Frames of
    nbins dB values are tested on the GPU; events come back as records of
    CFAR_EVENT_DTYPE. Keyword arguments override the default configuration
    (mode as "ca" / "os" / "go" / "so" or 0..3, rho2 as four values)."""

    def __init__(self, nbins, **cfg):
        self._lib = load()
        c = _CfarConfig()
        self._lib.atkdsp_cfar_defaults(C.byref(c))
        for k, v in cfg.items():
            if k == "mode":
                v = CFAR_MODES[v] if isinstance(v, str) else int(v)
            if k == "rho2":
                c.rho2[:] = [float(x) for x in v]
            elif k in dict(_CfarConfig._fields_) and k != "reserved":
                setattr(c, k, v)
            else:
                raise AtkDspError(f"cfar: unknown setting {k}")
        self.nbins = int(nbins)
        self._p = self._lib.atkdsp_cfar_create(self.nbins, C.byref(c))
        if not self._p:
            raise AtkDspError("cfar_create: a setting is out of range (or there is no GPU)")

    def __del__(self):
        p, self._p = getattr(self, "_p", None), None
        if p:
            self._lib.atkdsp_cfar_destroy(p)

    def reset(self):
        self._lib.atkdsp_cfar_reset(self._p)

    def set_mask(self, mask=None):
        m = None if mask is None else np.ascontiguousarray(mask, dtype=np.uint8)
        if m is not None and m.size != self.nbins:
            raise AtkDspError("cfar mask needs nbins values")
        _check(self._lib.atkdsp_cfar_set_mask(self._p, None if m is None else m.ctypes.data), "cfar_set_mask")

    def events_max(self, nframes):
        return int(self._lib.atkdsp_cfar_events_max(self._p, int(nframes)))

    def process(self, frames_db, t0=0.0, dt=1.0) -> np.ndarray:
        f = np.ascontiguousarray(frames_db, dtype=np.float32).reshape(-1)
        if f.size % self.nbins:
            raise AtkDspError("frames_db must hold whole frames of nbins values")
        nf = f.size // self.nbins
        ev = np.empty(max(self.events_max(nf), 1), CFAR_EVENT_DTYPE)
        n = _check(int(self._lib.atkdsp_cfar_process(self._p, _fp(f), nf, float(t0), float(dt), ev.ctypes.data,
                                                     ev.size if nf else 0)), "cfar_process")
        return ev[:n].copy()

    def flush(self) -> np.ndarray:
        ev = np.empty(65536, CFAR_EVENT_DTYPE)
        n = _check(int(self._lib.atkdsp_cfar_flush(self._p, ev.ctypes.data, ev.size)), "cfar_flush")
        return ev[:n].copy()

    def last(self):
        """(average dB, threshold dB, hits) of the last tested frame"""
        a = np.empty(self.nbins, np.float32)
        t = np.empty(self.nbins, np.float32)
        h = np.empty(self.nbins, np.uint8)
        _check(self._lib.atkdsp_cfar_last(self._p, _fp(a), _fp(t), h.ctypes.data), "cfar_last")
        return a, t, h

    def info(self) -> dict:
        a, aca, et, c = C.c_double(), C.c_double(), C.c_double(), _CfarConfig()
        _check(self._lib.atkdsp_cfar_info(self._p, C.byref(a), C.byref(aca), C.byref(et), C.byref(c)), "cfar_info")
        d = {k: getattr(c, k) for k, _ in _CfarConfig._fields_ if k not in ("reserved", "rho2")}
        d.update(rho2=tuple(c.rho2), alpha=a.value, alpha_ca=aca.value, et=et.value)
        return d

    def describe(self) -> str:
        buf = C.create_string_buffer(512)
        _check(self._lib.atkdsp_cfar_describe(self._p, buf, 512), "cfar_describe")
        return buf.value.decode()

    def stats(self) -> dict:
        st = _CfarCounters()
        _check(self._lib.atkdsp_cfar_stats(self._p, C.byref(st)), "cfar_stats")
        return {k: int(getattr(st, k)) for k, _ in _CfarCounters._fields_}


def stitch_max(seg, seg_lo_hz: float, seg_hz_per_bin: float, keep_lo_hz: float,
               keep_hi_hz: float, out_lo_hz: float, out_hz_per_bin: float,
               out: np.ndarray) -> np.ndarray:
    """Max-accumulate one segment into ``out`` (float32, NaN = unmeasured), in place."""
    seg = _f32(seg, "seg")
    if out.dtype != np.float32 or not out.flags.c_contiguous:
        raise AtkDspError("out must be a contiguous float32 array")
    load().atkdsp_stitch_max(_fp(seg), seg.size, float(seg_lo_hz), float(seg_hz_per_bin),
                             float(keep_lo_hz), float(keep_hi_hz), float(out_lo_hz),
                             float(out_hz_per_bin), _fp(out), out.size)
    return out


# ---- polyphase filterbank --------------------------------------------------------------
class Nco:
    """Phase-continuous complex mixer (reference python/atkdsp/__init__.py
    Nco): ``mix`` shifts a channel at +freq to DC; the FP64 phase is carried
    in the struct across calls, the rotation runs on the GPU."""

    def __init__(self, freq_hz: float, sample_rate: float):
        self._lib = load()
        self._s = _Nco()
        self._lib.atkdsp_nco_init(C.byref(self._s), float(freq_hz), float(sample_rate))

    @property
    def phase(self) -> float:
        return float(self._s.phase)

    def set_freq(self, freq_hz: float, sample_rate: float) -> None:
        self._lib.atkdsp_nco_set_freq(C.byref(self._s), float(freq_hz), float(sample_rate))

    def mix(self, x, out: np.ndarray | None = None) -> np.ndarray:
        x = _cf32(x, "x")
        if out is None:
            out = np.empty_like(x)
        self._lib.atkdsp_nco_mix(C.byref(self._s), x.ctypes.data_as(C.POINTER(_Cf32)),
                                 out.ctypes.data_as(C.POINTER(_Cf32)), x.size)
        return out


class Pfb:
    """Oversampled polyphase filterbank: nchan channels spaced fs/nchan,
    `oversample` x; process_n runs N element streams on one block clock."""

    def __init__(self, fs, nchan, oversample=2, taps_per_poly=12, atten_db=80.0, max_nel=1):
        self._lib = load()
        self._p = self._lib.atkdsp_pfb_create(float(fs), int(nchan), int(oversample),
                                              int(taps_per_poly), float(atten_db), int(max_nel))
        if not self._p:
            raise AtkDspError("pfb_create: check fs>0, nchan>=2, nchan%oversample==0 (and a GPU)")
        self.nchan = int(nchan)
        self.max_nel = int(max_nel)

    def close(self) -> None:
        p, self._p = getattr(self, "_p", None), None
        if p:
            self._lib.atkdsp_pfb_destroy(p)

    def __del__(self):
        self.close()

    def reset(self):
        self._lib.atkdsp_pfb_reset(self._p)

    def out_rate(self):
        return self._lib.atkdsp_pfb_out_rate(self._p)

    def group_delay(self):
        return self._lib.atkdsp_pfb_group_delay(self._p)

    def block_origin(self, hop_index):
        return int(self._lib.atkdsp_pfb_block_origin(self._p, int(hop_index)))

    def out_hops(self, n):
        return int(self._lib.atkdsp_pfb_out_hops(self._p, int(n)))

    def set_cal(self, table, nel):
        if table is None:
            _check(self._lib.atkdsp_pfb_set_cal(self._p, None, 0), "pfb_set_cal")
            return
        t = np.ascontiguousarray(np.asarray(table, np.complex64)).reshape(int(nel), self.nchan)
        _check(self._lib.atkdsp_pfb_set_cal(self._p, _cfp(t), int(nel)), "pfb_set_cal")

    def describe(self):
        buf = C.create_string_buffer(256)
        n = self._lib.atkdsp_pfb_describe(self._p, buf, 256)
        return buf.value.decode() if n > 0 else ""

    def process_n(self, ins, nel):
        nel = int(nel)
        arrs = [np.ascontiguousarray(np.asarray(ins[e]).ravel(), np.complex64) for e in range(nel)]
        n = arrs[0].size
        for a in arrs:
            if a.size != n:
                raise AtkDspError("pfb: elements must be equal length")
        m = self._lib.atkdsp_pfb_out_hops(self._p, n)
        out = np.zeros(nel * self.nchan * m, np.complex64)
        inarr = (C.POINTER(_Cf32) * nel)(*[_cfp(a) for a in arrs])
        got = _check(self._lib.atkdsp_pfb_process_n(self._p, inarr, nel, n, _cfp(out), out.size),
                     "pfb_process_n")
        return out[: nel * self.nchan * got].reshape(nel, self.nchan, got)

    def process(self, x):
        return self.process_n([x], 1)[0]


def pfb_cov_accumulate(out, nel, nchan, gate, R, counts):
    """W1: R[k] += sum_i y_k y_k^H over gated channels; counts[k] += m (in place)."""
    lib = load()
    o = np.ascontiguousarray(out, np.complex64)
    nel, nchan = int(nel), int(nchan)
    m = o.shape[2] if o.ndim == 3 else 0
    if not (R.flags["C_CONTIGUOUS"] and R.dtype == np.complex64):
        raise AtkDspError("pfb_cov_accumulate: R must be contiguous complex64")
    if not (counts.flags["C_CONTIGUOUS"] and counts.dtype == np.uint32):
        raise AtkDspError("pfb_cov_accumulate: counts must be contiguous uint32")
    gp = None
    if gate is not None:
        g = np.ascontiguousarray(gate, np.uint8)
        gp = g.ctypes.data_as(C.POINTER(C.c_uint8))
    lib.atkdsp_pfb_cov_accumulate(_cfp(o.reshape(-1)), nel, nchan, m, gp, _cfp(R.reshape(-1)),
                                  counts.ctypes.data_as(C.POINTER(C.c_uint32)))


def pfb_cov_reset(R, counts, nel, nchan, gate=None):
    lib = load()
    gp = None
    if gate is not None:
        g = np.ascontiguousarray(gate, np.uint8)
        gp = g.ctypes.data_as(C.POINTER(C.c_uint8))
    lib.atkdsp_pfb_cov_reset(_cfp(R.reshape(-1)), counts.ctypes.data_as(C.POINTER(C.c_uint32)),
                             int(nel), int(nchan), gp)


def pfb_channel_power(out, nel, nchan):
    """W2: returns (pwr[nel, nchan], mean_pwr[nchan]) float32."""
    lib = load()
    o = np.ascontiguousarray(out, np.complex64)
    nel, nchan = int(nel), int(nchan)
    m = o.shape[2] if o.ndim == 3 else 0
    pwr = np.zeros(nel * nchan, np.float32)
    meanp = np.zeros(nchan, np.float32)
    lib.atkdsp_pfb_channel_power(_cfp(o.reshape(-1)), nel, nchan, m, _fp(pwr), _fp(meanp))
    return pwr.reshape(nel, nchan), meanp


def design_lowpass(fp_hz: float, fs_hz: float, atten_db: float, rate: float) -> np.ndarray:
    """Kaiser-windowed sinc lowpass (include/atkdsp.h:505-506): odd length, unit DC gain."""
    lib = load()
    n = _check(int(lib.atkdsp_design_lowpass(float(fp_hz), float(fs_hz), float(atten_db), float(rate), None, 0)),
               "design_lowpass")
    out = np.zeros(n, np.float32)
    _check(int(lib.atkdsp_design_lowpass(float(fp_hz), float(fs_hz), float(atten_db), float(rate), _fp(out), n)),
           "design_lowpass")
    return out


class ArbResampler:
    """Resample by any positive out/in ratio (atkdsp_arb_resampler, reference
    include/atkdsp.h:368-377): a polyphase bank with linear interpolation
    between phases; the position and history carry across calls."""

    def __init__(self, in_rate: float, out_rate: float, atten_db: float = 60.0, nphase: int = 64):
        self._lib = load()
        self._h = self._lib.atkdsp_arb_resampler_create(float(in_rate), float(out_rate), float(atten_db),
                                                        int(nphase))
        if not self._h:
            raise AtkDspError("arb_resampler_create failed (no GPU or bad rates)")
        self.in_rate, self.out_rate = float(in_rate), float(out_rate)

    def __del__(self):
        h, self._h = getattr(self, "_h", None), None
        if h:
            self._lib.atkdsp_arb_resampler_destroy(h)

    def reset(self) -> None:
        self._lib.atkdsp_arb_resampler_reset(self._h)

    def set_ratio(self, ratio: float) -> None:
        self._lib.atkdsp_arb_resampler_set_ratio(self._h, float(ratio))
        self.out_rate = self.in_rate * float(ratio)

    def ratio(self) -> float:
        return float(self._lib.atkdsp_arb_resampler_ratio(self._h))

    def out_max(self, n_in: int) -> int:
        return int(self._lib.atkdsp_arb_resampler_out_max(self._h, int(n_in)))

    def process(self, x, out: np.ndarray | None = None) -> np.ndarray:
        x = _cf32(x, "x")
        if out is None:
            out = np.empty(self.out_max(x.size), np.complex64)
        got = _check(int(self._lib.atkdsp_arb_resampler_process(self._h, _cfp(x), x.size, _cfp(out), out.size)),
                     "arb_resampler_process")
        return out[:got]


class Osx:
    """Overlap-save extractor (atkdsp_osx, reference include/atkdsp.h:3252-3289):
    arbitrary channels at arbitrary centres from one forward FFT per block,
    each landed on its own out_rate; add/retune/remove between calls without
    disturbing the others; process_n runs N elements phase-coherently."""

    def __init__(self, fs, nfft=4096, atten_db=80.0, max_nel=1):
        self._lib = load()
        self._p = self._lib.atkdsp_osx_create(float(fs), int(nfft), float(atten_db), int(max_nel))
        if not self._p:
            raise AtkDspError("osx_create: check fs>0, nfft a power of two >= 64 (and a GPU)")
        self.max_nel = int(max_nel)

    def __del__(self):
        p, self._p = getattr(self, "_p", None), None
        if p:
            self._lib.atkdsp_osx_destroy(p)

    def reset(self):
        self._lib.atkdsp_osx_reset(self._p)

    def add(self, offset_hz, bw_hz, out_rate):
        ch = int(self._lib.atkdsp_osx_add(self._p, float(offset_hz), float(bw_hz), float(out_rate)))
        return _check(ch, "osx_add")

    def retune(self, ch, offset_hz):
        _check(self._lib.atkdsp_osx_retune(self._p, int(ch), float(offset_hz)), "osx_retune")

    def remove(self, ch):
        _check(self._lib.atkdsp_osx_remove(self._p, int(ch)), "osx_remove")

    def channels(self):
        return int(self._lib.atkdsp_osx_channels(self._p))

    def chan_rate(self, ch):
        return float(self._lib.atkdsp_osx_chan_rate(self._p, int(ch)))

    def out_max(self, ch, n):
        return int(self._lib.atkdsp_osx_out_max(self._p, int(ch), int(n)))

    def describe(self):
        buf = C.create_string_buffer(256)
        n = self._lib.atkdsp_osx_describe(self._p, buf, 256)
        return buf.value.decode() if n > 0 else ""

    def process_n(self, ins, nel):
        """outs[ch] is [nel][count] complex64 (None for a free slot); counts[ch]."""
        nel = int(nel)
        arrs = [_cf32(np.asarray(ins[e]).ravel(), "in") for e in range(nel)]
        n = arrs[0].size
        ns = self.channels()
        active = [self.chan_rate(ci) > 0 for ci in range(ns)]
        cap_each = max([self.out_max(ci, n) for ci in range(ns) if active[ci]], default=0)
        bufs = [[(np.zeros(cap_each, np.complex64) if active[ci] else None) for ci in range(ns)]
                for _ in range(nel)]
        ptrs = (C.POINTER(_Cf32) * max(1, nel * ns))()
        for e in range(nel):
            for ci in range(ns):
                ptrs[e * ns + ci] = _cfp(bufs[e][ci]) if active[ci] else None
        counts = (C.c_size_t * max(1, ns))()
        inarr = (C.POINTER(_Cf32) * nel)(*[_cfp(a) for a in arrs])
        _check(int(self._lib.atkdsp_osx_process_n(self._p, inarr, nel, n, ptrs, counts, cap_each)),
               "osx_process_n")
        cnts = [int(counts[ci]) for ci in range(ns)]
        outs = [None] * ns
        for ci in range(ns):
            if active[ci]:
                m = cnts[ci]
                outs[ci] = (np.array([bufs[e][ci][:m] for e in range(nel)], np.complex64)
                            if m else np.zeros((nel, 0), np.complex64))
        return outs, cnts

    def process(self, x):
        outs, cnts = self.process_n([x], 1)
        return [o[0] if o is not None else None for o in outs], cnts


def _meta_dict(m) -> dict:
    return {k: (float(getattr(m, k)) if isinstance(getattr(m, k), float) else int(getattr(m, k)))
            for k, _ in _RingMeta._fields_}


def ring_bytes(block_bytes: int, nblocks: int, max_consumers: int = 4) -> int:
    """Bytes of caller memory a ring of this shape needs; 0 if impossible."""
    return int(load().atkdsp_ring_bytes(int(block_bytes), int(nblocks), int(max_consumers)))


class Ring:
    """The block ring (atkdsp_ring, reference include/atkdsp.h:2020-2107) in
    caller memory: one producer that never waits, up to 16 consumers, a
    seqlock per block. `mem` (a writable uint8 array or any object with a
    buffer) attaches to an existing ring, made by this library or another;
    otherwise a new ring is formatted in a fresh 64-byte-aligned buffer."""

    def __init__(self, block_bytes: int = 0, nblocks: int = 0, max_consumers: int = 4, mem=None):
        self._lib = load()
        if mem is not None:
            self._buf = np.frombuffer(mem, np.uint8)
            if not self._buf.flags.writeable:        # consumers and the producer write into the ring
                raise AtkDspError("ring_attach: the ring's memory must be writable")
            self._r = self._lib.atkdsp_ring_attach(self._buf.ctypes.data, self._buf.size)
            if not self._r:
                raise AtkDspError("ring_attach: not a ring of this version")
        else:
            need = ring_bytes(block_bytes, nblocks, max_consumers)
            if not need:
                raise AtkDspError("ring: block_bytes >= 1, nblocks >= 2, max_consumers 1..16")
            raw = np.zeros(need + 64, np.uint8)
            off = (-raw.ctypes.data) % 64
            self._buf = raw[off:off + need]
            self._r = self._lib.atkdsp_ring_init(self._buf.ctypes.data, need, int(block_bytes), int(nblocks),
                                                 int(max_consumers))
            if not self._r:
                raise AtkDspError("ring_init failed")
        st = self.stats()
        self.block_bytes, self.nblocks = st["block_bytes"], st["nblocks"]
        self._rbuf = np.empty(self.block_bytes, np.uint8)

    @property
    def handle(self) -> int:
        """the atkdsp_ring pointer (for atkgpu_server_submit_ring)"""
        return int(self._r)

    @property
    def memory(self) -> np.ndarray:
        return self._buf

    def claim(self) -> np.ndarray:
        p = self._lib.atkdsp_ring_claim(self._r)
        return np.ctypeslib.as_array((C.c_uint8 * self.block_bytes).from_address(p))

    def commit(self, nbytes: int, **meta) -> None:
        """publish the claimed block; meta: first_sample, t_utc_ns, fc_hz, fs_hz,
        gain_db, format (an enum value or a name such as "ci16"), flags"""
        if isinstance(meta.get("format"), str):
            meta["format"] = _fmt(meta["format"])
        m = _RingMeta(**{k: v for k, v in meta.items() if k != "seq"})
        m.nbytes = int(nbytes)
        _check(self._lib.atkdsp_ring_commit(self._r, C.byref(m)), "ring_commit")

    def write(self, data, **meta) -> None:
        b = np.frombuffer(data, np.uint8) if not isinstance(data, np.ndarray) else data.view(np.uint8).ravel()
        self.claim()[: b.size] = b
        self.commit(b.size, **meta)

    def open(self, from_oldest: bool = False) -> int:
        return _check(int(self._lib.atkdsp_ring_open(self._r, int(bool(from_oldest)))), "ring_open")

    def close_consumer(self, cid: int) -> None:
        _check(self._lib.atkdsp_ring_close(self._r, int(cid)), "ring_close")

    def read(self, cid: int):
        """(bytes, meta, lost), or (None, None, lost) when there is nothing new"""
        m, lost = _RingMeta(), C.c_uint64(0)
        n = _check(int(self._lib.atkdsp_ring_read(self._r, int(cid), self._rbuf.ctypes.data, self._rbuf.size,
                                                  C.byref(m), C.byref(lost))), "ring_read")
        if n == 0:
            return None, None, int(lost.value)
        return self._rbuf[:n].copy(), _meta_dict(m), int(lost.value)

    def peek(self, cid: int):
        """(view into the ring, meta, lost), or (None, None, lost)"""
        m, lost = _RingMeta(), C.c_uint64(0)
        p = self._lib.atkdsp_ring_peek(self._r, int(cid), C.byref(m), C.byref(lost))
        if not p:
            return None, None, int(lost.value)
        return (np.ctypeslib.as_array((C.c_uint8 * m.nbytes).from_address(p)), _meta_dict(m), int(lost.value))

    def release(self, cid: int) -> bool:
        """True if the peeked block survived the use; False if it was overwritten"""
        rc = int(self._lib.atkdsp_ring_release(self._r, int(cid)))
        if rc == -6:
            return False
        _check(rc, "ring_release")
        return True

    def copy_range(self, seq0: int, seq1: int, cap: int | None = None, mcap: int | None = None):
        """whole blocks [seq0, seq1) still held, packed: (bytes, [meta])"""
        cap = self.block_bytes * max(0, seq1 - seq0) if cap is None else cap
        mcap = max(0, seq1 - seq0) if mcap is None else mcap
        out = np.empty(max(cap, 1), np.uint8)
        metas = (_RingMeta * max(mcap, 1))()
        n = _check(int(self._lib.atkdsp_ring_copy_range(self._r, int(seq0), int(seq1), out.ctypes.data, cap,
                                                        metas, mcap)), "ring_copy_range")
        ms = [_meta_dict(metas[i]) for i in range(n)]
        return out[: sum(m["nbytes"] for m in ms)], ms

    def stats(self) -> dict:
        st = _RingCounters()
        _check(self._lib.atkdsp_ring_stats(self._r, C.byref(st)), "ring_stats")
        d = {k: getattr(st, k) for k, _ in _RingCounters._fields_ if k not in ("lost", "read", "lag")}
        for k in ("lost", "read", "lag"):
            d[k] = list(getattr(st, k))
        return d

    def describe(self) -> str:
        buf = C.create_string_buffer(256)
        self._lib.atkdsp_ring_describe(self._r, buf, 256)
        return buf.value.decode()


__all__ = [
    "ABI_VERSION", "AtkDspError", "load", "available", "lib_path", "version", "build_info",
    "set_threads", "get_threads", "FMT", "DET", "WIN", "bytes_per_sample",
    "unpack", "unpack_dc", "Fft", "window", "power_db", "spectrum_reduce", "spectrum_stats",
    "window_stats", "db_to_pixels", "decimate_max", "stitch_max", "median", "detect_channels", "Channel",
    "cfar_defaults", "cfar_alpha", "cfar_n_eff", "cfar_window_rho2", "CFAR_MODES", "Cfar", "CFAR_EVENT",
    "CFAR_EVENT_DTYPE",
    "Pfb", "pfb_cov_accumulate", "pfb_cov_reset", "pfb_channel_power",
    "design_lowpass", "ArbResampler", "Osx", "Ring", "ring_bytes",
]
