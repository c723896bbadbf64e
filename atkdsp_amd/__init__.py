"""atkdsp_amd — the wideband front end of atkdsp on one MI355X.

A ctypes binding over ``atkdsp_amd/atkdsp.so``. Its public names, argument
meaning and error behaviour are those of the reference binding's front-end
subset (reference python/atkdsp/__init__.py: ``load`` :238-268, ``unpack``
:337, ``unpack_dc`` :361, ``Fft`` :531-599, ``window`` :602, ``db_to_pixels``
:673, ``decimate_max`` :686, ``stitch_max`` :719, ``Pfb`` :4159-4219,
``pfb_cov_accumulate`` :4222, ``pfb_channel_power`` :4252), so code written
against ``atkdsp`` runs against this package. The library is also a drop-in
for the reference binding itself: point ``ATKDSP_LIB`` at ``atkdsp_amd/atkdsp.so``.

How this binding is put together:
  - every C entry point is declared once, in ``_SIGNATURES``, as a line
    ``name  result:arguments`` in one-letter type codes (``_CODES``);
  - ``_Library`` opens the shared object, checks its ABI and applies the
    declarations;
  - ``_arr`` / ``_out`` turn Python values into C arguments and allocate
    results; ``_check`` turns a negative status into :class:`AtkDspError`.

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


# ---- C record types ----------------------------------------------------------------------
class _Cf32(C.Structure):
    _fields_ = [("re", C.c_float), ("im", C.c_float)]


class _ChannelC(C.Structure):
    _fields_ = [("start_bin", C.c_int), ("end_bin", C.c_int), ("centroid_bin", C.c_float),
                ("peak_db", C.c_float), ("snr_db", C.c_float)]


class _CfarConfig(C.Structure):
    _fields_ = [("mode", C.c_int32), ("n_ref", C.c_int32), ("n_guard", C.c_int32), ("k_frames", C.c_int32),
                ("pfa", C.c_double), ("looks", C.c_double), ("os_k", C.c_int32), ("open_m", C.c_int32),
                ("open_n", C.c_int32), ("close_k", C.c_int32), ("merge_gap", C.c_int32), ("min_bins", C.c_int32),
                ("max_tracks", C.c_int32), ("reserved", C.c_int32), ("rho2", C.c_double * 4)]


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


CFAR_MODES = {"ca": 0, "os": 1, "go": 2, "so": 3}
CFAR_EVENT = {1: "open", 2: "update", 3: "close"}
# the 64-byte event record (atkdsp_cfar_event) as a numpy record type
CFAR_EVENT_DTYPE = np.dtype([("t0", "<f8"), ("t1", "<f8"), ("fc_bin", "<f8"), ("type", "<i4"), ("id", "<u4"),
                             ("flags", "<u4"), ("lo_bin", "<i4"), ("hi_bin", "<i4"), ("n_hits", "<i4"),
                             ("level_db", "<f4"), ("noise_db", "<f4"), ("snr_db", "<f4"),
                             ("band_snr_db", "<f4")])
assert CFAR_EVENT_DTYPE.itemsize == 64

# ---- declarations ---------------------------------------------------------------------------
_P = C.POINTER
_CODES = {
    "v": None, "i": C.c_int, "u": C.c_uint, "z": C.c_size_t, "s": C.c_ssize_t, "d": C.c_double,
    "f": C.c_float, "p": C.c_void_p, "S": C.c_char_p, "L": C.c_int64, "Q": C.c_uint64,
    "c": _P(_Cf32), "F": _P(C.c_float), "D": _P(C.c_double), "U": _P(C.c_uint32), "B": _P(C.c_uint8),
    "q": _P(C.c_uint64), "Z": _P(C.c_size_t), "C": _P(_P(_Cf32)),
    "K": _P(_CfarConfig), "X": _P(_CfarCounters), "H": _P(_ChannelC), "M": _P(_RingMeta),
    "R": _P(_RingCounters), "N": _P(_Nco),
}
# name  result:arguments (codes above)
_SIGNATURES = """
atkdsp_abi_version i:
atkdsp_version_string S:
atkdsp_build_info S:
atkdsp_bytes_per_sample i:i
atkdsp_set_threads i:i
atkdsp_get_threads i:
atkdsp_unpack s:pzicz fc
atkdsp_unpack_dc s:pzicz ffDD
atkdsp_fft_create p:z
atkdsp_fft_destroy v:p
atkdsp_fft_length z:p
atkdsp_fft_exec i:pcci
atkdsp_window i:izF
atkdsp_power_db i:pcFF
atkdsp_spectrum_reduce s:pczzFiF
atkdsp_spectrum_stats s:pczzFFFFF
atkdsp_window_stats i:FzDD
atkdsp_db_to_pixels v:FzffUU
atkdsp_decimate_max v:FzFz
atkdsp_stitch_max v:FzddddddFz
atkdsp_median f:FzF
atkdsp_detect_channels s:FzffiiFHz
atkdsp_cfar_defaults v:K
atkdsp_cfar_alpha i:iiddiDDD
atkdsp_cfar_n_eff d:iD
atkdsp_cfar_window_rho2 i:FzD
atkdsp_cfar_create p:zK
atkdsp_cfar_destroy v:p
atkdsp_cfar_reset v:p
atkdsp_cfar_set_mask i:pp
atkdsp_cfar_events_max z:pz
atkdsp_cfar_process s:pFzddpz
atkdsp_cfar_flush s:ppz
atkdsp_cfar_last i:pFFp
atkdsp_cfar_info i:pDDDK
atkdsp_cfar_describe i:pSz
atkdsp_cfar_stats i:pX
atkdsp_pfb_create p:duuudu
atkdsp_pfb_destroy v:p
atkdsp_pfb_reset v:p
atkdsp_pfb_nchan u:p
atkdsp_pfb_out_rate d:p
atkdsp_pfb_out_hops z:pz
atkdsp_pfb_process s:pczcz
atkdsp_pfb_process_n s:pCuzcz
atkdsp_pfb_block_origin L:pL
atkdsp_pfb_group_delay d:p
atkdsp_pfb_set_cal i:pcu
atkdsp_pfb_describe i:pSz
atkdsp_pfb_cov_accumulate v:cuuzBcU
atkdsp_pfb_cov_reset v:cUuuB
atkdsp_pfb_channel_power v:cuuzFF
atkdsp_nco_init v:Ndd
atkdsp_nco_set_freq v:Ndd
atkdsp_nco_mix v:Nccz
atkdsp_design_lowpass s:ddddFz
atkdsp_arb_resampler_create p:dddu
atkdsp_arb_resampler_destroy v:p
atkdsp_arb_resampler_reset v:p
atkdsp_arb_resampler_set_ratio v:pd
atkdsp_arb_resampler_ratio d:p
atkdsp_arb_resampler_out_max z:pz
atkdsp_arb_resampler_process s:pczcz
atkdsp_osx_create p:dzdu
atkdsp_osx_destroy v:p
atkdsp_osx_reset v:p
atkdsp_osx_add i:pddd
atkdsp_osx_retune i:pid
atkdsp_osx_remove i:pi
atkdsp_osx_channels i:p
atkdsp_osx_chan_rate d:pi
atkdsp_osx_out_max z:piz
atkdsp_osx_process s:pczCZz
atkdsp_osx_process_n s:pCuzCZz
atkdsp_osx_describe i:pSz
atkdsp_ring_bytes z:zzi
atkdsp_ring_init p:pzzzi
atkdsp_ring_attach p:pz
atkdsp_ring_claim p:p
atkdsp_ring_commit i:pM
atkdsp_ring_open i:pi
atkdsp_ring_close i:pi
atkdsp_ring_read s:pipzMq
atkdsp_ring_peek p:piMq
atkdsp_ring_release i:pi
atkdsp_ring_copy_range s:pQQpzMz
atkdsp_ring_stats i:pR
atkdsp_ring_describe i:pSz
"""


def _declare(lib) -> None:
    for line in _SIGNATURES.split("\n"):
        if not line.strip():
            continue
        name, sig = line.split(None, 1)
        res, args = sig.replace(" ", "").split(":")
        f = getattr(lib, name)
        f.restype = _CODES[res]
        f.argtypes = [_CODES[k] for k in args]


_STATUS = {-1: "bad argument", -2: "out of memory", -3: "output buffer too small",
           -4: "FFT backend or device failure", -6: "overwritten while it was read (overrun)",
           -7: "nothing to return yet"}


def _check(rc: int, what: str) -> int:
    """rc if it is a count or ATKDSP_OK; AtkDspError for a negative status"""
    rc = int(rc)
    if rc >= 0:
        return rc
    raise AtkDspError(f"{what}: {_STATUS.get(rc, f'error {rc}')}")


class _Library:
    """The opened shared object, declared and ABI-checked (one per process)."""
    handle = None
    path: Path | None = None

    @classmethod
    def search(cls) -> list[Path]:
        env = os.environ.get("ATKDSP_AMD_LIB")
        return ([Path(env)] if env else []) + [_HERE / LIB_NAME]

    @classmethod
    def open(cls, path=None):
        if cls.handle is not None and path is None:
            return cls.handle
        why = []
        for cand in [Path(path)] if path else cls.search():
            if not cand.is_file():
                why.append(f"{cand}: missing (build it: python -m atkdsp_amd.build)")
                continue
            try:
                lib = C.CDLL(str(cand))
            except OSError as e:
                why.append(f"{cand}: {e}")
                continue
            lib.atkdsp_abi_version.restype = C.c_int
            abi = lib.atkdsp_abi_version()
            if abi != ABI_VERSION:
                why.append(f"{cand}: built for ABI {abi}, not {ABI_VERSION}")
                continue
            _declare(lib)
            cls.handle, cls.path = lib, cand
            return lib
        raise AtkDspError("atkdsp_amd library not loaded: " + "; ".join(why))


def load(path: str | os.PathLike | None = None):
    """The library handle (opened on first use). Raises AtkDspError."""
    return _Library.open(path)


def available() -> bool:
    """Whether the library opens here (a GPU may still be missing). Never raises."""
    try:
        _Library.open()
    except AtkDspError:
        return False
    return True


def lib_path() -> Path | None:
    return _Library.path


def version() -> str:
    return load().atkdsp_version_string().decode()


def build_info() -> str:
    return load().atkdsp_build_info().decode()


def set_threads(n: int) -> int:
    return int(load().atkdsp_set_threads(int(n)))


def get_threads() -> int:
    return int(load().atkdsp_get_threads())


def _fmt(fmt) -> int:
    """a sample format code from its name or number"""
    if not isinstance(fmt, str):
        return int(fmt)
    try:
        return FMT[fmt]
    except KeyError:
        raise AtkDspError(f"unknown sample format {fmt!r}") from None


def bytes_per_sample(fmt) -> int:
    return int(load().atkdsp_bytes_per_sample(_fmt(fmt)))


# ---- marshalling ------------------------------------------------------------------------------
def _arr(a, dtype, name="array") -> np.ndarray:
    """a as a contiguous one-dimensional array of dtype (copied only if needed)"""
    v = np.ascontiguousarray(a, dtype=dtype)
    if v.ndim != 1:
        raise AtkDspError(f"{name}: one-dimensional data expected, got shape {v.shape}")
    return v


def _cf32(a, name="array") -> np.ndarray:
    return _arr(a, np.complex64, name)


def _f32(a, name="array") -> np.ndarray:
    return _arr(a, np.float32, name)


def _out(out, n, dtype, name="out") -> np.ndarray:
    """a caller's output array (checked) or a new one of n elements"""
    if out is None:
        return np.empty(n, dtype=dtype)
    if out.dtype != np.dtype(dtype) or not out.flags.c_contiguous:
        raise AtkDspError(f"{name} must be a contiguous {np.dtype(dtype).name} array")
    return out


def _cfp(a: np.ndarray):
    return a.ctypes.data_as(_CODES["c"])


def _fp(a: np.ndarray):
    return a.ctypes.data_as(_CODES["F"])


def _up(a: np.ndarray):
    return a.ctypes.data_as(_CODES["U"])


def _bp(a):
    return None if a is None else a.ctypes.data_as(_CODES["B"])


def _opt_fp(a):
    return None if a is None else _fp(a)


def _raw_bytes(raw) -> np.ndarray:
    if isinstance(raw, np.ndarray):
        return np.ascontiguousarray(raw).reshape(-1).view(np.uint8)
    return np.frombuffer(raw, dtype=np.uint8)


# ---- unpack ----------------------------------------------------------------------------------
def _samples_in(src: np.ndarray, fmt) -> int:
    bps = bytes_per_sample(fmt)
    return src.size // bps if bps > 0 else 0


def unpack(raw, fmt, dc_alpha: float = 0.0, dc_state: np.ndarray | None = None,
           out: np.ndarray | None = None) -> np.ndarray:
    """Receiver bytes to complex64 samples. A positive ``dc_alpha`` runs the
    one-pole DC blocker, whose state the caller keeps in ``dc_state`` (a
    one-element complex64 array)."""
    src = _raw_bytes(raw)
    y = _out(out, _samples_in(src, fmt), np.complex64)
    state = None
    if dc_alpha > 0:
        ok = isinstance(dc_state, np.ndarray) and dc_state.dtype == np.complex64 and dc_state.size == 1
        if not ok:
            raise AtkDspError("dc_state must be a 1-element complex64 array")
        state = _cfp(dc_state)
    n = _check(load().atkdsp_unpack(src.ctypes.data, src.size, _fmt(fmt), _cfp(y), y.size, float(dc_alpha), state),
               "unpack")
    return y[:n]


def unpack_dc(raw, fmt, offset: complex = 0j, out: np.ndarray | None = None,
              want_mean: bool = True):
    """Receiver bytes to complex64 samples less ``offset``; returns
    ``(samples, mean)``, the mean being that of the samples before the offset."""
    src = _raw_bytes(raw)
    y = _out(out, _samples_in(src, fmt), np.complex64)
    acc = (C.c_double(), C.c_double())
    ptrs = (C.byref(acc[0]), C.byref(acc[1])) if want_mean else (None, None)
    n = _check(load().atkdsp_unpack_dc(src.ctypes.data, src.size, _fmt(fmt), _cfp(y), y.size, float(offset.real),
                                       float(offset.imag), *ptrs), "unpack_dc")
    return y[:n], complex(acc[0].value, acc[1].value)


# ---- FFT and spectrum lines ------------------------------------------------------------------
def _det(detector) -> int:
    return DET[detector] if isinstance(detector, str) else int(detector)


class Fft:
    """An FFT plan of one length n >= 2: forward unscaled, inverse scaled by 1/n."""

    def __init__(self, n: int):
        self._lib = load()
        self.n = int(n)
        self._h = self._lib.atkdsp_fft_create(self.n)
        if not self._h:
            raise AtkDspError(f"fft_create({self.n}) failed (no GPU or out of device memory)")

    def close(self) -> None:
        h, self._h = getattr(self, "_h", None), None
        if h:
            self._lib.atkdsp_fft_destroy(h)

    def __del__(self):
        self.close()

    def _frame(self, x, what):
        v = _cf32(x, "x")
        if v.size != self.n:
            raise AtkDspError(f"{what}: expected {self.n} samples, got {v.size}")
        return v

    def exec(self, x, inverse: bool = False, out: np.ndarray | None = None) -> np.ndarray:
        v = self._frame(x, "fft")
        y = _out(out, self.n, np.complex64)
        _check(self._lib.atkdsp_fft_exec(self._h, _cfp(v), _cfp(y), int(bool(inverse))), "fft_exec")
        return y

    def power_db(self, x, win=None, out: np.ndarray | None = None) -> np.ndarray:
        v = self._frame(x, "power_db")
        w = None if win is None else _f32(win, "win")
        y = _out(out, self.n, np.float32)
        _check(self._lib.atkdsp_power_db(self._h, _cfp(v), _opt_fp(w), _fp(y)), "power_db")
        return y

    def spectrum_reduce(self, x, hop: int | None = None, win=None, detector="max",
                        out: np.ndarray | None = None):
        """All frames of x (hop apart) reduced bin by bin to one dB line by the
        detector. Returns (line, frames)."""
        v = _cf32(x, "x")
        w = None if win is None else _f32(win, "win")
        y = _out(out, self.n, np.float32)
        nf = _check(self._lib.atkdsp_spectrum_reduce(self._h, _cfp(v), v.size, int(hop or self.n), _opt_fp(w),
                                                      _det(detector), _fp(y)), "spectrum_reduce")
        return y, nf

    def spectrum_stats(self, x, hop: int | None = None, win=None,
                       want=("max", "avg", "min", "sk")):
        """The per-bin max, average and min (dB) and spectral kurtosis of the
        frames, from one pass; a dict of the planes asked for, plus ``frames``."""
        v = _cf32(x, "x")
        w = None if win is None else _f32(win, "win")
        planes = {k: np.empty(self.n, np.float32) for k in ("max", "avg", "min", "sk") if k in set(want)}
        args = [_opt_fp(planes.get(k)) for k in ("max", "avg", "min", "sk")]
        nf = _check(self._lib.atkdsp_spectrum_stats(self._h, _cfp(v), v.size, int(hop or self.n), _opt_fp(w), *args),
                    "spectrum_stats")
        return dict(planes, frames=nf)


def window(kind, n: int) -> np.ndarray:
    y = np.empty(int(n), np.float32)
    _check(load().atkdsp_window(WIN[kind] if isinstance(kind, str) else int(kind), y.size, _fp(y)), "window")
    return y


def power_db(x, win=None) -> np.ndarray:
    v = _cf32(x, "x")
    return Fft(v.size).power_db(v, win)


def spectrum_reduce(x, n: int, hop: int | None = None, win=None, detector="max"):
    return Fft(n).spectrum_reduce(x, hop, win, detector)


def spectrum_stats(x, n: int, hop: int | None = None, win=None, want=("max", "avg", "min", "sk")):
    return Fft(n).spectrum_stats(x, hop, win, want)


def window_stats(win) -> tuple:
    """(coherent gain, equivalent noise bandwidth in bins) of a window"""
    w = _f32(win, "win")
    cg, enbw = C.c_double(), C.c_double()
    _check(load().atkdsp_window_stats(_fp(w), w.size, C.byref(cg), C.byref(enbw)), "window_stats")
    return cg.value, enbw.value


# ---- display lines -------------------------------------------------------------------------------
def db_to_pixels(db, lo: float, hi: float, lut256, out: np.ndarray | None = None) -> np.ndarray:
    """dB values to colours: [lo, hi] spread over the 256 entries of lut256."""
    v = _f32(db, "db")
    lut = _arr(lut256, np.uint32, "lut256")
    if lut.size != 256:
        raise AtkDspError(f"lut256 must have 256 entries, not {lut.size}")
    y = _out(out, v.size, np.uint32)
    load().atkdsp_db_to_pixels(_fp(v), v.size, float(lo), float(hi), _up(lut), _up(y))
    return y


def decimate_max(x, m: int, out: np.ndarray | None = None) -> np.ndarray:
    """x reduced to m columns, each the maximum of its bins (a shorter x is
    widened with its last value)."""
    v = _f32(x, "x")
    y = _out(out, int(m), np.float32)
    load().atkdsp_decimate_max(_fp(v), v.size, _fp(y), int(m))
    return y


def stitch_max(seg, seg_lo_hz: float, seg_hz_per_bin: float, keep_lo_hz: float,
               keep_hi_hz: float, out_lo_hz: float, out_hz_per_bin: float,
               out: np.ndarray) -> np.ndarray:
    """Fold one segment's bins into the wider trace ``out`` by frequency,
    keeping the larger value where one is already there (NaN in ``out`` means
    not yet measured). ``out`` is updated in place and returned."""
    v = _f32(seg, "seg")
    trace = _out(out, 0, np.float32)
    hz = [float(a) for a in (seg_lo_hz, seg_hz_per_bin, keep_lo_hz, keep_hi_hz, out_lo_hz, out_hz_per_bin)]
    load().atkdsp_stitch_max(_fp(v), v.size, *hz, _fp(trace), trace.size)
    return trace


def median(x, scratch: np.ndarray | None = None) -> float:
    """The median of x (numpy's rule for an even count): a robust noise floor."""
    v = _f32(x, "x")
    work = scratch if scratch is not None and scratch.size >= v.size else np.empty(max(v.size, 1), np.float32)
    return float(load().atkdsp_median(_fp(v), v.size, _fp(work)))


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


def _channels(buf, n) -> list:
    return [Channel(*(getattr(r, f) for f, _ in _ChannelC._fields_)) for r in buf[:n]]


def detect_channels(line, floor_db: float, threshold_db: float, gap_bins: int, min_run: int,
                    max_channels: int = 256, scratch: np.ndarray | None = None) -> list:
    """Runs of the 5-bin-smoothed line above floor + threshold (gaps under
    gap_bins bridged, runs under min_run dropped), strongest first."""
    v = _f32(line, "line")
    work = scratch if scratch is not None and scratch.size >= v.size else np.empty(max(v.size, 1), np.float32)
    cap = int(max_channels)
    recs = (_ChannelC * max(cap, 1))()
    n = _check(load().atkdsp_detect_channels(_fp(v), v.size, float(floor_db), float(threshold_db), int(gap_bins),
                                             int(min_run), _fp(work), recs, cap), "detect_channels")
    return _channels(recs, n)


# ---- CFAR --------------------------------------------------------------------------------------
def _struct_dict(s, skip=()) -> dict:
    return {k: getattr(s, k) for k, _ in s._fields_ if k not in skip}


def _rho2_arg(rho2):
    if rho2 is None:
        return None
    vals = [float(v) for v in rho2]
    if len(vals) != 4:
        raise AtkDspError("rho2 needs four correlations (bins 1..4 apart)")
    return (C.c_double * 4)(*vals)


def _cfar_mode(mode) -> int:
    return CFAR_MODES[mode] if isinstance(mode, str) else int(mode)


def cfar_defaults() -> dict:
    """The default CFAR configuration as a dict."""
    cfg = _CfarConfig()
    load().atkdsp_cfar_defaults(C.byref(cfg))
    return dict(_struct_dict(cfg, ("reserved", "rho2")), rho2=tuple(cfg.rho2))


def cfar_alpha(mode, n_ref: int, pfa: float, looks: float = 1.0, os_k: int = 0, rho2=None) -> tuple:
    """(alpha, E[T]): the threshold factor for a false-alarm rate pfa per
    cell, given the window's bin correlation rho2 (four values, or None)."""
    a, et = C.c_double(), C.c_double()
    _check(load().atkdsp_cfar_alpha(_cfar_mode(mode), int(n_ref), float(pfa), float(looks), int(os_k),
                                    _rho2_arg(rho2), C.byref(a), C.byref(et)), "cfar_alpha")
    return a.value, et.value


def cfar_n_eff(n_ref: int, rho2=None) -> float:
    """Equivalent independent cells among 2 n_ref window-correlated ones."""
    return float(load().atkdsp_cfar_n_eff(int(n_ref), _rho2_arg(rho2)))


def cfar_window_rho2(window) -> tuple:
    """The power correlation of periodogram bins 1..4 apart through window."""
    w = _f32(window, "window")
    rho = (C.c_double * 4)()
    _check(load().atkdsp_cfar_window_rho2(_fp(w), w.size, rho), "cfar_window_rho2")
    return tuple(rho)


class Cfar:
    """CFAR detection over spectrum frames and the burst tracker (atkdsp_cfar,
    reference include/atkdsp.h:1733-1925, src/cfar.c:270-819). Frames of
    nbins dB values are tested on the GPU; events come back as records of
    CFAR_EVENT_DTYPE. Keyword arguments override the default configuration
    (mode as "ca" / "os" / "go" / "so" or 0..3, rho2 as four values)."""

    def __init__(self, nbins, **settings):
        self._lib = load()
        cfg = _CfarConfig()
        self._lib.atkdsp_cfar_defaults(C.byref(cfg))
        known = {k for k, _ in _CfarConfig._fields_} - {"reserved"}
        for key, val in settings.items():
            if key not in known:
                raise AtkDspError(f"cfar: unknown setting {key}")
            if key == "rho2":
                cfg.rho2[:] = [float(r) for r in val]
            else:
                setattr(cfg, key, _cfar_mode(val) if key == "mode" else val)
        self.nbins = int(nbins)
        self._p = self._lib.atkdsp_cfar_create(self.nbins, C.byref(cfg))
        if not self._p:
            raise AtkDspError("cfar_create: a setting is out of range (or there is no GPU)")

    def __del__(self):
        p, self._p = getattr(self, "_p", None), None
        if p:
            self._lib.atkdsp_cfar_destroy(p)

    def reset(self):
        self._lib.atkdsp_cfar_reset(self._p)

    def set_mask(self, mask=None):
        m = None if mask is None else _arr(mask, np.uint8, "mask")
        if m is not None and m.size != self.nbins:
            raise AtkDspError(f"cfar mask needs {self.nbins} values")
        _check(self._lib.atkdsp_cfar_set_mask(self._p, None if m is None else m.ctypes.data), "cfar_set_mask")

    def events_max(self, nframes):
        return int(self._lib.atkdsp_cfar_events_max(self._p, int(nframes)))

    def process(self, frames_db, t0=0.0, dt=1.0) -> np.ndarray:
        f = np.ascontiguousarray(frames_db, dtype=np.float32).reshape(-1)
        nf, rest = divmod(f.size, self.nbins)
        if rest:
            raise AtkDspError("frames_db must hold whole frames of nbins values")
        ev = np.empty(max(self.events_max(nf), 1), CFAR_EVENT_DTYPE)
        n = _check(self._lib.atkdsp_cfar_process(self._p, _fp(f), nf, float(t0), float(dt), ev.ctypes.data,
                                                 ev.size if nf else 0), "cfar_process")
        return ev[:n].copy()

    def flush(self) -> np.ndarray:
        ev = np.empty(65536, CFAR_EVENT_DTYPE)
        n = _check(self._lib.atkdsp_cfar_flush(self._p, ev.ctypes.data, ev.size), "cfar_flush")
        return ev[:n].copy()

    def last(self):
        """(average dB, threshold dB, hits) of the last tested frame"""
        avg, thr = np.empty(self.nbins, np.float32), np.empty(self.nbins, np.float32)
        hits = np.empty(self.nbins, np.uint8)
        _check(self._lib.atkdsp_cfar_last(self._p, _fp(avg), _fp(thr), hits.ctypes.data), "cfar_last")
        return avg, thr, hits

    def info(self) -> dict:
        vals = [C.c_double() for _ in range(3)]
        cfg = _CfarConfig()
        _check(self._lib.atkdsp_cfar_info(self._p, *[C.byref(v) for v in vals], C.byref(cfg)), "cfar_info")
        out = dict(_struct_dict(cfg, ("reserved", "rho2")), rho2=tuple(cfg.rho2))
        out.update(zip(("alpha", "alpha_ca", "et"), (v.value for v in vals)))
        return out

    def describe(self) -> str:
        text = C.create_string_buffer(512)
        _check(self._lib.atkdsp_cfar_describe(self._p, text, 512), "cfar_describe")
        return text.value.decode()

    def stats(self) -> dict:
        st = _CfarCounters()
        _check(self._lib.atkdsp_cfar_stats(self._p, C.byref(st)), "cfar_stats")
        return {k: int(v) for k, v in _struct_dict(st).items()}


# ---- NCO ---------------------------------------------------------------------------------------
class Nco:
    """A phase-continuous complex mixer (reference python/atkdsp/__init__.py
    Nco): ``mix`` brings a channel at +freq to DC. The FP64 phase lives in the
    struct between calls; the rotation runs on the GPU."""

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
        v = _cf32(x, "x")
        y = _out(out, v.size, np.complex64)
        self._lib.atkdsp_nco_mix(C.byref(self._s), _cfp(v), _cfp(y), v.size)
        return y


# ---- polyphase filterbank --------------------------------------------------------------------
class Pfb:
    """The oversampled polyphase filterbank: nchan channels fs/nchan apart,
    ``oversample`` times critically sampled; process_n runs several element
    streams on one block clock."""

    def __init__(self, fs, nchan, oversample=2, taps_per_poly=12, atten_db=80.0, max_nel=1):
        self._lib = load()
        self.nchan, self.max_nel = int(nchan), int(max_nel)
        self._p = self._lib.atkdsp_pfb_create(float(fs), self.nchan, int(oversample), int(taps_per_poly),
                                              float(atten_db), self.max_nel)
        if not self._p:
            raise AtkDspError("pfb_create: check fs>0, nchan>=2, nchan%oversample==0 (and a GPU)")

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
        """Per-element, per-channel complex factors applied to every output
        (None clears them)."""
        if table is None:
            _check(self._lib.atkdsp_pfb_set_cal(self._p, None, 0), "pfb_set_cal")
            return
        t = np.ascontiguousarray(table, dtype=np.complex64).reshape(int(nel) * self.nchan)
        _check(self._lib.atkdsp_pfb_set_cal(self._p, _cfp(t), int(nel)), "pfb_set_cal")

    def describe(self):
        text = C.create_string_buffer(256)
        return text.value.decode() if self._lib.atkdsp_pfb_describe(self._p, text, 256) > 0 else ""

    def process_n(self, ins, nel):
        """nel equal-length element streams in; outputs [nel][nchan][hops]."""
        nel = int(nel)
        streams = [np.ascontiguousarray(np.asarray(ins[e]).reshape(-1), dtype=np.complex64) for e in range(nel)]
        n = streams[0].size
        if any(s.size != n for s in streams):
            raise AtkDspError("pfb: elements must be equal length")
        hops = self.out_hops(n)
        y = np.zeros(nel * self.nchan * hops, np.complex64)
        ptrs = (_CODES["c"] * nel)(*map(_cfp, streams))
        got = _check(self._lib.atkdsp_pfb_process_n(self._p, ptrs, nel, n, _cfp(y), y.size), "pfb_process_n")
        return y[: nel * self.nchan * got].reshape(nel, self.nchan, got)

    def process(self, x):
        return self.process_n([x], 1)[0]


def _hops_of(y: np.ndarray) -> int:
    return y.shape[2] if y.ndim == 3 else 0


def _inplace(a: np.ndarray, dtype, what: str) -> np.ndarray:
    if a.dtype != np.dtype(dtype) or not a.flags.c_contiguous:
        raise AtkDspError(f"{what} must be contiguous {np.dtype(dtype).name}")
    return a


def pfb_cov_accumulate(out, nel, nchan, gate, R, counts):
    """W1 over a filterbank output [nel][nchan][hops]: for each channel k the
    gate lets through, R[k] += the sum over hops of y y^H and counts[k] +=
    hops; R and counts are updated in place."""
    y = np.ascontiguousarray(out, np.complex64)
    _inplace(R, np.complex64, "pfb_cov_accumulate: R")
    _inplace(counts, np.uint32, "pfb_cov_accumulate: counts")
    g = None if gate is None else np.ascontiguousarray(gate, np.uint8)
    load().atkdsp_pfb_cov_accumulate(_cfp(y.reshape(-1)), int(nel), int(nchan), _hops_of(y), _bp(g),
                                     _cfp(R.reshape(-1)), _up(counts))


def pfb_cov_reset(R, counts, nel, nchan, gate=None):
    """Zero R and counts for the gated channels (all when gate is None)."""
    g = None if gate is None else np.ascontiguousarray(gate, np.uint8)
    load().atkdsp_pfb_cov_reset(_cfp(R.reshape(-1)), _up(counts), int(nel), int(nchan), _bp(g))


def pfb_channel_power(out, nel, nchan):
    """W2: the mean power of each element's channels and its mean over the
    elements: (pwr[nel][nchan], mean_pwr[nchan]), float32."""
    y = np.ascontiguousarray(out, np.complex64)
    pwr, mean = np.zeros(int(nel) * int(nchan), np.float32), np.zeros(int(nchan), np.float32)
    load().atkdsp_pfb_channel_power(_cfp(y.reshape(-1)), int(nel), int(nchan), _hops_of(y), _fp(pwr), _fp(mean))
    return pwr.reshape(int(nel), int(nchan)), mean


# ---- lowpass design, arbitrary resampler, overlap-save extractor ----------------------------------------
def design_lowpass(fp_hz: float, fs_hz: float, atten_db: float, rate: float) -> np.ndarray:
    """Kaiser-windowed sinc lowpass (include/atkdsp.h:505-506): odd length, unit DC gain."""
    lib = load()
    args = (float(fp_hz), float(fs_hz), float(atten_db), float(rate))
    n = _check(lib.atkdsp_design_lowpass(*args, None, 0), "design_lowpass")
    taps = np.zeros(n, np.float32)
    _check(lib.atkdsp_design_lowpass(*args, _fp(taps), n), "design_lowpass")
    return taps


class ArbResampler:
    """Resample by any positive out/in ratio (atkdsp_arb_resampler, reference
    include/atkdsp.h:368-377): a polyphase bank with linear interpolation
    between phases; the position and history carry across calls."""

    def __init__(self, in_rate: float, out_rate: float, atten_db: float = 60.0, nphase: int = 64):
        self._lib = load()
        self.in_rate, self.out_rate = float(in_rate), float(out_rate)
        self._h = self._lib.atkdsp_arb_resampler_create(self.in_rate, self.out_rate, float(atten_db), int(nphase))
        if not self._h:
            raise AtkDspError("arb_resampler_create failed (no GPU or bad rates)")

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
        v = _cf32(x, "x")
        y = _out(out, self.out_max(v.size), np.complex64)
        n = _check(self._lib.atkdsp_arb_resampler_process(self._h, _cfp(v), v.size, _cfp(y), y.size),
                   "arb_resampler_process")
        return y[:n]


class Osx:
    """Overlap-save extractor (atkdsp_osx, reference include/atkdsp.h:3252-3289):
    arbitrary channels at arbitrary centres from one forward FFT per block,
    each landed on its own out_rate; add/retune/remove between calls without
    disturbing the others; process_n runs N elements phase-coherently."""

    def __init__(self, fs, nfft=4096, atten_db=80.0, max_nel=1):
        self._lib = load()
        self.max_nel = int(max_nel)
        self._p = self._lib.atkdsp_osx_create(float(fs), int(nfft), float(atten_db), self.max_nel)
        if not self._p:
            raise AtkDspError("osx_create: check fs>0, nfft a power of two >= 64 (and a GPU)")

    def __del__(self):
        p, self._p = getattr(self, "_p", None), None
        if p:
            self._lib.atkdsp_osx_destroy(p)

    def reset(self):
        self._lib.atkdsp_osx_reset(self._p)

    def add(self, offset_hz, bw_hz, out_rate):
        return _check(self._lib.atkdsp_osx_add(self._p, float(offset_hz), float(bw_hz), float(out_rate)), "osx_add")

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
        text = C.create_string_buffer(256)
        return text.value.decode() if self._lib.atkdsp_osx_describe(self._p, text, 256) > 0 else ""

    def process_n(self, ins, nel):
        """outs[ch] is [nel][count] complex64 (None for a free slot); counts[ch]."""
        nel = int(nel)
        streams = [_cf32(np.asarray(ins[e]).reshape(-1), "in") for e in range(nel)]
        n = streams[0].size
        slots = self.channels()
        live = [self.chan_rate(ch) > 0 for ch in range(slots)]
        per = max((self.out_max(ch, n) for ch in range(slots) if live[ch]), default=0)
        bufs = [[np.zeros(per, np.complex64) if live[ch] else None for ch in range(slots)] for _ in range(nel)]
        ptrs = (_CODES["c"] * max(1, nel * slots))()
        for e in range(nel):
            for ch in range(slots):
                if live[ch]:
                    ptrs[e * slots + ch] = _cfp(bufs[e][ch])
        counts = (C.c_size_t * max(1, slots))()
        _check(self._lib.atkdsp_osx_process_n(self._p, (_CODES["c"] * nel)(*map(_cfp, streams)), nel, n, ptrs,
                                              counts, per), "osx_process_n")
        got = [int(counts[ch]) for ch in range(slots)]
        outs = []
        for ch in range(slots):
            if not live[ch]:
                outs.append(None)
            else:
                outs.append(np.stack([bufs[e][ch][:got[ch]] for e in range(nel)]).astype(np.complex64)
                            if got[ch] else np.zeros((nel, 0), np.complex64))
        return outs, got

    def process(self, x):
        outs, counts = self.process_n([x], 1)
        return [None if o is None else o[0] for o in outs], counts


# ---- the block ring ----------------------------------------------------------------------------
def _meta_dict(m) -> dict:
    return {k: (float(v) if isinstance(v, float) else int(v)) for k, v in _struct_dict(m).items()}


def ring_bytes(block_bytes: int, nblocks: int, max_consumers: int = 4) -> int:
    """Bytes of caller memory a ring of this shape needs; 0 if impossible."""
    return int(load().atkdsp_ring_bytes(int(block_bytes), int(nblocks), int(max_consumers)))


class Ring:
    """The block ring (atkdsp_ring, reference include/atkdsp.h:2020-2107) in
    caller memory: one producer that never waits, up to 16 consumers, a
    seqlock per block. ``mem`` (writable bytes: a uint8 array or any buffer)
    attaches to a ring made by this library or another; without it a new ring
    is formatted in a fresh 64-byte-aligned buffer."""

    def __init__(self, block_bytes: int = 0, nblocks: int = 0, max_consumers: int = 4, mem=None):
        self._lib = load()
        if mem is None:
            self._buf = self._new_memory(int(block_bytes), int(nblocks), int(max_consumers))
            self._r = self._lib.atkdsp_ring_init(self._buf.ctypes.data, self._buf.size, int(block_bytes),
                                                 int(nblocks), int(max_consumers))
            if not self._r:
                raise AtkDspError("ring_init failed")
        else:
            self._buf = np.frombuffer(mem, np.uint8)
            if not self._buf.flags.writeable:
                raise AtkDspError("ring_attach: the ring's memory must be writable (readers move their cursors)")
            self._r = self._lib.atkdsp_ring_attach(self._buf.ctypes.data, self._buf.size)
            if not self._r:
                raise AtkDspError("ring_attach: not a ring of this version")
        shape = self.stats()
        self.block_bytes, self.nblocks = shape["block_bytes"], shape["nblocks"]
        self._rbuf = np.empty(self.block_bytes, np.uint8)

    @staticmethod
    def _new_memory(block_bytes, nblocks, max_consumers) -> np.ndarray:
        need = ring_bytes(block_bytes, nblocks, max_consumers)
        if not need:
            raise AtkDspError("ring: block_bytes >= 1, nblocks >= 2, max_consumers 1..16")
        pool = np.zeros(need + 64, np.uint8)
        skip = -pool.ctypes.data % 64
        return pool[skip:skip + need]

    @property
    def handle(self) -> int:
        """the atkdsp_ring pointer (for atkgpu_server_submit_ring)"""
        return int(self._r)

    @property
    def memory(self) -> np.ndarray:
        return self._buf

    def claim(self) -> np.ndarray:
        addr = self._lib.atkdsp_ring_claim(self._r)
        return np.ctypeslib.as_array((C.c_uint8 * self.block_bytes).from_address(addr))

    def commit(self, nbytes: int, **meta) -> None:
        """Publish the claimed block. meta: first_sample, t_utc_ns, fc_hz,
        fs_hz, gain_db, format (a code or a name such as "ci16"), flags."""
        if isinstance(meta.get("format"), str):
            meta["format"] = _fmt(meta["format"])
        m = _RingMeta(**{k: v for k, v in meta.items() if k != "seq"})
        m.nbytes = int(nbytes)
        _check(self._lib.atkdsp_ring_commit(self._r, C.byref(m)), "ring_commit")

    def write(self, data, **meta) -> None:
        b = data.view(np.uint8).reshape(-1) if isinstance(data, np.ndarray) else np.frombuffer(data, np.uint8)
        self.claim()[: b.size] = b
        self.commit(b.size, **meta)

    def open(self, from_oldest: bool = False) -> int:
        return _check(self._lib.atkdsp_ring_open(self._r, int(bool(from_oldest))), "ring_open")

    def close_consumer(self, cid: int) -> None:
        _check(self._lib.atkdsp_ring_close(self._r, int(cid)), "ring_close")

    def read(self, cid: int):
        """(bytes, meta, lost), or (None, None, lost) when nothing new is there"""
        m, lost = _RingMeta(), C.c_uint64()
        n = _check(self._lib.atkdsp_ring_read(self._r, int(cid), self._rbuf.ctypes.data, self._rbuf.size,
                                              C.byref(m), C.byref(lost)), "ring_read")
        return (self._rbuf[:n].copy(), _meta_dict(m), lost.value) if n else (None, None, lost.value)

    def peek(self, cid: int):
        """(a view into the ring, meta, lost), or (None, None, lost)"""
        m, lost = _RingMeta(), C.c_uint64()
        addr = self._lib.atkdsp_ring_peek(self._r, int(cid), C.byref(m), C.byref(lost))
        if not addr:
            return None, None, lost.value
        return np.ctypeslib.as_array((C.c_uint8 * m.nbytes).from_address(addr)), _meta_dict(m), lost.value

    def release(self, cid: int) -> bool:
        """True if the peeked block was not overwritten while it was used"""
        rc = int(self._lib.atkdsp_ring_release(self._r, int(cid)))
        return rc != -6 and _check(rc, "ring_release") == 0

    def copy_range(self, seq0: int, seq1: int, cap: int | None = None, mcap: int | None = None):
        """Whole blocks [seq0, seq1) the ring still holds, packed: (bytes, [meta])."""
        span = max(0, seq1 - seq0)
        cap = self.block_bytes * span if cap is None else cap
        mcap = span if mcap is None else mcap
        data = np.empty(max(cap, 1), np.uint8)
        metas = (_RingMeta * max(mcap, 1))()
        n = _check(self._lib.atkdsp_ring_copy_range(self._r, int(seq0), int(seq1), data.ctypes.data, cap, metas,
                                                    mcap), "ring_copy_range")
        ms = [_meta_dict(metas[i]) for i in range(n)]
        return data[: sum(m["nbytes"] for m in ms)], ms

    def stats(self) -> dict:
        st = _RingCounters()
        _check(self._lib.atkdsp_ring_stats(self._r, C.byref(st)), "ring_stats")
        per = ("lost", "read", "lag")
        return dict(_struct_dict(st, per), **{k: list(getattr(st, k)) for k in per})

    def describe(self) -> str:
        text = C.create_string_buffer(256)
        self._lib.atkdsp_ring_describe(self._r, text, 256)
        return text.value.decode()


__all__ = [
    "ABI_VERSION", "AtkDspError", "load", "available", "lib_path", "version", "build_info",
    "set_threads", "get_threads", "FMT", "DET", "WIN", "bytes_per_sample",
    "unpack", "unpack_dc", "Fft", "window", "power_db", "spectrum_reduce", "spectrum_stats",
    "window_stats", "db_to_pixels", "decimate_max", "stitch_max", "median", "detect_channels", "Channel",
    "cfar_defaults", "cfar_alpha", "cfar_n_eff", "cfar_window_rho2", "CFAR_MODES", "Cfar", "CFAR_EVENT",
    "CFAR_EVENT_DTYPE", "Nco",
    "Pfb", "pfb_cov_accumulate", "pfb_cov_reset", "pfb_channel_power",
    "design_lowpass", "ArbResampler", "Osx", "Ring", "ring_bytes",
]
