"""The serving loop (include/atkgpu.h atkgpu_server_*) from Python.

    srv = Server(hbm_budget=32 << 30, max_streams=128)
    sid = srv.add_stream(fmt="ci16", fs=40e6, nchan=1024, taps_per_poly=16, chunk_max=65536)
    t = srv.submit(sid, raw_bytes)        # returns a ticket at once
    y = srv.read(sid, t)                  # [nchan][m] once the chunk is done
    t, meta, lost = srv.submit_ring(sid, ring, consumer)   # or from a block ring
    srv.stats()                           # chunks, samples, p50/p99/p99.9 latency

A stream may carry a display (display=dict(nfft=4096, hop=2048, frames=977,
cols=1024, ...)): line L is spectrum_reduce over frames [L*frames,
(L+1)*frames) of the stream, decimated to cols columns; srv.read_line(sid, L).
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import CFAR_EVENT_DTYPE, AtkDspError, _ChannelC, _Cf32, _RingMeta, _channels, _check, _fmt, _meta_dict, load

RING_EMPTY = -64               # atkgpu.h ATKGPU_RING_EMPTY: the ring consumer has nothing new


class _StreamCfg(C.Structure):
    _fields_ = [("fmt", C.c_int), ("fs", C.c_double), ("nchan", C.c_uint), ("oversample", C.c_uint),
                ("taps_per_poly", C.c_uint), ("atten_db", C.c_double), ("off_re", C.c_float),
                ("off_im", C.c_float), ("chunk_max", C.c_size_t), ("out_slots", C.c_uint), ("nel", C.c_uint),
                ("coherent", C.c_int), ("disp_nfft", C.c_uint), ("disp_hop", C.c_uint),
                ("disp_window", C.c_int), ("disp_detector", C.c_int), ("disp_frames", C.c_uint),
                ("disp_cols", C.c_uint), ("disp_slots", C.c_uint)]

_WINDOWS = {"rect": 0, "hann": 1, "hamming": 2, "blackmanharris": 3, "bh": 3}
_DETECTORS = {"max": 0, "min": 1, "avg": 2}


class _Stats(C.Structure):
    _fields_ = [("chunks", C.c_longlong), ("samples", C.c_longlong), ("batches", C.c_longlong),
                ("p50_ms", C.c_double), ("p99_ms", C.c_double), ("p999_ms", C.c_double),
                ("max_ms", C.c_double), ("mean_ms", C.c_double), ("busy_s", C.c_double),
                ("hbm_used", C.c_size_t), ("hbm_budget", C.c_size_t), ("streams", C.c_uint),
                ("lat_count", C.c_longlong), ("lat_rel_res", C.c_double)]


def _bind(lib) -> None:
    if getattr(lib, "_serve_bound", False):
        return
    P, vp, sz = C.POINTER, C.c_void_p, C.c_size_t
    sig = {
        "atkgpu_server_create": (vp, [sz, C.c_uint, C.c_uint, sz]),
        "atkgpu_server_destroy": (None, [vp]),
        "atkgpu_server_add_stream": (C.c_int, [vp, P(_StreamCfg)]),
        "atkgpu_server_remove_stream": (C.c_int, [vp, C.c_int]),
        "atkgpu_server_submit": (C.c_longlong, [vp, C.c_int, vp, sz]),
        "atkgpu_server_submit_ring": (C.c_longlong, [vp, C.c_int, vp, C.c_int, P(_RingMeta),
                                                     P(C.c_uint64)]),
        "atkgpu_server_wait": (C.c_int, [vp, C.c_int, C.c_longlong]),
        "atkgpu_server_read": (C.c_ssize_t, [vp, C.c_int, C.c_longlong, P(_Cf32), sz]),
        "atkgpu_server_drain": (C.c_int, [vp]),
        "atkgpu_server_read_cov": (C.c_int, [vp, C.c_int, P(_Cf32), P(C.c_uint32)]),
        "atkgpu_server_read_power": (C.c_int, [vp, C.c_int, P(C.c_float), P(C.c_float)]),
        "atkgpu_server_stats": (C.c_int, [vp, P(_Stats), C.c_int]),
        "atkgpu_server_lines": (C.c_longlong, [vp, C.c_int]),
        "atkgpu_server_read_line": (C.c_int, [vp, C.c_int, C.c_longlong, P(C.c_float), sz]),
        "atkgpu_server_detect": (C.c_ssize_t, [vp, C.c_int, C.c_longlong, C.c_uint, C.c_float, C.c_float,
                                               C.c_int, C.c_int, P(_ChannelC), sz, P(C.c_float)]),
        "atkgpu_server_detect_latest": (C.c_int, [vp, P(C.c_int), C.c_uint, C.c_float, C.c_int, C.c_int,
                                                  P(_ChannelC), sz, P(C.c_ssize_t), P(C.c_longlong),
                                                  P(C.c_float)]),
        "atkgpu_server_cfar": (C.c_ssize_t, [vp, C.c_int, C.c_uint, vp, C.c_longlong, sz, C.c_double, C.c_double,
                                             vp, sz]),
    }
    for name, (res, args) in sig.items():
        f = getattr(lib, name)
        f.restype, f.argtypes = res, args
    lib._serve_bound = True


class Server:
    def __init__(self, hbm_budget: int = 32 << 30, max_streams: int = 128, max_batch: int = 256,
                 staging_bytes: int = 256 << 20):
        self._lib = load()
        _bind(self._lib)
        self._h = self._lib.atkgpu_server_create(int(hbm_budget), int(max_streams), int(max_batch),
                                                  int(staging_bytes))
        if not self._h:
            raise AtkDspError("atkgpu_server_create failed (no GPU or out of memory)")
        self._nchan = {}
        self._cols = {}

    def close(self):
        h, self._h = getattr(self, "_h", None), None
        if h:
            self._lib.atkgpu_server_destroy(h)

    def __del__(self):
        self.close()

    def add_stream(self, fmt="ci16", fs=40e6, nchan=1024, oversample=2, taps_per_poly=16, atten_db=80.0,
                   offset: complex = 0j, chunk_max=65536, out_slots=4, nel=1, coherent=False,
                   display: dict | None = None) -> int:
        """nel > 1: an array of phase-coherent elements (a chunk is nel element
        blocks, element-major); coherent: W1 / W2 accumulated on the device.
        display: dict(nfft, hop, window, detector, frames, cols, slots) for
        display lines (window and detector by name or enum value)."""
        d = dict(display or {})
        nfft = int(d.get("nfft", 0))
        win = d.get("window", "hann")
        det = d.get("detector", "max")
        cfg = _StreamCfg(_fmt(fmt), float(fs), int(nchan), int(oversample), int(taps_per_poly), float(atten_db),
                         float(offset.real), float(offset.imag), int(chunk_max), int(out_slots), int(nel),
                         int(bool(coherent)), nfft, int(d.get("hop", 0)),
                         _WINDOWS[win] if isinstance(win, str) else int(win),
                         _DETECTORS[det] if isinstance(det, str) else int(det),
                         int(d.get("frames", 1)), int(d.get("cols", 0)), int(d.get("slots", 0)))
        sid = self._lib.atkgpu_server_add_stream(self._h, C.byref(cfg))
        _check(sid, "server_add_stream")
        self._nchan[sid] = (int(nchan), int(nel))
        if nfft:
            self._cols[sid] = int(d.get("cols", 0)) or nfft
        return sid

    def try_add_stream(self, **kw):
        """add_stream, returning None when admission is refused."""
        try:
            return self.add_stream(**kw)
        except AtkDspError:
            return None

    def remove_stream(self, sid: int) -> None:
        _check(self._lib.atkgpu_server_remove_stream(self._h, int(sid)), "server_remove_stream")
        self._nchan.pop(sid, None)

    def submit(self, sid: int, raw) -> int:
        b = np.frombuffer(raw, np.uint8) if not isinstance(raw, np.ndarray) else raw.view(np.uint8).ravel()
        t = self._lib.atkgpu_server_submit(self._h, int(sid), b.ctypes.data, b.size)
        return _check(int(t), "server_submit")

    def submit_ring(self, sid: int, ring, consumer: int):
        """The next intact block of a ring consumer (atkdsp_amd.Ring) as one
        chunk: (ticket, block meta, blocks lost before it), or (None, None,
        lost) when the consumer has nothing new."""
        m, lost = _RingMeta(), C.c_uint64()
        t = int(self._lib.atkgpu_server_submit_ring(self._h, int(sid), ring.handle, int(consumer), C.byref(m),
                                                    C.byref(lost)))
        if t == RING_EMPTY:
            return None, None, lost.value
        return _check(t, "server_submit_ring"), _meta_dict(m), lost.value

    def wait(self, sid: int, ticket: int) -> None:
        _check(self._lib.atkgpu_server_wait(self._h, int(sid), int(ticket)), "server_wait")

    def read(self, sid: int, ticket: int, max_hops: int = 1 << 17) -> np.ndarray:
        """[nchan][m] for a single stream, [nel][nchan][m] for an array."""
        M, nel = self._nchan[sid]
        out = np.empty(nel * M * max_hops, np.complex64)
        m = _check(self._lib.atkgpu_server_read(self._h, int(sid), int(ticket),
                                                out.ctypes.data_as(C.POINTER(_Cf32)), out.size), "server_read")
        y = out[: nel * M * m]
        return y.reshape(M, m) if nel == 1 else y.reshape(nel, M, m)

    def read_cov(self, sid: int):
        """W1 of a coherent array over every chunk submitted so far: R [nchan][nel][nel], counts [nchan]."""
        M, nel = self._nchan[sid]
        R = np.empty((M, nel, nel), np.complex64)
        cnt = np.empty(M, np.uint32)
        _check(self._lib.atkgpu_server_read_cov(self._h, int(sid), R.ctypes.data_as(C.POINTER(_Cf32)),
                                                cnt.ctypes.data_as(C.POINTER(C.c_uint32))), "server_read_cov")
        return R, cnt

    def read_power(self, sid: int):
        """W2 of a coherent array's latest completed chunk: pwr [nel][nchan], mean_pwr [nchan]."""
        M, nel = self._nchan[sid]
        p = np.empty((nel, M), np.float32)
        mp = np.empty(M, np.float32)
        _check(self._lib.atkgpu_server_read_power(self._h, int(sid), p.ctypes.data_as(C.POINTER(C.c_float)),
                                                  mp.ctypes.data_as(C.POINTER(C.c_float))), "server_read_power")
        return p, mp

    def lines(self, sid: int) -> int:
        """display lines of the stream complete so far"""
        return _check(int(self._lib.atkgpu_server_lines(self._h, int(sid))), "server_lines")

    def read_line(self, sid: int, line: int) -> np.ndarray:
        """display line `line` in dB: [cols] for a single stream, [nel][cols] for an array"""
        _, nel = self._nchan[sid]
        cols = self._cols[sid]
        out = np.empty(nel * cols, np.float32)
        _check(self._lib.atkgpu_server_read_line(self._h, int(sid), int(line),
                                                 out.ctypes.data_as(C.POINTER(C.c_float)), out.size),
               "server_read_line")
        return out if nel == 1 else out.reshape(nel, cols)

    def detect(self, sid: int, line: int, element: int = 0, floor_db: float | None = None,
               threshold_db: float = 6.0, gap_bins: int = 4, min_run: int = 2, max_channels: int = 256):
        """detect_channels over a served line on the device: (channels strongest
        first, the floor used). floor_db None = the median of the smoothed line."""
        buf = (_ChannelC * max(int(max_channels), 1))()
        fl = C.c_float()
        n = _check(int(self._lib.atkgpu_server_detect(self._h, int(sid), int(line), int(element),
                                                      float("nan") if floor_db is None else float(floor_db),
                                                      float(threshold_db), int(gap_bins), int(min_run), buf,
                                                      int(max_channels), C.byref(fl))), "server_detect")
        return _channels(buf, n), float(fl.value)

    def cfar(self, sid: int, cfar, first: int, n: int, t0: float = 0.0, dt: float = 1.0, element: int = 0):
        """CFAR (an atkdsp_amd.Cfar over the stream's display columns) over
        served lines [first, first + n) where they live on the device: the
        events, as Cfar.process would give for the same lines."""
        ev = np.empty(max(cfar.events_max(int(n)), 1), CFAR_EVENT_DTYPE)
        got = _check(int(self._lib.atkgpu_server_cfar(self._h, int(sid), int(element), cfar._p, int(first), int(n),
                                                      float(t0), float(dt), ev.ctypes.data, ev.size)),
                     "server_cfar")
        return ev[:got].copy()

    def detect_latest(self, sids, threshold_db: float = 6.0, gap_bins: int = 4, min_run: int = 2,
                      max_channels: int = 64):
        """detect over the newest completed line of every listed stream in one
        launch: per stream {line, channels (strongest first), floor, capped}
        (capped: more than max_channels found; channels are then the first
        ones in bin order); line -1 = no line yet."""
        n = len(sids)
        ids = (C.c_int * n)(*[int(x) for x in sids])
        buf = (_ChannelC * (n * max(int(max_channels), 1)))()
        cnt = (C.c_ssize_t * n)()
        lin = (C.c_longlong * n)()
        fl = (C.c_float * n)()
        _check(int(self._lib.atkgpu_server_detect_latest(self._h, ids, n, float(threshold_db), int(gap_bins),
                                                         int(min_run), buf, int(max_channels), cnt, lin, fl)),
               "server_detect_latest")
        out = []
        for i in range(n):
            capped = cnt[i] == -3
            k = int(max_channels) if capped else _check(int(cnt[i]), "server_detect_latest")
            base = i * int(max_channels)
            out.append({"line": int(lin[i]), "channels": _channels(buf[base:base + k], k),
                        "floor": float(fl[i]), "capped": capped})
        return out
    def drain(self) -> None:
        _check(self._lib.atkgpu_server_drain(self._h), "server_drain")

    def stats(self, reset: bool = False) -> dict:
        st = _Stats()
        _check(self._lib.atkgpu_server_stats(self._h, C.byref(st), int(reset)), "server_stats")
        return {k: getattr(st, k) for k, _ in _Stats._fields_}
