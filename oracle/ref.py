"""ctypes access to the reference library compiled from its own sources by
oracle/build_ref.sh (oracle/_ref/atkdsp_ref.so).

TEST INFRASTRUCTURE ONLY (golden vectors, pinning the numpy oracle, and the
CPU baseline timing). Independent of the product binding on purpose.
"""
from __future__ import annotations

import ctypes as C
from pathlib import Path

import numpy as np

REF_SO = Path(__file__).resolve().parent / "_ref" / "atkdsp_ref.so"


class ChanC(C.Structure):
    _fields_ = [("start_bin", C.c_int), ("end_bin", C.c_int), ("centroid_bin", C.c_float),
                ("peak_db", C.c_float), ("snr_db", C.c_float)]


class Cf(C.Structure):
    _fields_ = [("re", C.c_float), ("im", C.c_float)]


class NcoState(C.Structure):
    _fields_ = [("phase", C.c_double), ("step", C.c_double)]


class CfarCfg(C.Structure):
    _fields_ = [("mode", C.c_int32), ("n_ref", C.c_int32), ("n_guard", C.c_int32), ("k_frames", C.c_int32),
                ("pfa", C.c_double), ("looks", C.c_double), ("os_k", C.c_int32), ("open_m", C.c_int32),
                ("open_n", C.c_int32), ("close_k", C.c_int32), ("merge_gap", C.c_int32), ("min_bins", C.c_int32),
                ("max_tracks", C.c_int32), ("reserved", C.c_int32), ("rho2", C.c_double * 4)]


CFAR_EVENT = np.dtype([("t0", "<f8"), ("t1", "<f8"), ("fc_bin", "<f8"), ("type", "<i4"), ("id", "<u4"),
                       ("flags", "<u4"), ("lo_bin", "<i4"), ("hi_bin", "<i4"), ("n_hits", "<i4"),
                       ("level_db", "<f4"), ("noise_db", "<f4"), ("snr_db", "<f4"), ("band_snr_db", "<f4")])


def available() -> bool:
    return REF_SO.exists()


class Ref:
    def __init__(self, path: Path = REF_SO):
        L = C.CDLL(str(path))
        P, cf, fp, vp, sz, u, d = C.POINTER, C.POINTER(Cf), C.POINTER(C.c_float), C.c_void_p, C.c_size_t, C.c_uint, C.c_double
        sigs = {
            "atkdsp_abi_version": (C.c_int, []),
            "atkdsp_set_threads": (C.c_int, [C.c_int]),
            "atkdsp_unpack": (C.c_ssize_t, [vp, sz, C.c_int, cf, sz, C.c_float, cf]),
            "atkdsp_unpack_dc": (C.c_ssize_t, [vp, sz, C.c_int, cf, sz, C.c_float, C.c_float, P(d), P(d)]),
            "atkdsp_fft_create": (vp, [sz]),
            "atkdsp_fft_destroy": (None, [vp]),
            "atkdsp_fft_exec": (C.c_int, [vp, cf, cf, C.c_int]),
            "atkdsp_window": (C.c_int, [C.c_int, sz, fp]),
            "atkdsp_power_db": (C.c_int, [vp, cf, fp, fp]),
            "atkdsp_spectrum_reduce": (C.c_ssize_t, [vp, cf, sz, sz, fp, C.c_int, fp]),
            "atkdsp_spectrum_stats": (C.c_ssize_t, [vp, cf, sz, sz, fp, fp, fp, fp, fp]),
            "atkdsp_window_stats": (C.c_int, [fp, sz, P(d), P(d)]),
            "atkdsp_db_to_pixels": (None, [fp, sz, C.c_float, C.c_float, P(C.c_uint32), P(C.c_uint32)]),
            "atkdsp_decimate_max": (None, [fp, sz, fp, sz]),
            "atkdsp_stitch_max": (None, [fp, sz, d, d, d, d, d, d, fp, sz]),
            "atkdsp_pfb_create": (vp, [d, u, u, u, d, u]),
            "atkdsp_pfb_destroy": (None, [vp]),
            "atkdsp_pfb_out_hops": (sz, [vp, sz]),
            "atkdsp_pfb_process_n": (C.c_ssize_t, [vp, P(cf), u, sz, cf, sz]),
            "atkdsp_pfb_set_cal": (C.c_int, [vp, cf, u]),
            "atkdsp_pfb_cov_accumulate": (None, [cf, u, u, sz, P(C.c_uint8), cf, P(C.c_uint32)]),
            "atkdsp_pfb_channel_power": (None, [cf, u, u, sz, fp, fp]),
            "atkdsp_nco_init": (None, [P(NcoState), d, d]),
            "atkdsp_nco_mix": (None, [P(NcoState), cf, cf, sz]),
            "atkdsp_design_lowpass": (C.c_ssize_t, [d, d, d, d, fp, sz]),
            "atkdsp_arb_resampler_create": (vp, [d, d, d, u]),
            "atkdsp_arb_resampler_destroy": (None, [vp]),
            "atkdsp_arb_resampler_out_max": (sz, [vp, sz]),
            "atkdsp_arb_resampler_set_ratio": (None, [vp, d]),
            "atkdsp_arb_resampler_process": (C.c_ssize_t, [vp, cf, sz, cf, sz]),
            "atkdsp_osx_create": (vp, [d, sz, d, u]),
            "atkdsp_osx_destroy": (None, [vp]),
            "atkdsp_osx_add": (C.c_int, [vp, d, d, d]),
            "atkdsp_osx_retune": (C.c_int, [vp, C.c_int, d]),
            "atkdsp_osx_remove": (C.c_int, [vp, C.c_int]),
            "atkdsp_osx_channels": (C.c_int, [vp]),
            "atkdsp_osx_chan_rate": (d, [vp, C.c_int]),
            "atkdsp_osx_out_max": (sz, [vp, C.c_int, sz]),
            "atkdsp_osx_process_n": (C.c_ssize_t, [vp, P(cf), u, sz, P(cf), P(sz), sz]),
            "atkdsp_median": (C.c_float, [fp, sz, fp]),
            "atkdsp_cfar_defaults": (None, [vp]),
            "atkdsp_cfar_alpha": (C.c_int, [C.c_int, C.c_int, d, d, C.c_int, vp, P(d), P(d)]),
            "atkdsp_cfar_n_eff": (d, [C.c_int, vp]),
            "atkdsp_cfar_window_rho2": (C.c_int, [fp, sz, P(d)]),
            "atkdsp_cfar_create": (vp, [sz, P(CfarCfg)]),
            "atkdsp_cfar_destroy": (None, [vp]),
            "atkdsp_cfar_reset": (None, [vp]),
            "atkdsp_cfar_set_mask": (C.c_int, [vp, vp]),
            "atkdsp_cfar_events_max": (sz, [vp, sz]),
            "atkdsp_cfar_process": (C.c_ssize_t, [vp, fp, sz, d, d, vp, sz]),
            "atkdsp_cfar_flush": (C.c_ssize_t, [vp, vp, sz]),
            "atkdsp_cfar_last": (C.c_int, [vp, fp, fp, vp]),
            "atkdsp_cfar_info": (C.c_int, [vp, P(d), P(d), P(d), P(CfarCfg)]),
            "atkdsp_cfar_describe": (C.c_int, [vp, C.c_char_p, sz]),
            "atkdsp_cfar_stats": (C.c_int, [vp, vp]),
            "atkdsp_detect_channels": (C.c_ssize_t, [fp, sz, C.c_float, C.c_float, C.c_int, C.c_int, fp,
                                                     P(ChanC), sz]),
        }
        for name, (res, args) in sigs.items():
            f = getattr(L, name)
            f.restype, f.argtypes = res, args
        self.L = L

    def nco_blocks(self, x: np.ndarray, f: float, fs: float, blocks):
        """The reference NCO over x in consecutive blocks (sizes cycled);
        returns (output, final phase)."""
        st = NcoState()
        self.L.atkdsp_nco_init(C.byref(st), f, fs)
        out = np.empty_like(x)
        a, k = 0, 0
        while a < x.size:
            b = min(x.size, a + blocks[k % len(blocks)])
            xi = np.ascontiguousarray(x[a:b])
            yi = np.empty_like(xi)
            self.L.atkdsp_nco_mix(C.byref(st), self._cf(xi), self._cf(yi), xi.size)
            out[a:b] = yi
            a, k = b, k + 1
        return out, st.phase

    @staticmethod
    def _cf(a):
        return a.ctypes.data_as(C.POINTER(Cf))

    @staticmethod
    def _f(a):
        return a.ctypes.data_as(C.POINTER(C.c_float))

    def set_threads(self, n: int) -> int:
        return self.L.atkdsp_set_threads(n)

    def unpack_dc(self, raw: bytes, fmt: int, off: complex = 0j):
        b = np.frombuffer(raw, np.uint8)
        bps = {0: 2, 1: 2, 2: 4, 3: 4, 4: 8}[fmt]
        out = np.empty(b.size // bps, np.complex64)
        mr, mi = C.c_double(), C.c_double()
        n = self.L.atkdsp_unpack_dc(b.ctypes.data, b.size, fmt, self._cf(out), out.size,
                                    float(off.real), float(off.imag), C.byref(mr), C.byref(mi))
        return out[:n], complex(mr.value, mi.value)

    def unpack(self, raw: bytes, fmt: int, alpha: float, state: np.ndarray):
        b = np.frombuffer(raw, np.uint8)
        bps = {0: 2, 1: 2, 2: 4, 3: 4, 4: 8}[fmt]
        out = np.empty(b.size // bps, np.complex64)
        n = self.L.atkdsp_unpack(b.ctypes.data, b.size, fmt, self._cf(out), out.size, float(alpha),
                                 self._cf(state) if state is not None else None)
        return out[:n]

    def fft(self, x: np.ndarray, inverse=False):
        x = np.ascontiguousarray(x, np.complex64)
        p = self.L.atkdsp_fft_create(x.size)
        out = np.empty_like(x)
        self.L.atkdsp_fft_exec(p, self._cf(x), self._cf(out), int(inverse))
        self.L.atkdsp_fft_destroy(p)
        return out

    def window(self, kind, n):
        out = np.empty(n, np.float32)
        self.L.atkdsp_window(kind, n, self._f(out))
        return out

    def window_stats(self, w):
        w = np.ascontiguousarray(w, np.float32)
        a, b = C.c_double(), C.c_double()
        self.L.atkdsp_window_stats(self._f(w), w.size, C.byref(a), C.byref(b))
        return a.value, b.value

    def power_db(self, x, win):
        x = np.ascontiguousarray(x, np.complex64)
        p = self.L.atkdsp_fft_create(x.size)
        out = np.empty(x.size, np.float32)
        self.L.atkdsp_power_db(p, self._cf(x), None if win is None else self._f(win), self._f(out))
        self.L.atkdsp_fft_destroy(p)
        return out

    def spectrum_reduce(self, x, n, hop, win, det):
        x = np.ascontiguousarray(x, np.complex64)
        p = self.L.atkdsp_fft_create(n)
        out = np.empty(n, np.float32)
        fr = self.L.atkdsp_spectrum_reduce(p, self._cf(x), x.size, hop or n,
                                           None if win is None else self._f(win), det, self._f(out))
        self.L.atkdsp_fft_destroy(p)
        return out, fr

    def spectrum_stats(self, x, n, hop, win):
        x = np.ascontiguousarray(x, np.complex64)
        p = self.L.atkdsp_fft_create(n)
        o = {k: np.empty(n, np.float32) for k in ("max", "avg", "min", "sk")}
        fr = self.L.atkdsp_spectrum_stats(p, self._cf(x), x.size, hop or n,
                                          None if win is None else self._f(win),
                                          self._f(o["max"]), self._f(o["avg"]), self._f(o["min"]),
                                          self._f(o["sk"]))
        self.L.atkdsp_fft_destroy(p)
        o["frames"] = fr
        return o

    def db_to_pixels(self, db, lo, hi, lut):
        db = np.ascontiguousarray(db, np.float32)
        lut = np.ascontiguousarray(lut, np.uint32)
        out = np.empty(db.size, np.uint32)
        self.L.atkdsp_db_to_pixels(self._f(db), db.size, lo, hi, lut.ctypes.data_as(C.POINTER(C.c_uint32)),
                                   out.ctypes.data_as(C.POINTER(C.c_uint32)))
        return out

    def decimate_max(self, x, m):
        x = np.ascontiguousarray(x, np.float32)
        out = np.empty(m, np.float32)
        self.L.atkdsp_decimate_max(self._f(x), x.size, self._f(out), m)
        return out

    def stitch_max(self, seg, a, b, c, d, e, f, out):
        seg = np.ascontiguousarray(seg, np.float32)
        self.L.atkdsp_stitch_max(self._f(seg), seg.size, a, b, c, d, e, f, self._f(out), out.size)
        return out

    def median(self, x):
        x = np.ascontiguousarray(x, np.float32)
        scratch = np.empty(max(x.size, 1), np.float32)
        return float(self.L.atkdsp_median(self._f(x), x.size, self._f(scratch)))

    def detect_channels(self, line, floor_db, threshold_db, gap_bins, min_run, cap=4096):
        """(records as (start, end, centroid, peak, snr) tuples, return code)"""
        line = np.ascontiguousarray(line, np.float32)
        scratch = np.empty(max(line.size, 1), np.float32)
        buf = (ChanC * max(cap, 1))()
        n = self.L.atkdsp_detect_channels(self._f(line), line.size, floor_db, threshold_db, gap_bins, min_run,
                                          self._f(scratch), buf, cap)
        kept = cap if n == -3 else max(n, 0)                 # E_CAP: the first cap found, bin order
        recs = [(c.start_bin, c.end_bin, c.centroid_bin, c.peak_db, c.snr_db) for c in buf[:kept]]
        return recs, int(n)

    def cfar_alpha(self, mode, n_ref, pfa, looks=1.0, os_k=0, rho2=None):
        """(return code, alpha, E[T])"""
        a, et = C.c_double(), C.c_double()
        r = (C.c_double * 4)(*rho2) if rho2 is not None else None
        rc = self.L.atkdsp_cfar_alpha(int(mode), int(n_ref), float(pfa), float(looks), int(os_k),
                                      C.cast(r, C.c_void_p) if r is not None else None, C.byref(a), C.byref(et))
        return int(rc), a.value, et.value

    def cfar_n_eff(self, n_ref, rho2=None):
        r = (C.c_double * 4)(*rho2) if rho2 is not None else None
        return float(self.L.atkdsp_cfar_n_eff(int(n_ref), C.cast(r, C.c_void_p) if r is not None else None))

    def cfar_window_rho2(self, w):
        w = np.ascontiguousarray(w, np.float32)
        out = (C.c_double * 4)()
        rc = self.L.atkdsp_cfar_window_rho2(self._f(w), w.size, out)
        return int(rc), tuple(out)

    def cfar_defaults_bytes(self):
        buf = (C.c_uint8 * 96)()
        self.L.atkdsp_cfar_defaults(C.cast(buf, C.c_void_p))
        return bytes(buf)

    def cfar_run(self, nbins, cfg: dict, ops):
        """Runs a CFAR handle through ops: ("process", frames (nf x nbins
        float32), t0, dt), ("mask", bytes or None), ("reset",), ("flush",).
        Returns (events per op, the 11 counters, last() as (avg, thr, hits) or
        None, (alpha, alpha_ca, et), describe text); None if create fails."""
        c = CfarCfg()
        self.L.atkdsp_cfar_defaults(C.cast(C.pointer(c), C.c_void_p))
        for k, v in cfg.items():
            if k == "rho2":
                c.rho2[:] = list(v)
            else:
                setattr(c, k, v)
        h = self.L.atkdsp_cfar_create(nbins, C.byref(c))
        if not h:
            return None
        try:
            out = []
            for op in ops:
                if op[0] == "process":
                    fr = np.ascontiguousarray(op[1], np.float32)
                    nf = fr.size // nbins
                    ev = np.empty(max(self.L.atkdsp_cfar_events_max(h, nf), 1), CFAR_EVENT)
                    n = self.L.atkdsp_cfar_process(h, self._f(fr), nf, op[2], op[3], ev.ctypes.data, ev.size)
                    assert n >= 0, n
                    out.append(ev[:n].copy())
                elif op[0] == "flush":
                    ev = np.zeros(65536, CFAR_EVENT)
                    n = self.L.atkdsp_cfar_flush(h, ev.ctypes.data, ev.size)
                    assert n >= 0, n
                    out.append(ev[:n].copy())
                elif op[0] == "mask":
                    m = None if op[1] is None else np.ascontiguousarray(op[1], np.uint8)
                    assert self.L.atkdsp_cfar_set_mask(h, None if m is None else m.ctypes.data) == 0
                    out.append(np.zeros(0, CFAR_EVENT))
                elif op[0] == "reset":
                    self.L.atkdsp_cfar_reset(h)
                    out.append(np.zeros(0, CFAR_EVENT))
            st = np.zeros(11, np.uint64)
            assert self.L.atkdsp_cfar_stats(h, st.ctypes.data) == 0
            a, t, hh = np.empty(nbins, np.float32), np.empty(nbins, np.float32), np.empty(nbins, np.uint8)
            last = (a, t, hh) if self.L.atkdsp_cfar_last(h, self._f(a), self._f(t), hh.ctypes.data) == 0 else None
            al, aca, et = C.c_double(), C.c_double(), C.c_double()
            self.L.atkdsp_cfar_info(h, C.byref(al), C.byref(aca), C.byref(et), None)
            buf = C.create_string_buffer(512)
            self.L.atkdsp_cfar_describe(h, buf, 512)
            return out, st, last, (al.value, aca.value, et.value), buf.value.decode()
        finally:
            self.L.atkdsp_cfar_destroy(h)

    def pfb(self, fs, nchan, os_, K, atten, max_nel=1):
        return RefPfb(self, fs, nchan, os_, K, atten, max_nel)

    def cov_accumulate(self, out, nel, nchan, gate, R, counts):
        o = np.ascontiguousarray(out, np.complex64)
        g = None if gate is None else np.ascontiguousarray(gate, np.uint8).ctypes.data_as(C.POINTER(C.c_uint8))
        self.L.atkdsp_pfb_cov_accumulate(self._cf(o), nel, nchan, o.shape[2], g, self._cf(R),
                                         counts.ctypes.data_as(C.POINTER(C.c_uint32)))

    def channel_power(self, out, nel, nchan):
        o = np.ascontiguousarray(out, np.complex64)
        p = np.zeros(nel * nchan, np.float32)
        mp = np.zeros(nchan, np.float32)
        self.L.atkdsp_pfb_channel_power(self._cf(o), nel, nchan, o.shape[2], self._f(p), self._f(mp))
        return p.reshape(nel, nchan), mp


class RefPfb:
    def __init__(self, ref: Ref, fs, nchan, os_, K, atten, max_nel):
        self.r, self.M = ref, nchan
        self.p = ref.L.atkdsp_pfb_create(fs, nchan, os_, K, atten, max_nel)

    def set_cal(self, table, nel):
        t = np.ascontiguousarray(table, np.complex64)
        self.r.L.atkdsp_pfb_set_cal(self.p, Ref._cf(t), nel)

    def process_n(self, ins, nel):
        arrs = [np.ascontiguousarray(x, np.complex64) for x in ins]
        n = arrs[0].size
        m = self.r.L.atkdsp_pfb_out_hops(self.p, n)
        out = np.zeros(nel * self.M * m, np.complex64)
        ptrs = (C.POINTER(Cf) * nel)(*[Ref._cf(a) for a in arrs])
        got = self.r.L.atkdsp_pfb_process_n(self.p, ptrs, nel, n, Ref._cf(out), out.size)
        return out[: nel * self.M * got].reshape(nel, self.M, got)

    def process(self, x):
        return self.process_n([x], 1)[0]

    def __del__(self):
        if getattr(self, "p", None):
            self.r.L.atkdsp_pfb_destroy(self.p)
            self.p = None


class RefArb:
    """The reference arbitrary resampler (src/resample.c:130-252)."""

    def __init__(self, ref: Ref, in_rate, out_rate, atten=60.0, nphase=64):
        self.L = ref.L
        self.h = self.L.atkdsp_arb_resampler_create(float(in_rate), float(out_rate), float(atten), int(nphase))
        assert self.h

    def set_ratio(self, ratio):
        self.L.atkdsp_arb_resampler_set_ratio(self.h, float(ratio))

    def process(self, x):
        x = np.ascontiguousarray(x, np.complex64)
        out = np.empty(self.L.atkdsp_arb_resampler_out_max(self.h, x.size), np.complex64)
        got = self.L.atkdsp_arb_resampler_process(self.h, Ref._cf(x), x.size, Ref._cf(out), out.size)
        assert got >= 0, got
        return out[:got]

    def __del__(self):
        if getattr(self, "h", None):
            self.L.atkdsp_arb_resampler_destroy(self.h)
            self.h = None


class RefOsx:
    """The reference overlap-save extractor (src/channelizer.c:473-788)."""

    def __init__(self, ref: Ref, fs, nfft=4096, atten=80.0, max_nel=1):
        self.L = ref.L
        self.h = self.L.atkdsp_osx_create(float(fs), int(nfft), float(atten), int(max_nel))
        assert self.h

    def add(self, off, bw, rate):
        ch = self.L.atkdsp_osx_add(self.h, float(off), float(bw), float(rate))
        assert ch >= 0, ch
        return ch

    def retune(self, ch, off):
        assert self.L.atkdsp_osx_retune(self.h, int(ch), float(off)) == 0

    def remove(self, ch):
        assert self.L.atkdsp_osx_remove(self.h, int(ch)) == 0

    def process_n(self, ins, nel):
        arrs = [np.ascontiguousarray(np.asarray(ins[e]).ravel(), np.complex64) for e in range(nel)]
        n = arrs[0].size
        ns = self.L.atkdsp_osx_channels(self.h)
        active = [self.L.atkdsp_osx_chan_rate(self.h, ci) > 0 for ci in range(ns)]
        cap = max([self.L.atkdsp_osx_out_max(self.h, ci, n) for ci in range(ns) if active[ci]], default=0)
        bufs = [[np.zeros(cap, np.complex64) if active[ci] else None for ci in range(ns)] for _ in range(nel)]
        ptrs = (C.POINTER(Cf) * max(1, nel * ns))()
        for e in range(nel):
            for ci in range(ns):
                ptrs[e * ns + ci] = Ref._cf(bufs[e][ci]) if active[ci] else None
        counts = (C.c_size_t * max(1, ns))()
        inarr = (C.POINTER(Cf) * nel)(*[Ref._cf(a) for a in arrs])
        got = self.L.atkdsp_osx_process_n(self.h, inarr, nel, n, ptrs, counts, cap)
        assert got >= 0, got
        outs = [None] * ns
        for ci in range(ns):
            if active[ci]:
                outs[ci] = np.array([bufs[e][ci][:counts[ci]] for e in range(nel)], np.complex64)
        return outs, [int(counts[ci]) for ci in range(ns)]

    def process(self, x):
        outs, cnts = self.process_n([x], 1)
        return [o[0] if o is not None else None for o in outs], cnts

    def __del__(self):
        if getattr(self, "h", None):
            self.L.atkdsp_osx_destroy(self.h)
            self.h = None


def design_lowpass(ref: Ref, fp, fs_, atten, rate):
    n = ref.L.atkdsp_design_lowpass(float(fp), float(fs_), float(atten), float(rate), None, 0)
    out = np.zeros(n, np.float32)
    assert ref.L.atkdsp_design_lowpass(float(fp), float(fs_), float(atten), float(rate), Ref._f(out), n) == n
    return out
