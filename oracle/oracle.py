"""The oracle: a CPU restatement of the reference's wideband front end.

TEST INFRASTRUCTURE ONLY. Only tests/, __graft_entry__.smoke() and bench.py's
CPU-baseline leg may import this module, and only as the checker. The product
(atkdsp_amd) never imports it.

Each function restates the reference C (file:line given) in numpy:
  - byte/integer work and the display helpers are restated exactly in float32
    (bit-identical to the reference: unpack, decimate_max, db_to_pixels,
    stitch_max, median, detect_channels);
  - the filterbank, FFT and W1/W2 are restated in FP64, the "FP64 twin" the
    reference's own tests measure against (tests/test_channelizer.py:397-398);
  - the spectrum line restates the reference's mixed arithmetic: window and
    FFT in FP64, power and detector in float32 (src/fft.c:91-108, :177-191),
    stats planes in FP64 (src/fft.c:317-331);
  - the CFAR frame test in FP64 and its burst tracker (src/cfar.c:336-365,
    :523-726), with alpha, CA's alpha and E[T] given (the threshold design is
    pinned on its own against closed forms and the reference's values).

How each function is pinned (tests/test_oracle.py): against golden vectors
written by the compiled reference (tests/golden/, made by
tests/golden/make_golden.py from oracle/_ref), against the reference's own
known-answer tests (tests/test_smoke.c lines cited there), and against closed
forms. No function here is unpinned.
"""
from __future__ import annotations

import math

import numpy as np

F32 = np.float32
BPS = {0: 2, 1: 2, 2: 4, 3: 4, 4: 8}
CU8, CI8, CI16, CI16Q11, CF32 = 0, 1, 2, 3, 4


# ---- unpack: src/convert.c:33-111 ------------------------------------------------
def unpack_consts(fmt: int, off: complex):
    """The folded per-format constants, formed in float32 as src/convert.c:46-68."""
    ore, oim = F32(off.real), F32(off.imag)
    if fmt == CU8:
        k = F32(1.0) / F32(127.5)
        return F32(127.5) + ore / k, F32(127.5) + oim / k, k
    if fmt == CI8:
        k = F32(1.0) / F32(128.0)
    elif fmt == CI16:
        k = F32(1.0) / F32(32768.0)
    elif fmt == CI16Q11:
        k = F32(1.0) / F32(2048.0)
    else:
        return F32(0), F32(0), F32(1)
    return ore / k, oim / k, k


def unpack_dc(raw: bytes | np.ndarray, fmt: int, off: complex = 0j, out_cap: int | None = None):
    """Returns (out complex64, mean complex) exactly as atkdsp_unpack_dc."""
    b = np.frombuffer(bytes(raw), np.uint8) if not isinstance(raw, np.ndarray) else raw.view(np.uint8).ravel()
    n = b.size // BPS[fmt]
    if out_cap is not None:
        n = min(n, out_cap)
    if fmt == CU8:
        v = b[: 2 * n].astype(F32)
    elif fmt == CI8:
        v = b[: 2 * n].view(np.int8).astype(F32)
    elif fmt in (CI16, CI16Q11):
        v = b[: 4 * n].view("<i2").astype(F32)
    elif fmt == CF32:
        v = b[: 8 * n].view("<f4").copy()
    else:
        raise ValueError("format")
    if fmt == CF32:
        re, im = v[0::2], v[1::2]
        if not (off.real == 0 and off.imag == 0):
            re = re - F32(off.real)
            im = im - F32(off.imag)
    else:
        cre, cim, k = unpack_consts(fmt, off)
        re = (v[0::2] - cre) * k
        im = (v[1::2] - cim) * k
    out = np.empty(n, np.complex64)
    out.real = re
    out.imag = im
    return out, mean_of(out, off)


def mean_of(out: np.ndarray, off: complex = 0j) -> complex:
    """Eight FP64 lane sums, tail into lane 0, lanes added in order, then
    sum/n + off (src/convert.c:91-108)."""
    n = out.size
    if n == 0:
        return complex(float(F32(off.real)), float(F32(off.imag)))
    m = n & ~7
    parts = []
    for comp in (out.real.astype(np.float64), out.imag.astype(np.float64)):
        lanes = np.zeros(8)
        if m:
            lanes = np.cumsum(comp[:m].reshape(-1, 8), axis=0)[-1].copy()
        for x in comp[m:]:
            lanes[0] += x
        s = 0.0
        for l in range(8):
            s += lanes[l]
        parts.append(s)
    inv = 1.0 / n
    return complex(parts[0] * inv + float(F32(off.real)), parts[1] * inv + float(F32(off.imag)))


def unpack(raw, fmt: int, dc_alpha: float = 0.0, state: complex = 0j):
    """atkdsp_unpack with the one-pole DC block restated in FP64
    (src/convert.c:7-17, :113-120). Returns (out complex128, new state)."""
    x, _ = unpack_dc(raw, fmt, 0j)
    if not dc_alpha > 0:
        return x.astype(np.complex128), state
    from scipy.signal import lfilter
    a = float(F32(dc_alpha))
    xd = x.astype(np.complex128)
    # s_i = (1-a) s_{i-1} + a x_i, y_i = x_i - s_i
    s, _ = lfilter([a], [1.0, -(1.0 - a)], xd, zi=np.array([(1.0 - a) * state]))
    return xd - s, complex(s[-1]) if s.size else state


# ---- windows: src/fft.c:133-152, :433-444 ---------------------------------------------
def window(kind: int, n: int) -> np.ndarray:
    if n == 1:
        return np.ones(1, F32)
    x = 2.0 * math.pi * np.arange(n) / (n - 1)
    if kind == 0:
        w = np.ones(n)
    elif kind == 1:
        w = 0.5 - 0.5 * np.cos(x)
    elif kind == 2:
        w = 0.54 - 0.46 * np.cos(x)
    elif kind == 3:
        w = 0.35875 - 0.48829 * np.cos(x) + 0.14128 * np.cos(2 * x) - 0.01168 * np.cos(3 * x)
    else:
        raise ValueError("kind")
    return w.astype(F32)


def window_stats(w: np.ndarray):
    w = w.astype(np.float64)
    s, s2 = w.sum(), (w * w).sum()
    return s / w.size, (w.size * s2 / (s * s) if s > 0 else 0.0)


# ---- FFT: FP64 (pocketfft semantics, src/fft.c:91-131) --------------------------------
def fft(x: np.ndarray, inverse: bool = False) -> np.ndarray:
    xd = np.asarray(x, np.complex128)
    return np.fft.ifft(xd) if inverse else np.fft.fft(xd)


def power_db(x: np.ndarray, win: np.ndarray | None) -> np.ndarray:
    """src/fft.c:155-174: FP64 window and FFT, float32 magnitude and dB, fftshifted."""
    n = x.size
    xd = np.asarray(x, np.complex64).astype(np.complex128)
    if win is not None:
        xd = xd * win.astype(np.float64)
    X = np.fft.fft(xd)
    k = (np.arange(n) + n // 2) % n
    re, im = X.real[k].astype(F32), X.imag[k].astype(F32)
    mag = np.sqrt(re * re + im * im) * (F32(1.0) / F32(n))
    return (F32(20.0) * np.log10(mag + F32(1e-12))).astype(F32)


def _frames(x: np.ndarray, n: int, hop: int) -> int:
    if hop == 0 or hop > n:
        hop = n
    if x.size < n:
        return 0, hop
    return (x.size - n) // hop + 1, hop


def _frame_ffts(x, n, hop, win, frames, chunk=256):
    xd = np.asarray(x, np.complex64)
    wd = None if win is None else win.astype(np.float64)
    for f0 in range(0, frames, chunk):
        f1 = min(frames, f0 + chunk)
        idx = (np.arange(f0, f1)[:, None] * hop) + np.arange(n)[None, :]
        blk = xd[idx].astype(np.complex128)
        if wd is not None:
            blk = blk * wd
        yield np.fft.fft(blk, axis=1)


def spectrum_reduce(x: np.ndarray, n: int, hop: int | None, win, detector: int):
    """src/fft.c:177-295: returns (line float32, frames)."""
    frames, hop = _frames(x, n, hop or n)
    line = np.zeros(n, F32)
    if frames == 0:
        return line, 0
    k = (np.arange(n) + n // 2) % n
    inv2 = F32(1.0) / (F32(n) * F32(n))
    acc = None
    for X in _frame_ffts(x, n, hop, win, frames):
        re = X.real[:, k].astype(F32)
        im = X.imag[:, k].astype(F32)
        pw = (re * re + im * im) * inv2
        for row in pw:
            if acc is None:
                acc = row.copy()
            elif detector == 0:
                acc = np.where(row > acc, row, acc)
            elif detector == 1:
                acc = np.where(row < acc, row, acc)
            else:
                acc = acc + row
    scale = F32(1.0) / F32(frames) if detector == 2 else F32(1.0)
    mag = np.sqrt(acc * scale)
    return (F32(20.0) * np.log10(mag + F32(1e-12))).astype(F32), frames


def spectrum_stats(x: np.ndarray, n: int, hop: int | None, win):
    """src/fft.c:297-431: FP64 planes; returns dict max/avg/min/sk and frames."""
    frames, hop = _frames(x, n, hop or n)
    if frames == 0:
        return {"frames": 0}
    k = (np.arange(n) + n // 2) % n
    inv2 = 1.0 / (float(n) * float(n))
    mx = mn = su = s2 = None
    for X in _frame_ffts(x, n, hop, win, frames):
        pw = (X.real[:, k] ** 2 + X.imag[:, k] ** 2) * inv2
        bmx, bmn, bsu, bs2 = pw.max(0), pw.min(0), np.cumsum(pw, 0)[-1], np.cumsum(pw * pw, 0)[-1]
        if mx is None:
            mx, mn, su, s2 = bmx, bmn, bsu, bs2
        else:
            mx, mn, su, s2 = np.maximum(mx, bmx), np.minimum(mn, bmn), su + bsu, s2 + bs2
    M = float(frames)
    db = lambda v: (F32(10.0) * np.log10(v.astype(F32) + F32(1e-12))).astype(F32)
    den = su * su
    sk = np.where((frames < 2) | (den <= 0), 1.0, M * s2 / np.where(den > 0, den, 1.0)).astype(F32)
    return {"max": db(mx), "min": db(mn), "avg": db(su / M), "sk": sk, "frames": frames}


# ---- display: src/display.c:4-37, stitch: src/detect.c:116-134 -----------------------------
def db_to_pixels(db: np.ndarray, lo: float, hi: float, lut: np.ndarray) -> np.ndarray:
    db = np.asarray(db, F32)
    lo, hi = F32(lo), F32(hi)
    span = F32(hi - lo)
    k = F32(255.0) / span if span > 0 else F32(0.0)
    with np.errstate(invalid="ignore"):
        t = (db - lo) * k
        t = np.where(t < 0, F32(0), t)
        t = np.where(t > 255, F32(255), t)
        idx = np.where(np.isnan(db), 0, np.floor(np.where(np.isnan(t), 0, t) + F32(0.5))).astype(np.int64)
    return np.asarray(lut, np.uint32)[idx]


def decimate_max(x: np.ndarray, m: int) -> np.ndarray:
    x = np.asarray(x, F32)
    n = x.size
    out = np.empty(m, F32)
    if m >= n:
        out[:n] = x
        out[n:] = x[n - 1]
        return out
    per = n // m
    for c in range(m):
        a = c * per
        b = n if c + 1 == m else a + per
        best = x[a]
        for v in x[a + 1:b]:
            if v > best:
                best = v
        out[c] = best
    return out


def stitch_max(seg, seg_lo, seg_step, keep_lo, keep_hi, out_lo, out_step, out: np.ndarray) -> np.ndarray:
    seg = np.asarray(seg, F32)
    nout = out.size
    if nout == 0 or seg_step <= 0 or out_step <= 0:
        return out
    out_hi = out_lo + out_step * float(nout)
    for i in range(seg.size):
        f = seg_lo + (float(i) + 0.5) * seg_step
        if f < keep_lo or f > keep_hi or f < out_lo or f > out_hi:
            continue
        k = int((f - out_lo) / out_step)       # truncation toward zero, as the C cast
        k = min(max(k, 0), nout - 1)
        v, cur = seg[i], out[k]
        if cur != cur or v > cur:
            out[k] = v
    return out


# ---- detection on a spectrum line: src/detect.c:4-114 ---------------------------------------
def median(x) -> float:
    """src/detect.c:25-34: the middle order statistic, or for even n the
    float32 0.5 * (lower middle + upper middle); -120 for an empty input."""
    x = np.asarray(x, F32)
    n = x.size
    if n == 0:
        return -120.0
    s = np.sort(x, kind="stable")
    if n & 1:
        return float(s[n // 2])
    return float(F32(0.5) * (s[n // 2 - 1] + s[n // 2]))


def smooth5(x) -> np.ndarray:
    """src/detect.c:38-50: the 5-bin moving average with edge padding, summed
    from 0.0f in index order, then times 0.2f, all in float32."""
    x = np.asarray(x, F32)
    n = x.size
    if n < 5:
        return x.copy()
    idx = np.arange(n)
    s = np.zeros(n, F32)
    for d in (-2, -1, 0, 1, 2):
        s = s + x[np.clip(idx + d, 0, n - 1)]
    return s * F32(0.2)


def detect_channels(line, floor_db: float, threshold_db: float, gap_bins: int, min_run: int):
    """src/detect.c:58-114: runs of the smoothed line above floor + threshold
    (float32 sum), interior gaps shorter than gap_bins bridged, runs of at
    least min_run kept; each run's peak (scan from its first bin), the FP64
    power-weighted centroid summed in bin order with the C library's pow,
    snr = peak - floor. Records strongest first by a stable sort (the
    reference's qsort is glibc's merge sort here). Returns a list of
    (start, end, centroid, peak, snr) with float32 fields."""
    line = np.asarray(line, F32)
    n = line.size
    if n == 0:
        return []
    fl = F32(floor_db)
    thr = fl + F32(threshold_db)
    gap = max(2, int(gap_bins))
    mr = max(2, int(min_run))
    m = (smooth5(line) > thr).copy()
    i = 0
    while i < n:
        if not m[i]:
            j = i
            while j < n and not m[j]:
                j += 1
            if i > 0 and j < n and j - i < gap:
                m[i:j] = True
            i = j
        else:
            i += 1
    out = []
    i = 0
    while i < n:
        if not m[i]:
            i += 1
            continue
        j = i
        while j < n and m[j]:
            j += 1
        if j - i >= mr:
            wsum = csum = 0.0
            peak = line[i]
            for k in range(i, j):
                w = math.pow(10.0, float(line[k]) / 10.0)
                wsum += w
                csum += float(k - i) * w
                if line[k] > peak:
                    peak = line[k]
            out.append((i, j, float(F32(float(i) + csum / (wsum + 1e-12))), float(peak), float(F32(peak - fl))))
        i = j
    out.sort(key=lambda c: -c[4])
    return out


# ---- polyphase filterbank: src/channelizer.c:33-43, :79-99, :218-402 --------------------------
def kaiser_beta(a: float) -> float:
    """src/design.c:18-22."""
    if a > 50.0:
        return 0.1102 * (a - 8.7)
    if a >= 21.0:
        return 0.5842 * (a - 21.0) ** 0.4 + 0.07886 * (a - 21.0)
    return 0.0


def _i0_series(x: np.ndarray) -> np.ndarray:
    """I0 by its power series with the reference's per-element stopping rule."""
    x = np.asarray(x, np.float64)
    s = np.ones_like(x)
    term = np.ones_like(x)
    y = x * 0.5
    live = np.ones(x.shape, bool)
    for k in range(1, 500):
        term = np.where(live, term * (y / k), term)
        t2 = term * term
        s = np.where(live, s + t2, s)
        live &= ~(t2 < s * 1e-17)
        if not live.any():
            break
    return s


def pfb_prototype(nchan: int, taps_per_poly: int, atten_db: float) -> np.ndarray:
    """Kaiser-windowed sinc, fc = 1/(2 nchan), unit DC gain times nchan (FP64)."""
    atten_db = max(atten_db, 20.0)
    L = nchan * taps_per_poly
    fc = 1.0 / (2.0 * nchan)
    Mid = (L - 1) * 0.5
    beta = kaiser_beta(atten_db)
    t = np.arange(L) - Mid
    x = 2.0 * fc * t
    with np.errstate(invalid="ignore", divide="ignore"):
        sinc = np.where(t == 0.0, 1.0, np.sin(math.pi * x) / (math.pi * x))
    r = t / Mid if Mid > 0 else np.zeros(L)
    w = _i0_series(beta * np.sqrt(np.where(r * r < 1.0, 1.0 - r * r, 0.0))) / _i0_series(np.array(beta))
    h = 2.0 * fc * sinc * w
    s = h.sum()
    if s == 0.0:
        s = 1.0
    return h / s * nchan


class Pfb:
    """FP64 streaming filterbank with the reference's state machine:
    s = hist(ov) ++ leftover(lc) ++ in; hop i's window s[i hop, i hop + L);
    fold over k ascending; forward FFT; rotation by exp(-2 pi i k shift/M),
    shift = ((b + i + 1) hop) mod M; cal; output [el][chan][hop]."""

    def __init__(self, fs, nchan, oversample=2, taps_per_poly=12, atten_db=80.0, max_nel=1):
        self.M, self.os, self.K = int(nchan), int(oversample), int(taps_per_poly)
        self.hop = self.M // self.os
        self.L = self.M * self.K
        self.ov = self.L - self.hop
        self.h = pfb_prototype(self.M, self.K, float(atten_db))
        self.max_nel = int(max_nel)
        self.hist = np.zeros((self.max_nel, self.ov), np.complex128)
        self.left = np.zeros((self.max_nel, 0), np.complex128)
        self.b = 0
        self.cal = None
        self.fs = fs

    def set_cal(self, table, nel):
        self.cal = None if table is None else np.asarray(table, np.complex128).reshape(nel, self.M)

    def out_rate(self):
        return self.os * self.fs / self.M

    def process_n(self, ins, nel):
        M, hop, L, ov, K = self.M, self.hop, self.L, self.ov, self.K
        xs = [np.asarray(ins[e], np.complex64).astype(np.complex128).ravel() for e in range(nel)]
        n = xs[0].size
        lc = self.left.shape[1]
        m = (lc + n) // hop
        outs = []
        hk = self.h.reshape(K, M)
        newhist, newleft = [], []
        for e in range(nel):
            s = np.concatenate([self.hist[e], self.left[e], xs[e]])
            if m:
                st = s.strides[0]
                W = np.lib.stride_tricks.as_strided(s, shape=(m, K, M), strides=(hop * st, M * st, st))
                fold = np.einsum("ikq,kq->iq", W, hk)
                X = np.fft.fft(fold, axis=1)
                shift = (((self.b + np.arange(m) + 1) * hop) % M)[:, None]
                idx = (np.arange(M)[None, :] * shift) % M
                X = X * np.exp(-2j * np.pi * idx / M)
                if self.cal is not None and e < self.cal.shape[0]:
                    X = X * self.cal[e][None, :]
                outs.append(X.T)
            else:
                outs.append(np.zeros((M, 0), np.complex128))
            newhist.append(s[m * hop: m * hop + ov])
            newleft.append(s[m * hop + ov:])
        for e in range(nel):
            self.hist[e] = newhist[e]
        self.left = np.zeros((self.max_nel, len(newleft[0])), np.complex128)
        for e in range(nel):
            self.left[e] = newleft[e]
        self.b += m
        return np.stack(outs)

    def process(self, x):
        return self.process_n([x], 1)[0]


def pfb_cov_accumulate(out, nel, nchan, gate, R, counts):
    """W1 in FP64 (src/channelizer.c:420-440): R[k] += y y^H over hops."""
    o = np.asarray(out, np.complex128)
    m = o.shape[2]
    for k in range(nchan):
        if gate is not None and not gate[k]:
            continue
        y = o[:, k, :]
        R[k] += y @ y.conj().T
        counts[k] += m


def pfb_channel_power(out, nel, nchan):
    """W2 in FP64 (src/channelizer.c:453-471)."""
    o = np.asarray(out, np.complex128)
    m = o.shape[2]
    p = (np.abs(o) ** 2).sum(axis=2) / m if m else np.zeros((nel, nchan))
    return p, p.mean(axis=0)


# ---- NCO: src/nco.c:27-98, src/internal.h:36-41 (twin python/atkdsp/reference.py:95-108) ----
def wrap(ph: float) -> float:
    """atk_wrap: fmod by 2 pi, then into [-pi, pi)."""
    ph = math.fmod(ph, 2.0 * math.pi)
    if ph >= math.pi:
        ph -= 2.0 * math.pi
    if ph < -math.pi:
        ph += 2.0 * math.pi
    return ph


class Nco:
    """out[n] = in[n] * exp(j (phase + n step)), step = -2 pi f / fs, the
    phase carried wrapped in FP64 (the reference's float multiply is done
    here in FP64, one rounding to complex64 at the end)."""

    def __init__(self, freq_hz: float, fs: float):
        self.phase = 0.0
        self.step = -2.0 * math.pi * freq_hz / fs if fs > 0 else 0.0

    def mix(self, x: np.ndarray) -> np.ndarray:
        x = np.asarray(x, np.complex64)
        if self.step == 0.0:
            return x.copy()
        n = np.arange(x.size, dtype=np.float64)
        y = (x.astype(np.complex128) * np.exp(1j * (self.phase + self.step * n))).astype(np.complex64)
        self.phase = wrap(self.phase + self.step * float(x.size))
        return y


def nco_exact_phase(n: int, p: int, q: int) -> float:
    """The exact NCO phase after n samples at f = fs * p / q (step -2 pi p/q):
    -2 pi ((n p) mod q) / q, wrapped: integer arithmetic, no accumulation."""
    return wrap(-2.0 * math.pi * ((n * p) % q) / q)


# ---- Kaiser lowpass: src/design.c:24-62 ------------------------------------------------
def kaiser_length(atten_db: float, transition_hz: float, rate: float) -> int:
    """src/design.c:24-32."""
    if transition_hz <= 0.0 or rate <= 0.0:
        return 0
    dw = 2.0 * math.pi * transition_hz / rate
    n = (atten_db - 8.0 if atten_db > 8.0 else 1.0) / (2.285 * dw)
    N = int(math.ceil(n)) + 1
    N = max(N, 3)
    return N + 1 if N % 2 == 0 else N


def design_lowpass(fp_hz: float, fs_hz: float, atten_db: float, rate: float) -> np.ndarray:
    """src/design.c:34-62: FP64 Kaiser-windowed sinc, normalised by the
    sequential FP64 sum of the taps, each tap cast once to float32."""
    if rate <= 0.0 or fp_hz < 0.0 or fs_hz <= fp_hz or fs_hz > rate * 0.5 + 1e-9:
        raise ValueError("design_lowpass: bad edges")
    N = kaiser_length(atten_db, fs_hz - fp_hz, rate)
    beta = kaiser_beta(atten_db)
    i0b = _i0_series(np.array([beta]))[0]
    fc = 0.5 * (fp_hz + fs_hz) / rate
    M = (N - 1) * 0.5
    t = np.arange(N, dtype=np.float64) - M
    x = 2.0 * fc * t
    with np.errstate(invalid="ignore", divide="ignore"):
        sinc = np.where(t == 0.0, 1.0, np.sin(math.pi * x) / (math.pi * x))
    r = t / M
    rr = r * r
    w = _i0_series(beta * np.sqrt(np.where(rr < 1.0, 1.0 - rr, 0.0))) / i0b
    h = 2.0 * fc * sinc * w
    total = float(np.cumsum(h)[-1])               # the reference's sequential sum
    if total == 0.0:
        total = 1.0
    return (h / total).astype(F32)


# ---- arbitrary-ratio resampler: src/resample.c:100-252 -------------------------------
class ArbResampler:
    """The FP64 twin of the reference resampler: the same prototype (gain P,
    forward differences), positions base + K*step - in_off carried across
    calls, the tap c = p + mu*d formed in float32 as the reference does, and
    the 2Q-tap FIR accumulated in FP64."""

    def __init__(self, in_rate: float, out_rate: float, atten_db: float = 60.0, nphase: int = 64):
        if atten_db <= 0.0:
            atten_db = 60.0
        P = nphase or 64
        nyq = 0.5 * min(in_rate, out_rate)
        taps = design_lowpass(0.8 * nyq, nyq, atten_db, in_rate * P)
        Q = (taps.size + P - 1) // P
        p = np.zeros(P * Q, F32)
        p[: taps.size] = taps
        p = (p * F32(P)).astype(F32)
        d = np.empty_like(p)
        d[:-1] = p[1:] - p[:-1]
        d[-1] = -p[-1]
        self.P, self.Q, self.p, self.d = P, Q, p, d
        self.step = in_rate / out_rate
        self.base, self.K, self.in_off = 0.0, 0, 0
        self.hist = np.zeros(Q, np.complex128)

    def set_ratio(self, ratio: float) -> None:
        """src/resample.c:185-192: re-anchor so the next output stays put."""
        self.base = self.base + float(self.K) * self.step
        self.K = 0
        self.step = 1.0 / ratio

    def _local(self, K: int) -> float:
        return self.base + float(K) * self.step - float(self.in_off)

    def process(self, x) -> np.ndarray:
        x = np.asarray(x, np.complex64).astype(np.complex128)
        n, P, Q = x.size, self.P, self.Q
        ext = np.concatenate([self.hist, x])       # input m at ext[Q + m]
        q = np.arange(Q)
        out = []
        K = self.K
        local = self._local(K)
        while local < n:
            i = math.floor(local)
            frac = local - i
            ph = frac * P
            b = min(int(ph), P - 1)
            mu = F32(ph - b)
            idx = b + q * P
            c = self.p[idx] + F32(mu) * self.d[idx]             # float32, as src/resample.c:237
            m = i - q
            ok = Q + m >= 0
            out.append(np.sum(c[ok].astype(np.float64) * ext[Q + m[ok]]))
            K += 1
            local = self._local(K)
        self.K = K
        self.in_off += n
        self.hist = ext[-Q:].copy()
        return np.array(out, np.complex64)


# ---- overlap-save extractor: src/channelizer.c:473-788 -------------------------------------
class Osx:
    """The FP64 twin of the reference extractor: the block FFT, the fold with
    the channel's float32 frequency response, the M-point inverse FFT and the
    absolute-time mix in FP64, rounded to float32 where the reference stores
    (the pre-resampler samples), then the resampler twin."""

    def __init__(self, fs: float, nfft: int = 4096, atten_db: float = 80.0, max_nel: int = 1):
        self.fs, self.nfft, self.atten = float(fs), int(nfft), max(float(atten_db), 20.0)
        self.ov = self.nfft // 2
        self.hop = self.nfft - self.ov
        self.max_nel = max_nel
        self.hist = np.zeros((max_nel, self.ov), np.complex64)
        self.left = np.zeros((max_nel, 0), np.complex64)
        self.blk = 0
        self.chans = []

    def _decim(self, bw, out_rate):
        target = out_rate if out_rate > 1.25 * bw else 1.25 * bw
        decim = 1
        while True:
            nd = decim * 2
            if self.fs / nd < target or self.ov % nd or self.nfft % nd:
                return decim
            decim = nd

    def _tune(self, c, off):
        s = math.floor(off / self.fs * self.nfft + 0.5)
        c["shift"] = s % self.nfft
        c["delta"] = off - s * self.fs / self.nfft
        c["offset"] = off

    def add(self, off, bw, out_rate) -> int:
        decim = self._decim(bw, out_rate)
        chan_rate = self.fs / decim
        fp, fstop = 0.5 * bw, min(chan_rate - 0.5 * bw, 0.5 * self.fs)
        taps = design_lowpass(fp, fstop, self.atten, self.fs)
        if taps.size > self.ov + 1:
            taps = taps[: self.ov + 1]
            tot = float(np.cumsum(taps.astype(np.float64))[-1]) or 1.0
            taps = (taps.astype(np.float64) / tot).astype(F32)
        hin = np.zeros(self.nfft, np.complex128)
        hin[: taps.size] = taps
        M = self.nfft // decim
        c = dict(decim=decim, M=M, ovd=self.ov // decim, chan_rate=chan_rate, out_rate=out_rate,
                 H=np.fft.fft(hin).astype(np.complex64).astype(np.complex128),
                 rs=[ArbResampler(chan_rate, out_rate, self.atten, 64) for _ in range(self.max_nel)]
                 if abs(chan_rate - out_rate) > 1e-6 * out_rate else None)
        self._tune(c, off)
        for i, old in enumerate(self.chans):
            if old is None:
                self.chans[i] = c
                return i
        self.chans.append(c)
        return len(self.chans) - 1

    def retune(self, ch, off):
        self._tune(self.chans[ch], off)

    def remove(self, ch):
        self.chans[ch] = None

    def process_n(self, ins, nel):
        nfft, ov, hop = self.nfft, self.ov, self.hop
        xs = [np.concatenate([self.hist[e], self.left[e], np.asarray(ins[e], np.complex64)]) for e in range(nel)]
        avail = self.left.shape[1] + np.asarray(ins[0]).size
        m = avail // hop
        outs = [[] if c is not None else None for c in self.chans]
        for i in range(m):
            bi = self.blk + i
            X = [np.fft.fft(xs[e][i * hop: i * hop + nfft].astype(np.complex128)) for e in range(nel)]
            for ci, c in enumerate(self.chans):
                if c is None:
                    continue
                M, decim, sh = c["M"], c["decim"], c["shift"]
                idx = np.arange(M)[:, None] + np.arange(decim)[None, :] * M
                start = bi * hop - ov
                a0 = ((sh * start) % nfft) / nfft
                nvalid = M - c["ovd"]
                nabs = bi * hop + np.arange(nvalid) * decim
                ph = -2.0 * math.pi * (a0 + c["delta"] * nabs.astype(np.float64) / self.fs)
                rot = np.exp(1j * ph)
                row = []
                for e in range(nel):
                    Yf = np.sum(X[e][(idx + sh) % nfft] * c["H"][idx], axis=1)
                    ys = np.fft.ifft(Yf)
                    v = (ys[c["ovd"]:] / decim * rot).astype(np.complex64)
                    row.append(c["rs"][e].process(v) if c["rs"] is not None else v)
                outs[ci].append(row)
        res = []
        for ci, c in enumerate(self.chans):
            if c is None:
                res.append(None)
            elif outs[ci]:
                res.append(np.array([np.concatenate([blk[e] for blk in outs[ci]]) for e in range(nel)],
                                    np.complex64))
            else:
                res.append(np.zeros((nel, 0), np.complex64))
        # next history and leftover
        for e in range(nel):
            tail = xs[e][m * hop:]
            self.hist[e] = tail[:ov]
        self.left = np.array([xs[e][m * hop + ov:] for e in range(nel)], np.complex64).reshape(nel, -1)
        self.blk += m
        return res, [0 if r is None else r.shape[1] for r in res]

    def process(self, x):
        outs, cnts = self.process_n([x], 1)
        return [o[0] if o is not None else None for o in outs], cnts


# ---- CFAR detection and the burst tracker: src/cfar.c:270-819 -------------------
CFAR_EVENT = np.dtype([("t0", "<f8"), ("t1", "<f8"), ("fc_bin", "<f8"), ("type", "<i4"), ("id", "<u4"),
                       ("flags", "<u4"), ("lo_bin", "<i4"), ("hi_bin", "<i4"), ("n_hits", "<i4"),
                       ("level_db", "<f4"), ("noise_db", "<f4"), ("snr_db", "<f4"), ("band_snr_db", "<f4")])
CFAR_DEFAULTS = dict(mode=0, n_ref=16, n_guard=4, k_frames=1, pfa=1e-6, looks=1.0, os_k=0, open_m=2, open_n=3,
                     close_k=3, merge_gap=1, min_bins=1, max_tracks=256, rho2=(0.0, 0.0, 0.0, 0.0))


def cfar_tables(nbins: int, n_ref: int, n_guard: int, mask):
    """Each bin's reference cells (src/cfar.c:336-365): up to n_ref unmasked
    bins beyond the guard on the left (nearest first), then on the right, then
    a short side's shortfall taken beyond the other side's cells (left's
    first). Returns (cells (nbins x 2 n_ref), ok, sided)."""
    N = 2 * n_ref
    m = np.zeros(nbins, bool) if mask is None else np.asarray(mask) != 0
    cells = np.zeros((nbins, N), np.int64)
    ok = np.zeros(nbins, bool)
    sided = np.zeros(nbins, bool)
    free = np.flatnonzero(~m)
    for i in range(nbins):
        if m[i]:
            continue
        lo_side = free[free < i - n_guard][::-1]          # nearest first
        hi_side = free[free > i + n_guard]
        left, right = lo_side[:n_ref], hi_side[:n_ref]
        more = list(hi_side[n_ref:][:n_ref - len(left)]) + list(lo_side[n_ref:][:n_ref - len(right)])
        if len(left) + len(right) + len(more) < N:
            continue
        cells[i] = np.concatenate([left, right, np.array(more, np.int64)])
        ok[i] = True
        sided[i] = not more
    return cells, ok, sided


def cfar_test(X, cells, ok, sided, mode, n_ref, os_k, alpha, alpha_ca, et):
    """The test of one averaged frame X (src/cfar.c:547-590): the statistic
    T over each bin's cells (CA the mean, summed in table order; GO / SO the
    larger / smaller side mean, CA with CA's alpha and unit scale where a side
    borrowed; OS the os_k-th smallest, NaN if any cell is NaN); thr = a T,
    noise = T / E[T], hit X > thr. Returns (thr, noise, hits)."""
    N = 2 * n_ref
    V = X[cells]                                          # nbins x N
    a = np.full(X.size, alpha)
    scale = np.full(X.size, et)
    if mode == 1:
        nan = np.isnan(V).any(axis=1)
        Z = np.partition(np.where(np.isnan(V), 0.0, V), os_k - 1, axis=1)[:, os_k - 1]
        Z[nan] = np.nan
    else:
        tot = np.zeros(X.size)
        for q in range(N):
            tot = tot + V[:, q]
        Z = tot / N
        if mode in (2, 3):
            L = np.zeros(X.size)
            R = np.zeros(X.size)
            for q in range(n_ref):
                L = L + V[:, q]
                R = R + V[:, n_ref + q]
            L, R = L / n_ref, R / n_ref
            side = np.maximum(L, R) if mode == 2 else np.minimum(L, R)
            Z = np.where(sided, side, Z)
            a = np.where(sided, alpha, alpha_ca)
            scale = np.where(sided, et, 1.0)
    Z = np.where(ok, Z, np.nan)
    thr = a * Z
    with np.errstate(invalid="ignore"):
        hits = ok & (X > thr)
    return thr, Z / scale, hits


def _gamma_beta(nb: int, looks: float, rho2) -> float:
    """beta with P(Gamma(s) / s > beta) = 1e-3, s the equivalent independent
    looks over nb contiguous bins (src/cfar.c:253-268), by scipy's inverse."""
    from scipy.special import gammainccinv
    pairs = sum((nb - m) * rho2[m - 1] for m in range(1, 5) if m < nb)
    s = nb / (1.0 + 2.0 * pairs / nb) * looks
    return float(gammainccinv(s, 1e-3) / s)


class Cfar:
    """The reference's CFAR handle on the CPU (src/cfar.c:270-819). alpha,
    alpha_ca and et are given (the reference's values for the
    configuration)."""

    def __init__(self, nbins, cfg, alpha, alpha_ca, et):
        c = dict(CFAR_DEFAULTS)
        c.update(cfg)
        self.c = c
        self.nb = nbins
        self.N = 2 * c["n_ref"]
        if c["mode"] == 1 and c["os_k"] == 0:
            c["os_k"] = (3 * self.N + 3) // 4
        self.alpha, self.alpha_ca, self.et = alpha, alpha_ca, et
        self.mask = np.zeros(nbins, bool)
        self.cells, self.ok, self.sided = cfar_tables(nbins, c["n_ref"], c["n_guard"], None)
        self.k = c["k_frames"]
        self.hist = np.zeros((self.k, nbins))
        self.ht = np.zeros(self.k)
        self.seen = self.frame = 0
        self.tracks = []
        self.next_id = 1
        self.beta = {}
        self.st = dict(frames=0, warmup=0, cells_tested=0, cell_hits=0, untestable=0, opened=0, closed=0, updates=0,
                       dropped=0, resets=0, tracks=0)
        self.last = None

    def set_mask(self, mask):
        self.mask = np.zeros(self.nb, bool) if mask is None else np.asarray(mask) != 0
        self.cells, self.ok, self.sided = cfar_tables(self.nb, self.c["n_ref"], self.c["n_guard"], self.mask)

    def reset(self):
        self.tracks = []
        self.seen = self.frame = 0
        self.hist[:] = 0.0
        self.last = None
        self.st["resets"] += 1

    def _band(self, raw, noise, lo, hi):
        sel = np.flatnonzero(self.ok[lo:hi + 1]) + lo
        if not sel.size:
            return False
        sr = sz = 0.0
        for b in sel:
            sr += raw[b]
            sz += noise[b]
        if sel.size not in self.beta:
            self.beta[sel.size] = _gamma_beta(sel.size, self.c["looks"], self.c["rho2"])
        return sr > self.beta[sel.size] * sz

    def _flags(self, t):
        f = t["flags"]
        sp = slice(t["ulo"], t["uhi"] + 1)
        if (~self.sided[sp]).any() or (~self.ok[sp]).any():
            f |= 4
        m = self.c["merge_gap"] + 1
        if self.mask[max(t["ulo"] - m, 0):min(t["uhi"] + m + 1, self.nb)].any():
            f |= 8
        return f

    def _event(self, t, typ, t1, extra=0):
        e = np.zeros(1, CFAR_EVENT)
        e["type"], e["id"], e["flags"] = typ, t["id"], self._flags(t) | extra
        e["lo_bin"], e["hi_bin"], e["n_hits"] = t["ulo"], t["uhi"], t["n_hits"]
        e["t0"], e["t1"] = t["t0"], t1
        with np.errstate(invalid="ignore"):
            e["fc_bin"] = t["sbe"] / t["se"] if t["se"] > 0 else 0.5 * (t["ulo"] + t["uhi"])
        lv = np.float32(10.0 * math.log10(t["px"])) if t["px"] > 0 else np.float32(-np.inf)
        nz = np.float32(10.0 * math.log10(t["pn"])) if t["pn"] > 0 else np.float32(-np.inf)
        e["level_db"], e["noise_db"] = lv, nz
        e["snr_db"] = lv - nz if t["pn"] > 0 else np.inf
        e["band_snr_db"] = np.float32(10.0 * math.log10(t["band"])) if t["band"] > 0 else -np.inf
        return e

    def _close(self, t, extra=0):
        k = self.k
        if k == 1:
            t1 = t["t_hit_end"]
        elif t["last_raw"] >= t["last_hit"] - (k - 1):
            t1 = t["t_raw_end"]
        else:
            t1 = t["t_hit_end"] - 0.5 * (k - 1) * t["dt_hit"]
            t["flags"] |= 2
        self.st["closed"] += 1
        return self._event(t, 3, t1, extra)

    def _frame(self, row, tf, dt):
        c, nb, k = self.c, self.nb, self.k
        self.st["frames"] += 1
        F = self.frame
        slot = F % k
        v = np.maximum(row.astype(np.float64), -300.0)
        v = np.where(np.isnan(row), np.nan, v)
        self.hist[slot] = np.exp(v * 0.23025850929940458)
        self.ht[slot] = tf
        self.seen += 1
        self.frame += 1
        if self.seen < k:
            self.st["warmup"] += 1
            return []
        raw = self.hist[slot]
        tot = np.zeros(nb)
        for r in range(k):
            tot = tot + self.hist[(F + 1 + r) % k]
        X = tot / k
        thr, noise, hits = cfar_test(X, self.cells, self.ok, self.sided, c["mode"], c["n_ref"], c["os_k"],
                                     self.alpha, self.alpha_ca, self.et)
        self.st["cells_tested"] += int(self.ok.sum())
        self.st["cell_hits"] += int(hits.sum())
        self.st["untestable"] += int((~self.ok & ~self.mask).sum())
        self.last = (X, thr, hits)
        # blobs: hit bins no more than merge_gap non-hit bins apart, min_bins wide
        g = c["merge_gap"]
        idx = np.flatnonzero(hits)
        blobs = []
        if idx.size:
            cuts = np.flatnonzero(np.diff(idx) - 1 > g) + 1
            for grp in np.split(idx, cuts):
                lo, hi = int(grp[0]), int(grp[-1])
                if hi - lo + 1 < c["min_bins"]:
                    continue
                b = dict(lo=lo, hi=hi, sbe=0.0, se=0.0, sx=0.0, sn=0.0, track=-1)
                pk = -1
                for q in grp:
                    e = max(X[q] - noise[q], 0.0) if not np.isnan(X[q] - noise[q]) else X[q] - noise[q]
                    b["sbe"] += q * e
                    b["se"] += e
                    b["sx"] += X[q]
                    b["sn"] += noise[q]
                    if pk < 0 or X[q] > X[pk]:
                        pk = q
                b["px"], b["pn"] = X[pk], noise[pk]
                blobs.append(b)
        # association to the existing track overlapping most (ties: the older)
        existing = len(self.tracks)
        for t in self.tracks:
            t["has_obs"] = False
        m = g + 1
        for b in blobs:
            best, bov = -1, 0
            for ti in range(existing):
                t = self.tracks[ti]
                if b["lo"] <= t["cur_hi"] + m and b["hi"] >= t["cur_lo"] - m:
                    ov = min(b["hi"], t["cur_hi"]) - max(b["lo"], t["cur_lo"]) + 1
                    if best < 0 or ov > bov:
                        best, bov = ti, ov
            b["track"] = best
        for b in blobs:
            if b["track"] >= 0:
                continue
            if len(self.tracks) >= c["max_tracks"]:
                self.st["dropped"] += 1
                continue
            self.tracks.append(dict(id=self.next_id, state=0, hist=0, misses=0, is_new=True, n_hits=0,
                                    cur_lo=b["lo"], cur_hi=b["hi"], ulo=b["lo"], uhi=b["hi"], sbe=0.0, se=0.0,
                                    px=-1.0, pn=0.0, band=0.0, flags=0, t0=0.0, last_hit=-1, last_raw=-1,
                                    t_hit_end=0.0, dt_hit=dt, t_raw_end=0.0, has_obs=False))
            self.next_id = (self.next_id + 1) & 0xFFFFFFFF or 1
            b["track"] = len(self.tracks) - 1
        for b in blobs:
            if b["track"] < 0:
                continue
            t = self.tracks[b["track"]]
            if not t["has_obs"]:
                t.update(has_obs=True, olo=b["lo"], ohi=b["hi"], fsx=0.0, fsn=0.0)
            else:
                t["olo"], t["ohi"] = min(t["olo"], b["lo"]), max(t["ohi"], b["hi"])
            t["sbe"] += b["sbe"]
            t["se"] += b["se"]
            t["fsx"] += b["sx"]
            t["fsn"] += b["sn"]
            if b["px"] > t["px"]:
                t["px"], t["pn"] = b["px"], b["pn"]
        # each track in id order: window, state, event
        wmask = (1 << c["open_n"]) - 1
        events, keep = [], []
        for t in self.tracks:
            grew = False
            if t["has_obs"]:
                lo, hi = t["olo"], t["ohi"]
                t["hist"] = ((t["hist"] << 1) | 1) & wmask
                t["misses"] = 0
                t["n_hits"] += 1
                t["cur_lo"], t["cur_hi"] = lo, hi
                if not t["is_new"]:
                    grew = lo < t["ulo"] or hi > t["uhi"]
                t["ulo"], t["uhi"] = min(t["ulo"], lo), max(t["uhi"], hi)
                if t["fsn"] > 0 and t["fsx"] / t["fsn"] > t["band"]:
                    t["band"] = t["fsx"] / t["fsn"]
                t["last_hit"] = F
                t["t_hit_end"] = tf + dt
                t["dt_hit"] = dt
                if t["is_new"]:
                    t["is_new"] = False
                    if k == 1:
                        t["t0"] = tf
                    else:
                        for r in range(k):
                            s = (F + 1 + r) % k
                            if self._band(self.hist[s], noise, lo, hi):
                                t["t0"] = self.ht[s]
                                break
                        else:
                            t["t0"] = tf - 0.5 * (k - 1) * dt
                            t["flags"] |= 1
            else:
                t["hist"] = (t["hist"] << 1) & wmask
                t["misses"] += 1
            t["has_obs"] = False
            if k > 1 and self._band(raw, noise, t["cur_lo"], t["cur_hi"]):
                t["last_raw"] = F
                t["t_raw_end"] = tf + dt
            if t["state"] == 0:
                if bin(t["hist"]).count("1") >= c["open_m"]:
                    t["state"] = 1
                    self.st["opened"] += 1
                    events.append(self._event(t, 1, np.nan))
                elif t["hist"] == 0:
                    continue
            elif t["misses"] >= c["close_k"]:
                events.append(self._close(t))
                continue
            elif grew:
                self.st["updates"] += 1
                events.append(self._event(t, 2, np.nan))
            keep.append(t)
        self.tracks = keep
        return events

    def process(self, frames, t0, dt):
        frames = np.asarray(frames, np.float32).reshape(-1, self.nb)
        ev = []
        for j in range(frames.shape[0]):
            ev += self._frame(frames[j], t0 + j * dt, dt)
        return np.concatenate(ev) if ev else np.zeros(0, CFAR_EVENT)

    def flush(self):
        ev = [self._close(t, 16) for t in self.tracks if t["state"] == 1]
        self.tracks = []
        return np.concatenate(ev) if ev else np.zeros(0, CFAR_EVENT)

    def stats(self):
        st = dict(self.st)
        st["tracks"] = len(self.tracks)
        return st


def rel_err(a, b) -> float:
    """max |a - b| / max |b| over the whole output (the reference's measure,
    tests/test_channelizer.py:88)."""
    a = np.asarray(a)
    b = np.asarray(b)
    den = np.max(np.abs(b)) if b.size else 0.0
    return float(np.max(np.abs(a - b)) / den) if den > 0 else float(np.max(np.abs(a - b)) if a.size else 0.0)
