"""Seeded synthetic receiver data for the wideband front end's workloads.

  c1: 2.4 MSPS cu8, 120,000 samples: noise (sigma 12 counts about 127.5) plus
      3 tones at -20..-3 dBFS within +-1.1 MHz.
  c2: 40 MSPS ci16q11 (12-bit in int16), 2,000,000 samples: complex Gaussian
      noise at -30 dBFS, 20 tones at -60..-10 dBFS within +-19 MHz, 4 on/off
      bursts of 10-200 us at -20 dBFS.
  c4 streams: per-stream format/rate/content drawn from a seed.
  c5: 8 phase-coherent cf32 elements at 20 MSPS: common background, 5 plane
      waves with per-element phase ramps, per-element noise 20 dB down.
"""
from __future__ import annotations

import numpy as np

CU8, CI8, CI16, CI16Q11, CF32 = 0, 1, 2, 3, 4
BPS = {CU8: 2, CI8: 2, CI16: 4, CI16Q11: 4, CF32: 8}


def _tones(n, fs, freqs, levels_db, phases):
    t = np.arange(n, dtype=np.float64)
    z = np.zeros(n, np.complex128)
    for f, l, p in zip(freqs, levels_db, phases):
        z += 10 ** (l / 20.0) * np.exp(1j * (2 * np.pi * f / fs * t + p))
    return z


def c1_bytes(seed: int = 1001, n: int = 120_000, fs: float = 2.4e6) -> bytes:
    rng = np.random.default_rng(seed)
    freqs = rng.uniform(-1.1e6, 1.1e6, 3)
    lv = rng.uniform(-20, -3, 3)
    ph = rng.uniform(0, 2 * np.pi, 3)
    z = 127.5 * _tones(n, fs, freqs, lv, ph)
    re = np.clip(np.round(127.5 + 12 * rng.standard_normal(n) + z.real), 0, 255)
    im = np.clip(np.round(127.5 + 12 * rng.standard_normal(n) + z.imag), 0, 255)
    out = np.empty(2 * n, np.uint8)
    out[0::2], out[1::2] = re, im
    return out.tobytes()


def wideband_iq(seed: int, n: int, fs: float, noise_dbfs=-30.0, ntones=20, nbursts=4,
                tone_db=(-60.0, -10.0), burst_db=-20.0) -> np.ndarray:
    """Complex baseband at full scale 1: noise, tones, bursts."""
    rng = np.random.default_rng(seed)
    sig = 10 ** (noise_dbfs / 20.0) / np.sqrt(2.0)
    z = sig * (rng.standard_normal(n) + 1j * rng.standard_normal(n))
    fr = rng.uniform(-0.475 * fs, 0.475 * fs, ntones)
    z += _tones(n, fs, fr, rng.uniform(*tone_db, ntones), rng.uniform(0, 2 * np.pi, ntones))
    for _ in range(nbursts):
        ln = int(rng.uniform(10e-6, 200e-6) * fs)
        st = int(rng.integers(0, max(1, n - ln)))
        f = rng.uniform(-0.45 * fs, 0.45 * fs)
        t = np.arange(ln)
        z[st:st + ln] += 10 ** (burst_db / 20.0) * np.exp(2j * np.pi * f / fs * t)
    return z


def quantize(z: np.ndarray, fmt: int) -> bytes:
    """Complex full-scale-1 samples to raw bytes of a receiver format."""
    n = z.size
    if fmt == CU8:
        v = np.clip(np.round(127.5 + 127.5 * np.stack([z.real, z.imag], 1)), 0, 255).astype(np.uint8)
    elif fmt == CI8:
        v = np.clip(np.round(128 * np.stack([z.real, z.imag], 1)), -128, 127).astype(np.int8)
    elif fmt == CI16:
        v = np.clip(np.round(32768 * np.stack([z.real, z.imag], 1)), -32768, 32767).astype("<i2")
    elif fmt == CI16Q11:
        v = np.clip(np.round(2048 * np.stack([z.real, z.imag], 1)), -2048, 2047).astype("<i2")
    elif fmt == CF32:
        v = np.stack([z.real, z.imag], 1).astype("<f4")
    else:
        raise ValueError(fmt)
    return v.reshape(2 * n).tobytes()


def c2_bytes(seed: int = 1002, n: int = 2_000_000, fs: float = 40e6) -> bytes:
    return quantize(wideband_iq(seed, n, fs), CI16Q11)


def c5_elements(seed: int = 1005, n: int = 1_000_000, nel: int = 8, fs: float = 20e6) -> list:
    rng = np.random.default_rng(seed)
    base = 0.05 * (rng.standard_normal(n) + 1j * rng.standard_normal(n))
    t = np.arange(n)
    els = [base.copy() for _ in range(nel)]
    for _ in range(5):
        f = rng.uniform(-0.45 * fs, 0.45 * fs)
        a = 10 ** (rng.uniform(-30, -10) / 20)
        dphi = rng.uniform(-np.pi, np.pi)
        w = a * np.exp(2j * np.pi * f / fs * t)
        for e in range(nel):
            els[e] += w * np.exp(1j * e * dphi)
    for e in range(nel):
        els[e] += 0.005 * (rng.standard_normal(n) + 1j * rng.standard_normal(n))
    return [x.astype(np.complex64) for x in els]
