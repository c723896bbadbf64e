"""Write the golden vectors in tests/golden/*.npz from the reference library
compiled from its own sources (oracle/build_ref.sh -> oracle/_ref/atkdsp_ref.so).

Inputs are seeded and small; outputs are what the reference C produced. The
fixtures pin the numpy oracle (tests/test_oracle.py, CPU) and the HIP path
(tests/test_gpu_*.py). Run in a container that has /root/reference:

    sh oracle/build_ref.sh && python tests/golden/make_golden.py

Reference build flags are recorded in each file (build.sh's -std=c99: no FMA
contraction, so the FP32 fold is multiply-then-add).
"""
from __future__ import annotations

import sys
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parents[2]
sys.path.insert(0, str(ROOT))
from oracle import oracle  # noqa: E402
from oracle.ref import Ref, RefArb, RefOsx, design_lowpass  # noqa: E402
from atkdsp_amd import synth  # noqa: E402

OUT = Path(__file__).resolve().parent
FLAGS = "gcc -std=c99 -O3 -fno-math-errno -mavx2 -mfma -fopenmp (oracle/build_ref.sh)"


def noise(rng, n, scale=1.0):
    return ((rng.standard_normal(n) + 1j * rng.standard_normal(n)) * scale).astype(np.complex64)


def save(name, **arrs):
    np.savez_compressed(OUT / f"{name}.npz", flags=np.array(FLAGS), **arrs)
    print(name, {k: getattr(v, "shape", None) for k, v in arrs.items()})


def main():
    r = Ref()
    r.set_threads(1)
    rng = np.random.default_rng(20261203)

    # -- unpack / unpack_dc, every format, torn tail, with and without an offset
    d = {}
    for fmt in range(5):
        if fmt == 0:
            raw = rng.integers(0, 256, 4000, dtype=np.uint8).tobytes()
        elif fmt == 1:
            raw = rng.integers(-128, 128, 4000, dtype=np.int8).tobytes()
        elif fmt == 4:
            raw = rng.standard_normal(1000).astype("<f4").tobytes()
        else:
            raw = rng.integers(-2048, 2048, 2000, dtype="<i2").tobytes()
        raw = raw[:-1]
        for j, off in enumerate((0j, 0.0123 - 0.0456j)):
            out, mean = r.unpack_dc(raw, fmt, off)
            d[f"raw{fmt}"] = np.frombuffer(raw, np.uint8)
            d[f"out{fmt}_{j}"] = out
            d[f"mean{fmt}_{j}"] = np.array([mean.real, mean.imag])
    d["offsets"] = np.array([0j, 0.0123 - 0.0456j])
    save("unpack", **d)

    # -- DC pole continuous across blocks (tests/test_cross_check.py:107-116 shape)
    raw = (rng.integers(-2048, 2048, 20000, dtype="<i2") + 300).astype("<i2").tobytes()
    st = np.zeros(1, np.complex64)
    outs, states = [], []
    for chunk in (raw[:4000], raw[4000:4004], raw[4004:]):
        outs.append(r.unpack(chunk, 3, 0.01, st))
        states.append(st[0])
    save("pole", raw=np.frombuffer(raw, np.uint8), alpha=np.float32(0.01), out=np.concatenate(outs),
         states=np.array(states), splits=np.array([4000, 4004]))

    # -- FFT
    d = {}
    for n in (16, 1000, 1024, 4096):
        x = noise(rng, n)
        d[f"x{n}"] = x
        d[f"fwd{n}"] = r.fft(x)
        d[f"inv{n}"] = r.fft(x, inverse=True)
    save("fft", **d)

    # -- windows, window stats, power_db
    d = {}
    for k in range(4):
        w = r.window(k, 2048)
        d[f"win{k}"] = w
        d[f"wstats{k}"] = np.array(r.window_stats(w))
    x = noise(rng, 4096) + 0.5 * np.exp(2j * np.pi * 300 / 4096 * np.arange(4096)).astype(np.complex64)
    w = r.window(1, 4096)
    d["pdb_x"] = x.astype(np.complex64)
    d["pdb_hann"] = r.power_db(x.astype(np.complex64), w)
    d["pdb_rect"] = r.power_db(x.astype(np.complex64), None)
    save("window", **d)

    # -- spectrum reduce / stats
    x = noise(rng, 40 * 1024 + 37)
    x[5000:5300] += (3.0 * np.exp(2j * np.pi * 0.2 * np.arange(300))).astype(np.complex64)
    w = r.window(1, 1024)
    d = {"x": x, "w": w}
    for det in range(3):
        for hop in (1024, 512):
            line, fr = r.spectrum_reduce(x, 1024, hop, w, det)
            d[f"red{det}_{hop}"] = line
            d[f"fr{det}_{hop}"] = np.int64(fr)
    st = r.spectrum_stats(x, 1024, 512, w)
    for k in ("max", "avg", "min", "sk"):
        d[f"stats_{k}"] = st[k]
    d["stats_frames"] = np.int64(st["frames"])
    z = synth.wideband_iq(7, 60000, 40e6).astype(np.complex64)
    w4 = r.window(1, 4096)
    line, fr = r.spectrum_reduce(z, 4096, 2048, w4, 0)
    d.update(c2x=z, c2w=w4, c2red=line, c2fr=np.int64(fr))
    st = r.spectrum_stats(z, 4096, 2048, w4)
    for k in ("max", "avg", "min", "sk"):
        d[f"c2stats_{k}"] = st[k]
    save("spectrum", **d)

    # -- display helpers and stitch (exact)
    d = {}
    line = (rng.standard_normal(4096) * 10 - 60).astype(np.float32)
    line[[0, 7, 100, 101, 4095]] = np.nan
    line[5] = np.nan
    line[2000:2004] = [0.0, -0.0, 0.0, -0.0]
    d["line"] = line
    for m in (1024, 1000, 4096, 5000, 3, 1):
        d[f"dec{m}"] = r.decimate_max(line, m)
    lut = rng.integers(0, 2 ** 32, 256, dtype=np.uint64).astype(np.uint32)
    vals = np.concatenate([line, np.array([-np.inf, np.inf, -120.0, 0.0, -60.0, -59.99], np.float32)])
    d["lut"] = lut
    d["pixvals"] = vals
    d["pix"] = r.db_to_pixels(vals, -120.0, 0.0, lut)
    # a zero span maps every finite value to entry 0; +-inf there is undefined
    # behaviour in the reference ((int) of NaN), so the case feeds finite values
    d["pix_flat"] = r.db_to_pixels(line, -50.0, -50.0, lut)
    seg = (rng.standard_normal(1024) * 5 - 70).astype(np.float32)
    seg[[3, 500]] = np.nan
    trace = np.full(8192, np.nan, np.float32)
    trace[100:200] = -80.0
    d["seg"], d["trace0"] = seg, trace.copy()
    cases = [(100e6, 40e6 / 1024, 100.5e6, 139e6, 90e6, 10e3),
             (120e6, 40e6 / 1024, 0.0, 1e12, 90e6, 39062.5),
             (95e6, 40e6 / 1024, 96e6, 97e6, 90e6, 2e3),
             (80e6, 40e6 / 1024, 0.0, 1e12, 90e6, 10e3)]
    t = trace.copy()
    for j, c in enumerate(cases):
        t = r.stitch_max(seg, *c, t)
        d[f"stitch{j}"] = t.copy()
    d["stitch_cases"] = np.array(cases)
    save("display", **d)

    # -- PFB
    d = {}
    c1 = synth.c1_bytes()
    x1, _ = r.unpack_dc(c1, 0, 0j)
    p = r.pfb(2.4e6, 64, 1, 8, 81.2953, 1)
    d["c1_x"], d["c1_y"] = x1, p.process(x1)
    fs = 2.4e6
    k = np.arange(192 * 80)
    xt = (np.exp(2j * np.pi * (40 * fs / 192) / fs * k) + 0.02 * (rng.standard_normal(k.size)
                                                                    + 1j * rng.standard_normal(k.size)))
    xt = xt.astype(np.complex64)
    d["t192_x"] = xt
    d["t192_y"] = r.pfb(fs, 192, 2, 12, 80.0, 1).process(xt)
    for os_ in (3, 4):
        xs = noise(rng, 24 * 300 + 17)
        d[f"t24_{os_}_x"] = xs
        d[f"t24_{os_}_y"] = r.pfb(fs, 24, os_, 12, 80.0, 1).process(xs)
    for M, K, nmul in ((1026, 12, 40), (1024, 16, 48), (2048, 12, 24)):
        xs = noise(rng, M * nmul)
        d[f"m{M}_x"] = xs
        d[f"m{M}_y"] = r.pfb(fs, M, 2, K, 80.0, 1).process(xs)
    xs = [noise(rng, 60 * 300 + 17) for _ in range(2)]
    d["t60_x"] = np.stack(xs)
    d["t60_y"] = r.pfb(fs, 60, 3, 12, 80.0, 2).process_n(xs, 2)
    # five elements of a plane wave with a cal table on elements 0..3
    base = xt[: 192 * 60]
    phi = np.deg2rad(37.0)
    ins = [(base * np.exp(1j * e * phi)).astype(np.complex64) for e in range(5)]
    cal = np.exp(1j * rng.uniform(0, 2 * np.pi, (4, 192))).astype(np.complex64)
    pc = r.pfb(fs, 192, 2, 12, 80.0, 5)
    pc.set_cal(cal, 4)
    y5 = pc.process_n(ins, 5)
    d["el5_x"], d["el5_cal"], d["el5_y"] = np.stack(ins), cal, y5
    # W1 / W2 on that output, gated and not
    R = np.zeros((192, 5, 5), np.complex64)
    cnt = np.zeros(192, np.uint32)
    r.cov_accumulate(y5, 5, 192, None, R, cnt)
    d["w1_R"], d["w1_cnt"] = R.copy(), cnt.copy()
    gate = np.zeros(192, np.uint8)
    gate[[3, 40, 41, 191]] = 1
    r.cov_accumulate(y5, 5, 192, gate, R, cnt)
    d["w1_gate"], d["w1_R2"], d["w1_cnt2"] = gate, R, cnt
    pw, mp = r.channel_power(y5, 5, 192)
    d["w2_p"], d["w2_m"] = pw, mp
    save("pfb", **d)


def nco_golden():
    """The reference NCO on noise in the block pattern of its own continuity
    test (tests/test_cross_check.py:120-130), and a long tone run."""
    r = Ref()
    rng = np.random.default_rng(20261204)
    x = noise(rng, 30000)
    y, ph = r.nco_blocks(x, 173_125.0, 2.4e6, [1, 999, 4096, 7, 13000])
    y1, ph1 = r.nco_blocks(x, 173_125.0, 2.4e6, [30000])
    # 1,000,000 samples of a rational frequency (p/q of fs) in 65,536-sample calls
    fs, p, q = 40e6, 1234, 10007
    t = np.ones(1_000_000, np.complex64)
    yt, pht = r.nco_blocks(t, fs * p / q, fs, [65536])
    save("nco", x=x, y=y, phase=np.array(ph), y1=y1, phase1=np.array(ph1), tone_y_tail=yt[-4096:],
         tone_phase=np.array(pht), pq=np.array([p, q]), fs=np.array(fs))


OSX_FS = 2.4e6
# channels of the extractor case: (offset, bandwidth, out_rate); the first
# three are resampled (75 kHz channel rate onto 50 kHz), the last is not
OSX_CHANS = [(300e3, 50e3, 50e3), (513 * OSX_FS / 4096, 50e3, 50e3), (-411737.3, 50e3, 50e3),
             (-150e3, 50e3, 75e3)]
OSX_CHUNKS = [7000, 1, 12000]          # then the rest
OSX_PHI = np.deg2rad(31.0)


def osx_input(n=4096 * 8, seed=31):
    """tones 7 kHz above each channel centre plus noise, complex64"""
    k = np.arange(n)
    x = sum(np.exp(2j * np.pi * (off + 7e3) / OSX_FS * k) for off, _, _ in OSX_CHANS)
    rng = np.random.default_rng(seed)
    x = x + 0.01 * (rng.standard_normal(n) + 1j * rng.standard_normal(n))
    return x.astype(np.complex64)


def osx_elements(x, nel=3):
    return [(x * np.exp(1j * e * OSX_PHI)).astype(np.complex64) for e in range(nel)]


def osx_golden():
    """The reference lowpass design, arbitrary resampler and overlap-save
    extractor (src/design.c, src/resample.c, src/channelizer.c:473-788)."""
    r = Ref()
    designs = [(25e3, 50e3, 80.0, 2.4e6), (30e3, 37.5e3, 80.0, 4.8e6), (30e3, 40e3, 40.0, 1e6),
               (1e3, 3e3, 95.0, 48e3)]
    taps = [design_lowpass(r, *a) for a in designs]
    rng = np.random.default_rng(20261205)
    rx = noise(rng, 5000)
    rcases = [(75e3, 50e3, 80.0, 64), (1.0, 1.37, 60.0, 64), (48e3, 44.1e3, 80.0, 32)]
    ry = []
    for ri, ro, at, ph in rcases:
        a = RefArb(r, ri, ro, at, ph)
        ry.append(np.concatenate([a.process(rx[:1000]), a.process(rx[1000:1001]), a.process(rx[1001:])]))
    x = osx_input()
    o = RefOsx(r, OSX_FS, 4096)
    chans = [o.add(*c) for c in OSX_CHANS]
    outs = [[] for _ in chans]
    a = 0
    for c in OSX_CHUNKS + [x.size - sum(OSX_CHUNKS)]:
        y, _ = o.process(x[a:a + c])
        for i, ch in enumerate(chans):
            outs[i].append(y[ch])
        a += c
    ys = [np.concatenate(v) for v in outs]
    # three phase-coherent elements, two channels, one call
    oe = RefOsx(r, OSX_FS, 4096, 80.0, 3)
    c0 = oe.add(*OSX_CHANS[0])
    c1 = oe.add(*OSX_CHANS[3])
    ye, _ = oe.process_n(osx_elements(x), 3)
    save("osx", design_args=np.array(designs), design_len=np.array([t.size for t in taps]),
         design_taps=np.concatenate(taps), rs_args=np.array(rcases), rs_x=rx,
         rs_len=np.array([y.size for y in ry]), rs_y=np.concatenate(ry), osx_x=x,
         osx_chans=np.array(OSX_CHANS), osx_chunks=np.array(OSX_CHUNKS),
         osx_len=np.array([y.size for y in ys]), osx_y=np.concatenate(ys),
         osx_el_y0=ye[c0], osx_el_y1=ye[c1])


def detect_lines(seed=41):
    """Spectrum lines in dB for detect_channels and median: each is
    (line, threshold_db, gap_bins, min_run, cap); the floor is the reference
    median of the reference-smoothed line, as signal_id computes it."""
    rng = np.random.default_rng(seed)
    cases = []
    # a busy band (the reference's own test shape, tests/test_cross_check.py:366-383)
    n = 16384
    line = (-100.0 + 10 * np.log10(rng.exponential(1.0, n) + 1e-12)).astype(np.float32)
    for c, w, lvl in [(2000, 12, -60), (2030, 6, -70), (9000, 80, -50), (15000, 3, -65), (16380, 4, -55)]:
        line[c:c + w] = lvl + rng.standard_normal(min(w, n - c)).astype(np.float32)
    cases.append((line, 6.0, 4, 2, 4096))
    # noise only at 12 dB: nothing
    cases.append(((-100 + 10 * np.log10(rng.exponential(1.0, 8192))).astype(np.float32), 12.0, 4, 3, 256))
    # a display-sized line with plateaus of equal level (SNR ties), runs at both ends,
    # gaps of 1..6 bins between plateaus, and gap / min_run below their minimum
    line = np.full(4096, -110.0, np.float32)
    pos = 0
    for k in range(60):
        w = 2 + (k * 7) % 13
        line[pos:pos + w] = -70.0 if k % 3 else -60.0
        pos += w + 1 + (k % 6)
        if pos > 4000:
            break
    line[:5] = -65.0
    line[-4:] = -65.0
    cases.append((line, 10.0, 3, 2, 4096))
    cases.append((line, 10.0, 0, 0, 4096))
    # out_cap too small: ATKDSP_E_CAP with the first records in bin order
    cases.append((line, 10.0, 3, 2, 5))
    # fewer than five bins: no smoothing
    cases.append((np.array([-50.0, -51.0, -120.0], np.float32), -1.5, 2, 2, 16))
    return cases


def detect_golden():
    r = Ref()
    lines, params, recs, nrec, rcs, floors = [], [], [], [], [], []
    for line, thr, gap, mr, cap in detect_lines():
        fl = r.median(oracle.smooth5(line))             # an input: the floor as signal_id makes it
        got, rc = r.detect_channels(line, fl, thr, gap, mr, cap)
        lines.append(line)
        params.append((line.size, thr, gap, mr, cap))
        floors.append(fl)
        recs += got
        nrec.append(len(got))
        rcs.append(rc)
    rng = np.random.default_rng(43)
    med_x = [rng.standard_normal(n).astype(np.float32) for n in (5, 6, 1023, 1024, 131072)]
    save("detect", d_line=np.concatenate(lines), d_params=np.array(params, np.float64),
         d_floor=np.array(floors, np.float32), d_rc=np.array(rcs, np.int64), d_nrec=np.array(nrec),
         d_rec=np.array([rc_[:2] for rc_ in recs], np.int64).reshape(-1, 2),
         d_recf=np.array([rc_[2:] for rc_ in recs], np.float32).reshape(-1, 3),
         m_len=np.array([x.size for x in med_x]), m_x=np.concatenate(med_x),
         m_y=np.array([r.median(x) for x in med_x], np.float32))


CFAR_CASES = [(mode, n, pfa, looks, k, rho)
              for mode in (0, 1, 2, 3) for n in (1, 2, 8, 16, 64) for pfa in (0.5, 1e-3, 1e-6, 1e-12, 1e-15)
              for looks in (0.25, 1.0, 2.5, 10.0) for k in ((0, 1) if mode == 1 else (0,))
              for rho in (0, 1)]


def cfar_golden():
    """The reference's CFAR threshold design: alpha and E[T] over modes, cell
    counts, rates, looks, OS ranks and window correlation (Hann's, or none);
    n_eff; window correlations; the default configuration's bytes."""
    r = Ref()
    hann = (4.0 / 9.0, 1.0 / 36.0, 0.0, 0.0)
    rc, al, et = [], [], []
    for mode, n, pfa, looks, k, rho in CFAR_CASES:
        c, a, e = r.cfar_alpha(mode, n, pfa, looks, k, hann if rho else None)
        rc.append(c)
        al.append(a)
        et.append(e)
    rho2 = {f"rho2_{kind}_{n}": np.array(r.cfar_window_rho2(oracle.window(kind, n))[1])
            for kind in (0, 1, 2, 3) for n in (1000, 4096)}
    save("cfar", c_args=np.array([c[:5] for c in CFAR_CASES], np.float64), c_rho=np.array([c[5] for c in CFAR_CASES]),
         c_rc=np.array(rc), c_alpha=np.array(al), c_et=np.array(et),
         n_eff=np.array([[r.cfar_n_eff(n, hann), r.cfar_n_eff(n, None)] for n in range(1, 40)]),
         defaults=np.frombuffer(r.cfar_defaults_bytes(), np.uint8), **rho2)


HANN = [4.0 / 9.0, 1.0 / 36.0, 0.0, 0.0]


def cfar_scene(seed, nbins, nframes, looks, bursts, floor_db=-100.0, ripple_db=6.0):
    """Spectrum frames in dB: gamma(looks) noise power of unit mean over a
    slowly rippling floor, plus bursts (f0, f1, b0, b1, snr_db) of the same
    statistics; stored as float16 so the fixture stays small (the frames are
    defined as the float16 values)."""
    rng = np.random.default_rng(seed)
    p = rng.gamma(looks, 1.0 / looks, (nframes, nbins))
    for f0, f1, b0, b1, snr in bursts:
        f1 = min(f1, nframes)
        if f0 >= f1:
            continue
        p[f0:f1, b0:b1] += 10.0 ** (snr / 10.0) * rng.gamma(looks, 1.0 / looks, (f1 - f0, b1 - b0))
    floor = floor_db + ripple_db * np.sin(np.arange(nbins) * (2.0 * np.pi / nbins) * 3.0)
    return (10.0 * np.log10(p) + floor).astype(np.float16)


def cfar_scenarios():
    """(name, nbins, config, frames, ops, masks): ops index rows of frames"""
    out = []
    B = [(5, 30, 200, 206, 15.0), (10, 12, 300, 301, 18.0), (20, 28, 500, 504, 14.0), (28, 40, 500, 511, 14.0),
         (8, 50, 2, 7, 16.0), (15, 45, 700, 704, 13.0), (15, 45, 706, 709, 13.0), (40, 60, 900, 930, 12.0),
         (0, 60, 1015, 1024, 20.0)]
    fr = cfar_scene(101, 1024, 60, 1.0, B)
    out.append(("ca_default", 1024, {}, fr,
                [("process", 0, 1), ("process", 1, 7), ("process", 8, 20), ("process", 28, 32), ("flush",)], {}))
    fr = cfar_scene(102, 1024, 60, 1.0, B)
    out.append(("os_k4", 1024, dict(mode=1, k_frames=4, merge_gap=2, min_bins=2, open_m=3, open_n=4, close_k=4), fr,
                [("process", 0, 3), ("process", 3, 50), ("process", 53, 7), ("flush",)], {}))
    fr = cfar_scene(103, 1024, 70, 1.0, B + [(10, 50, 420, 425, 15.0), (10, 50, 395, 399, 15.0)])
    m1 = np.zeros(1024, np.uint8)
    m1[400:420] = 1
    m1[0:10] = 1
    m2 = np.zeros(1024, np.uint8)
    m2[690:702] = 1
    out.append(("go_mask", 1024, dict(mode=2, k_frames=2), fr,
                [("mask", "m1"), ("process", 0, 30), ("mask", "m2"), ("process", 30, 30), ("mask", None),
                 ("process", 60, 10), ("flush",)], {"m1": m1, "m2": m2}))
    fr = cfar_scene(104, 512, 48, 1.0, [(4, 40, 100, 110, 12.0), (20, 30, 300, 302, 15.0)])
    out.append(("so_busy", 512, dict(mode=3, pfa=1e-2, max_tracks=8, open_m=1, open_n=1, close_k=1), fr,
                [("process", i, 16) for i in range(0, 48, 16)] + [("flush",)], {}))
    fr = cfar_scene(105, 2048, 50, 4.0, [(5, 40, 1000, 1012, 8.0), (12, 20, 1500, 1501, 10.0),
                                         (30, 48, 60, 70, 9.0)])
    out.append(("ca_hann_looks", 2048, dict(rho2=HANN, looks=4.0, k_frames=3, open_m=3, open_n=5, close_k=5), fr,
                [("process", 0, 25), ("reset",), ("process", 25, 25), ("flush",)], {}))
    fr = cfar_scene(106, 1024, 40, 1.0, B).astype(np.float32)
    fr[5:9, 150:160] = np.nan
    fr[12:14, 205:207] = np.nan
    fr[20:22, 600:640] = -np.inf
    fr[25, 800:802] = 250.0
    fr[30:33, 350] = np.inf
    fr = fr.astype(np.float16)
    out.append(("ca_nan_inf", 1024, dict(k_frames=2), fr, [("process", 0, 40), ("flush",)], {}))
    fr = cfar_scene(107, 64, 80, 1.0, [(10, 60, 30, 32, 28.0), (20, 70, 0, 2, 28.0), (5, 15, 50, 51, 30.0)])
    m3 = np.ones(64, np.uint8)
    m3[[0, 1, 33]] = 0
    out.append(("os_small", 64, dict(mode=1, n_ref=2, n_guard=2, os_k=3, pfa=1e-3, merge_gap=0), fr,
                [("process", 0, 40), ("mask", "m3"), ("process", 40, 5), ("mask", None), ("process", 45, 35),
                 ("flush",)], {"m3": m3}))
    # narrow 5 dB bursts under an 8-frame average: many are found only on the
    # average, so t0 / t1 fall back to their coarse estimates
    rng = np.random.default_rng(5)
    B = []
    for i in range(30):
        f0, d = int(rng.integers(0, 60)), int(rng.integers(15, 40))
        B.append((f0, f0 + d, 40 + i * 32, 41 + i * 32, 5.0))
    fr = cfar_scene(108, 1024, 100, 1.0, B)
    out.append(("ca_weak_k8", 1024, dict(k_frames=8, pfa=1e-4, close_k=3), fr,
                [("process", 0, 37), ("process", 37, 63), ("flush",)], {}))
    return out


def cfar_events_golden():
    """The reference's CFAR detector and tracker over the scenarios above:
    every event of every call, the counters, the last frame's average /
    threshold / hits, alpha / alpha_ca / E[T] and the description."""
    import json
    r = Ref()
    arrs, meta = {}, []
    for name, nbins, cfg, fr, ops, masks in cfar_scenarios():
        f32 = fr.astype(np.float32)
        calls = []
        for op in ops:
            if op[0] == "process":
                calls.append(("process", f32[op[1]:op[1] + op[2]], op[1] * 0.01, 0.01))
            elif op[0] == "mask":
                calls.append(("mask", None if op[1] is None else masks[op[1]]))
            else:
                calls.append(op)
        res = r.cfar_run(nbins, cfg, calls)
        assert res is not None, name
        evs, st, last, info, desc = res
        arrs[f"{name}_frames"] = fr
        for k, v in masks.items():
            arrs[f"{name}_{k}"] = v
        arrs[f"{name}_events"] = np.concatenate(evs)
        arrs[f"{name}_evop"] = np.concatenate([np.full(len(e), i, np.int32) for i, e in enumerate(evs)])
        arrs[f"{name}_stats"] = st
        arrs[f"{name}_last"] = np.stack([last[0], last[1], last[2].astype(np.float32)])
        arrs[f"{name}_info"] = np.array(info)
        meta.append(dict(name=name, nbins=nbins, cfg=cfg, ops=ops, describe=desc))
        print(name, {k: int(v) for k, v in zip(("frames", "warmup", "tested", "hits", "untestable", "opened",
                                                   "closed", "updates", "dropped", "resets", "tracks"), st)})
    save("cfar_events", meta=np.array(json.dumps(meta)), **arrs)


if __name__ == "__main__":
    import sys
    if sys.argv[1:] == ["nco"]:
        nco_golden()
    elif sys.argv[1:] == ["osx"]:
        osx_golden()
    elif sys.argv[1:] == ["detect"]:
        detect_golden()
    elif sys.argv[1:] == ["cfar"]:
        cfar_golden()
    elif sys.argv[1:] == ["cfar_events"]:
        cfar_events_golden()
    else:
        main()
        nco_golden()
        osx_golden()
        detect_golden()
        cfar_golden()
        cfar_events_golden()
