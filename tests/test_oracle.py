"""Pins the numpy oracle (oracle/oracle.py), CPU only.

Every oracle function is checked against at least one of: the golden vectors
the compiled reference wrote (tests/golden/), the reference's own known-answer
tests (reference tests/test_smoke.c and tests/test_cross_check.py lines cited
per test), or a closed form chosen so that a plausible slip (a wrong sign, a
dropped term, a transposed index, an off-by-one hop) fails it.
"""
import math

import numpy as np
import pytest

from conftest import golden
from oracle import oracle as O


# ---- unpack -----------------------------------------------------------------
@pytest.mark.parametrize("fmt", range(5))
def test_unpack_is_bit_exact_with_the_reference(fmt):
    g = golden("unpack")
    raw = g[f"raw{fmt}"].tobytes()
    for j, off in enumerate(g["offsets"]):
        out, mean = O.unpack_dc(raw, fmt, complex(off))
        ref = g[f"out{fmt}_{j}"]
        assert out.size == ref.size == len(raw) // O.BPS[fmt]
        assert np.array_equal(out.view(np.uint32), ref.view(np.uint32))
        # the mean restates the reference's eight FP64 lanes: equal to the bit
        assert mean.real == g[f"mean{fmt}_{j}"][0] and mean.imag == g[f"mean{fmt}_{j}"][1]


def test_unpack_known_answers():
    # reference tests/test_smoke.c:26-34: q11 2048 -> 1.0, -2048 -> -1.0, torn tail ignored
    raw = np.array([2048, -2048, 0, 1024], "<i2").tobytes() + b"\x01"
    out, _ = O.unpack_dc(raw, 3)
    assert out.tolist() == [1 - 1j, 0.5j]
    # cu8 255 -> (255-127.5)/127.5 = 1; 0 -> -1 (closed form)
    out, _ = O.unpack_dc(bytes([255, 0]), 0)
    assert out[0] == 1 - 1j
    # the offset is removed and the mean is BEFORE the offset
    # (reference tests/test_cross_check.py:566-591)
    raw = np.array([100, -50] * 64, "<i2").tobytes()
    out, mean = O.unpack_dc(raw, 3, 0.01 + 0.02j)
    # the mean is formed from the float32 outputs plus the offset, so it holds
    # to float32 resolution of the samples (reference gate: 1e-6)
    assert abs(mean - (100 / 2048 - 50j / 2048)) < 1e-8
    assert np.allclose(out, (100 / 2048 - 0.01) + 1j * (-50 / 2048 - 0.02), atol=1e-7)


def test_dc_pole_matches_the_reference_across_blocks():
    g = golden("pole")
    raw = g["raw"].tobytes()
    a, b = g["splits"]
    st = 0j
    outs = []
    for chunk in (raw[:a], raw[a:b], raw[b:]):
        y, st = O.unpack(chunk, 3, float(g["alpha"]), st)
        outs.append(y)
    y = np.concatenate(outs)
    # the reference runs the pole in FP32, the oracle in FP64: tolerance of
    # reference tests/test_cross_check.py:115
    assert np.max(np.abs(y - g["out"])) < 2e-6
    assert abs(st - g["states"][-1]) < 2e-6
    assert abs(y[-500:].mean()) < 0.01          # the DC is gone


# ---- windows / FFT ----------------------------------------------------------------
@pytest.mark.parametrize("kind", range(4))
def test_window_and_stats_match_the_reference(kind):
    g = golden("window")
    w = O.window(kind, 2048)
    assert np.max(np.abs(w - g[f"win{kind}"])) <= 6e-8
    cg, enbw = O.window_stats(g[f"win{kind}"])
    assert np.allclose([cg, enbw], g[f"wstats{kind}"], rtol=1e-12)


def test_window_stats_textbook_values():
    # reference tests/test_cross_check.py:1492-1506: rect 1/1, Hann 0.5/1.5,
    # 4-term Blackman-Harris 0.35875/2.0044 (Harris 1978, table 1)
    n = 1 << 16
    for kind, cg, enbw, tol in ((0, 1.0, 1.0, 1e-12), (1, 0.5, 1.5, 1e-3), (3, 0.35875, 2.0044, 1e-3)):
        a, b = O.window_stats(O.window(kind, n))
        assert abs(a - cg) < tol and abs(b - enbw) < tol


@pytest.mark.parametrize("n", [16, 1000, 1024, 4096])
def test_fft_matches_the_reference_and_a_brute_force_dft(n):
    g = golden("fft")
    x = g[f"x{n}"]
    X = O.fft(x)
    assert O.rel_err(X, g[f"fwd{n}"]) < 1e-7
    assert O.rel_err(O.fft(x, True), g[f"inv{n}"]) < 1e-7
    if n <= 1024:
        j = np.arange(n)
        D = np.exp(-2j * np.pi * np.outer(j, j) / n) @ x.astype(np.complex128)
        assert O.rel_err(X, D) < 1e-12


def test_power_db_matches_the_reference():
    g = golden("window")
    assert np.max(np.abs(O.power_db(g["pdb_x"], g["win1"][:0] if False else O.window(1, 4096)) - g["pdb_hann"])) < 1e-3
    assert np.max(np.abs(O.power_db(g["pdb_x"], None) - g["pdb_rect"])) < 1e-3


def test_power_db_tone_known_answer():
    # reference tests/test_smoke.c:102-121: a tone at bin +100 lands in shifted bin n/2+100
    n = 1024
    x = np.exp(2j * np.pi * 100 / n * np.arange(n)).astype(np.complex64)
    line = O.power_db(x, None)
    assert int(np.argmax(line)) == n // 2 + 100
    assert abs(line[n // 2 + 100]) < 1e-4          # |X|/n = 1 -> 0 dB


@pytest.mark.parametrize("det", range(3))
@pytest.mark.parametrize("hop", [1024, 512])
def test_spectrum_reduce_matches_the_reference(det, hop):
    g = golden("spectrum")
    line, fr = O.spectrum_reduce(g["x"], 1024, hop, g["w"], det)
    assert fr == int(g[f"fr{det}_{hop}"])
    assert np.max(np.abs(line - g[f"red{det}_{hop}"])) < 1e-3


def test_spectrum_stats_match_the_reference():
    g = golden("spectrum")
    st = O.spectrum_stats(g["x"], 1024, 512, g["w"])
    assert st["frames"] == int(g["stats_frames"])
    for k in ("max", "avg", "min"):
        assert np.max(np.abs(st[k] - g[f"stats_{k}"])) < 1e-3
    assert np.allclose(st["sk"], g["stats_sk"], rtol=1e-4)
    st = O.spectrum_stats(g["c2x"], 4096, 2048, g["c2w"])
    for k in ("max", "avg", "min"):
        assert np.max(np.abs(st[k] - g[f"c2stats_{k}"])) < 1e-3
    line, fr = O.spectrum_reduce(g["c2x"], 4096, 2048, g["c2w"], 0)
    assert fr == int(g["c2fr"]) and np.max(np.abs(line - g["c2red"])) < 1e-3


def test_spectrum_reduce_of_identical_frames_is_power_db():
    # reference tests/test_smoke.c:110-121
    n = 256
    fr = (np.exp(2j * np.pi * 17 / n * np.arange(n)) + 0.1).astype(np.complex64)
    x = np.tile(fr, 6)
    line, frames = O.spectrum_reduce(x, n, n, None, 0)
    assert frames == 6
    assert np.max(np.abs(line - O.power_db(fr, None))) < 1e-3


def test_spectral_kurtosis_closed_forms():
    # reference tests/test_cross_check.py:1453-1466: ~2 for noise, ~1 for a steady tone
    rng = np.random.default_rng(5)
    n = 512
    x = (rng.standard_normal(n * 400) + 1j * rng.standard_normal(n * 400)).astype(np.complex64)
    sk = O.spectrum_stats(x, n, n, None)["sk"]
    assert abs(np.median(sk) - 2.0) < 0.1
    t = np.exp(2j * np.pi * 64 / n * np.arange(n * 50)).astype(np.complex64)
    sk = O.spectrum_stats(t, n, n, None)["sk"]
    assert sk[n // 2 + 64] < 1.01


# ---- display ----------------------------------------------------------------------
def test_display_helpers_are_bit_exact_with_the_reference():
    g = golden("display")
    for m in (1024, 1000, 4096, 5000, 3, 1):
        a = O.decimate_max(g["line"], m)
        assert np.array_equal(a.view(np.uint32), g[f"dec{m}"].view(np.uint32)), m
    assert np.array_equal(O.db_to_pixels(g["pixvals"], -120.0, 0.0, g["lut"]), g["pix"])
    assert np.array_equal(O.db_to_pixels(g["line"], -50.0, -50.0, g["lut"]), g["pix_flat"])
    t = g["trace0"].copy()
    for j, c in enumerate(g["stitch_cases"]):
        O.stitch_max(g["seg"], *c, t)
        assert np.array_equal(t.view(np.uint32), g[f"stitch{j}"].view(np.uint32)), j


def test_display_known_answers():
    # reference tests/test_smoke.c:141-153 style placements, NaN first sticks
    x = np.array([np.nan, 1, 2, 3, 4, np.nan, 0, 9], np.float32)
    assert np.isnan(O.decimate_max(x, 2)[0]) and O.decimate_max(x, 2)[1] == 9
    assert O.decimate_max(np.arange(10, dtype=np.float32), 3).tolist() == [2, 5, 9]
    out = np.full(4, np.nan, np.float32)
    O.stitch_max(np.array([1, 2, 3, 4], np.float32), 0.0, 1.0, 0.0, 10.0, 0.0, 2.0, out)
    assert out[:2].tolist() == [2, 4] and np.isnan(out[2:]).all()


# ---- PFB ----------------------------------------------------------------------------
def test_prototype_closed_form_properties():
    h = O.pfb_prototype(64, 8, 81.2953)
    assert h.size == 512
    assert abs(h.sum() - 64) < 1e-9                      # unit DC gain x nchan
    assert np.allclose(h, h[::-1], atol=1e-15)           # linear phase
    assert abs(O.kaiser_beta(81.2953) - 8.0) < 1e-4      # beta = 8 at 81.2953 dB


def _pfb_rel(M, os_, K, atten, x, y_ref, nel=1, cal=None):
    p = O.Pfb(2.4e6, M, os_, K, atten, nel)
    if cal is not None:
        p.set_cal(cal, cal.shape[0])
    y = p.process_n(x if nel > 1 else [x], nel)
    y = y if nel > 1 else y[0]
    m = min(y.shape[-1], y_ref.shape[-1])
    return O.rel_err(y[..., :m], y_ref[..., :m])


def test_pfb_fp64_oracle_matches_the_reference():
    # the reference differs from its FP64 twin by <= 8.8e-8 of full scale
    # (reference CHANGELOG.md:155-157) and its tests hold it to 5e-7
    # (tests/test_channelizer.py:397-398): the oracle must sit inside that
    g = golden("pfb")
    assert _pfb_rel(64, 1, 8, 81.2953, g["c1_x"], g["c1_y"]) < 5e-7
    assert _pfb_rel(192, 2, 12, 80.0, g["t192_x"], g["t192_y"]) < 5e-7
    for os_ in (3, 4):
        assert _pfb_rel(24, os_, 12, 80.0, g[f"t24_{os_}_x"], g[f"t24_{os_}_y"]) < 5e-7
    assert _pfb_rel(1026, 2, 12, 80.0, g["m1026_x"], g["m1026_y"]) < 5e-7
    assert _pfb_rel(1024, 2, 16, 80.0, g["m1024_x"], g["m1024_y"]) < 5e-7
    assert _pfb_rel(2048, 2, 12, 80.0, g["m2048_x"], g["m2048_y"]) < 5e-7
    assert _pfb_rel(60, 3, 12, 80.0, list(g["t60_x"]), g["t60_y"], nel=2) < 5e-7
    assert _pfb_rel(192, 2, 12, 80.0, list(g["el5_x"]), g["el5_y"], nel=5, cal=g["el5_cal"]) < 5e-7


def test_pfb_oracle_chunk_invariance_and_bookkeeping():
    rng = np.random.default_rng(3)
    x = (rng.standard_normal(192 * 50 + 7) + 1j * rng.standard_normal(192 * 50 + 7)).astype(np.complex64)
    y1 = O.Pfb(2.4e6, 192, 2, 12).process(x)
    p = O.Pfb(2.4e6, 192, 2, 12)
    parts = [p.process(x[a:b]) for a, b in ((0, 5), (5, 1000), (1000, 1001), (1001, x.size))]
    y2 = np.concatenate(parts, axis=1)
    assert y1.shape == y2.shape and np.allclose(y1, y2, atol=1e-12, rtol=0)


def test_pfb_tone_lands_in_its_channel_as_true_baseband():
    # reference tests/test_smoke.c:674-700 (tone at channel 10 of 64) and
    # tests/test_channelizer.py:299-319 (true baseband at 2x/3x/4x)
    fs, M = 2.4e6, 24
    for os_ in (2, 3, 4):
        for b, df in ((1, 3e3), (5, -4e3)):
            n = M * 2500
            x = np.exp(2j * np.pi * (b * fs / M + df) / fs * np.arange(n)).astype(np.complex64)
            p = O.Pfb(fs, M, os_, 12)
            y = p.process(x)[b]
            rate = p.out_rate()
            k = 1 << int(np.floor(np.log2(len(y))))
            Y = np.abs(np.fft.fft(y[:k] * np.blackman(k)))
            f = np.fft.fftfreq(k, 1 / rate)[np.argmax(Y)]
            assert abs(f - df) <= rate / k
    x = np.exp(2j * np.pi * 10 * (2.4e6 / 64) / 2.4e6 * np.arange(64 * 200)).astype(np.complex64)
    y = O.Pfb(2.4e6, 64, 2, 12).process(x)[:, 40:]
    assert int(np.argmax((np.abs(y) ** 2).mean(1))) == 10


def test_w1_w2_oracle_matches_the_reference():
    g = golden("pfb")
    y = g["el5_y"]
    R = np.zeros((192, 5, 5), np.complex128)
    c = np.zeros(192, np.uint32)
    O.pfb_cov_accumulate(y, 5, 192, None, R, c)
    assert O.rel_err(R, g["w1_R"]) < 1e-5 and np.array_equal(c, g["w1_cnt"])
    O.pfb_cov_accumulate(y, 5, 192, g["w1_gate"], R, c)
    assert O.rel_err(R, g["w1_R2"]) < 1e-5 and np.array_equal(c, g["w1_cnt2"])
    # Hermitian by construction
    assert np.allclose(R, np.conj(np.transpose(R, (0, 2, 1))))
    p, mp = O.pfb_channel_power(y, 5, 192)
    assert O.rel_err(p, g["w2_p"]) < 1e-6 and O.rel_err(mp, g["w2_m"]) < 1e-6


def test_oracle_detects_planted_mistakes():
    """A wrong rotation sign, a dropped tap or an off-by-one hop in the oracle
    would fail the golden comparison: plant each on the CPU and check."""
    g = golden("pfb")
    x, yref = g["t192_x"], g["t192_y"]
    good = O.Pfb(2.4e6, 192, 2, 12).process(x)
    m = min(good.shape[1], yref.shape[1])
    # conjugated rotation
    bad = good.copy()
    bad[1::2, 1::2] *= -1
    assert O.rel_err(bad[:, :m], yref[:, :m]) > 1e-2
    # hop off by one
    assert O.rel_err(good[:, 1:m], yref[:, :m - 1]) > 1e-2
    # dropped last tap
    p = O.Pfb(2.4e6, 192, 2, 12)
    p.h = p.h.copy()
    p.h[-192:] = 0
    assert O.rel_err(p.process(x)[:, :m], yref[:, :m]) > 1e-6


# ---- NCO ---------------------------------------------------------------------------
def _nco_blocks(nco, x, blocks):
    out, a, k = [], 0, 0
    while a < x.size:
        b = min(x.size, a + blocks[k % len(blocks)])
        out.append(nco.mix(x[a:b]))
        a, k = b, k + 1
    return np.concatenate(out)


def test_nco_oracle_matches_the_reference():
    """The reference's own continuity pattern (tests/test_cross_check.py:120-130)
    against golden vectors written by the compiled reference."""
    g = golden("nco")
    o = O.Nco(173_125.0, 2.4e6)
    y = _nco_blocks(o, g["x"], [1, 999, 4096, 7, 13000])
    assert np.max(np.abs(y - g["y"])) < 1e-6
    assert abs(o.phase - float(g["phase"])) < 1e-12
    o1 = O.Nco(173_125.0, 2.4e6)
    assert np.max(np.abs(o1.mix(g["x"]) - g["y1"])) < 1e-6
    # and a long rational-frequency run: the reference's phasors are on the exact phase
    p, q = (int(v) for v in g["pq"])
    fs = float(g["fs"])
    n_end = 1_000_000
    exact = np.array([O.nco_exact_phase(n, p, q) for n in range(n_end - 4096, n_end)])
    assert np.max(np.abs(g["tone_y_tail"] - np.exp(1j * exact))) < 1e-6
    assert abs(O.wrap(float(g["tone_phase"]) - O.nco_exact_phase(n_end, p, q))) < 1e-8
    ot = O.Nco(fs * p / q, fs)
    _nco_blocks(ot, np.ones(n_end, np.complex64), [65536])
    assert abs(O.wrap(ot.phase - O.nco_exact_phase(n_end, p, q))) < 1e-8


def test_nco_known_answers():
    """A tone at f0 mixed by an NCO at f0 is DC, continuously across blocks
    (tests/test_smoke.c:60-76); zero frequency is the identity; the exact
    phase closed form is periodic in q and linear in n."""
    fs, f0, N = 1e6, 12345.0, 4096
    n = np.arange(N)
    x = (np.cos(2 * np.pi * f0 * n / fs) + 1j * np.sin(2 * np.pi * f0 * n / fs)).astype(np.complex64)
    o = O.Nco(f0, fs)
    y = np.concatenate([o.mix(x[: N // 2]), o.mix(x[N // 2:])])
    assert np.max(np.abs(y.real - 1.0) + np.abs(y.imag)) < 2e-4
    z = (np.arange(10) + 1j).astype(np.complex64)
    assert np.array_equal(O.Nco(0.0, 1e6).mix(z), z)
    p, q = 1234, 10007
    assert O.nco_exact_phase(q, p, q) == 0.0
    assert abs(O.wrap(O.nco_exact_phase(1, p, q) + 2 * np.pi * p / q)) < 1e-15
    # a planted mistake (a conjugated mixer) fails the golden comparison
    g = golden("nco")
    o = O.Nco(-173_125.0, 2.4e6)
    assert np.max(np.abs(o.mix(g["x"]) - g["y1"])) > 0.1


# ---- the lowpass design, the arbitrary resampler and the overlap-save extractor --------
OSX_FS = 2.4e6


def _osx_gold():
    g = golden("osx")
    lens = g["osx_len"]
    ys = np.split(g["osx_y"], np.cumsum(lens)[:-1])
    return g, ys


def test_lowpass_design_is_bit_exact_with_the_reference():
    g = golden("osx")
    taps = np.split(g["design_taps"], np.cumsum(g["design_len"])[:-1])
    for args, t in zip(g["design_args"], taps):
        o = O.design_lowpass(*args)
        assert np.array_equal(o.view(np.uint32), t.view(np.uint32)), args


def test_lowpass_design_closed_form_properties():
    h = O.design_lowpass(25e3, 50e3, 80.0, 2.4e6).astype(np.float64)
    assert h.size % 2 == 1 and np.allclose(h, h[::-1])            # odd, symmetric
    assert abs(h.sum() - 1.0) < 1e-6                              # unit DC gain
    H = np.abs(np.fft.rfft(h, 1 << 16))
    f = np.fft.rfftfreq(1 << 16, 1.0 / 2.4e6)
    assert 20 * np.log10(H[f >= 50e3].max()) < -79.0               # the stopband edge holds 80 dB
    assert np.max(np.abs(20 * np.log10(H[f <= 20e3]))) < 0.01      # flat passband


def test_resampler_oracle_matches_the_reference():
    g = golden("osx")
    ys = np.split(g["rs_y"], np.cumsum(g["rs_len"])[:-1])
    x = g["rs_x"]
    for (ri, ro, at, ph), y in zip(g["rs_args"], ys):
        r = O.ArbResampler(ri, ro, at, int(ph))
        o = np.concatenate([r.process(x[:1000]), r.process(x[1000:1001]), r.process(x[1001:])])
        assert o.size == y.size, (ri, ro)
        assert O.rel_err(o, y) < 1e-6, (ri, ro, O.rel_err(o, y))


def test_resampler_oracle_float32_form_is_bit_exact():
    """The reference's float32 operation order (c = p + mu*d, acc += c*x,
    q ascending) restated per output reproduces the reference's values."""
    g = golden("osx")
    ri, ro, at, ph = g["rs_args"][0]
    ref = g["rs_y"][:150]
    r = O.ArbResampler(ri, ro, at, int(ph))
    F = np.float32
    x = g["rs_x"]
    ext = np.concatenate([np.zeros(r.Q, np.complex64), x])
    for K in range(ref.size):
        local = float(K) * r.step
        i = math.floor(local)
        b = min(int((local - i) * r.P), r.P - 1)
        mu = F((local - i) * r.P - b)
        ar = ai = F(0.0)
        for q in range(r.Q):
            if r.Q + i - q < 0:
                continue
            xv = ext[r.Q + i - q]
            c = F(r.p[b + q * r.P] + F(mu * r.d[b + q * r.P]))
            ar = F(ar + F(c * F(xv.real)))
            ai = F(ai + F(c * F(xv.imag)))
        assert F(ar) == ref[K].real and F(ai) == ref[K].imag, K


def test_resampler_oracle_tone_and_set_ratio():
    """A tone through a 3:2 down-resampler comes out at the same frequency
    (phase advancing by 2 pi f/out per output); after set_ratio the next
    output keeps its position and the rate changes."""
    fi, fo, f = 75e3, 50e3, 7e3
    n = 30000
    x = np.exp(2j * np.pi * f / fi * np.arange(n)).astype(np.complex64)
    r = O.ArbResampler(fi, fo, 80.0)
    y = r.process(x)
    k = np.arange(y.size)
    # output K sits at input K*fi/fo; the FIR's group delay is (Q*P-1)/2/P inputs
    d = np.angle(y[200:] * np.exp(-2j * np.pi * f / fo * k[200:]))
    assert np.std(np.unwrap(d)) < 1e-3
    assert abs(np.abs(y[200:]).mean() - 1.0) < 1e-3
    r2 = O.ArbResampler(fi, fo, 80.0)
    a = r2.process(x[:10000])
    r2.set_ratio(fo / fi * 1.001)
    b = r2.process(x[10000:])
    assert abs((a.size + b.size) - (10000 * fo / fi + 20000 * fo / fi * 1.001)) < 3


def test_osx_oracle_matches_the_reference():
    g, ys = _osx_gold()
    x = g["osx_x"]
    o = O.Osx(OSX_FS, 4096)
    chans = [o.add(*c) for c in g["osx_chans"]]
    outs = [[] for _ in chans]
    a = 0
    for c in list(g["osx_chunks"]) + [x.size - int(g["osx_chunks"].sum())]:
        y, _ = o.process(x[a:a + c])
        for i, ch in enumerate(chans):
            outs[i].append(y[ch])
        a += c
    for i, ch in enumerate(chans):
        got = np.concatenate(outs[i])
        assert got.size == ys[i].size, i
        assert O.rel_err(got, ys[i]) < 2e-6, (i, O.rel_err(got, ys[i]))


def test_osx_oracle_elements_match_the_reference():
    g, _ = _osx_gold()
    x = g["osx_x"]
    phi = np.deg2rad(31.0)
    els = [(x * np.exp(1j * e * phi)).astype(np.complex64) for e in range(3)]
    o = O.Osx(OSX_FS, 4096, 80.0, 3)
    c0 = o.add(*g["osx_chans"][0])
    c1 = o.add(*g["osx_chans"][3])
    y, _ = o.process_n(els, 3)
    assert O.rel_err(y[c0], g["osx_el_y0"]) < 2e-6
    assert O.rel_err(y[c1], g["osx_el_y1"]) < 2e-6


def test_osx_oracle_tone_lands_on_centre_with_unit_gain_and_stopband():
    """the reference's own properties (tests/test_channelizer.py:203-219, :319-330)"""
    n = 4096 * 30
    for off in (300e3, 513 * OSX_FS / 4096, -411737.3):
        x = np.exp(2j * np.pi * (off + 7e3) / OSX_FS * np.arange(n)).astype(np.complex64)
        o = O.Osx(OSX_FS, 4096)
        c = o.add(off, 50e3, 50e3)
        y = o.process(x)[0][c]
        assert abs(np.abs(y[300:800]).mean() - 1.0) < 1e-3
        k = np.arange(y.size)
        d = np.unwrap(np.angle(y * np.exp(-2j * np.pi * 7e3 / 50e3 * k)))[300:-50]
        assert np.std(d) < 1e-3, off
    x = np.exp(2j * np.pi * (300e3 + 150e3) / OSX_FS * np.arange(n)).astype(np.complex64)
    o = O.Osx(OSX_FS, 4096)
    c = o.add(300e3, 50e3, 50e3)
    y = o.process(x)[0][c]
    assert 20 * np.log10(np.abs(y[300:800]).mean() + 1e-20) < -80.0


def test_osx_oracle_catches_planted_mistakes():
    """a roll one bin off and a residual with the wrong sign each fail the
    comparison with the reference"""
    g, ys = _osx_gold()
    x = g["osx_x"]
    for plant in ("shift", "delta"):
        o = O.Osx(OSX_FS, 4096)
        c = o.add(*g["osx_chans"][2])
        if plant == "shift":
            o.chans[c]["shift"] = (o.chans[c]["shift"] + 1) % 4096
        else:
            o.chans[c]["delta"] = -o.chans[c]["delta"]
        y = o.process(x)[0][c]
        m = min(y.size, ys[2].size)
        assert O.rel_err(y[:m], ys[2][:m]) > 1e-2, plant


# ---- detection on a spectrum line ----------------------------------------------------------
def _detect_cases():
    g = golden("detect")
    p = g["d_params"]
    lines = np.split(g["d_line"], np.cumsum(p[:, 0].astype(int))[:-1])
    idx = np.split(np.arange(len(g["d_rec"])), np.cumsum(g["d_nrec"])[:-1])
    out = []
    for line, pp, fl, rc, ri in zip(lines, p, g["d_floor"], g["d_rc"], idx):
        recs = [(int(a), int(b), float(c), float(d), float(e))
                for (a, b), (c, d, e) in zip(g["d_rec"][ri], g["d_recf"][ri])]
        out.append((line, float(fl), float(pp[1]), int(pp[2]), int(pp[3]), int(pp[4]), int(rc), recs))
    return out


def test_detect_channels_oracle_is_the_reference():
    """every golden line (a busy band, noise only, SNR ties, runs at both
    ends, gap / min_run below their minimum, fewer than five bins) gives the
    reference's records exactly, in the reference's order; when out_cap is
    short the reference returns E_CAP with the first out_cap in bin order"""
    for line, fl, thr, gap, mr, cap, rc, recs in _detect_cases():
        got = O.detect_channels(line, fl, thr, gap, mr)
        if rc == -3:
            assert len(got) > cap and sorted(got)[:cap] == recs
        else:
            assert got == recs and rc == len(recs)


def test_median_oracle_is_the_reference_and_numpy():
    g = golden("detect")
    for x, y in zip(np.split(g["m_x"], np.cumsum(g["m_len"])[:-1]), g["m_y"]):
        assert O.median(x) == y
        assert O.median(x) == np.float32(np.median(x))      # the reference's own check, exact here
    assert O.median([3.0, 1.0, 2.0]) == 2.0 and O.median([4.0, 1.0, 3.0, 2.0]) == 2.5
    assert O.median([]) == -120.0


def _runs_brute(line, floor, thr, gap, mr):
    """an independent statement on a short line: regular expressions over the
    0/1 mask string (interior zero runs shorter than gap become ones, then
    runs of ones at least min_run wide)"""
    import re
    sm = O.smooth5(line)
    mask = "".join("1" if v > np.float32(floor) + np.float32(thr) else "0" for v in sm)
    gap, mr = max(gap, 2), max(mr, 2)
    mask = re.sub(r"(?<=1)0+(?=1)", lambda m: "1" * len(m.group()) if len(m.group()) < gap else m.group(), mask)
    return [(m.start(), m.end()) for m in re.finditer(r"1+", mask) if m.end() - m.start() >= mr]


def test_detect_channels_oracle_runs_match_a_brute_force_on_random_masks():
    rng = np.random.default_rng(5)
    for trial in range(300):
        n = int(rng.integers(1, 60))
        line = rng.choice(np.array([-100.0, -90.0, -40.0, -30.0], np.float32), n)
        gap, mr = int(rng.integers(0, 6)), int(rng.integers(0, 5))
        got = sorted((c[0], c[1]) for c in O.detect_channels(line, -100.0, 20.0, gap, mr))
        assert got == _runs_brute(line, -100.0, 20.0, gap, mr), (trial, line, gap, mr)


def test_detect_channels_oracle_record_fields_closed_form():
    """a flat run: peak = the level, snr = level - floor, centroid = the
    run's middle (equal weights); a run with one loud bin: the centroid sits
    at that bin to within its power share"""
    line = np.full(64, -120.0, np.float32)
    line[10:20] = -50.0
    # threshold 60 dB over a -120 floor: the smoothed ramps (-78 and -64 dB at
    # the plateau's first two and last two bins) fall below it, so the run is
    # the plateau less two bins each side
    (c,) = O.detect_channels(line, -120.0, 60.0, 2, 2)
    assert c[:2] == (12, 18) and c[3] == -50.0 and c[4] == np.float32(70.0)
    assert abs(c[2] - 14.5) < 1e-4
    # at 10 dB the 5-bin smoothing widens the run by two bins each side
    assert O.detect_channels(line, -120.0, 10.0, 2, 2)[0][:2] == (8, 22)
    line[15] = -20.0
    (c,) = O.detect_channels(line, -120.0, 60.0, 2, 2)
    # weights 1e-5 at five bins, 1e-2 at bin 15: 12 + 0.03012 / 0.01005
    assert abs(c[2] - 14.99701) < 1e-4 and c[3] == -20.0
