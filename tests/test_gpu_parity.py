"""Parity of the HIP path with the reference, through the C ABI (ctypes).

Against the golden vectors the compiled reference wrote (tests/golden/), the
FP64 oracle (oracle/oracle.py) and the reference's own invariants. Bars:
unpack, display helpers, stitch and W1 bit-exact; FFT rtol 1e-5 vs numpy
(reference tests/test_cross_check.py:265); power/reduce/stats within 3e-3 dB
(:294-295, :1445-1450); the filterbank within 5e-7 of FP64 (max-normalised,
tests/test_channelizer.py:397-398) and 1e-5 of the reference C; chunk
invariance exact (tests/test_channelizer.py:92-104).
"""
import threading

import numpy as np
import pytest

from conftest import golden
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def u32(a):
    return np.ascontiguousarray(a, np.float32).view(np.uint32)


# ---- unpack --------------------------------------------------------------------
@pytest.mark.parametrize("fmt", range(5))
def test_unpack_dc_bit_exact(gpu, fmt):
    g = golden("unpack")
    raw = g[f"raw{fmt}"].tobytes()
    for j, off in enumerate(g["offsets"]):
        out, mean = gpu.unpack_dc(raw, fmt, complex(off))
        ref = g[f"out{fmt}_{j}"]
        assert out.size == ref.size
        assert np.array_equal(out.view(np.uint32), ref.view(np.uint32))
        assert abs(mean - complex(*g[f"mean{fmt}_{j}"])) < 1e-12
        assert np.array_equal(gpu.unpack(raw, fmt).view(np.uint32), O.unpack_dc(raw, fmt)[0].view(np.uint32))


def test_unpack_large_block_and_mean(gpu):
    from atkdsp_amd import synth
    raw = synth.c2_bytes(n=2_000_000)
    out, mean = gpu.unpack_dc(raw, "ci16q11", 0.001 - 0.002j)
    ref, rmean = O.unpack_dc(raw, 3, 0.001 - 0.002j)
    assert np.array_equal(out.view(np.uint32), ref.view(np.uint32))
    assert abs(mean - rmean) < 1e-12 * max(1.0, abs(rmean))


def test_unpack_errors(gpu):
    with pytest.raises(gpu.AtkDspError):
        gpu.unpack(b"\x00\x01", 9)
    assert gpu.unpack(b"\x01", "cu8").size == 0          # torn: nothing whole


def test_dc_pole_continuous_across_blocks(gpu):
    g = golden("pole")
    raw = g["raw"].tobytes()
    a, b = g["splits"]
    st = np.zeros(1, np.complex64)
    outs = [gpu.unpack(c, "ci16q11", float(g["alpha"]), st) for c in (raw[:a], raw[a:b], raw[b:])]
    y = np.concatenate(outs)
    assert np.max(np.abs(y - g["out"])) < 2e-6
    assert abs(st[0] - g["states"][-1]) < 2e-6
    yo, _ = O.unpack(raw, 3, float(g["alpha"]))
    assert np.max(np.abs(y - yo)) < 2e-6


# ---- FFT ------------------------------------------------------------------------------
@pytest.mark.parametrize("n", [2, 4, 8, 16, 32, 64, 128, 256, 512, 1000, 1024, 1026, 2048, 4093, 4096,
                               8192, 65536, 12289])
def test_fft_matches_numpy(gpu, n):
    rng = np.random.default_rng(n)
    x = ((rng.standard_normal(n) + 1j * rng.standard_normal(n))).astype(np.complex64)
    p = gpu.Fft(n)
    np.testing.assert_allclose(p.exec(x), np.fft.fft(x.astype(np.complex128)), rtol=1e-5,
                               atol=1e-4 * np.sqrt(n))
    np.testing.assert_allclose(p.exec(x, inverse=True), np.fft.ifft(x.astype(np.complex128)), rtol=1e-5,
                               atol=1e-6)
    # and against the FP64 oracle to the max-normalised 5e-7 used for the filterbank
    assert O.rel_err(p.exec(x), O.fft(x)) < 5e-7 if n <= 8192 else True


@pytest.mark.parametrize("n", [16, 1000, 1024, 4096])
def test_fft_matches_reference_golden(gpu, n):
    g = golden("fft")
    p = gpu.Fft(n)
    assert O.rel_err(p.exec(g[f"x{n}"]), g[f"fwd{n}"]) < 5e-7
    assert O.rel_err(p.exec(g[f"x{n}"], inverse=True), g[f"inv{n}"]) < 5e-7


def test_fft_plan_shared_by_threads(gpu):
    """Two threads hammer one plan with exec and spectrum_reduce (reference
    tests/test_cross_check.py:1360-1421); no corruption allowed."""
    n = 1024
    p = gpu.Fft(n)
    rng = np.random.default_rng(1)
    xs = [(rng.standard_normal(n) + 1j * rng.standard_normal(n)).astype(np.complex64) for _ in range(4)]
    want = [p.exec(x) for x in xs]
    big = (rng.standard_normal(n * 20) + 1j * rng.standard_normal(n * 20)).astype(np.complex64)
    line0, _ = p.spectrum_reduce(big, 512, None, "max")
    errs = []

    def worker(k):
        for it in range(200):
            if (it + k) % 5 == 0:
                line, _ = p.spectrum_reduce(big, 512, None, "max")
                if not np.array_equal(line, line0):
                    errs.append("reduce")
            else:
                x = xs[(it + k) % 4]
                if not np.array_equal(p.exec(x), want[(it + k) % 4]):
                    errs.append("exec")

    ts = [threading.Thread(target=worker, args=(k,)) for k in range(3)]
    for t in ts:
        t.start()
    for t in ts:
        t.join()
    assert not errs, errs[:5]


# ---- power, reduce, stats -------------------------------------------------------------------
def test_power_db_matches_reference(gpu):
    g = golden("window")
    p = gpu.Fft(4096)
    assert np.max(np.abs(p.power_db(g["pdb_x"], gpu.window("hann", 4096)) - g["pdb_hann"])) < 2e-3
    assert np.max(np.abs(p.power_db(g["pdb_x"]) - g["pdb_rect"])) < 2e-3


@pytest.mark.parametrize("det", ["max", "min", "avg"])
@pytest.mark.parametrize("hop", [1024, 512])
def test_spectrum_reduce_matches_reference(gpu, det, hop):
    g = golden("spectrum")
    d = gpu.DET[det]
    line, fr = gpu.Fft(1024).spectrum_reduce(g["x"], hop, g["w"], det)
    assert fr == int(g[f"fr{d}_{hop}"])
    assert np.max(np.abs(line - g[f"red{d}_{hop}"])) < 3e-3


def line_close(a_db, b_db, peak_db, tol_db, floor=1e-9):
    """The display-line tolerance: |dB error| <= tol_db on every bin within
    60 dB of the line's peak; below that the FP32 transform's error is set by
    the frame's energy, so the bound is on power: |dP| <= floor * P_peak."""
    a, b = np.asarray(a_db, np.float64), np.asarray(b_db, np.float64)
    strong = b >= peak_db - 60.0
    ok_strong = np.all(np.abs(a - b)[strong] <= tol_db)
    dp = np.abs(10 ** (a / 10) - 10 ** (b / 10))
    ok_weak = np.all(dp[~strong] <= floor * 10 ** (peak_db / 10))
    return ok_strong and ok_weak


def test_spectrum_reduce_c2_shape_and_stats(gpu):
    """The C2 content spans about 80 dB (tones at -10 dBFS over -30 dBFS
    noise spread across 4096 bins): the stated line tolerance applies."""
    g = golden("spectrum")
    p = gpu.Fft(4096)
    line, fr = p.spectrum_reduce(g["c2x"], 2048, g["c2w"], "max")
    assert fr == int(g["c2fr"])
    assert line_close(line, g["c2red"], float(g["c2red"].max()), 3e-3)
    st = p.spectrum_stats(g["c2x"], 2048, g["c2w"])
    pk = float(g["c2stats_max"].max())
    for k in ("max", "avg", "min"):
        assert line_close(st[k], g[f"c2stats_{k}"], pk, 2e-3), k
    strong = g["c2stats_avg"] >= pk - 60
    np.testing.assert_allclose(st["sk"][strong], g["c2stats_sk"][strong], rtol=2e-3)
    st = gpu.Fft(1024).spectrum_stats(g["x"], 512, g["w"])
    for k in ("max", "avg", "min"):
        assert np.max(np.abs(st[k] - g[f"stats_{k}"])) < 2e-3, k
    np.testing.assert_allclose(st["sk"], g["stats_sk"], rtol=2e-3)


def test_spectrum_edge_cases(gpu):
    p = gpu.Fft(1024)
    line, fr = p.spectrum_reduce(np.zeros(100, np.complex64))
    assert fr == 0                                         # short input: 0 frames
    with pytest.raises(gpu.AtkDspError):
        p.spectrum_reduce(np.zeros(2048, np.complex64), None, None, 7)
    st = p.spectrum_stats(np.zeros(2048, np.complex64), want=("max",))
    assert st["frames"] == 2 and set(st) == {"max", "frames"}
    assert np.allclose(st["max"], -120.0)                   # silent bins: the stats floor
    line, _ = p.spectrum_reduce(np.zeros(2048, np.complex64))
    assert np.allclose(line, -240.0)                        # the reduce floor
    # stats.avg equals reduce.avg (reference tests/test_cross_check.py:1479-1489)
    rng = np.random.default_rng(2)
    x = (rng.standard_normal(1024 * 30) + 1j * rng.standard_normal(1024 * 30)).astype(np.complex64)
    a = p.spectrum_stats(x, 512, gpu.window("hann", 1024))["avg"]
    b, _ = p.spectrum_reduce(x, 512, gpu.window("hann", 1024), "avg")
    assert np.max(np.abs(a - b)) < 2e-3


def test_spectrum_reduce_long_input_in_chunks(gpu):
    """More frames than one device chunk: the line must equal the oracle's."""
    rng = np.random.default_rng(9)
    n = 4096
    x = (rng.standard_normal(n * 700) + 1j * rng.standard_normal(n * 700)).astype(np.complex64)
    x[123456:125000] += 4 * np.exp(2j * np.pi * 0.1 * np.arange(1544))
    w = gpu.window("hann", n)
    for det in ("max", "avg"):
        line, fr = gpu.Fft(n).spectrum_reduce(x, 2048, w, det)
        lo, fo = O.spectrum_reduce(x, n, 2048, w, gpu.DET[det])
        assert fr == fo and np.max(np.abs(line - lo)) < 3e-3


# ---- display and stitch ----------------------------------------------------------------------------
def test_display_bit_exact(gpu):
    g = golden("display")
    for m in (1024, 1000, 4096, 5000, 3, 1):
        assert np.array_equal(u32(gpu.decimate_max(g["line"], m)), u32(g[f"dec{m}"])), m
    assert np.array_equal(gpu.db_to_pixels(g["pixvals"], -120.0, 0.0, g["lut"]), g["pix"])
    assert np.array_equal(gpu.db_to_pixels(g["line"], -50.0, -50.0, g["lut"]), g["pix_flat"])
    t = g["trace0"].copy()
    for j, c in enumerate(g["stitch_cases"]):
        gpu.stitch_max(g["seg"], *c, t)
        assert np.array_equal(u32(t), u32(g[f"stitch{j}"])), j


def test_decimate_long_columns_match_oracle(gpu):
    rng = np.random.default_rng(4)
    x = rng.standard_normal(65536 * 3 + 5).astype(np.float32)
    x[::977] = np.nan
    x[0] = np.nan
    x[65536] = 7.0
    x[65537] = 7.0
    for m in (16, 3, 7):
        assert np.array_equal(u32(gpu.decimate_max(x, m)), u32(O.decimate_max(x, m))), m


# ---- PFB ------------------------------------------------------------------------------------------
def pfb_case(gpu, M, os_, K, atten, x, nel=1, cal=None):
    p = gpu.Pfb(2.4e6, M, os_, K, atten, nel)
    if cal is not None:
        p.set_cal(cal, cal.shape[0])
    return p.process_n(x if nel > 1 else [x], nel) if nel > 1 else p.process(x)


@pytest.mark.parametrize("case", ["c1", "t192", "t24_3", "t24_4", "m1026", "m1024", "m2048", "t60", "el5"])
def test_pfb_matches_reference_and_fp64(gpu, case):
    g = golden("pfb")
    spec = {"c1": (64, 1, 8, 81.2953), "t192": (192, 2, 12, 80.0), "t24_3": (24, 3, 12, 80.0),
            "t24_4": (24, 4, 12, 80.0), "m1026": (1026, 2, 12, 80.0), "m1024": (1024, 2, 16, 80.0),
            "m2048": (2048, 2, 12, 80.0), "t60": (60, 3, 12, 80.0), "el5": (192, 2, 12, 80.0)}[case]
    nel = {"t60": 2, "el5": 5}.get(case, 1)
    x = g[f"{case}_x"]
    x = list(x) if nel > 1 else x
    cal = g["el5_cal"] if case == "el5" else None
    y = pfb_case(gpu, *spec, x, nel, cal)
    yr = g[f"{case}_y"]
    assert y.shape == yr.shape
    assert O.rel_err(y, yr) < 1e-5                        # vs the reference C
    po = O.Pfb(2.4e6, *spec, max_nel=nel)
    if cal is not None:
        po.set_cal(cal, cal.shape[0])
    yo = po.process_n(x if nel > 1 else [x], nel)
    yo = yo if nel > 1 else yo[0]
    assert O.rel_err(y, yo) < 5e-7                        # vs FP64


@pytest.mark.parametrize("blk", [1, 7, 999, 4096])
def test_pfb_streaming_continuity_c1(gpu, blk):
    """C1: blocks of 1, 7, 999 and 4096 samples give exactly the unbroken call."""
    g = golden("pfb")
    x = g["c1_x"]
    whole = gpu.Pfb(2.4e6, 64, 1, 8, 81.2953).process(x)
    p = gpu.Pfb(2.4e6, 64, 1, 8, 81.2953)
    parts = []
    for i in range(0, x.size, blk):
        o = p.process(x[i:i + blk])
        if o.shape[1]:
            parts.append(o)
    y = np.concatenate(parts, axis=1)
    assert y.shape == whole.shape
    assert np.array_equal(y, whole)


def test_pfb_chunk_invariance_and_identical_elements(gpu):
    rng = np.random.default_rng(4)
    x = (np.exp(2j * np.pi * 40 / 192 * np.arange(192 * 80))
         + 0.02 * (rng.standard_normal(192 * 80) + 1j * rng.standard_normal(192 * 80))).astype(np.complex64)
    y1 = gpu.Pfb(2.4e6, 192).process(x)
    p = gpu.Pfb(2.4e6, 192)
    outs = [p.process(s) for s in (x[:5000], x[5000:5003], x[5003:30000], x[30000:])]
    y2 = np.concatenate([o for o in outs if o.shape[1]], axis=1)
    assert np.array_equal(y1, y2)
    o = gpu.Pfb(2.4e6, 192, max_nel=5).process_n([x] * 5, 5)
    assert all(np.array_equal(o[0], o[e]) for e in range(5))


def test_pfb_long_block_spans_device_slabs(gpu):
    """C2 shape on a block longer than one device slab: equal to the call
    split in two (so the slab seam is invisible) and within 5e-7 of FP64."""
    from atkdsp_amd import synth
    raw = synth.c2_bytes(n=2_300_000)
    x, _ = O.unpack_dc(raw, 3)
    y = gpu.Pfb(40e6, 1024, 2, 16, 80.0).process(x)
    p = gpu.Pfb(40e6, 1024, 2, 16, 80.0)
    y2 = np.concatenate([p.process(x[:1_000_001]), p.process(x[1_000_001:])], axis=1)
    assert np.array_equal(y, y2)
    yo = O.Pfb(40e6, 1024, 2, 16, 80.0).process(x[:400_000])
    assert O.rel_err(y[:, :yo.shape[1]], yo) < 5e-7


def test_pfb_bookkeeping_cal_and_errors(gpu):
    p = gpu.Pfb(2.4e6, 192, 2)
    assert abs(p.out_rate() - 2 * 2.4e6 / 192) < 1e-6
    assert abs(p.group_delay() - (192 * 12 - 1) / 2.0) < 1e-9
    assert p.block_origin(0) == -(192 * 12 - 96)
    assert "192 chan" in p.describe()
    assert p.process(np.zeros(10, np.complex64)).shape == (192, 0)      # stash
    assert p.out_hops(86) == 1
    lib = gpu.load()
    out = np.zeros(10, np.complex64)
    x = np.zeros(1000, np.complex64)
    rc = lib.atkdsp_pfb_process(p._p, gpu._cfp(x), x.size, gpu._cfp(out), out.size)
    assert rc == -3                                                         # E_CAP, no state change
    assert p.out_hops(0) == 0 and p.out_hops(86) == 1
    with pytest.raises(gpu.AtkDspError):
        gpu.Pfb(2.4e6, 100, 3)                                              # nchan % os != 0
    # cal applied once: element 1 rotated by 90 degrees (reference tests/test_channelizer.py:181-193)
    rng = np.random.default_rng(9)
    base = (rng.standard_normal(192 * 40) + 1j * rng.standard_normal(192 * 40)).astype(np.complex64)
    y0 = gpu.Pfb(2.4e6, 192, max_nel=2).process_n([base, base], 2)
    p1 = gpu.Pfb(2.4e6, 192, max_nel=2)
    cal = np.ones((2, 192), np.complex64)
    cal[1, :] = np.exp(1j * np.deg2rad(90.0))
    p1.set_cal(cal, 2)
    y1 = p1.process_n([base, base], 2)
    assert np.allclose(y1[0], y0[0], atol=1e-5)
    assert np.allclose(y1[1], y0[1] * np.exp(1j * np.deg2rad(90.0)), atol=1e-4)
    p1.reset()
    p1.set_cal(None, 0)
    assert np.array_equal(p1.process_n([base, base], 2), y0)


def test_pfb_plane_wave_phase(gpu):
    phi = np.deg2rad(37.0)
    base = np.exp(2j * np.pi * 40 / 192 * np.arange(192 * 200))
    ins = [(base * np.exp(1j * e * phi)).astype(np.complex64) for e in range(5)]
    o = gpu.Pfb(2.4e6, 192, max_nel=5).process_n(ins, 5)[:, 40, 60:-5]
    rel = np.angle(o.mean(1))
    rel = rel - rel[0]
    err = np.rad2deg(np.max(np.abs(np.angle(np.exp(1j * (rel - np.arange(5) * phi))))))
    assert err < 0.1


def test_w1_w2_match_reference(gpu):
    g = golden("pfb")
    y = g["el5_y"]
    R = np.zeros((192, 5, 5), np.complex64)
    c = np.zeros(192, np.uint32)
    gpu.pfb_cov_accumulate(y, 5, 192, None, R, c)
    assert np.array_equal(R.view(np.uint32), g["w1_R"].view(np.uint32))   # same order: bit-exact
    assert np.array_equal(c, g["w1_cnt"])
    gpu.pfb_cov_accumulate(y, 5, 192, g["w1_gate"], R, c)
    assert np.array_equal(R.view(np.uint32), g["w1_R2"].view(np.uint32))
    assert np.array_equal(c, g["w1_cnt2"])
    gpu.pfb_cov_reset(R, c, 5, 192, g["w1_gate"])
    assert np.all(R[40] == 0) and c[40] == 0 and c[39] == g["w1_cnt2"][39]
    p, mp = gpu.pfb_channel_power(y, 5, 192)
    assert np.array_equal(p, g["w2_p"]) and np.array_equal(mp, g["w2_m"])
