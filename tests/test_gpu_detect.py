"""Detection on spectrum lines on the GPU (reference src/detect.c:4-114):
the drop-in atkdsp_median and atkdsp_detect_channels, and the served
display lines' consumer (atkgpu_server_detect), which finds channels where
the line lives and returns only the records.

The bar: the median bit-exact; each channel's bins, peak and SNR bit-exact
and its order the reference's (ties included); the FP64 centroid within one
float ulp (the device's pow against the C library's). Checked against the
reference's golden lines and against the oracle on random lines of many
sizes, and the server's on-device floor (median of the smoothed line)
against the oracle on the same served line."""
import ctypes as C

import numpy as np
import pytest

from conftest import golden
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def _ulps(a, b):
    return abs(int(np.float32(a).view(np.int32)) - int(np.float32(b).view(np.int32)))


def _same(got, exp, what=""):
    """got: Channel objects or raw records; exp: oracle tuples"""
    assert len(got) == len(exp), (what, len(got), len(exp))
    for c, e in zip(got, exp):
        assert (c.start_bin, c.end_bin) == e[:2], (what, c, e)
        assert c.peak_db == np.float32(e[3]) and c.snr_db == np.float32(e[4]), (what, c, e)
        assert _ulps(c.centroid_bin, e[2]) <= 1, (what, c.centroid_bin, e[2])


def _golden_cases():
    g = golden("detect")
    p = g["d_params"]
    lines = np.split(g["d_line"], np.cumsum(p[:, 0].astype(int))[:-1])
    idx = np.split(np.arange(len(g["d_rec"])), np.cumsum(g["d_nrec"])[:-1])
    for line, pp, fl, rc, ri in zip(lines, p, g["d_floor"], g["d_rc"], idx):
        recs = [(int(a), int(b), float(c), float(d), float(e))
                for (a, b), (c, d, e) in zip(g["d_rec"][ri], g["d_recf"][ri])]
        yield line, float(fl), float(pp[1]), int(pp[2]), int(pp[3]), int(pp[4]), int(rc), recs


def _raw_detect(gpu, line, fl, thr, gap, mr, cap):
    lib = gpu.load()
    buf = (gpu._ChannelC * max(cap, 1))()
    sc = np.empty(max(line.size, 1), np.float32)
    n = lib.atkdsp_detect_channels(gpu._fp(line), line.size, fl, thr, gap, mr, gpu._fp(sc), buf, cap)
    return int(n), buf


def test_detect_channels_matches_the_reference_golden(gpu):
    for line, fl, thr, gap, mr, cap, rc, recs in _golden_cases():
        n, buf = _raw_detect(gpu, line, fl, thr, gap, mr, cap)
        assert n == rc, (line.size, n, rc)
        _same(buf[:cap if rc == -3 else n], recs, f"golden n={line.size} gap={gap}")


def test_median_matches_the_reference_golden(gpu):
    g = golden("detect")
    for x, y in zip(np.split(g["m_x"], np.cumsum(g["m_len"])[:-1]), g["m_y"]):
        assert np.float32(gpu.median(x)).view(np.uint32) == np.float32(y).view(np.uint32), x.size


@pytest.mark.parametrize("n", [1, 2, 3, 4, 5, 6, 7, 100, 1023, 1024, 4096, 16384, 100_003, 1 << 21])
def test_detect_random_lines_match_the_oracle(gpu, n):
    rng = np.random.default_rng(n)
    line = (-100.0 + 10 * np.log10(rng.exponential(1.0, n) + 1e-12)).astype(np.float32)
    for _ in range(max(1, n // 200)):                        # plateaus of several levels and widths
        a = int(rng.integers(0, n))
        line[a:a + int(rng.integers(1, 40))] = rng.choice([-80.0, -70.0, -60.0])
    fl = O.median(O.smooth5(line))
    for gap, mr, thr in [(4, 2, 6.0), (0, 0, 3.0), (9, 5, 10.0)]:
        cap = n // 4 + 2
        got = gpu.detect_channels(line, fl, thr, gap, mr, max_channels=cap)
        _same(got, O.detect_channels(line, fl, thr, gap, mr), f"n={n} gap={gap} min_run={mr}")


@pytest.mark.parametrize("n", [1, 2, 5, 6, 1000, 65_537, 1 << 22])
def test_median_random_and_ties_match_the_oracle(gpu, n):
    rng = np.random.default_rng(n)
    x = rng.standard_normal(n).astype(np.float32)
    assert gpu.median(x) == O.median(x)
    t = rng.integers(-3, 4, n).astype(np.float32)            # heavy ties
    assert gpu.median(t) == O.median(t)
    assert gpu.median(-np.abs(x) - 1e30) == O.median(-np.abs(x) - 1e30)


def test_detect_arguments(gpu):
    line = np.full(100, -100.0, np.float32)
    line[40:60] = -40.0
    with pytest.raises(gpu.AtkDspError):
        gpu.detect_channels(line, float("nan"), 6, 2, 2)
    with pytest.raises(gpu.AtkDspError):                     # found 1 > max_channels 0
        gpu.detect_channels(line, -100.0, 6, 2, 2, max_channels=0)
    assert gpu.detect_channels(line[:0], -100.0, 6, 2, 2) == []
    # gap and min_run below two act as two
    a = gpu.detect_channels(line, -100.0, 6, -5, 0)
    b = gpu.detect_channels(line, -100.0, 6, 2, 2)
    assert [c.as_tuple() for c in a] == [c.as_tuple() for c in b] and len(a) == 1
    with pytest.raises(gpu.AtkDspError):                     # beyond the drop-in's line size
        gpu.detect_channels(np.zeros((1 << 21) + 1, np.float32), -100.0, 6, 2, 2)


def test_server_detect_on_the_device_lines(gpu):
    from atkdsp_amd import synth
    from atkdsp_amd.serve import Server
    fs = 40e6
    srv = Server(hbm_budget=2 << 30, max_streams=4, max_batch=16, staging_bytes=64 << 20)
    disp = dict(nfft=4096, hop=2048, window="hann", detector="max", frames=9, cols=1024, slots=64)
    sid = srv.add_stream(fmt="ci16", fs=fs, nchan=1024, oversample=2, taps_per_poly=16, chunk_max=65536,
                         out_slots=2, display=disp)
    aid = srv.add_stream(fmt="cf32", fs=fs, nchan=256, oversample=2, taps_per_poly=12, chunk_max=32768,
                         out_slots=2, nel=2, display=dict(disp, cols=512))
    n = 400_000
    raw = synth.quantize(synth.wideband_iq(7, n, fs, ntones=30), O.CI16)
    z = synth.wideband_iq(8, n, fs, ntones=12)
    arr = np.concatenate([z, z * np.exp(0.7j)]).astype(np.complex64)
    for a in range(0, n, 50_000):
        srv.submit(sid, raw[a * 4:(a + 50_000) * 4])
    for a in range(0, n, 25_000):
        blk = np.concatenate([arr[a:a + 25_000], arr[n + a:n + a + 25_000]])
        srv.submit(aid, blk)
    srv.drain()
    assert srv.lines(sid) >= 10 and srv.lines(aid) >= 10
    checked = 0
    for s, nel, cols in ((sid, 1, 1024), (aid, 2, 512)):
        for L in range(srv.lines(s)):
            full = srv.read_line(s, L).reshape(nel, cols)
            for e in range(nel):
                line = full[e]
                chans, fl = srv.detect(s, L, element=e, threshold_db=6.0, gap_bins=4, min_run=2)
                assert np.float32(fl) == np.float32(O.median(O.smooth5(line))), (s, L, e)
                exp = O.detect_channels(line, fl, 6.0, 4, 2)
                assert len(exp) >= 1
                _same(chans, exp, f"stream {s} line {L} element {e}")
                # a given floor is used as given
                chans2, fl2 = srv.detect(s, L, element=e, floor_db=-90.0, threshold_db=3.0)
                assert fl2 == -90.0
                _same(chans2, O.detect_channels(line, -90.0, 3.0, 4, 2), f"floor -90, stream {s} line {L}")
                checked += 1
    assert checked >= 30
    # the newest lines in one launch equal the one-line queries (a query holds
    # one line length: the 1024- and 512-column streams go separately)
    with pytest.raises(gpu.AtkDspError):
        srv.detect_latest([sid, aid])
    for group in ([sid, sid], [aid]):
        res = srv.detect_latest(group, threshold_db=6.0, gap_bins=4, min_run=2, max_channels=256)
        for s_, r in zip(group, res):
            one, fl1 = srv.detect(s_, r["line"], element=0, threshold_db=6.0, gap_bins=4, min_run=2)
            assert r["line"] == srv.lines(s_) - 1 and r["floor"] == fl1 and not r["capped"]
            assert [c.as_tuple() for c in r["channels"]] == [c.as_tuple() for c in one]
    one, _ = srv.detect(sid, srv.lines(sid) - 1, threshold_db=6.0, gap_bins=4, min_run=2)
    assert len(one) >= 2
    capped = srv.detect_latest([sid], threshold_db=6.0, gap_bins=4, min_run=2, max_channels=1)[0]
    assert capped["capped"] and len(capped["channels"]) == 1
    assert capped["channels"][0].start_bin == min(c.start_bin for c in one)    # the first in bin order
    with pytest.raises(gpu.AtkDspError):
        srv.detect(sid, srv.lines(sid) + 100)                # not completed by the chunks submitted
    with pytest.raises(gpu.AtkDspError):
        srv.detect(aid, 0, element=2)                        # no such element
    srv.close()
