"""Display lines of served streams: line L of a stream is spectrum_reduce
(reference src/fft.c:177-295) over frames [L*F, (L+1)*F) of the stream's
samples, decimated with decimate_max (src/display.c:19-37). The frames are
transformed as uneven chunks arrive, in batches with other streams; every
line must match the oracle on the same span within the display tolerance,
the line count must follow the frame grid, and a stream's lines must be
bit-identical alone and in the mix."""
import numpy as np
import pytest

from oracle import oracle as O

pytestmark = pytest.mark.gpu

FS = 40e6
# (fmt, M, K, offset, display); display None = filterbank only
STREAMS = [
    ("ci16", 1024, 16, 0j, dict(nfft=4096, hop=2048, window="hann", detector="max", frames=9, cols=1024)),
    ("cu8", 256, 12, 0j, dict(nfft=1024, hop=512, window="hamming", detector="avg", frames=7, cols=0)),
    ("cf32", 64, 8, 0j, None),
    ("ci8", 2048, 12, 0.01 - 0.02j, dict(nfft=2048, hop=2048, window="bh", detector="min", frames=3, cols=256)),
    ("ci16q11", 128, 9, 0.005j, dict(nfft=4096, hop=1000, window="rect", detector="avg", frames=1, cols=512)),
    ("cf32", 512, 10, 0j, dict(nfft=1024, hop=1024, window="hann", detector="max", frames=30, cols=100)),
]
WIN = {"rect": 0, "hann": 1, "hamming": 2, "bh": 3}
DET = {"max": 0, "min": 1, "avg": 2}


def line_close(a_db, b_db, tol_db=3e-3, floor=1e-9):
    """within tol_db where the oracle is within 60 dB of its peak, and within
    floor (relative to the peak power) in power below that"""
    a, b = np.asarray(a_db, np.float64), np.asarray(b_db, np.float64)
    pk = b.max()
    strong = b >= pk - 60.0
    dp = np.abs(10 ** (a / 10) - 10 ** (b / 10))
    return np.all(np.abs(a - b)[strong] <= tol_db) and np.all(dp[~strong] <= floor * 10 ** (pk / 10))


def _inputs(nsamp=120_000, seed=5):
    from atkdsp_amd import synth
    rng = np.random.default_rng(seed)
    out = []
    for i, (fmt, M, K, off, disp) in enumerate(STREAMS):
        f = {"cu8": 0, "ci8": 1, "ci16": 2, "ci16q11": 3, "cf32": 4}[fmt]
        raw = synth.quantize(synth.wideband_iq(300 + i, nsamp, FS), f)
        cuts = [0]
        while cuts[-1] < nsamp:
            cuts.append(min(nsamp, cuts[-1] + int(rng.integers(1, 20_000))))
        out.append((f, raw, cuts))
    return out


def _serve(inputs, which):
    from atkdsp_amd.serve import Server
    srv = Server(hbm_budget=4 << 30, max_streams=8, max_batch=64, staging_bytes=64 << 20)
    sids = {}
    for i in which:
        fmt, M, K, off, disp = STREAMS[i]
        d = dict(disp, slots=4096) if disp else None
        sids[i] = srv.add_stream(fmt=fmt, fs=FS, nchan=M, oversample=2, taps_per_poly=K, offset=off,
                                 chunk_max=20_000, out_slots=2, display=d)
    pos = {i: 0 for i in which}
    active = list(which)
    while active:                       # interleave the streams' chunks
        for i in list(active):
            f, raw, cuts = inputs[i]
            c = pos[i]
            if c + 1 >= len(cuts):
                active.remove(i)
                continue
            bps = O.BPS[f]
            srv.submit(sids[i], raw[cuts[c] * bps:cuts[c + 1] * bps])
            pos[i] = c + 1
    srv.drain()
    lines = {}
    for i in which:
        if STREAMS[i][4] is None:
            continue
        nl = srv.lines(sids[i])
        lines[i] = np.array([srv.read_line(sids[i], L) for L in range(nl)])
    st = srv.stats()
    srv.close()
    return lines, st


def test_served_display_lines_match_the_oracle(gpu):
    inputs = _inputs()
    which = list(range(len(STREAMS)))
    lines, st = _serve(inputs, which)
    assert st["chunks"] == sum(len(x[2]) - 1 for x in inputs)
    for i in which:
        fmt, M, K, off, disp = STREAMS[i]
        if disp is None:
            assert i not in lines
            continue
        f, raw, cuts = inputs[i]
        x, _ = O.unpack_dc(raw, f, off)
        N, hop, F = disp["nfft"], disp["hop"], disp["frames"]
        cols = disp["cols"] or N
        frames = (x.size - N) // hop + 1
        assert lines[i].shape == (frames // F, cols), i
        assert frames // F >= 3, i
        win = O.window(WIN[disp["window"]], N)
        for L in range(frames // F):
            span = x[L * F * hop: L * F * hop + (F - 1) * hop + N]
            ref, fr = O.spectrum_reduce(span, N, hop, win, DET[disp["detector"]])
            assert fr == F
            assert line_close(lines[i][L], O.decimate_max(ref, cols)), f"stream {i} line {L}"


def test_served_display_alone_equals_mixed(gpu):
    inputs = _inputs(nsamp=60_000, seed=9)
    mix, _ = _serve(inputs, list(range(len(STREAMS))))
    for i in (0, 4):
        alone, _ = _serve(inputs, [i])
        assert alone[i].shape == mix[i].shape
        assert np.array_equal(alone[i].view(np.uint32), mix[i].view(np.uint32)), i


def test_served_display_arguments(gpu):
    from atkdsp_amd.serve import Server
    srv = Server(hbm_budget=1 << 30, max_streams=4, max_batch=8, staging_bytes=16 << 20)
    with pytest.raises(gpu.AtkDspError):     # display FFT lengths are 1024, 2048 and 4096
        srv.add_stream(fmt="ci16", nchan=256, taps_per_poly=8, chunk_max=4096,
                       display=dict(nfft=512, hop=256, frames=4))
    with pytest.raises(gpu.AtkDspError):     # hop beyond the frame
        srv.add_stream(fmt="ci16", nchan=256, taps_per_poly=8, chunk_max=4096,
                       display=dict(nfft=1024, hop=2048, frames=4))
    sid = srv.add_stream(fmt="ci16", nchan=256, taps_per_poly=8, chunk_max=4096,
                         display=dict(nfft=1024, hop=1024, frames=2, cols=256))
    srv.submit(sid, np.zeros(3000 * 4, np.uint8))
    srv.drain()
    assert srv.lines(sid) == 1                # frames 0, 1 complete (2 x 1024 <= 3000)
    line = srv.read_line(sid, 0)
    assert line.shape == (256,) and np.allclose(line, -240.0, atol=1e-3)   # silence: 20 log10(1e-12)
    with pytest.raises(gpu.AtkDspError):     # not completed by what was submitted
        srv.read_line(sid, 1)
    srv.close()
