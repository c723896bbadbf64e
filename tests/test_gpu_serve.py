"""The serving loop: independent streams of mixed formats and filterbank
shapes, chunks of uneven size, continuous batching. Each stream's output
equals the drop-in filterbank fed the same samples (bit-identical: same
kernel, same per-hop arithmetic), equals its own output when served alone,
and admission refuses streams beyond the HBM budget."""
import numpy as np
import pytest

from oracle import oracle as O

pytestmark = pytest.mark.gpu

FS = 40e6
SHAPES = [("ci16", 1024, 16, 0j), ("cu8", 256, 12, 0j), ("ci8", 64, 8, 0.01 - 0.02j), ("cf32", 1024, 10, 0j),
          ("ci16q11", 256, 16, 0.005j), ("cu8", 64, 14, 0j), ("ci16", 2048, 12, 0j), ("cf32", 128, 9, 0j)]


def _streams(seed=11, nsamp=150_000):
    from atkdsp_amd import synth
    rng = np.random.default_rng(seed)
    out = []
    for i, (fmt, M, K, off) in enumerate(SHAPES):
        f = synth.FMT[fmt] if hasattr(synth, "FMT") else {"cu8": 0, "ci8": 1, "ci16": 2, "ci16q11": 3, "cf32": 4}[fmt]
        raw = synth.quantize(synth.wideband_iq(50 + i, nsamp, FS), f)
        cuts = [0]
        while cuts[-1] < nsamp:
            cuts.append(min(nsamp, cuts[-1] + int(rng.integers(1, 20_000))))
        out.append((fmt, f, M, K, off, raw, cuts))
    return out


def _serve(gpu, streams, which):
    from atkdsp_amd.serve import Server
    srv = Server(hbm_budget=4 << 30, max_streams=16, max_batch=64, staging_bytes=64 << 20)
    sids = {i: srv.add_stream(fmt=streams[i][0], fs=FS, nchan=streams[i][2], oversample=2,
                              taps_per_poly=streams[i][3], offset=streams[i][4], chunk_max=20_000,
                              out_slots=64) for i in which}
    tickets = {i: [] for i in which}
    # interleave the streams' chunks as they would arrive
    pos = {i: 0 for i in which}
    active = list(which)
    while active:
        for i in list(active):
            fmt, f, M, K, off, raw, cuts = streams[i]
            bps = O.BPS[f]
            c = pos[i]
            if c + 1 >= len(cuts):
                active.remove(i)
                continue
            a, b = cuts[c], cuts[c + 1]
            tickets[i].append(srv.submit(sids[i], raw[a * bps:b * bps]))
            pos[i] = c + 1
            if len(tickets[i]) % 32 == 0:       # read before the output ring wraps
                pass
    outs = {}
    for i in which:
        parts = [srv.read(sids[i], t) for t in tickets[i]]
        outs[i] = np.concatenate([p for p in parts if p.shape[1]], axis=1)
    st = srv.stats()
    srv.close()
    return outs, st


def test_serving_matches_dropin_and_alone(gpu):
    streams = _streams()
    which = list(range(len(SHAPES)))
    mix, st = _serve(gpu, streams, which)
    assert st["chunks"] == sum(len(s[6]) - 1 for s in streams)
    assert st["p50_ms"] > 0 and st["p99_ms"] >= st["p50_ms"] and st["max_ms"] >= st["p999_ms"] >= st["p99_ms"]
    assert st["lat_count"] == st["chunks"]                  # every chunk is in the latency figures
    for i in which:
        fmt, f, M, K, off, raw, cuts = streams[i]
        x, _ = O.unpack_dc(raw, f, off)
        ref = gpu.Pfb(FS, M, 2, K, 80.0).process(x)
        assert mix[i].shape == ref.shape, i
        assert np.array_equal(mix[i], ref), f"stream {i} ({fmt}, M={M}, K={K}) differs from the drop-in"
    alone, _ = _serve(gpu, streams, [0])
    assert np.array_equal(alone[0], mix[0])
    alone, _ = _serve(gpu, streams, [6])
    assert np.array_equal(alone[6], mix[6])
    # and the FP64 oracle: within the filterbank's 5e-7 of full scale
    x, _ = O.unpack_dc(streams[1][5], streams[1][1], streams[1][4])
    yo = O.Pfb(FS, 256, 2, 12, 80.0).process(x)
    assert O.rel_err(mix[1], yo) < 5e-7


def test_admission_against_the_hbm_budget(gpu):
    from atkdsp_amd.serve import Server
    # one 1024-channel, 16-tap stream with 64K-sample chunks needs about 3.5 MiB
    # (two history rows, taps, two output slots): the budget holds one, not two
    srv = Server(hbm_budget=6 << 20, max_streams=8, max_batch=8, staging_bytes=16 << 20)
    a = srv.try_add_stream(fmt="ci16", nchan=1024, taps_per_poly=16, chunk_max=65536, out_slots=2)
    assert a is not None
    used = srv.stats()["hbm_used"]
    assert 0 < used <= 6 << 20
    refused = srv.try_add_stream(fmt="ci16", nchan=1024, taps_per_poly=16, chunk_max=65536, out_slots=2)
    assert refused is None                                   # over the budget
    srv.remove_stream(a)
    assert srv.stats()["hbm_used"] == 0
    b = srv.try_add_stream(fmt="ci16", nchan=1024, taps_per_poly=16, chunk_max=65536, out_slots=2)
    assert b is not None                                     # memory back, admitted
    with pytest.raises(gpu.AtkDspError):
        srv.add_stream(fmt="ci16", nchan=1000)               # not a supported shape
    with pytest.raises(gpu.AtkDspError):
        srv.submit(b, np.zeros(65537 * 4, np.uint8))         # chunk over chunk_max
    srv.close()


def test_coherent_array_stream_in_the_mix(gpu):
    """An array stream (4 phase-coherent cf32 elements per chunk) served
    beside single streams: its outputs equal the drop-in filterbank over the
    same elements, and its W1 / W2 equal the drop-in calls on those outputs,
    chunk by chunk, to the bit."""
    from atkdsp_amd import synth
    from atkdsp_amd.serve import Server
    rng = np.random.default_rng(21)
    nel, M, K, n = 4, 256, 8, 60_000
    els = synth.c5_elements(n=n, nel=nel, fs=20e6)
    singles = _streams(nsamp=60_000)[:2]
    srv = Server(hbm_budget=4 << 30, max_streams=8, max_batch=64, staging_bytes=64 << 20)
    aid = srv.add_stream(fmt="cf32", fs=20e6, nchan=M, oversample=2, taps_per_poly=K, chunk_max=20_000,
                         out_slots=64, nel=nel, coherent=True)
    sids = [srv.add_stream(fmt=s[0], fs=FS, nchan=s[2], oversample=2, taps_per_poly=s[3], offset=s[4],
                           chunk_max=20_000, out_slots=64) for s in singles]
    cuts = [0]
    while cuts[-1] < n:
        cuts.append(min(n, cuts[-1] + int(rng.integers(1, 20_000))))
    tickets = []
    for c in range(len(cuts) - 1):
        a, b = cuts[c], cuts[c + 1]
        chunk = np.concatenate([e[a:b] for e in els])              # element-major
        tickets.append(srv.submit(aid, chunk))
        for i, sid in enumerate(sids):                              # interleave other traffic
            raw, sc = singles[i][5], singles[i][6]
            if c + 1 < len(sc):
                bps = O.BPS[singles[i][1]]
                srv.submit(sid, raw[sc[c] * bps:sc[c + 1] * bps])
    parts = [srv.read(aid, t) for t in tickets]
    y = np.concatenate([p for p in parts if p.shape[2]], axis=2)
    ref = gpu.Pfb(20e6, M, 2, K, 80.0, max_nel=nel).process_n(els, nel)
    assert y.shape == ref.shape and np.array_equal(y, ref)
    R_ref = np.zeros((M, nel, nel), np.complex64)
    c_ref = np.zeros(M, np.uint32)
    for p in parts:
        if p.shape[2]:
            gpu.pfb_cov_accumulate(p, nel, M, None, R_ref, c_ref)
    R, cnt = srv.read_cov(aid)
    assert np.array_equal(cnt, c_ref) and np.array_equal(R, R_ref)
    last = [p for p in parts if p.shape[2]][-1]
    pw, mp = srv.read_power(aid)
    pw_ref, mp_ref = gpu.pfb_channel_power(last, nel, M)
    assert np.array_equal(pw, pw_ref) and np.array_equal(mp, mp_ref)
    with pytest.raises(gpu.AtkDspError):
        srv.read_cov(sids[0])                                       # not an array
    srv.close()


def test_array_stream_with_odd_chunk_max(gpu):
    """An array stream whose chunk_max is odd (its element rows then need
    padding to stay aligned) served beside a single stream: both equal the
    drop-in filterbank, to the bit."""
    from atkdsp_amd import synth
    from atkdsp_amd.serve import Server
    rng = np.random.default_rng(5)
    nel, M, K, n, cmax = 3, 256, 8, 70_003, 20_001
    els = synth.c5_elements(n=n, nel=nel, fs=20e6)
    single = _streams(nsamp=70_003)[1]
    srv = Server(hbm_budget=4 << 30, max_streams=4, max_batch=16, staging_bytes=64 << 20)
    aid = srv.add_stream(fmt="cf32", fs=20e6, nchan=M, oversample=2, taps_per_poly=K, chunk_max=cmax,
                         out_slots=64, nel=nel)
    sid = srv.add_stream(fmt=single[0], fs=FS, nchan=single[2], oversample=2, taps_per_poly=single[3],
                         offset=single[4], chunk_max=cmax, out_slots=64)
    cuts = [0]
    while cuts[-1] < n:
        cuts.append(min(n, cuts[-1] + int(rng.integers(1, cmax + 1))))
    ta, ts = [], []
    bps = O.BPS[single[1]]
    for c in range(len(cuts) - 1):
        a, b = cuts[c], cuts[c + 1]
        ta.append(srv.submit(aid, np.concatenate([e[a:b] for e in els])))
        ts.append(srv.submit(sid, single[5][a * bps:b * bps]))
    ya = np.concatenate([p for p in (srv.read(aid, t) for t in ta) if p.shape[2]], axis=2)
    ys = np.concatenate([p for p in (srv.read(sid, t) for t in ts) if p.shape[1]], axis=1)
    srv.close()
    assert np.array_equal(ya, gpu.Pfb(20e6, M, 2, K, 80.0, max_nel=nel).process_n(els, nel))
    x, _ = O.unpack_dc(single[5], single[1], single[4])
    assert np.array_equal(ys, gpu.Pfb(FS, single[2], 2, single[3], 80.0).process(x))


def test_a_read_is_exact_or_refused(gpu):
    """With two output slots and later chunks already submitted, reading an
    older chunk either returns exactly the drop-in's output for it or is
    refused (its slot gone before the call, or reused during the copy):
    never a torn copy."""
    from atkdsp_amd.serve import Server
    fmt, f, M, K, off, raw, cuts = _streams(seed=3, nsamp=200_000)[0]
    bps = O.BPS[f]
    srv = Server(hbm_budget=1 << 30, max_streams=2, max_batch=8, staging_bytes=64 << 20)
    sid = srv.add_stream(fmt=fmt, fs=FS, nchan=M, oversample=2, taps_per_poly=K, offset=off, chunk_max=20_000,
                         out_slots=2)
    pfb = gpu.Pfb(FS, M, 2, K, 80.0)
    exact = refused = 0
    prev = None
    for c in range(len(cuts) - 1):
        seg = raw[cuts[c] * bps:cuts[c + 1] * bps]
        t = srv.submit(sid, seg)
        x, _ = O.unpack_dc(seg, f, off)
        ref = pfb.process(x)
        if prev is not None:
            pt, pref = prev
            try:
                y = srv.read(sid, pt)
            except gpu.AtkDspError as e:
                assert "overrun" in str(e) or "bad argument" in str(e), e
                refused += 1
            else:
                assert np.array_equal(y, pref), c
                exact += 1
        prev = (t, ref)
    srv.close()
    assert exact + refused == len(cuts) - 2 and exact > 0
