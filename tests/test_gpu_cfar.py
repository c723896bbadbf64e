"""CFAR detection and the burst tracker on the GPU (atkdsp_cfar, reference
include/atkdsp.h:1733-1925, src/cfar.c:270-819) against the reference build's
events (tests/golden/cfar_events.npz, made by tests/golden/make_golden.py).

Eight scenarios: CA with the defaults; OS with four-frame averaging, a merge
gap and a minimum width; GO with masks changed mid-stream; SO at a high
false-alarm rate with few track slots (blobs dropped); CA over four-look frames
with the Hann window's bin correlation, three-frame averaging and a reset; CA
with NaN, -inf, +inf and very large dB values; OS on 64 bins with a mask that
leaves bins untestable; narrow 5 dB bursts under an 8-frame average, found
only on the average, so that t0 and t1 take their coarse estimates. Each replays the same calls (frames split over calls
of 1 to 50 frames) and compares every event: type, id, flags, span and hit
count exactly; t0 / t1 exactly (both are frame times); the centroid within
1e-9 and the dB figures within 2e-6 relative (the GPU forms power with its
own exp, within an ulp of the C library's). The counters, the last frame's
hits and NaN pattern, alpha and the description match exactly.

Beyond the fixtures: 65536 bins x 150 frames through slabs of the device
buffer give the oracle's events and counters, and the same events in one call
as in calls of 1, 7 and 64 frames; refused
configurations, E_CAP before any frame is taken, E_EMPTY before a tested
frame."""
import ctypes as C
import json

import numpy as np
import pytest

import atkdsp_amd as A
from conftest import golden, same_cfar_events
from oracle import oracle as O

pytestmark = pytest.mark.gpu

G = golden("cfar_events")
META = json.loads(str(G["meta"]))


def replay(m, frames):
    c = A.Cfar(m["nbins"], **m["cfg"])
    out = []
    for op in m["ops"]:
        if op[0] == "process":
            out.append(c.process(frames[op[1]:op[1] + op[2]], op[1] * 0.01, 0.01))
        elif op[0] == "mask":
            c.set_mask(None if op[1] is None else G[f"{m['name']}_{op[1]}"])
            out.append(np.zeros(0, A.CFAR_EVENT_DTYPE))
        elif op[0] == "reset":
            c.reset()
            out.append(np.zeros(0, A.CFAR_EVENT_DTYPE))
        else:
            out.append(c.flush())
    return c, out


@pytest.mark.parametrize("m", META, ids=[m["name"] for m in META])
def test_events_match_the_reference(gpu, m):
    name = m["name"]
    frames = G[f"{name}_frames"].astype(np.float32)
    c, out = replay(m, frames)
    ev, evop = G[f"{name}_events"], G[f"{name}_evop"]
    for i, got in enumerate(out):
        same_cfar_events(got, ev[evop == i], f"{name} op {i} {m['ops'][i]}")
    st = c.stats()
    assert list(st.values()) == [int(v) for v in G[f"{name}_stats"]], (st, G[f"{name}_stats"])
    a, t, h = c.last()
    la = G[f"{name}_last"]
    assert np.array_equal(h, la[2].astype(np.uint8))
    assert np.array_equal(np.isnan(t), np.isnan(la[1])) and np.array_equal(np.isnan(a), np.isnan(la[0]))
    fin = np.isfinite(la[0])
    np.testing.assert_allclose(a[fin], la[0][fin], rtol=1e-6, atol=1e-4)
    fin = np.isfinite(la[1])
    np.testing.assert_allclose(t[fin], la[1][fin], rtol=1e-6, atol=1e-4)
    info = c.info()
    ra, rca, ret = G[f"{name}_info"]
    assert abs(info["alpha"] / ra - 1) < 1e-8 and abs(info["alpha_ca"] / rca - 1) < 1e-8
    assert abs(info["et"] / ret - 1) < 5e-6
    # the summary line carries the handle's facts (its wording is this library's)
    d = c.describe()
    assert d.startswith("cfar[") and f"bins={m['nbins']}" in d and f"alpha={info['alpha']:.6g}" in d, d
    assert f"tracks={st['tracks']}/" in d, d


def _wide_frames(nbins=65536, nframes=150, seed=9):
    rng = np.random.default_rng(seed)
    p = rng.exponential(1.0, (nframes, nbins))
    for b0, w, f0, f1 in [(1000, 8, 10, 120), (30000, 4, 3, 40), (30000, 12, 40, 90), (30050, 3, 61, 149),
                          (65530, 6, 0, 150), (50000, 1, 70, 75)]:
        p[f0:f1, b0:b0 + w] += 40.0
    return (10.0 * np.log10(p) - 90.0).astype(np.float32)


def test_slabs_and_call_sizes_give_the_same_events(gpu):
    fr = _wide_frames()
    cfg = dict(mode="ca", k_frames=4, n_ref=16, n_guard=4, open_m=2, open_n=3, close_k=3)
    whole = A.Cfar(fr.shape[1], **cfg)
    ref = np.concatenate([whole.process(fr, 0.0, 0.5), whole.flush()])
    assert len(ref) >= 8 and set(ref["type"]) == {1, 2, 3}
    info = whole.info()
    o = O.Cfar(fr.shape[1], dict(cfg, mode=0), info["alpha"], info["alpha_ca"], info["et"])
    same_cfar_events(ref, np.concatenate([o.process(fr, 0.0, 0.5), o.flush()]), "65536 bins against the oracle")
    assert whole.stats() == o.stats()
    for step in (1, 7, 64):
        c = A.Cfar(fr.shape[1], **cfg)
        got = [c.process(fr[i:i + step], i * 0.5, 0.5) for i in range(0, fr.shape[0], step)]
        got = np.concatenate(got + [c.flush()])
        assert got.tobytes() == ref.tobytes(), step
        assert c.stats() == whole.stats()


def test_refusals(gpu):
    with pytest.raises(A.AtkDspError):
        A.Cfar(32)                                   # fewer bins than 2 n_ref + 2 n_guard + 1
    for bad in (dict(mode=4), dict(n_ref=0), dict(k_frames=65), dict(pfa=0.6), dict(open_m=4, open_n=3),
                dict(mode="os", os_k=33), dict(rho2=(1.0, 0, 0, 0)), dict(max_tracks=0)):
        with pytest.raises(A.AtkDspError):
            A.Cfar(1024, **bad)
    c = A.Cfar(1024)
    with pytest.raises(A.AtkDspError):
        c.last()                                     # E_EMPTY before a tested frame
    lib = A.load()
    fr = np.full((3, 1024), -100.0, np.float32)
    ev = np.zeros(3 * 256, A.CFAR_EVENT_DTYPE)
    rc = lib.atkdsp_cfar_process(c._p, fr.ctypes.data_as(C.POINTER(C.c_float)), 3, 0.0, 1.0, ev.ctypes.data,
                                 3 * 256 - 1)
    assert rc == -3 and c.stats()["frames"] == 0
    assert c.events_max(3) == 3 * 256
    assert len(c.process(fr)) == 0 and c.stats()["frames"] == 3
    a, t, h = c.last()
    assert not h.any() and np.allclose(a, -100.0, atol=1e-4)


@pytest.mark.parametrize("n_ref", [3, 5, 32, 33, 64])
def test_os_in_registers_and_in_scratch(gpu, n_ref):
    """OS sorts up to 64 cells in registers (6 and 10 cells padded to 8 and
    16, 64 exactly) and ranks 66 and 128 cells in per-lane scratch: each
    against the oracle over 1536 bins x 40 frames."""
    rng = np.random.default_rng(n_ref)
    p = rng.exponential(1.0, (40, 1536))
    p[5:30, 700:706] += 60.0
    p[10:40, 100:102] += 80.0
    p[12:14, 1500:1503] += 200.0
    fr = (10.0 * np.log10(p) - 80.0).astype(np.float32)
    cfg = dict(mode=1, n_ref=n_ref, n_guard=2, k_frames=2, os_k=min(n_ref + 3, 2 * n_ref))
    c = A.Cfar(1536, **cfg)
    got = np.concatenate([c.process(fr[:17], 0.0, 0.25), c.process(fr[17:], 17 * 0.25, 0.25), c.flush()])
    info = c.info()
    o = O.Cfar(1536, cfg, info["alpha"], info["alpha_ca"], info["et"])
    exp = np.concatenate([o.process(fr[:17], 0.0, 0.25), o.process(fr[17:], 17 * 0.25, 0.25), o.flush()])
    assert len(exp) >= 4
    same_cfar_events(got, exp, f"OS n_ref {n_ref}")
    assert c.stats() == o.stats()
    assert np.array_equal(c.last()[2], o.last[2].astype(np.uint8))


def test_server_cfar_on_the_device_lines(gpu):
    """CFAR over served display lines where they live (atkgpu_server_cfar):
    the same events and counters as the drop-in handle given the same lines
    read back to the host, over calls of 1 to 8 lines; a 2-element array
    stream by element; refusals for lines not yet submitted, lines that have
    left the ring, a column count the handle was not made for, an element the
    stream lacks, and more lines than the ring keeps."""
    from atkdsp_amd import synth
    from atkdsp_amd.serve import Server
    fs = 40e6
    srv = Server(hbm_budget=2 << 30, max_streams=4, max_batch=16, staging_bytes=64 << 20)
    disp = dict(nfft=4096, hop=2048, window="hann", detector="max", frames=9, cols=1024, slots=64)
    sid = srv.add_stream(fmt="ci16", fs=fs, nchan=1024, oversample=2, taps_per_poly=16, chunk_max=65536,
                         out_slots=2, display=disp)
    small = srv.add_stream(fmt="ci16", fs=fs, nchan=1024, oversample=2, taps_per_poly=16, chunk_max=65536,
                           out_slots=2, display=dict(disp, slots=4))
    aid = srv.add_stream(fmt="cf32", fs=fs, nchan=256, oversample=2, taps_per_poly=12, chunk_max=32768,
                         out_slots=2, nel=2, display=dict(disp, cols=512))
    n = 600_000
    z = synth.wideband_iq(21, n, fs, ntones=8, nbursts=0)
    k = np.arange(n)
    for f, a, b in ((5.3e6, 90_000, 300_000), (-11.7e6, 200_000, 520_000), (1.1e6, 400_000, 430_000)):
        z[a:b] += 0.05 * np.exp(2j * np.pi * f / fs * k[a:b])
    raw = synth.quantize(z, O.CI16)
    for a in range(0, n, 50_000):
        srv.submit(sid, raw[a * 4:(a + 50_000) * 4])
        srv.submit(small, raw[a * 4:(a + 50_000) * 4])
    arr = np.concatenate([z, z * np.exp(0.4j)]).astype(np.complex64)
    for a in range(0, n, 25_000):
        srv.submit(aid, np.concatenate([arr[a:a + 25_000], arr[n + a:n + a + 25_000]]))
    srv.drain()
    L = srv.lines(sid)
    assert L >= 25 and srv.lines(small) == L
    cfg = dict(mode="ca", n_ref=8, n_guard=2, k_frames=2, pfa=1e-4, open_m=2, open_n=3, close_k=2)
    host = A.Cfar(1024, **cfg)
    frames = np.stack([srv.read_line(sid, i) for i in range(L)])
    dt = 0.0625                        # exact: calls' start times add up exactly
    exp = np.concatenate([host.process(frames, 0.0, dt), host.flush()])
    assert set(exp["type"]) == {1, 3} or set(exp["type"]) == {1, 2, 3}
    dev = A.Cfar(1024, **cfg)
    got, i, step = [], 0, 1
    while i < L:
        m = min(step, L - i)
        got.append(srv.cfar(sid, dev, i, m, i * dt, dt))
        i += m
        step = step % 8 + 1
    got = np.concatenate(got + [dev.flush()])
    assert got.tobytes() == exp.tobytes()
    assert dev.stats() == host.stats()
    # the array stream, element by element
    for e in range(2):
        h = A.Cfar(512, **cfg)
        fr = np.stack([srv.read_line(aid, i)[e] for i in range(srv.lines(aid))])
        want = h.process(fr, 0.0, dt)
        d = A.Cfar(512, **cfg)
        assert srv.cfar(aid, d, 0, fr.shape[0], 0.0, dt, element=e).tobytes() == want.tobytes()
    with pytest.raises(A.AtkDspError):
        srv.cfar(sid, A.Cfar(1024, **cfg), L + 50, 1)      # not completed by the chunks submitted
    with pytest.raises(A.AtkDspError):
        srv.cfar(small, A.Cfar(1024, **cfg), 0, 1)         # left the 4-line ring
    ok = srv.cfar(small, A.Cfar(1024, **cfg), L - 4, 4)    # the 4 lines it keeps
    assert ok.dtype == A.CFAR_EVENT_DTYPE
    with pytest.raises(A.AtkDspError):
        srv.cfar(small, A.Cfar(1024, **cfg), L - 5, 5)     # more lines than the ring keeps
    with pytest.raises(A.AtkDspError):
        srv.cfar(sid, A.Cfar(512, **cfg), 0, 1)            # columns differ from the handle's bins
    with pytest.raises(A.AtkDspError):
        srv.cfar(aid, A.Cfar(512, **cfg), 0, 1, element=2)  # no such element
    srv.close()
