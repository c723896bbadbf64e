"""The overlap-save extractor (reference src/channelizer.c:473-788) and the
arbitrary resampler (src/resample.c:100-252) on the GPU, through the C ABI.

- The resampler is bit-exact with the reference (its float operation order is
  restated exactly), chunked as the golden vectors were made.
- The extractor matches the reference within 1e-5 relative (its forward and
  inverse FFTs run in FP32 where the reference's run in FP64), with the
  reference's output counts, in the reference's own chunk pattern, for
  resampled and plain channels, even and odd bins and fractional offsets, and
  for three phase-coherent elements.
- The reference's own properties (tests/test_channelizer.py:203-357): unit
  gain in band, > 80 dB stopband, exact chunk invariance, add/remove/retune
  leaving other channels bit-identical, inter-element phase, landing on the
  centre and phase continuity at hard offsets."""
import numpy as np
import pytest

from conftest import golden
from oracle import oracle as O

pytestmark = pytest.mark.gpu

FS = 2.4e6
OFF, BW, RATE = 300e3, 50e3, 50e3


def _tone(f, n, seed=0, noise=0.0):
    k = np.arange(n)
    x = np.exp(2j * np.pi * f / FS * k)
    if noise:
        rng = np.random.default_rng(seed)
        x = x + noise * (rng.standard_normal(n) + 1j * rng.standard_normal(n))
    return x.astype(np.complex64)


def _peak_hz(y, rate):
    y = np.asarray(y)[200:-50]
    n = 1 << int(np.floor(np.log2(len(y))))
    Y = np.abs(np.fft.fft(y[:n] * np.blackman(n)))
    f = np.fft.fftfreq(n, 1.0 / rate)
    return f[np.argmax(Y)], rate / n


def test_resampler_is_bit_exact_with_the_reference(gpu):
    g = golden("osx")
    ys = np.split(g["rs_y"], np.cumsum(g["rs_len"])[:-1])
    x = g["rs_x"]
    for (ri, ro, at, ph), y in zip(g["rs_args"], ys):
        r = gpu.ArbResampler(ri, ro, at, int(ph))
        o = np.concatenate([r.process(x[:1000]), r.process(x[1000:1001]), r.process(x[1001:])])
        assert o.size == y.size, (ri, ro)
        assert np.array_equal(o.view(np.uint32), y.view(np.uint32)), (ri, ro, O.rel_err(o, y))


def test_resampler_set_ratio_reset_and_capacity(gpu):
    rng = np.random.default_rng(3)
    x = (rng.standard_normal(40000) + 1j * rng.standard_normal(40000)).astype(np.complex64)
    r = gpu.ArbResampler(75e3, 50e3, 80.0)
    t = O.ArbResampler(75e3, 50e3, 80.0)
    a = r.process(x[:15000])
    b = t.process(x[:15000])
    r.set_ratio(0.6671)
    t.set_ratio(0.6671)
    assert abs(r.ratio() - 0.6671) < 1e-12
    a2 = r.process(x[15000:])
    b2 = t.process(x[15000:])
    assert a.size == b.size and a2.size == b2.size
    assert O.rel_err(np.concatenate([a, a2]), np.concatenate([b, b2])) < 1e-6
    # reset clears the history and the position and keeps the ratio (src/resample.c:177-183)
    r.reset()
    again = r.process(x[:15000])
    fresh = gpu.ArbResampler(75e3, 50e3, 80.0)
    fresh.set_ratio(0.6671)
    assert np.array_equal(again, fresh.process(x[:15000]))
    r0 = gpu.ArbResampler(75e3, 50e3, 80.0)
    with pytest.raises(gpu.AtkDspError):                        # outputs beyond out_cap
        r0.process(x[:5000], out=np.empty(10, np.complex64))
    # the refused call left the state as it was: r0 still equals a fresh one
    r1 = gpu.ArbResampler(75e3, 50e3, 80.0)
    assert np.array_equal(r0.process(x[:5000]), r1.process(x[:5000]))


def _osx_gold():
    g = golden("osx")
    return g, np.split(g["osx_y"], np.cumsum(g["osx_len"])[:-1])


def test_osx_matches_the_reference(gpu):
    g, ys = _osx_gold()
    x = g["osx_x"]
    o = gpu.Osx(FS, 4096)
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
        assert O.rel_err(got, ys[i]) < 1e-5, (i, O.rel_err(got, ys[i]))


def test_osx_elements_match_the_reference(gpu):
    g, _ = _osx_gold()
    x = g["osx_x"]
    phi = np.deg2rad(31.0)
    els = [(x * np.exp(1j * e * phi)).astype(np.complex64) for e in range(3)]
    o = gpu.Osx(FS, 4096, 80.0, 3)
    c0 = o.add(*g["osx_chans"][0])
    c1 = o.add(*g["osx_chans"][3])
    y, _ = o.process_n(els, 3)
    assert y[c0].shape == g["osx_el_y0"].shape and y[c1].shape == g["osx_el_y1"].shape
    assert O.rel_err(y[c0], g["osx_el_y0"]) < 1e-5
    assert O.rel_err(y[c1], g["osx_el_y1"]) < 1e-5


def test_osx_in_band_unit_gain_stopband_and_rate(gpu):
    n = 4096 * 50
    o = gpu.Osx(FS, 4096)
    ci = o.add(OFF, BW, RATE)
    y, cnt = o.process(_tone(OFF + 7e3, n))
    assert abs(np.abs(y[ci][300:800]).mean() - 1.0) < 1e-3
    assert abs(cnt[ci] - n * RATE / FS) < 4 * (4096 / (FS / RATE))
    o2 = gpu.Osx(FS, 4096)
    c2 = o2.add(OFF, BW, RATE)
    y2, _ = o2.process(_tone(OFF + 150e3, n))
    assert 20 * np.log10(np.abs(y2[c2][300:800]).mean() + 1e-20) < -80.0
    assert "4096-pt FFT" in o.describe()


def test_osx_chunk_invariance_is_exact(gpu):
    x = _tone(OFF + 7e3, 4096 * 50, noise=0.01, seed=12)
    w = gpu.Osx(FS, 4096)
    cw = w.add(OFF, BW, RATE)
    whole = w.process(x)[0][cw]
    p = gpu.Osx(FS, 4096)
    pc = p.add(OFF, BW, RATE)
    outs = []
    for seg in (x[:7000], x[7000:7001], x[7001:50000], x[50000:]):
        r, _ = p.process(seg)
        if r[pc] is not None and r[pc].size:
            outs.append(r[pc])
    split = np.concatenate(outs)
    assert split.size == whole.size
    assert np.array_equal(whole, split)


def test_osx_add_remove_retune_leave_others_bit_identical(gpu):
    x = _tone(OFF + 7e3, 4096 * 40, noise=0.01, seed=13)
    ref = gpu.Osx(FS, 4096)
    r0 = ref.add(OFF, BW, RATE)
    y_ref = ref.process(x)[0][r0]
    o = gpu.Osx(FS, 4096)
    c0 = o.add(OFF, BW, RATE)
    c1 = o.add(OFF + 200e3, BW, RATE)
    y_pre = o.process(x[: x.size // 2])[0][c0]
    o.retune(c1, OFF - 250e3)
    y_mid = o.process(x[x.size // 2: 3 * x.size // 4])[0][c0]
    o.remove(c1)
    assert o.chan_rate(c1) == 0.0 and o.channels() == 2
    y_post = o.process(x[3 * x.size // 4:])
    assert y_post[0][c1] is None
    y = np.concatenate([y_pre, y_mid, y_post[0][c0]])
    assert np.array_equal(y_ref, y)
    c2 = o.add(OFF + 100e3, BW, RATE)                            # the freed slot is reused
    assert c2 == c1


def test_osx_multi_element_phase_coherent(gpu):
    phi = np.deg2rad(31.0)
    base = _tone(OFF + 7e3, 4096 * 60)
    o = gpu.Osx(FS, 4096, max_nel=5)
    ci = o.add(OFF, BW, RATE)
    outs, _ = o.process_n([(base * np.exp(1j * e * phi)).astype(np.complex64) for e in range(5)], 5)
    y = outs[ci]
    ph = np.angle(y[:, 200:-50].mean(1))
    ph = ph - ph[0]
    err = np.rad2deg(np.max(np.abs(np.angle(np.exp(1j * (ph - np.arange(5) * phi))))))
    assert err < 0.1, err
    o2 = gpu.Osx(FS, 4096, max_nel=5)
    o2.add(OFF, BW, RATE)
    yy = o2.process_n([base] * 5, 5)[0][0]
    assert all(np.array_equal(yy[0], yy[e]) for e in range(5))


@pytest.mark.parametrize("off", [300e3, 513 * FS / 4096, 300200.0, -411737.3])
def test_osx_lands_on_centre_and_is_phase_continuous(gpu, off):
    x = _tone(off + 7e3, 4096 * 50)
    o = gpu.Osx(FS, 4096)
    c = o.add(off, BW, RATE)
    y = o.process(x)[0][c]
    f, res = _peak_hz(y, RATE)
    assert abs(f - 7e3) <= res, (off, f)
    k = np.arange(len(y))
    d = np.unwrap(np.angle(y * np.exp(-2j * np.pi * 7e3 / RATE * k)))[300:-50]
    assert np.std(d) < 1e-3, off


def test_osx_retune_to_an_odd_bin_stays_on_centre(gpu):
    o = gpu.Osx(FS, 4096)
    c = o.add(300e3, BW, RATE)
    o.process(_tone(300e3 + 7e3, 4096 * 10))
    off = 513 * FS / 4096 + 123.0
    o.retune(c, off)
    y = o.process(_tone(off + 7e3, 4096 * 50))[0][c]
    f, res = _peak_hz(y[400:], RATE)
    assert abs(f - 7e3) <= res


def test_osx_arguments_and_long_input(gpu):
    with pytest.raises(gpu.AtkDspError):
        gpu.Osx(FS, 1000)                                        # not a power of two
    o = gpu.Osx(FS, 1024)
    with pytest.raises(gpu.AtkDspError):
        o.add(0.0, 0.0, RATE)                                    # no bandwidth
    c = o.add(-200e3, 100e3, 150e3)                              # channel rate 150 kHz: no resampler
    # an input longer than one device slab equals the same input in pieces
    x = _tone(-200e3 + 11e3, 700_001, noise=0.01, seed=4)
    y = o.process(x)[0][c]
    p = gpu.Osx(FS, 1024)
    pc = p.add(-200e3, 100e3, 150e3)
    y2 = np.concatenate([p.process(x[a:a + 99_999])[0][pc] for a in range(0, x.size, 99_999)])
    assert np.array_equal(y, y2)
    t = O.Osx(FS, 1024)
    tc = t.add(-200e3, 100e3, 150e3)
    yt = t.process(x)[0][tc]
    assert y.size == yt.size and O.rel_err(y, yt) < 1e-5
