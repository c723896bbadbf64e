"""The drop-in calls' copies of caller memory: large ones go through pinned
staging in 4 MiB pieces (a copy team on the host, DMA per piece), unpack_dc
and the NCO mix pipelined in pieces of 2^19 samples over two streams, small
ones straight through the runtime. Results must not depend on the route: sizes
either side of each threshold, odd lengths, in-place mixing, and several
threads driving the engines at once all give the single-threaded result bit
for bit."""
import threading

import numpy as np
import pytest

from oracle import oracle as O

pytestmark = pytest.mark.gpu

PASS = 1 << 19                    # the pipelined piece, one pass of the mean
STAGE_MIN = 1 << 20               # bytes below which copies are not staged
PIECE = 4 << 20                   # bytes a staging piece holds


def _raw(code, n, seed):
    rng = np.random.default_rng(seed)
    if code == 4:
        return (rng.standard_normal(2 * n) * 0.4).astype(np.float32).view(np.uint8)
    bps = {0: 2, 1: 2, 2: 4, 3: 4}[code]
    return rng.integers(0, 256, n * bps, dtype=np.uint8)


SIZES = sorted({STAGE_MIN // 8 - 1, STAGE_MIN // 8, STAGE_MIN // 8 + 1, PASS - 1, PASS, PASS + 1,
                PIECE // 8 + 3, 3 * PASS + 17, (1 << 22) + 5})


@pytest.mark.parametrize("fmt,code", [("cu8", 0), ("ci16q11", 3), ("cf32", 4)])
def test_unpack_dc_across_the_staging_thresholds(gpu, fmt, code):
    off = 0.0031 - 0.0047j
    for i, n in enumerate(SIZES):
        raw = _raw(code, n, 100 * code + i)
        y, mean = gpu.unpack_dc(raw, fmt, off)
        ref, rmean = O.unpack_dc(raw, code, off)
        assert y.size == n
        assert np.array_equal(y.view(np.uint32), ref.view(np.uint32)), (fmt, n)
        assert abs(mean - rmean) < 1e-12 * max(1.0, abs(rmean)), (fmt, n)


def test_nco_mix_in_place_equals_out_of_place(gpu):
    n = 3 * PASS + 12_345                # pipelined pieces and a ragged tail
    x = (np.random.default_rng(3).standard_normal(n) + 1j * np.random.default_rng(4).standard_normal(n))
    x = x.astype(np.complex64)
    a = gpu.Nco(1.234e6, 40e6)
    y = a.mix(x)
    b = gpu.Nco(1.234e6, 40e6)
    z = x.copy()
    b.mix(z, out=z)                      # in place: a piece is written only after it was read
    assert np.array_equal(y.view(np.uint32), z.view(np.uint32))
    assert a.phase == b.phase
    ref = O.Nco(1.234e6, 40e6).mix(x)
    assert float(np.max(np.abs(y - ref))) < 2e-6 * float(np.max(np.abs(x)))


def test_threads_share_the_staging_engines(gpu):
    """Four threads, each its own filterbank handle and FFT plan, driving
    unpack_dc, spectrum_reduce and pfb_process at the wideband shape at once:
    every result equals the one computed alone."""
    from atkdsp_amd import synth
    n = 1_000_000
    raws = [synth.quantize(synth.wideband_iq(40 + t, n, 40e6), O.CI16Q11) for t in range(4)]
    win = gpu.window("hann", 4096)

    def work(raw):
        x, mean = gpu.unpack_dc(raw, "ci16q11", 0.001 + 0.002j)
        line, _ = gpu.Fft(4096).spectrum_reduce(x, 2048, win, "max")
        y = gpu.Pfb(40e6, 1024, 2, 16, 80.0).process(x)
        return x, mean, line, y

    alone = [work(r) for r in raws]
    got = [None] * 4
    errors = []

    def run(t):
        try:
            for _ in range(2):
                got[t] = work(raws[t])
        except Exception as e:          # noqa: BLE001 (reported below)
            errors.append(e)

    threads = [threading.Thread(target=run, args=(t,)) for t in range(4)]
    for th in threads:
        th.start()
    for th in threads:
        th.join()
    assert not errors, errors
    for t in range(4):
        for a, b in zip(alone[t], got[t]):
            if isinstance(a, complex):
                assert a == b
            else:
                assert np.array_equal(np.asarray(a).view(np.uint32), np.asarray(b).view(np.uint32)), t


def test_filterbank_rows_longer_than_a_piece(gpu):
    """Four channels, hop 4: one slab's output row holds over a million hops
    (8 MiB, more than a staging piece), so the copy back cuts each row into
    pieces and lands them in the caller's rows of the whole call. It equals
    the same input run in blocks small enough to bypass staging."""
    n = 5_000_003
    rng = np.random.default_rng(9)
    x = (rng.standard_normal(n) + 1j * rng.standard_normal(n)).astype(np.complex64) * np.float32(0.3)
    whole = gpu.Pfb(1e6, 4, 1, 4, 60.0).process(x)
    p = gpu.Pfb(1e6, 4, 1, 4, 60.0)
    blk = 100_003                            # 0.8 MB in, 0.8 MB out: not staged
    y = np.concatenate([p.process(x[a:a + blk]) for a in range(0, n, blk)], axis=1)
    assert whole.shape == y.shape == (4, n // 4)
    assert np.array_equal(whole.view(np.uint32), y.view(np.uint32))


@pytest.mark.parametrize("nchan,os_,K,cal", [(1024, 2, 16, False), (1024, 2, 16, True), (64, 1, 8, False),
                                             (2048, 2, 12, False)])
def test_filterbank_output_in_ranges(gpu, nchan, os_, K, cal):
    """A one-element call with more hops than one staging piece of output holds
    runs in ranges of hops, each range's output coming back while the next
    range's input goes in. It equals the same samples in blocks too small for
    ranges bit for bit, with a leftover carried into the call, a cal table,
    and a call of one range and one hop more."""
    hop = nchan // os_
    rows = (4 << 20) // (nchan * 8) // 16 * 16          # hops a range holds
    rng = np.random.default_rng(nchan + K)
    n1 = 2_000_003
    n2 = rows * hop + hop                               # one range and one hop more, with the leftover
    x = (rng.standard_normal(n1 + n2 + 777) + 1j * rng.standard_normal(n1 + n2 + 777)).astype(np.complex64)
    x *= np.float32(0.2)
    tab = (np.exp(1j * rng.uniform(0, 2 * np.pi, nchan)) * rng.uniform(0.5, 2.0, nchan)).astype(np.complex64)

    def make():
        p = gpu.Pfb(40e6, nchan, os_, K, 80.0)
        if cal:
            p.set_cal(tab, 1)
        return p

    p = make()
    parts = [p.process(x[:777]), p.process(x[777:777 + n1]), p.process(x[777 + n1:])]
    assert parts[1].shape[1] > rows and parts[2].shape[1] > rows
    whole = np.concatenate([a for a in parts if a.shape[1]], axis=1)
    q = make()
    blk = 20 * hop + 13                                 # a few dozen hops a call: one range, no pipeline
    y = np.concatenate([b for b in (q.process(x[a:a + blk]) for a in range(0, x.size, blk)) if b.shape[1]], axis=1)
    assert whole.shape == y.shape
    assert np.array_equal(whole.view(np.uint32), y.view(np.uint32))
