"""Streaming continuity at the wideband shape (40 MSPS, 1024 channels, 16
taps/branch, 2x oversampled): blocks of 1, 7, 999 and 4096 samples give the
one-call result to the bit, through the drop-in filterbank and through the
device-resident chain. At 2x the rotation is (-1)^k on alternate hops, so a
block boundary that lands on an odd hop checks the carried hop parity
(reference tests/test_channelizer.py:92-104, :367-394)."""
import numpy as np
import pytest

from oracle import oracle as O

pytestmark = pytest.mark.gpu

FS, M, OS, K = 40e6, 1024, 2, 16


def _samples(n):
    from atkdsp_amd import synth
    return synth.quantize(synth.wideband_iq(77, n, FS), O.CI16Q11)


@pytest.mark.parametrize("blk", [1, 7, 999, 4096])
def test_dropin_filterbank_blocks_equal_one_call(gpu, blk):
    n = 20_000 if blk < 100 else 150_000
    x, _ = O.unpack_dc(_samples(n), O.CI16Q11, 0j)
    whole = gpu.Pfb(FS, M, OS, K, 80.0).process(x)
    p = gpu.Pfb(FS, M, OS, K, 80.0)
    parts = [p.process(x[a:a + blk]) for a in range(0, n, blk)]
    y = np.concatenate([q for q in parts if q.shape[1]], axis=1)
    assert y.shape == whole.shape and whole.shape[1] > 2
    assert np.array_equal(y.view(np.uint32), whole.view(np.uint32))
    # block boundaries landed on odd hops (4096 samples are 8 whole hops:
    # every boundary of that case falls on an even hop)
    if blk % (2 * (M // OS)):
        assert any(q.shape[1] % 2 for q in parts if q.shape[1])


@pytest.mark.parametrize("blk", [1, 7, 999, 4096])
def test_chain_blocks_equal_one_call(gpu, blk):
    """The chain with a fixed offset (dc_mode 0: a per-block re-estimate is
    not chunk-invariant, by design): every block's filterbank output,
    concatenated, equals one call over the whole input."""
    from atkdsp_amd.chain import Chain
    n = 12_000 if blk < 100 else 150_000
    raw = _samples(n)

    def chain(bmax):
        return Chain(fmt="ci16q11", fs=FS, nchan=M, oversample=OS, taps_per_poly=K, nfft=4096, fft_hop=2048,
                     line_cols=1024, trace_bins=0, dc_mode=0, streams=2, block_max=bmax)

    one = chain(n)
    for s in range(2):
        one.upload(s, raw)
    one.run(n)
    whole = [one.read_pfb(s) for s in range(2)]
    ch = chain(blk)
    parts = [[], []]
    for a in range(0, n, blk):
        b = min(n, a + blk)
        for s in range(2):
            ch.upload(s, raw[4 * a:4 * b])
        if ch.run(b - a):
            for s in range(2):
                parts[s].append(ch.read_pfb(s))
    for s in range(2):
        y = np.concatenate(parts[s], axis=1)
        assert y.shape == whole[s].shape
        assert np.array_equal(y.view(np.uint32), whole[s].view(np.uint32)), s
