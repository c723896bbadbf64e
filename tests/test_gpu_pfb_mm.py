"""The filterbank's FP32 fold on the matrix cores ("fp32mm") against the FP32
VALU fold ("fp32") and the FP64 oracle.

The 4x4x1 FP32 matrix product rounds each step as one fused multiply-add and
the band feeds it the taps in ascending order, so for finite input the two
forms agree to the bit; against FP64 the FP32 path's bound (5e-7 of full
scale) holds. Runs of uneven length cross the kept history, the leftover and
the 2x hop parity; several streams of different content share each launch.
"""
import numpy as np
import pytest

from oracle import oracle as O

pytestmark = pytest.mark.gpu

# fs, channels, oversampling, taps/branch, format
SHAPES = [
    (40e6, 1024, 2, 16, "ci16q11"),     # the wideband shape
    (40e6, 1024, 2, 12, "ci16q11"),
    (10e6, 1024, 2, 8, "cu8"),
    (20e6, 512, 2, 16, "ci8"),
    (10e6, 256, 2, 12, "cu8"),
    (2.4e6, 256, 2, 5, "cf32"),
    (2.4e6, 512, 1, 8, "cu8"),
    (2.4e6, 256, 1, 16, "ci16q11"),
]
FMT = {"cu8": 0, "ci8": 1, "ci16q11": 3, "cf32": 4}


def _raws(fmt, n, S, fs):
    from atkdsp_amd import synth
    return [synth.quantize(synth.wideband_iq(500 + s, n, fs, noise_dbfs=-25.0, ntones=8), FMT[fmt])
            for s in range(S)]


@pytest.mark.parametrize("fs,nchan,os_,K,fmt", SHAPES)
def test_matrix_fold_equals_fp32_fold(gpu, fs, nchan, os_, K, fmt):
    from atkdsp_amd.pfbx import PfbX
    S = 3
    runs = [100_003, 7, 65_536, 1, 999, 40_001]
    n = sum(runs)
    bps = O.BPS[FMT[fmt]]
    raws = _raws(fmt, n, S, fs)
    off = 0.002 - 0.001j
    cfg = dict(fmt=fmt, fs=fs, nchan=nchan, oversample=os_, taps_per_poly=K, streams=S,
               block_max=(max(runs) + 7) // 8 * 8, offset=off)
    a, b = PfbX(form="fp32", **cfg), PfbX(form="fp32mm", **cfg)
    got = [[] for _ in range(S)]
    pos = 0
    for r in runs:
        for s in range(S):
            seg = raws[s][bps * pos:bps * (pos + r)]
            a.upload(s, seg)
            b.upload(s, seg)
        ma, mb = a.run(r), b.run(r)
        assert ma == mb
        if ma:
            for s in range(S):
                ya, yb = a.read(s), b.read(s)
                assert np.array_equal(ya.view(np.uint32), yb.view(np.uint32)), (r, s)
                got[s].append(yb)
        pos += r
    x, _ = O.unpack_dc(raws[1], FMT[fmt], off)
    y64 = O.Pfb(fs, nchan, os_, K, 80.0).process(x)
    y = np.concatenate(got[1], 1)
    assert y.shape == y64.shape
    err = float(np.max(np.abs(y - y64)) / np.max(np.abs(y64)))
    assert err < 5e-7, err
    a.close()
    b.close()


@pytest.mark.parametrize("blk", [1, 7, 999, 4096])
def test_matrix_fold_block_continuity_wideband(gpu, blk):
    """Blocks of 1, 7, 999 and 4096 samples through the 1024-channel 2x shape
    give the one-call result bit for bit: history, leftover and the (-1)^k
    rotation of odd hops carry across runs."""
    from atkdsp_amd.pfbx import PfbX
    fs, nchan, os_, K = 40e6, 1024, 2, 16
    nsamp = 24_000 if blk < 100 else 120_000
    raw = _raws("ci16q11", nsamp, 1, fs)[0]
    off = 0.001 + 0.0005j
    one = PfbX(form="fp32mm", fmt="ci16q11", fs=fs, nchan=nchan, oversample=os_, taps_per_poly=K,
               block_max=(nsamp + 7) // 8 * 8, offset=off)
    whole = one.process([raw], 4)[0]
    px = PfbX(form="fp32mm", fmt="ci16q11", fs=fs, nchan=nchan, oversample=os_, taps_per_poly=K,
              block_max=(blk + 7) // 8 * 8, offset=off)
    parts = []
    for a in range(0, nsamp, blk):
        b = min(nsamp, a + blk)
        px.upload(0, raw[4 * a:4 * b])
        if px.run(b - a):
            parts.append(px.read(0))
    y = np.concatenate(parts, 1)
    assert y.shape == whole.shape
    assert np.array_equal(y.view(np.uint32), whole.view(np.uint32))


def test_matrix_fold_refuses_uncovered_shapes(gpu):
    from atkdsp_amd.pfbx import PfbX
    with pytest.raises(gpu.AtkDspError):
        PfbX(form="fp32mm", fmt="cu8", fs=2.4e6, nchan=64, oversample=1, taps_per_poly=8)       # too few channels
    with pytest.raises(gpu.AtkDspError):
        PfbX(form="fp32mm", fmt="cu8", fs=40e6, nchan=2048, oversample=2, taps_per_poly=12)     # too many
    with pytest.raises(gpu.AtkDspError):
        PfbX(form="fp32mm", fmt="cu8", fs=40e6, nchan=1024, oversample=2, taps_per_poly=24)     # taps
    with pytest.raises(gpu.AtkDspError):
        PfbX(form="fp32mm", fmt="cu8", fs=40e6, nchan=1024, oversample=1, taps_per_poly=8)      # 1x at 1024
