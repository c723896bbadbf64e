"""The coherent 8-element form: eight phase-coherent cf32 elements at 20 MSPS
through one shared-prototype filterbank launch (2048 channels, 12 taps/branch,
2x), then W1 (per-channel 8x8 cross-spectra) and W2 on the device output.

Bars: filterbank within 5e-7 of full scale of the FP64 oracle; W1 and W2
bit-identical with the drop-in calls (atkdsp_pfb_cov_accumulate /
atkdsp_pfb_channel_power, which follow the reference's summation order) on
the same outputs, run by run; W1 within 1e-5 of the largest entry of an FP64
W1 over the FP64 outputs, W2 within 1e-5 relative per channel (power from
FP32 outputs: twice their relative error, larger in the weakest channels;
1.8e-6 measured).
"""
import numpy as np
import pytest

from oracle import oracle as O

pytestmark = pytest.mark.gpu

FS, M, OS, K, NEL = 20e6, 2048, 2, 12, 8


@pytest.mark.parametrize("nsamp,blocks", [(200_000, 2), (131_072, 3)])
def test_coherent_array_pfb_w1_w2(gpu, nsamp, blocks):
    from atkdsp_amd import synth
    from atkdsp_amd.pfbx import PfbX
    els = synth.c5_elements(n=nsamp)
    cut = np.linspace(0, nsamp, blocks + 1).astype(int) // 8 * 8
    cut[-1] = nsamp
    px = PfbX(form="fp32", fmt="cf32", fs=FS, nchan=M, oversample=OS, taps_per_poly=K, streams=NEL,
              block_max=int(np.diff(cut).max() + 8), coherent=True)
    outs = [[] for _ in range(NEL)]
    R_ref = np.zeros((M, NEL, NEL), np.complex64)
    c_ref = np.zeros(M, np.uint32)
    for a, b in zip(cut[:-1], cut[1:]):
        for e in range(NEL):
            px.upload(e, els[e][a:b].tobytes())
        m = px.run(b - a)
        ys = np.stack([px.read(e) for e in range(NEL)])
        for e in range(NEL):
            outs[e].append(ys[e])
        if m:
            gpu.pfb_cov_accumulate(ys, NEL, M, None, R_ref, c_ref)      # the drop-in on the same outputs
    y = np.stack([np.concatenate(o, 1) for o in outs])
    R, cnt = px.read_cov()
    assert np.array_equal(cnt, c_ref) and int(cnt[0]) == y.shape[2]
    assert np.array_equal(R, R_ref)                                      # same kernel order: bit-identical
    pw, mp = px.read_power()
    pw_ref, mp_ref = gpu.pfb_channel_power(ys, NEL, M)                   # W2 of the last run
    assert np.array_equal(pw, pw_ref) and np.array_equal(mp, mp_ref)
    # against FP64
    po = O.Pfb(FS, M, OS, K, 80.0, max_nel=NEL)
    y64 = po.process_n([e.astype(np.complex64) for e in els], NEL)
    assert y.shape == y64.shape
    assert O.rel_err(y, y64) < 5e-7
    R64 = np.zeros((M, NEL, NEL), np.complex128)
    c64 = np.zeros(M, np.int64)
    O.pfb_cov_accumulate(y64, NEL, M, None, R64, c64)
    assert np.max(np.abs(R - R64)) / np.max(np.abs(R64)) < 1e-5
    p64, m64 = O.pfb_channel_power(y64[:, :, -ys.shape[2]:], NEL, M)
    assert np.max(np.abs(pw - p64) / p64) < 1e-5
    # the plane waves are coherent across elements: the strongest channel's
    # cross-spectrum has the element-to-element phase ramp (|R_ab| ~ sqrt(R_aa R_bb))
    k = int(np.argmax(np.real(np.trace(R64, axis1=1, axis2=2))))
    coh = np.abs(R[k, 0, 1]) / np.sqrt(np.real(R[k, 0, 0]) * np.real(R[k, 1, 1]))
    assert coh > 0.9


def test_cov_reset_and_fmt_checks(gpu):
    from atkdsp_amd.pfbx import PfbX
    px = PfbX(form="fp32", fmt="cf32", fs=FS, nchan=64, oversample=2, taps_per_poly=8, streams=3,
              block_max=4096, coherent=True)
    x = (np.random.default_rng(5).standard_normal(4096) + 0j).astype(np.complex64)
    for e in range(3):
        px.upload(e, x.tobytes())
    px.run(4096)
    R, cnt = px.read_cov()
    assert cnt[0] > 0 and np.abs(R).max() > 0
    px.cov_reset()
    R, cnt = px.read_cov()
    assert cnt.max() == 0 and np.abs(R).max() == 0
    # a non-coherent object has no W1
    q = PfbX(form="fp32", fmt="cf32", fs=FS, nchan=64, oversample=2, taps_per_poly=8, streams=3, block_max=4096)
    with pytest.raises(gpu.AtkDspError):
        q.read_cov()
