"""The filterbank's matrix-core form against the FP64 oracle.

Samples are exact in BF16/FP16 and the taps are split into hi + lo terms (or
one term), so the only new rounding is the taps'. Each form's bound is stated
here, from the split's precision and measured with margin:
    fp32    5e-7 of full scale (the FP32 path's bound) and bit-exact with the drop-in
    fp32raw the FP32 form reading the raw rows: bit-exact with fp32
    fp16x2  2e-6   (taps to about 22 bits)
    bf16x2  5e-5   (taps to about 16 bits)
    fp16    2e-3   (11 bits)
    bf16    1e-2   (8 bits)
Full-scale relative error is max|y - y64| / max|y64| over every output; the dB
error is over outputs within 60 dB of the peak.
"""
import json
from pathlib import Path

import numpy as np
import pytest

from oracle import oracle as O

pytestmark = pytest.mark.gpu

C1 = dict(fs=2.4e6, nchan=64, oversample=1, taps_per_poly=8, atten_db=81.2953)
BOUND = {"fp32": 5e-7, "fp32raw": 5e-7, "fp16x2": 2e-6, "bf16x2": 5e-5, "fp16": 2e-3, "bf16": 1e-2}


def _errors(y, y64):
    full = float(np.max(np.abs(y - y64)) / np.max(np.abs(y64)))
    a64 = np.abs(y64)
    strong = a64 >= 1e-3 * a64.max()
    db = float(np.max(np.abs(20 * np.log10(np.abs(y[strong])) - 20 * np.log10(a64[strong]))))
    return full, db


def _oracle(raw, fmt, off, fs, nchan, oversample, taps_per_poly, atten_db):
    x, _ = O.unpack_dc(raw, fmt, off)
    return x, O.Pfb(fs, nchan, oversample, taps_per_poly, atten_db).process(x)


def _record(name, rows):
    out = Path("gpurun_out")
    if out.is_dir():
        (out / f"{name}.json").write_text(json.dumps(rows, indent=1))


def test_forms_on_c1_against_fp64(gpu):
    from atkdsp_amd import synth
    from atkdsp_amd.pfbx import PfbX
    raw = synth.c1_bytes()
    _, mean = O.unpack_dc(raw, O.CU8, 0j)
    off = complex(np.float32(mean.real), np.float32(mean.imag))      # a fixed block offset, as unpack_dc takes it
    x, y64 = _oracle(raw, O.CU8, off, **C1)
    n = len(raw) // 2
    rows = {}
    outs = {}
    for form in ["fp32", "fp32raw", "bf16x2", "fp16x2", "bf16", "fp16"]:
        px = PfbX(form=form, fmt="cu8", streams=1, block_max=n, offset=off, **C1)
        y = px.process([raw], 2)[0]
        assert y.shape == y64.shape
        full, db = _errors(y, y64)
        rows[form] = {"rel_full_scale": full, "max_db": db}
        assert full < BOUND[form], (form, full)
        if form == "fp32":
            # the FP32 form is the drop-in's arithmetic
            ref = gpu.Pfb(**C1).process(x)
            assert np.array_equal(y, ref)
        outs[form] = y
        px.close()
    # converting at load is the unpack's arithmetic: the same outputs bit for bit
    assert np.array_equal(outs["fp32raw"].view(np.uint32), outs["fp32"].view(np.uint32))
    _record("precision_errors_c1", rows)
    print(json.dumps(rows))
    # the hi+lo split is what buys the accuracy
    assert rows["bf16x2"]["rel_full_scale"] < rows["bf16"]["rel_full_scale"] / 10
    assert rows["fp16x2"]["rel_full_scale"] < rows["fp16"]["rel_full_scale"] / 10


MX_CONTINUITY = [
    ("bf16x2", "cu8", C1),
    ("fp16x2", "cu8", C1),
    ("bf16", "cu8", C1),
    ("bf16x2", "cu8", dict(fs=2.4e6, nchan=64, oversample=2, taps_per_poly=8, atten_db=80.0)),
    ("fp16x2", "ci16q11", dict(fs=10e6, nchan=128, oversample=2, taps_per_poly=8, atten_db=80.0)),
]


@pytest.mark.parametrize("blk", [1, 7, 999, 4096])
@pytest.mark.parametrize("form,fmt,shape", MX_CONTINUITY, ids=lambda v: v if isinstance(v, str) else None)
def test_block_continuity(gpu, blk, form, fmt, shape):
    """Blocks of 1, 7, 999 and 4096 samples give the one-call result bit for
    bit: history, leftover and hop parity carry across runs, the zero history
    only at the start, and every hop sits at the same place in its matrix tile
    (groups start on absolute hop multiples) whatever the block boundaries."""
    from atkdsp_amd import synth
    from atkdsp_amd.pfbx import PfbX
    bps = {"cu8": 2, "ci16q11": 4}[fmt]
    nsamp = 6000 if blk < 100 else 60_000
    raw = (synth.c1_bytes() if fmt == "cu8"
           else synth.quantize(synth.wideband_iq(31, nsamp, shape["fs"]), O.CI16Q11))[: bps * nsamp]
    off = 0.01 - 0.02j
    one = PfbX(form=form, fmt=fmt, block_max=nsamp, offset=off, **shape)
    whole = one.process([raw], bps)[0]
    px = PfbX(form=form, fmt=fmt, block_max=((blk + 7) // 8) * 8, offset=off, **shape)
    parts = []
    for a in range(0, nsamp, blk):
        b = min(nsamp, a + blk)
        px.upload(0, raw[bps * a:bps * b])
        if px.run(b - a):
            parts.append(px.read(0))
    y = np.concatenate(parts, 1)
    assert y.shape == whole.shape
    assert np.array_equal(y.view(np.uint32), whole.view(np.uint32))
    if shape is C1:
        _, y64 = _oracle(raw, O.CU8, off, **C1)
        assert _errors(y, y64)[0] < BOUND[form]


@pytest.mark.parametrize("fmt,nchan,os_,K,form", [
    ("cu8", 64, 2, 8, "bf16x2"), ("ci16q11", 64, 2, 8, "fp16x2"), ("ci16q11", 128, 2, 8, "fp16x2"),
    ("cu8", 128, 2, 8, "bf16x2"), ("cu8", 64, 1, 16, "bf16x2"), ("cu8", 64, 1, 32, "bf16x2"),
    ("ci8", 64, 1, 8, "bf16x2"), ("ci8", 64, 1, 8, "fp16x2"), ("ci16q11", 64, 1, 8, "fp16")])
def test_shapes_formats_and_streams(gpu, fmt, nchan, os_, K, form):
    """Other shapes (2x oversampling, two k-steps), formats with offsets, and
    several streams of different content in one launch."""
    from atkdsp_amd import synth
    from atkdsp_amd.pfbx import PfbX
    f = {"cu8": 0, "ci8": 1, "ci16q11": 3}[fmt]
    n = 40_000
    off = 0.003 + 0.001j
    raws = [synth.quantize(synth.wideband_iq(70 + s, n, 10e6, noise_dbfs=-25.0, ntones=6), f) for s in range(3)]
    px = PfbX(form=form, fmt=fmt, fs=10e6, nchan=nchan, oversample=os_, taps_per_poly=K, streams=3,
              block_max=n, offset=off)
    # two runs: the second crosses the kept history
    half = (n // 2) & ~7
    ys = [[], [], []]
    for a, b in [(0, half), (half, n)]:
        for s in range(3):
            px.upload(s, raws[s][O.BPS[f] * a:O.BPS[f] * b])
        px.run(b - a)
        for s in range(3):
            ys[s].append(px.read(s))
    for s in range(3):
        y = np.concatenate(ys[s], 1)
        _, y64 = _oracle(raws[s], f, off, 10e6, nchan, os_, K, 80.0)
        assert y.shape == y64.shape
        assert _errors(y, y64)[0] < BOUND[form], (s, _errors(y, y64))


RAW_SHAPES = [
    ("cu8", C1), ("ci8", C1), ("ci16q11", C1), ("ci16", C1),
    ("cu8", dict(fs=2.4e6, nchan=64, oversample=2, taps_per_poly=8, atten_db=80.0)),
    ("ci16q11", dict(fs=10e6, nchan=128, oversample=2, taps_per_poly=16, atten_db=80.0)),
    ("ci8", dict(fs=10e6, nchan=128, oversample=1, taps_per_poly=32, atten_db=80.0)),
]


@pytest.mark.parametrize("blk", [1, 7, 999, 4096])
@pytest.mark.parametrize("fmt,shape", RAW_SHAPES, ids=lambda v: v if isinstance(v, str) else None)
def test_fp32raw_equals_fp32(gpu, blk, fmt, shape):
    """The FP32 form on raw rows (each sample converted as the fold loads it,
    the stream's zero history before its first sample) equals the FP32 form
    on unpacked rows bit for bit, one call against blocks of 1, 7, 999 and
    4096 samples, with an offset."""
    from atkdsp_amd import synth
    from atkdsp_amd.pfbx import PfbX
    f = {"cu8": 0, "ci8": 1, "ci16": 2, "ci16q11": 3}[fmt]
    nsamp = 6000 if blk < 100 else 60_000
    raw = synth.quantize(synth.wideband_iq(41 + f, nsamp, shape["fs"], noise_dbfs=-25.0, ntones=6), f)
    off = 0.004 - 0.003j
    ref = PfbX(form="fp32", fmt=fmt, block_max=nsamp, offset=off, **shape).process([raw], O.BPS[f])[0]
    px = PfbX(form="fp32raw", fmt=fmt, block_max=((blk + 7) // 8) * 8, offset=off, **shape)
    parts = []
    for a in range(0, nsamp, blk):
        b = min(nsamp, a + blk)
        px.upload(0, raw[O.BPS[f] * a:O.BPS[f] * b])
        if px.run(b - a):
            parts.append(px.read(0))
    y = np.concatenate(parts, 1)
    assert y.shape == ref.shape
    assert np.array_equal(y.view(np.uint32), ref.view(np.uint32))


def test_unsupported_forms_fail_loudly(gpu):
    from atkdsp_amd.pfbx import PfbX
    with pytest.raises(gpu.AtkDspError):
        PfbX(form="bf16x2", fmt="ci16q11", **C1)        # 12-bit samples are not exact in BF16
    with pytest.raises(gpu.AtkDspError):
        PfbX(form="fp16x2", fmt="ci16", **C1)           # 16-bit samples are not exact in FP16
    with pytest.raises(gpu.AtkDspError):
        PfbX(form="bf16x2", fmt="cu8", fs=40e6, nchan=1024, oversample=2, taps_per_poly=16)
    with pytest.raises(gpu.AtkDspError):
        PfbX(form="fp32", fmt="cu8", block_max=1001, **C1)   # block_max must be a multiple of 8
    with pytest.raises(gpu.AtkDspError):
        PfbX(form="fp32raw", fmt="cf32", **C1)               # raw rows hold integer formats only
    with pytest.raises(gpu.AtkDspError):
        PfbX(form="fp32raw", fmt="ci16q11", fs=40e6, nchan=1024, oversample=2, taps_per_poly=16)
