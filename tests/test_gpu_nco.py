"""The NCO on the GPU: the drop-in mixer against the reference's golden
vectors and its continuity pattern, and the FP64 phase of a 10 s run at
40 MSPS against the exact rational phase.

Tolerances: outputs within 1e-5 (the reference's own NCO tests,
tests/test_cross_check.py:120-130); the carried phase within 1e-12 rad of the
reference's after a call sequence (the same FP64 arithmetic); the long run's
phase within 1e-6 rad of exact at every chunk boundary (the FP64 step's own
rounding over 4e8 samples is about 1e-7 rad).
"""
import math

import numpy as np
import pytest

from oracle import oracle as O

pytestmark = pytest.mark.gpu


def _blocks(nco, x, blocks):
    out, a, k = [], 0, 0
    while a < x.size:
        b = min(x.size, a + blocks[k % len(blocks)])
        out.append(nco.mix(x[a:b]))
        a, k = b, k + 1
    return np.concatenate(out)


def test_nco_mix_matches_the_reference_across_blocks(gpu, G):
    g = G("nco")
    a = gpu.Nco(173_125.0, 2.4e6)
    y = _blocks(a, g["x"], [1, 999, 4096, 7, 13000])
    assert np.max(np.abs(y - g["y"])) < 1e-5
    assert abs(a.phase - float(g["phase"])) < 1e-12
    y1 = gpu.Nco(173_125.0, 2.4e6).mix(g["x"])
    assert np.max(np.abs(y1 - g["y1"])) < 1e-5
    assert np.max(np.abs(y - y1)) < 1e-5
    # against the FP64 oracle, tighter
    # against the FP64 oracle, relative to each sample's size (FP32 rotation)
    d = np.abs(y1 - O.Nco(173_125.0, 2.4e6).mix(g["x"]))
    assert np.max(d / np.maximum(np.abs(g["x"]), 1e-3)) < 5e-7


def test_nco_edge_cases(gpu):
    z = (np.arange(100) + 1j).astype(np.complex64)
    assert np.array_equal(gpu.Nco(0.0, 1e6).mix(z), z)          # zero frequency: identity
    x = np.ones(5000, np.complex64)
    n = gpu.Nco(1000.0, 1e6)
    y = n.mix(x, out=x)                                          # in and out may alias
    assert y is x and abs(x[1] - np.exp(-2j * np.pi * 1e-3)) < 1e-6
    assert gpu.Nco(1000.0, 1e6).mix(np.zeros(0, np.complex64)).size == 0


def test_ten_seconds_at_40msps_fp64_phase(gpu):
    """4e8 samples in 65,536-sample runs; f = fs * p / q with q = 10007, so
    the exact phase after n samples is -2 pi ((n p) mod q) / q."""
    from atkdsp_amd.nco import NcoRun
    fs, p, q = 40e6, 1234, 10007
    chunk = 65536
    nrun = int(math.ceil(4e8 / chunk))                           # 6104 runs, 400,031,744 samples
    run = NcoRun(fs * p / q, fs, chunk_max=chunk, log_max=nrun)
    run.upload(np.ones(chunk, np.complex64))
    for _ in range(nrun):
        run.run(chunk)
    run.sync()
    ph, ns = run.phase()
    assert ns == nrun * chunk
    log = run.log()
    assert log.size == nrun and log[-1] == ph
    errs = np.array([abs(O.wrap(log[k] - O.nco_exact_phase((k + 1) * chunk, p, q))) for k in range(nrun)])
    print(f"phase error: final {errs[-1]:.3e} rad, max {errs.max():.3e} rad over {nrun} runs")
    assert errs.max() < 1e-6
    # the last run's samples are on the exact phase
    y = run.read(chunk)
    n0 = (nrun - 1) * chunk
    exact = np.array([O.nco_exact_phase(n0 + i, p, q) for i in range(0, chunk, 97)])
    assert np.max(np.abs(y[::97] - np.exp(1j * exact))) < 1e-6
