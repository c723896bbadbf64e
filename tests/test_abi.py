"""The drop-in boundary, CPU only: the library loads, reports ABI 26, exports
every name of the reference ABI and every function include/atkdsp.h declares,
and its stubs return the documented failure values. No compute calls."""
import ctypes as C
import math
import re
import subprocess

import pytest

from conftest import ROOT, has_gpu

LIB = ROOT / "atkdsp_amd" / "atkdsp.so"


def exported():
    r = subprocess.run(["nm", "-D", "--defined-only", str(LIB)], capture_output=True, text=True, check=True)
    return {ln.split()[-1] for ln in r.stdout.splitlines() if " T " in ln}


def header_functions(path):
    src = (ROOT / "include" / path).read_text()
    return set(re.findall(r"ATK\w*API\s+[^;(]*?\b(atk\w+)\s*\(", src))


def test_library_is_built():
    assert LIB.exists(), "run python -m atkdsp_amd.build"


def test_exports_every_reference_abi_name():
    from atkdsp_amd.build import abi_symbols
    names = {n for n, _ in abi_symbols()}
    assert len(names) == 334
    missing = names - exported()
    assert not missing, sorted(missing)[:10]


def test_exports_every_declared_function():
    ex = exported()
    for h in ("atkdsp.h", "atkgpu.h"):
        if not (ROOT / "include" / h).exists():
            continue
        decl = header_functions(h)
        assert decl, h
        assert not (decl - ex), (h, sorted(decl - ex))


def test_identity_without_a_device():
    import atkdsp_amd as A
    lib = A.load()
    assert lib.atkdsp_abi_version() == 26
    assert A.version() == "0.25.1"
    assert "abi 26" in A.build_info() and "gfx950" in A.build_info()
    assert A.bytes_per_sample("cu8") == 2 and A.bytes_per_sample("ci16q11") == 4
    assert A.bytes_per_sample("cf32") == 8 and A.bytes_per_sample(9) == 0
    assert A.set_threads(3) == 3 and A.get_threads() == 3
    assert A.set_threads(0) >= 1


def test_stubs_return_failure_values():
    lib = C.CDLL(str(LIB))
    f = lib.atkdsp_fir_create
    f.restype = C.c_void_p
    assert f() is None                                  # a *_create stub: NULL
    g = lib.atkdsp_fir_reset
    g.restype = None
    g()                                                 # void: nothing
    h = lib.atkdsp_ddc_out_rate
    h.restype = C.c_double
    assert h() != h()                                   # floating: NaN
    k = lib.atkdsp_cfar_process
    k.restype = C.c_ssize_t
    assert k() == -1                                    # signed count: ATKDSP_E_ARG


def test_lowpass_design_is_host_code_and_bit_exact():
    """atkdsp_design_lowpass runs on the host (a design-time helper) and
    reproduces the reference's taps bit for bit (src/design.c:34-62)."""
    import numpy as np
    import atkdsp_amd as A
    from conftest import golden
    g = golden("osx")
    taps = np.split(g["design_taps"], np.cumsum(g["design_len"])[:-1])
    for args, t in zip(g["design_args"], taps):
        h = A.design_lowpass(*args)
        assert np.array_equal(h.view(np.uint32), t.view(np.uint32)), args
    with pytest.raises(A.AtkDspError):
        A.design_lowpass(50e3, 25e3, 80.0, 2.4e6)              # stopband below the passband


def test_window_and_window_stats_are_host_design_helpers():
    import numpy as np
    import atkdsp_amd as A
    from oracle import oracle as O
    for kind in ("rectangular", "hann", "hamming", "blackmanharris"):
        k = A.WIN[kind]
        assert np.array_equal(A.window(kind, 513), O.window(k, 513))
    with pytest.raises(A.AtkDspError):
        A.window(7, 16)
    cg, enbw = A.window_stats(A.window("hann", 4096))
    assert abs(cg - 0.5) < 1e-3 and abs(enbw - 1.5) < 1e-3


@pytest.mark.skipif(has_gpu(), reason="checks the behaviour without a device")
def test_no_device_fails_loudly():
    import numpy as np
    import atkdsp_amd as A
    with pytest.raises(A.AtkDspError):
        A.Fft(1024)
    with pytest.raises(A.AtkDspError):
        A.Pfb(2.4e6, 64)
    with pytest.raises(A.AtkDspError):
        A.unpack(b"\x00\x01\x02\x03", "cu8")
    with pytest.raises(A.AtkDspError):
        A.Osx(2.4e6, 4096)
    with pytest.raises(A.AtkDspError):
        A.ArbResampler(75e3, 50e3)
    # detection runs on the device: no GPU is E_FFT / NaN ("unmeasured"), after
    # the reference's own argument checks (a NaN floor is E_ARG, n = 0 is 0)
    line = np.full(64, -100.0, np.float32)
    with pytest.raises(A.AtkDspError):
        A.detect_channels(line, -100.0, 6.0, 2, 2)
    assert A.detect_channels(line[:0], -100.0, 6.0, 2, 2) == []
    lib = A.load()
    sc = np.empty(64, np.float32)
    assert lib.atkdsp_detect_channels(A._fp(line), 64, float("nan"), 6.0, 2, 2, A._fp(sc), None, 0) == -1
    assert math.isnan(A.median(line)) and A.median(line[:0]) == -120.0


def test_reference_binding_would_load_it():
    """The reference binding needs ABI 26 and every name it binds; both are
    checked above. Here: the binding's own identity sentence (build_info
    contains "abi 26", reference tests/test_cross_check.py:58-62)."""
    lib = C.CDLL(str(LIB))
    lib.atkdsp_build_info.restype = C.c_char_p
    assert b"abi 26" in lib.atkdsp_build_info()


def test_filterbank_launch_refuses_an_output_pitch_beyond_its_offsets():
    """The fused filterbank's store uses 32-bit lane offsets, so its launchers
    refuse an output pitch above 2^20 hops, before any device call (so this
    runs without a device); a call with no hops is a no-op that succeeds."""
    lib = C.CDLL(str(LIB))
    f = lib.atkgpu_pfb_fused
    f.restype = C.c_int
    f.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_size_t, C.c_uint, C.c_size_t, C.c_longlong, C.c_uint,
                  C.c_uint, C.c_uint, C.c_void_p, C.c_void_p, C.c_void_p, C.c_uint, C.c_void_p, C.c_size_t,
                  C.c_size_t]
    big = (1 << 20) + 1
    assert f(None, None, 1 << 22, 1 << 22, 1, 16, 0, 1024, 2, 32, None, None, None, 0, None, big, 0) == -1
    assert f(None, None, 1 << 22, 1 << 22, 1, 0, 0, 1024, 2, 32, None, None, None, 0, None, big, 0) == 0
    r = lib.atkgpu_pfb_fused_raw
    r.restype = C.c_int
    r.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_size_t, C.c_uint, C.c_size_t, C.c_longlong, C.c_uint,
                  C.c_uint, C.c_uint, C.c_int, C.c_float, C.c_float, C.c_int, C.c_void_p, C.c_void_p, C.c_void_p,
                  C.c_size_t]
    assert r(None, None, 1 << 22, 1 << 22, 1, 16, 0, 64, 1, 8, 0, 0.0, 0.0, 0, None, None, None, big) == -1
