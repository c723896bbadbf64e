"""Throughput of the wideband front end on one MI355X.

Workload (one step): one 50 ms block of the 40 MSPS ci16q11 receiver workload
(2,000,000 samples: -30 dBFS noise, 20 tones, 4 bursts) for each of S
independent streams (default 128, the >= 100 concurrent 40 MSPS streams the
path is sized for), already resident in HBM, through
    unpack with block DC removal -> polyphase filterbank (1024 channels,
    16 taps/branch, 2x oversampled, history in HBM) -> 4096-point Hann
    spectrum line at hop 2048 with max-hold (tail carried across blocks) ->
    decimate to 1024 columns -> NaN-aware stitch into an 8192-bin trace.

Prints one JSON line. `value` is aggregate input samples per second over all
streams; `roofline` is the filterbank kernel (the dominant one) timed with
HIP events on its own stream; `cpu_baseline` is the reference C library
(built from its sources into oracle/_ref, OpenMP on the host cores) running
the same chain on a bounded sample of the same blocks.
"""
from __future__ import annotations

import argparse
import ctypes as C
import json
import os
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

FS = 40e6
N_BLOCK = 2_000_000
M, OS, K = 1024, 2, 16
NFFT, FHOP, COLS, TRACE = 4096, 2048, 1024, 8192
HBM_PEAK_GBS = 8000.0


def make_streams(S: int):
    from atkdsp_amd import synth
    base = np.frombuffer(synth.c2_bytes(n=N_BLOCK), "<i2").reshape(-1, 2)
    # stream s: the same block content at a different phase (circular shift)
    return base, [np.roll(base, s * 7919, axis=0) for s in range(S)]


# the whole chain per input sample at the wideband shape: the compulsory bytes
# (the raw ci16 sample read once, the 2x cf32 filterbank output written once,
# the display lines negligible) and the useful FP32 work (fold 16 taps x 2
# branches x complex = 128, the 1024-point FFT per 512 samples = 100, its
# rotation 12, the 4096-point display FFT per 2048 samples = 120, window,
# power and max-hold 12)
CHAIN_BYTES_PER_SAMPLE = 4 + 16
CHAIN_FLOP_PER_SAMPLE = 128 + 100 + 12 + 120 + 12
VALU_FP32_PEAK_TFLOPS = 78.6      # MI355X FP32 vector, without packed FP32


def chain_roofline(samples_per_s: float) -> dict:
    gbs = samples_per_s * CHAIN_BYTES_PER_SAMPLE / 1e9
    tfl = samples_per_s * CHAIN_FLOP_PER_SAMPLE / 1e12
    return {"bytes_per_sample": CHAIN_BYTES_PER_SAMPLE, "achieved_GBps": round(gbs, 1),
            "hbm_frac": round(gbs / HBM_PEAK_GBS, 4), "flop_per_sample": CHAIN_FLOP_PER_SAMPLE,
            "achieved_TFLOPs": round(tfl, 2), "valu_frac": round(tfl / VALU_FP32_PEAK_TFLOPS, 4),
            "note": "the whole chain against its compulsory bytes and its useful FP32 operations per input "
                    "sample; the VALU peak is the non-packed FP32 rate (packed FP32 is not used)"}


def pfb_bytes_per_launch(S: int, n: int, m: int) -> float:
    """Compulsory bytes of one filterbank launch: every input sample and the
    history read once as cf32, every output written once as cf32."""
    ov = M * K - M // OS
    return S * ((n + ov) * 8 + m * M * 8)


def precision_forms(S: int = 256, n: int = 1_200_000, steps: int = 10):
    """The filterbank's matrix-core forms against the plain FP32 form on the
    C1 shape (2.4 MSPS cu8, 64 channels, 8 taps/branch, critically sampled):
    S streams of n samples each, resident in HBM. The FP32 form is unpack +
    the fused FP32 filterbank; the FP32 raw form is the fused FP32 filterbank
    reading the raw rows (bit-identical to the FP32 form); the matrix-core
    forms read the same raw rows. Speedups are stated against both FP32
    forms, whole run and filterbank kernel alone."""
    from atkdsp_amd import synth
    from atkdsp_amd.pfbx import PfbX
    base = np.frombuffer(synth.c1_bytes(n=n), np.uint8).reshape(-1, 2)
    raws = [np.roll(base, s * 7919, axis=0) for s in range(S)]
    out = {}
    for form in ["fp32", "fp32raw", "bf16x2", "fp16x2", "bf16"]:
        px = PfbX(form=form, fmt="cu8", fs=2.4e6, nchan=64, oversample=1, taps_per_poly=8, atten_db=81.2953,
                  streams=S, block_max=n, offset=0.001 + 0.002j)
        for s in range(S):
            px.upload(s, raws[s])
        px.sync()
        for _ in range(2):
            px.run(n)
        px.sync()
        t0 = time.perf_counter()
        for _ in range(steps):
            px.run(n)
        px.sync()
        wall = (time.perf_counter() - t0) * 1e3 / steps
        ev = []
        for _ in range(3):
            px.run(n)
            ev.append(px.timing())
        run_ms, unpack_ms, pfb_ms = (float(np.mean([e[i] for e in ev])) for i in range(3))
        out[form] = {"ms_per_run": round(wall, 4), "device_ms": round(run_ms, 4), "unpack_ms": round(unpack_ms, 4),
                     "filterbank_ms": round(pfb_ms, 4), "Msamples_per_s": round(S * n / (wall * 1e-3) / 1e6, 1)}
        px.close()
        print(f"[bench] precision form {form}: {out[form]}", file=sys.stderr, flush=True)
    fp, fr = out["fp32"]["device_ms"], out["fp32raw"]["device_ms"]
    kr = out["fp32raw"]["filterbank_ms"]
    mx = [f for f in out if f not in ("fp32", "fp32raw")]
    return {"workload": f"{S} streams x {n} cu8 samples (2.4 MSPS), PFB 64 ch, 8 taps/branch, 1x, 81.3 dB",
            "forms": out,
            "speedup_vs_fp32": {f: round(fp / out[f]["device_ms"], 3) for f in out if f != "fp32"},
            "speedup_vs_fp32raw": {f: round(fr / out[f]["device_ms"], 3) for f in mx},
            "kernel_speedup_vs_fp32raw": {f: round(kr / out[f]["filterbank_ms"], 3) for f in mx},
            **mx_busy(),
            "note": "device_ms: HIP events around one whole run (unpack where there is one, filterbank, "
                    "history copy); filterbank_ms: the filterbank kernel alone; fp32raw reads the raw rows the "
                    "matrix-core forms read; errors against the FP64 oracle are in tests/test_gpu_precision.py"}


def drop_in(base: np.ndarray, calls: int = 20):
    """The drop-in ABI as the reference's callers use it, at the wideband
    shape: one 50 ms block (2,000,000 ci16q11 samples) per call, host buffers
    in and out, every copy over PCIe inside the timed call. ms per call of
    unpack_dc, pfb_process (1024 ch, 16 taps, 2x) and spectrum_reduce
    (4096/2048 Hann max), and the three in a row."""
    import atkdsp_amd as A
    lib = A.load()
    raw = np.ascontiguousarray(base).view(np.uint8).reshape(-1)
    x = np.empty(N_BLOCK, np.complex64)
    pfb = A.Pfb(FS, M, OS, K, 80.0)
    y = np.empty(M * (N_BLOCK // (M // OS) + 2), np.complex64)
    fft, win = A.Fft(NFFT), A.window("hann", NFFT)
    line = np.empty(NFFT, np.float32)
    hops = pfb.out_hops(N_BLOCK)

    def t_unpack():
        A.unpack_dc(raw, "ci16q11", 0j, out=x)

    def t_pfb():
        assert A._check(lib.atkdsp_pfb_process(pfb._p, A._cfp(x), x.size, A._cfp(y), y.size), "pfb_process") >= hops - 1

    def t_reduce():
        fft.spectrum_reduce(x, FHOP, win, "max", out=line)

    res = {}
    for name, fn in (("unpack_dc", t_unpack), ("pfb_process", t_pfb), ("spectrum_reduce", t_reduce)):
        fn()
        t0 = time.perf_counter()
        for _ in range(calls):
            fn()
        res[name] = (time.perf_counter() - t0) * 1e3 / calls
    t0 = time.perf_counter()
    for _ in range(calls):
        t_unpack(); t_pfb(); t_reduce()
    res["block"] = (time.perf_counter() - t0) * 1e3 / calls
    pfb.close()
    fft.close()
    out = {k + "_ms": round(v, 3) for k, v in res.items()}
    out["realtime_x"] = round(50.0 / res["block"], 1)
    out["workload"] = ("one 50 ms block (2,000,000 ci16q11 samples at 40 MSPS) per call, host buffers in and out "
                       "(the copies over PCIe inside each call): unpack_dc, pfb_process 1024 ch 16 taps 2x, "
                       "spectrum_reduce 4096/2048 Hann max")
    return out


def filterbank_forms_c2(base: np.ndarray, S: int = 128, steps: int = 5):
    """The FP32 filterbank forms at the wideband shape: S streams of one
    2,000,000-sample ci16q11 block each, resident in HBM, through the unpack
    and then the fused VALU filterbank ("fp32") or the one whose fold runs on
    the matrix cores ("fp32mm", bit-identical). Kernel-only times from HIP
    events around each launch."""
    from atkdsp_amd.pfbx import PfbX
    out = {}
    for form in ("fp32", "fp32mm"):
        px = PfbX(form=form, fmt="ci16q11", fs=FS, nchan=M, oversample=OS, taps_per_poly=K, streams=S,
                  block_max=N_BLOCK, offset=0.001 + 0.002j)
        for s in range(S):
            px.upload(s, np.roll(base, s * 7919, axis=0))
        px.sync()
        px.run(N_BLOCK)
        ev = []
        for _ in range(steps):
            px.run(N_BLOCK)
            ev.append(px.timing())
        px.close()
        out[form] = {"unpack_ms": round(float(np.mean([e[1] for e in ev])), 4),
                     "filterbank_ms": round(float(np.mean([e[2] for e in ev])), 4)}
        print(f"[bench] filterbank form {form} at the wideband shape: {out[form]}", file=sys.stderr, flush=True)
    out["workload"] = f"{S} streams x {N_BLOCK} ci16q11 samples, PFB {M} ch, {K} taps/branch, {OS}x"
    out["matrix_over_valu"] = round(out["fp32"]["filterbank_ms"] / out["fp32mm"]["filterbank_ms"], 3)
    return out


def coherent_array(n: int = 1_000_000, nel: int = 8, steps: int = 10):
    """The coherent 8-element form: eight phase-coherent cf32 elements at
    20 MSPS, one block of n samples each per run, resident in HBM, through
    one filterbank launch (2048 channels, 12 taps/branch, 2x), then W1 and W2
    on the device output."""
    from atkdsp_amd import synth
    from atkdsp_amd.pfbx import PfbX
    els = synth.c5_elements(n=n, nel=nel)
    px = PfbX(form="fp32", fmt="cf32", fs=20e6, nchan=2048, oversample=2, taps_per_poly=12, streams=nel,
              block_max=n, coherent=True)
    for e in range(nel):
        px.upload(e, els[e].tobytes())
    px.sync()
    for _ in range(2):
        px.run(n)
    px.sync()
    t0 = time.perf_counter()
    for _ in range(steps):
        m = px.run(n)
    px.sync()
    wall = (time.perf_counter() - t0) * 1e3 / steps
    ev = []
    for _ in range(3):
        px.run(n)
        ev.append(px.timing())
    run_ms, unpack_ms, pfb_ms, w_ms = (float(np.mean([e[i] for e in ev])) for i in range(4))
    px.close()
    # compulsory bytes of the filterbank launch: every element's input and history once, output once
    pfb_bytes = nel * ((n + 2048 * 12 - 1024) * 8 + m * 2048 * 8)
    return {"workload": f"{nel} elements x {n} cf32 samples (20 MSPS), PFB 2048 ch, 12 taps/branch, 2x, W1 + W2",
            "ms_per_run": round(wall, 4), "device_ms": round(run_ms, 4), "unpack_ms": round(unpack_ms, 4),
            "filterbank_ms": round(pfb_ms, 4), "w1_w2_ms": round(w_ms, 4), "hops": m,
            "Msamples_per_s": round(nel * n / (wall * 1e-3) / 1e6, 1),
            "filterbank_GBps": round(pfb_bytes / (pfb_ms * 1e-3) / 1e9, 1),
            "realtime_x": round((n / 20e6) / (wall * 1e-3), 1)}


def osx_signal(n: int = 2_000_000, nchan: int = 32, fs: float = 40e6):
    rng = np.random.default_rng(7)
    offs = np.linspace(-15e6, 15e6, nchan)
    k = np.arange(n)
    x = (0.01 * (rng.standard_normal(n) + 1j * rng.standard_normal(n))).astype(np.complex64)
    for f in offs[::4]:
        x += (0.1 * np.exp(2j * np.pi * (f + 20e3) / fs * k)).astype(np.complex64)
    return x, offs


def osx_extract(n: int = 2_000_000, nchan: int = 32, calls: int = 10):
    """The overlap-save extractor through the drop-in ABI: 40 MSPS cf32 in
    50 ms calls (host buffers in and out, as the reference's callers use it),
    8192-point blocks, nchan channels of 100 kHz landed on 125 kHz (through the
    arbitrary resampler from the 156.25 kHz channel rate), spread over
    +-15 MHz. (The reference C on the same calls is timed in the CPU leg.)"""
    import atkdsp_amd as A
    fs = 40e6
    x, offs = osx_signal(n, nchan, fs)
    o = A.Osx(fs, 8192)
    for f in offs:
        o.add(float(f), 100e3, 125e3)
    o.process(x)                                   # warm-up
    t0 = time.perf_counter()
    outs = 0
    for _ in range(calls):
        y, cnt = o.process(x)
        outs += sum(cnt)
    wall = (time.perf_counter() - t0) / calls
    res = {"workload": f"40 MSPS cf32, 50 ms calls, 8192-pt blocks, {nchan} channels of 100 kHz onto 125 kHz "
                       "(resampled from 156.25 kHz); host buffers in and out",
           "ms_per_call": round(wall * 1e3, 3), "Msamples_per_s": round(n / wall / 1e6, 1),
           "realtime_x": round((n / fs) / wall, 1), "outputs_per_call": int(outs // calls)}
    return res


CFAR_CFGS = (("ca_k1", dict(mode=0)), ("os_k4", dict(mode=1, k_frames=4)), ("go_k2", dict(mode=2, k_frames=2)))


def cfar_signal(nbins: int = 4096, nframes: int = 512):
    rng = np.random.default_rng(1)
    p = rng.exponential(1.0, (nframes, nbins))
    p[:, 1000:1008] += 40.0
    p[100:300, 2000:2003] += 30.0
    return (10.0 * np.log10(p) - 90.0).astype(np.float32)


def cfar_frames(nbins: int = 4096, nframes: int = 512, calls: int = 5):
    """CFAR detection and burst tracking through the drop-in ABI: frames of
    4096 dB bins (host buffers, as the reference's callers pass them) in calls
    of 512 frames, for CA (one frame), OS (four-frame average) and GO (two),
    with the default 16 + 16 reference cells and 4 guard cells. (The
    reference C on the same frames is timed in the CPU leg.)"""
    import atkdsp_amd as A
    fr = cfar_signal(nbins, nframes)
    out = {"workload": f"{nbins}-bin dB frames in calls of {nframes}, 16+16 reference cells, 4 guard, "
                       "Pfa 1e-6, open 2 of 3, close after 3; host buffers in, events out"}
    for name, cfg in CFAR_CFGS:
        c = A.Cfar(nbins, **cfg)
        c.process(fr[:64], 0.0, 0.01)                # warm-up
        t0 = time.perf_counter()
        nev = 0
        for r in range(calls):
            nev += len(c.process(fr, float(r), 0.01))
        us = (time.perf_counter() - t0) / (calls * nframes) * 1e6
        out[name] = {"us_per_frame": round(us, 2), "frames_per_s": round(1e6 / us),
                     "events_per_call": nev // calls}
    return out


def nco_phase_run(seconds: float = 10.0, fs: float = 40e6, p: int = 1234, q: int = 10007, chunk: int = 65536):
    """10 s of 40 MSPS through the device-resident NCO in 65,536-sample runs,
    f = fs * p / q: the FP64 phase after every run against the exact phase
    -2 pi ((n p) mod q) / q (integer arithmetic, no accumulation)."""
    import math
    from atkdsp_amd.nco import NcoRun

    def wrap(ph):
        ph = math.fmod(ph, 2 * math.pi)
        return ph - 2 * math.pi if ph >= math.pi else (ph + 2 * math.pi if ph < -math.pi else ph)

    nrun = int(math.ceil(seconds * fs / chunk))
    run = NcoRun(fs * p / q, fs, chunk_max=chunk, log_max=nrun)
    run.upload(np.ones(chunk, np.complex64))
    run.sync()
    t0 = time.perf_counter()
    for _ in range(nrun):
        run.run(chunk)
    run.sync()
    ms = (time.perf_counter() - t0) * 1e3
    log = run.log()
    err = np.array([abs(wrap(log[k] - wrap(-2 * math.pi * (((k + 1) * chunk * p) % q) / q)))
                    for k in range(nrun)])
    run.close()
    return {"samples": nrun * chunk, "runs": nrun, "chunk": chunk, "f_over_fs": f"{p}/{q}", "ms": round(ms, 2),
            "Msamples_per_s": round(nrun * chunk / (ms * 1e-3) / 1e6, 1),
            "final_phase_err_rad": float(err[-1]), "max_phase_err_rad": float(err.max())}


def mx_busy():
    """Matrix-core busy fraction of each matrix-core form at this workload,
    from the committed rocprofv3 counter summary (the counters cannot be read
    from inside the timed run)."""
    f = ROOT / "profiles" / "pfb_mx_c1_mfma.json"
    if not f.exists():
        return {}
    d = json.loads(f.read_text())
    busy = {k: round(v["mfma_busy_cycles"] / (v["grbm_gui_active"] / d["xcds"] * d["simds"]), 4)
            for k, v in d["forms"].items()}
    return {"mfma_busy": busy, "mfma_busy_source": f"profiles/{f.name}: {d['source']}"}


def pfb_traffic():
    """HBM bytes per launch of the filterbank kernel from the committed
    rocprofv3 FETCH_SIZE / WRITE_SIZE summary (each scaled by what the counter
    reads for a plain copy kernel of known size)."""
    f = ROOT / "profiles" / "k_pfb_fused_c2_traffic.json"
    if not f.exists():
        return None, None
    d = json.loads(f.read_text())
    b = (d["fetch_kib"] * d["fetch_scale"] + d["write_kib"] * d["write_scale"]) * 1024
    return round(b), f"profiles/{f.name}: {d['source']}"


def _ref_library():
    """The reference C library built from its sources into oracle/_ref (the
    CPU baseline only), or the reason it is absent."""
    try:
        from oracle.ref import Ref, available
    except Exception as e:                       # noqa: BLE001
        return None, f"unavailable: {e}"
    if not available():
        return None, "oracle/_ref/atkdsp_ref.so not built"
    return Ref(), None


def cpu_baseline_calls(r, base: np.ndarray, threads: int, calls: int = 5):
    """The reference C on the drop-in calls bench.drop_in times: one 50 ms
    ci16q11 block per call, ms per call on `threads` host threads."""
    from oracle.ref import Cf
    r.set_threads(threads)
    L = r.L
    b = np.ascontiguousarray(base).view(np.uint8).reshape(-1)
    x = np.empty(N_BLOCK, np.complex64)
    pfb = L.atkdsp_pfb_create(FS, M, OS, K, 80.0, 1)
    plan = L.atkdsp_fft_create(NFFT)
    win = r.window(1, NFFT)
    y = np.empty(M * (N_BLOCK // (M // OS) + 2), np.complex64)
    line = np.empty(NFFT, np.float32)
    cf = lambda a: a.ctypes.data_as(C.POINTER(Cf))
    fp = lambda a: a.ctypes.data_as(C.POINTER(C.c_float))
    mr, mi = C.c_double(), C.c_double()
    fns = {
        "unpack_dc": lambda: L.atkdsp_unpack_dc(b.ctypes.data, b.size, 3, cf(x), x.size, 0.0, 0.0,
                                                C.byref(mr), C.byref(mi)),
        "pfb_process": lambda: L.atkdsp_pfb_process_n(pfb, (C.POINTER(Cf) * 1)(cf(x)), 1, x.size, cf(y), y.size),
        "spectrum_reduce": lambda: L.atkdsp_spectrum_reduce(plan, cf(x), x.size, FHOP, fp(win), 0, fp(line)),
    }
    out = {}
    for name, fn in fns.items():
        fn()
        t0 = time.perf_counter()
        for _ in range(calls):
            fn()
        out[name + "_ms"] = round((time.perf_counter() - t0) * 1e3 / calls, 3)
    L.atkdsp_pfb_destroy(pfb)
    L.atkdsp_fft_destroy(plan)
    out["block_ms"] = round(sum(out.values()), 3)
    out["threads"] = threads
    return out


def cpu_baseline_osx(r, seconds: float = 3.0):
    """The reference C extractor on the calls bench.osx_extract times (one
    host thread: the reference's extractor is serial)."""
    from oracle.ref import RefOsx
    fs = 40e6
    x, offs = osx_signal()
    o = RefOsx(r, fs, 8192)
    for f in offs:
        o.add(float(f), 100e3, 125e3)
    t0, c = time.perf_counter(), 0
    while time.perf_counter() - t0 < seconds:
        o.process(x)
        c += 1
    cw = (time.perf_counter() - t0) / c
    return {"ms_per_call": round(cw * 1e3, 2), "Msamples_per_s": round(x.size / cw / 1e6, 1), "threads": 1,
            "calls": c}


def cpu_baseline_cfar(r, nbins: int = 4096, nframes: int = 512):
    """The reference C detector and tracker on the frames bench.cfar_frames
    times (one host thread: its detector is serial)."""
    fr = cfar_signal(nbins, nframes)
    out = {}
    for name, cfg in CFAR_CFGS:
        t0 = time.perf_counter()
        r.cfar_run(nbins, cfg, [("process", fr, 0.0, 0.01)])
        out[name] = {"us_per_frame": round((time.perf_counter() - t0) / nframes * 1e6, 2), "threads": 1}
    return out


def cpu_baseline(seconds: float, base: np.ndarray, calls: bool = True):
    """The reference C library on the host cores: the same chain per block,
    streaming, blocks of the same content, until `seconds` elapse; then (when
    `calls`) the drop-in calls, the extractor and the CFAR frames the GPU
    side times, each on the same inputs."""
    r, why = _ref_library()
    if r is None:
        return {"value": None, "unit": "Msamples/s", "cores": 0, "kind": "reference", "sample": why}
    from oracle.ref import Cf
    cores = min(16, os.cpu_count() or 1)
    r.set_threads(cores)
    L = r.L
    raw = np.ascontiguousarray(base).tobytes()
    pfb = L.atkdsp_pfb_create(FS, M, OS, K, 80.0, 1)
    plan = L.atkdsp_fft_create(NFFT)
    win = r.window(1, NFFT)
    x = np.empty(N_BLOCK, np.complex64)
    out = np.empty(M * (N_BLOCK // (M // OS) + 2), np.complex64)
    line = np.empty(NFFT, np.float32)
    cols = np.empty(COLS, np.float32)
    trace = np.full(TRACE, np.nan, np.float32)
    tail = np.zeros(0, np.complex64)
    off = [0.0, 0.0]
    b = np.frombuffer(raw, np.uint8)
    cf = lambda a: a.ctypes.data_as(C.POINTER(Cf))
    fp = lambda a: a.ctypes.data_as(C.POINTER(C.c_float))
    blocks = 0
    t0 = time.perf_counter()
    while True:
        mr, mi = C.c_double(), C.c_double()
        L.atkdsp_unpack_dc(b.ctypes.data, b.size, 3, cf(x), x.size, off[0], off[1], C.byref(mr), C.byref(mi))
        off = [mr.value, mi.value]
        ins = (C.POINTER(Cf) * 1)(cf(x))
        L.atkdsp_pfb_process_n(pfb, ins, 1, x.size, cf(out), out.size)
        d = np.concatenate([tail, x])
        fr = L.atkdsp_spectrum_reduce(plan, cf(d), d.size, FHOP, fp(win), 0, fp(line))
        tail = d[fr * FHOP:].copy()
        L.atkdsp_decimate_max(fp(line), NFFT, fp(cols), COLS)
        L.atkdsp_stitch_max(fp(cols), COLS, -FS / 2, FS / COLS, -FS / 2, FS / 2, -4 * FS, FS / COLS * 8 / 8,
                            fp(trace), TRACE)
        blocks += 1
        el = time.perf_counter() - t0
        if el >= seconds and blocks >= 2:
            break
    L.atkdsp_pfb_destroy(pfb)
    L.atkdsp_fft_destroy(plan)
    rate = blocks * N_BLOCK / el / 1e6
    res = {"value": round(rate, 2), "unit": "Msamples/s", "cores": cores, "kind": "reference",
           "sample": f"{blocks} consecutive 2,000,000-sample ci16q11 blocks of one stream through "
                     f"unpack_dc, pfb_process (1024 ch, 16 taps, 2x), spectrum_reduce (4096/2048 Hann max, "
                     f"carried tail), decimate_max, stitch_max: {el:.1f} s wall on {cores} threads"}
    if calls:
        print("[bench] reference C on the drop-in calls, the extractor and the CFAR frames", file=sys.stderr,
              flush=True)
        res["drop_in_calls"] = cpu_baseline_calls(r, base, cores)
        res["osx_extract"] = cpu_baseline_osx(r)
        res["cfar"] = cpu_baseline_cfar(r)
    return res


def serving(mix: str, streams: int, seconds: float, extra=()):
    """The serving loop under Poisson chunk arrivals at the streams' real rates
    (the native load generator beside the library), as one JSON object.
    extra: more serve_load options (--detect 1: a consumer finding channels on
    every display line on the device; --ring 1: chunks fed from per-stream
    block rings)."""
    import subprocess
    exe = ROOT / "atkdsp_amd" / "serve_load"
    if not exe.exists():
        return {"error": "atkdsp_amd/serve_load not built"}
    print(f"[bench] serving {mix} x{streams} for {seconds:.0f} s", file=sys.stderr, flush=True)
    r = subprocess.run([str(exe), "--mix", mix, "--streams", str(streams), "--seconds", str(seconds), *extra],
                       capture_output=True, text=True, timeout=seconds + 120)
    if r.returncode != 0:
        return {"error": f"serve_load exited {r.returncode}: {r.stderr.strip()[-300:]}"}
    return json.loads(r.stdout.strip().splitlines()[-1])


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--streams", type=int, default=128)
    ap.add_argument("--cpu-seconds", type=float, default=8.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--serve-seconds", type=float, default=10.0, help="0 skips the serving runs")
    ap.add_argument("--no-extras", action="store_true",
                    help="skip the precision-form comparison and the NCO phase run")
    ap.add_argument("--cfar-only", action="store_true",
                    help="only time the CFAR detector and tracker (and the reference C beside it) and exit")
    ap.add_argument("--only", choices=["filterbank_forms_wideband", "precision_form", "drop_in", "osx_extract",
                                       "nco_phase_run", "coherent_array"],
                    help="only run this one measured section (for profiling it alone), print it and exit")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="only run a plain 1 GiB -> 1 GiB copy kernel 5 times (the reference point for "
                         "the FETCH_SIZE / WRITE_SIZE counters) and exit")
    args = ap.parse_args()
    if args.copy_calibration:
        import torch
        a = torch.ones(1 << 28, dtype=torch.float32, device="cuda")
        b = torch.empty_like(a)
        for _ in range(5):
            torch.mul(a, 1.0, out=b)
        torch.cuda.synchronize()
        print(json.dumps({"copy_calibration": {"bytes_read": a.numel() * 4, "bytes_written": b.numel() * 4,
                                               "launches": 5}}))
        return 0
    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU; --gpus {args.gpus} is not supported "
              "(streams are independent: several GPUs would each take a share of them)", file=sys.stderr)
        return 2
    import atkdsp_amd
    from atkdsp_amd.chain import Chain
    if args.only:
        if args.only in ("filterbank_forms_wideband", "drop_in"):
            from atkdsp_amd import synth
            base = np.frombuffer(synth.c2_bytes(n=N_BLOCK), "<i2").reshape(-1, 2)
            res = filterbank_forms_c2(base) if args.only == "filterbank_forms_wideband" else drop_in(base)
        else:
            res = {"precision_form": precision_forms, "osx_extract": osx_extract, "nco_phase_run": nco_phase_run,
                   "coherent_array": coherent_array}[args.only]()
        print(json.dumps({args.only: res}))
        return 0
    if args.cfar_only:
        r, why = _ref_library()
        print(json.dumps({"cfar": cfar_frames(), "cpu_baseline_cfar": cpu_baseline_cfar(r) if r else why}))
        return 0

    S = args.streams
    print(f"[bench] generating {S} streams of {N_BLOCK} samples", file=sys.stderr, flush=True)
    base, streams = make_streams(S)
    ch = Chain(fmt="ci16q11", fs=FS, nchan=M, oversample=OS, taps_per_poly=K, atten_db=80.0, nfft=NFFT,
               fft_hop=FHOP, window="hann", detector="max", line_cols=COLS, trace_bins=TRACE,
               fc0_hz=0.0, fc_step_hz=FS, trace_lo_hz=-FS / 2, trace_hz_per_bin=FS / 1024 * 8 / 8 * 4,
               dc_mode=1, streams=S, block_max=N_BLOCK)
    for s in range(S):
        ch.upload(s, streams[s])
    ch.sync()
    print(f"[bench] uploaded; warmup {args.warmup}", file=sys.stderr, flush=True)
    # the timed steps record no timing events (each costs the stream ~5 us);
    # the per-kernel loop below turns them back on
    ch.set_timing(False)
    for _ in range(args.warmup):
        ch.run(N_BLOCK)
    ch.sync()
    # the timed region: exactly `steps` blocks of every stream
    t0 = time.perf_counter()
    for _ in range(args.steps):
        m = ch.run(N_BLOCK)
    ch.sync()
    t1 = time.perf_counter()
    ms_step = (t1 - t0) * 1e3 / args.steps
    print(f"[bench] timed {args.steps} steps: {ms_step:.3f} ms/step", file=sys.stderr, flush=True)
    # per-kernel device times (HIP events on the chain's stream), one step at a time
    ch.set_timing(True)
    runs, pfbs, disps, lat = [], [], [], []
    for _ in range(args.steps):
        ts = time.perf_counter()
        ch.run(N_BLOCK)
        a, b, d = ch.timing()           # waits for the step
        lat.append((time.perf_counter() - ts) * 1e3)
        runs.append(a)
        pfbs.append(b)
        disps.append(d)
    ms_pfb = float(np.mean(pfbs))
    traffic, traffic_src = pfb_traffic()
    pfb_b = pfb_bytes_per_launch(S, N_BLOCK, m)
    achieved = pfb_b / (ms_pfb * 1e-3) / 1e9
    value = S * N_BLOCK / (ms_step * 1e-3)
    del ch
    serve = {}
    if args.serve_seconds > 0:
        serve["uniform"] = serving("uniform", S, args.serve_seconds, ("--detect", "1"))
        serve["c4"] = serving("c4", 128, 3 * args.serve_seconds)
        serve["uniform_from_rings"] = serving("uniform", S, args.serve_seconds, ("--ring", "1"))
    extras = {}
    if not args.no_extras:
        print("[bench] drop-in calls with host buffers", file=sys.stderr, flush=True)
        extras["drop_in"] = drop_in(base)
        extras["filterbank_forms_wideband"] = filterbank_forms_c2(base)
        extras["precision_form"] = precision_forms()
        extras["nco_phase_run"] = nco_phase_run()
        extras["coherent_array"] = coherent_array()
        extras["osx_extract"] = osx_extract()
        extras["cfar"] = cfar_frames()
    cpu = None if args.no_cpu else cpu_baseline(args.cpu_seconds, streams[0], calls=not args.no_extras)
    if cpu and "drop_in" in extras and "drop_in_calls" in cpu:
        extras["drop_in"]["vs_reference_c"] = {
            k: round(cpu["drop_in_calls"][k] / extras["drop_in"][k], 2)
            for k in ("unpack_dc_ms", "pfb_process_ms", "spectrum_reduce_ms", "block_ms")}
    uni = serve.get("uniform", {})
    out = {
        "metric": "channelized Msamples/s at M=1024 (1 GPU, %HBM peak); p99 chunk latency ms",
        "value": round(value / 1e6, 1),
        "unit": "Msamples/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_step, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": f"40 MSPS ci16q11 2,000,000-sample blocks x {S} streams: unpack+DC, "
                               f"PFB M=1024 K=16 2x, 4096/2048 Hann max line, 1024 cols, 8192-bin stitch",
                   "streams": S, "block_samples": N_BLOCK, "nchan": M, "taps_per_poly": K, "oversample": OS,
                   "nfft": NFFT, "fft_hop": FHOP, "parallelism": "single GPU"},
        "p50_chunk_latency_ms": uni.get("p50_ms"),
        "p99_chunk_latency_ms": uni.get("p99_ms"),
        "latency_note": (f"serving loop, {S} streams of 40 MSPS ci16 into M=1024 K=16 2x filterbanks and "
                         "4096-point display lines (hop 2048, Hann, max-hold over 50 ms, 1024 columns), chunks "
                         "of 1..65536 samples arriving as Poisson processes at the real rate; latency = host "
                         "submit call to the device completion of the chunk's batch (its filterbank output "
                         "and every display line it completes), host monotonic clock; a consumer finds the "
                         "channels of every completed line on the device meanwhile (serving.uniform.detect_*)"),
        "step_latency_ms": {"p50": round(float(np.percentile(lat, 50)), 3),
                            "p99": round(float(np.percentile(lat, 99)), 3),
                            "note": "one chained step (a 50 ms block of every stream) submitted to completed"},
        "serving": serve,
        "kernel_ms": {"run": round(float(np.mean(runs)), 4), "pfb_fused": round(ms_pfb, 4),
                      "display_with_unpack": round(float(np.mean(disps)), 4),
                      "note": "HIP events on the chain's stream: run = the whole step; display_with_unpack = "
                              "the display reduction that reads the raw blocks and writes the unpacked rows "
                              "on the way, with the block mean; the rest of run is the line post-processing"},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic,
                     "traffic_source": traffic_src,
                     "kernel": "k_pfb_fused<1024,2,32>",
                     "bytes_per_launch": pfb_b},
        "chain_roofline": chain_roofline(value),
        "cpu_baseline": cpu,
        **extras,
        "build": atkdsp_amd.build_info(),
    }
    print(json.dumps(out))
    return 0


if __name__ == "__main__":
    sys.exit(main())
