"""Build atkdsp_amd/atkdsp.so for gfx950.

Device code (csrc/dev/*.hip) is compiled by hipcc for gfx950 only; the host
layer (csrc/host/*.c) is C11 compiled by gcc; the library also exports every
other name of the reference ABI as a stub (generated from csrc/abi/symbols.txt)
so the reference ctypes binding, which binds every name eagerly, loads it.

Usage: python -m atkdsp_amd.build   (or build() from Python)
"""
from __future__ import annotations

import os
import re
import shutil
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor
from pathlib import Path

PKG = Path(__file__).resolve().parent
ROOT = PKG.parent
CSRC = PKG / "csrc"
BUILD = PKG / "_build"
LIB = PKG / "atkdsp.so"

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"
# packed FP32 instructions are switched off (the compiler would otherwise form
# them from ordinary float code); no fast-math anywhere: NaN means "unmeasured"
HIP_FLAGS = ["-O3", f"--offload-arch={ARCH}", "-fPIC", "-std=c++17", "-fno-fast-math",
             "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
             "-Wall", "-Wno-unused-function", "-Wno-unused-variable",
             "-I", str(CSRC / "dev"), "-I", str(ROOT / "include")]
C_FLAGS = ["-std=c11", "-O2", "-fPIC", "-Wall", "-Wextra", "-Wno-unused-parameter",
           "-fno-fast-math", "-D__HIP_PLATFORM_AMD__", "-I", str(ROOT / "include"),
           "-I", str(CSRC / "dev"), "-I", str(CSRC / "host")]

# the entry points this library implements; every other ABI name is a stub
IMPLEMENTED = {
    "atkdsp_abi_version", "atkdsp_version_string", "atkdsp_build_info",
    "atkdsp_bytes_per_sample", "atkdsp_set_threads", "atkdsp_get_threads",
    "atkdsp_unpack", "atkdsp_unpack_dc",
    "atkdsp_fft_create", "atkdsp_fft_destroy", "atkdsp_fft_length", "atkdsp_fft_exec",
    "atkdsp_window", "atkdsp_power_db", "atkdsp_spectrum_reduce", "atkdsp_spectrum_stats",
    "atkdsp_window_stats", "atkdsp_db_to_pixels", "atkdsp_decimate_max", "atkdsp_stitch_max",
    "atkdsp_pfb_create", "atkdsp_pfb_destroy", "atkdsp_pfb_reset", "atkdsp_pfb_nchan",
    "atkdsp_pfb_out_rate", "atkdsp_pfb_out_hops", "atkdsp_pfb_process", "atkdsp_pfb_process_n",
    "atkdsp_pfb_block_origin", "atkdsp_pfb_group_delay", "atkdsp_pfb_set_cal",
    "atkdsp_pfb_describe", "atkdsp_pfb_cov_accumulate", "atkdsp_pfb_cov_reset",
    "atkdsp_pfb_channel_power",
    "atkdsp_nco_init", "atkdsp_nco_set_freq", "atkdsp_nco_mix",
    "atkdsp_arb_resampler_create", "atkdsp_arb_resampler_destroy", "atkdsp_arb_resampler_reset",
    "atkdsp_arb_resampler_set_ratio", "atkdsp_arb_resampler_ratio", "atkdsp_arb_resampler_out_max",
    "atkdsp_arb_resampler_process", "atkdsp_design_lowpass",
    "atkdsp_osx_create", "atkdsp_osx_destroy", "atkdsp_osx_reset", "atkdsp_osx_add", "atkdsp_osx_retune",
    "atkdsp_osx_remove", "atkdsp_osx_channels", "atkdsp_osx_chan_rate", "atkdsp_osx_out_max",
    "atkdsp_osx_process", "atkdsp_osx_process_n", "atkdsp_osx_describe",
    "atkdsp_ring_bytes", "atkdsp_ring_init", "atkdsp_ring_attach", "atkdsp_ring_claim", "atkdsp_ring_commit",
    "atkdsp_ring_open", "atkdsp_ring_close", "atkdsp_ring_read", "atkdsp_ring_peek", "atkdsp_ring_release",
    "atkdsp_ring_copy_range", "atkdsp_ring_stats", "atkdsp_ring_describe",
    "atkdsp_median", "atkdsp_detect_channels",
    "atkdsp_cfar_defaults", "atkdsp_cfar_alpha", "atkdsp_cfar_n_eff", "atkdsp_cfar_window_rho2",
    "atkdsp_cfar_create", "atkdsp_cfar_destroy", "atkdsp_cfar_reset", "atkdsp_cfar_set_mask",
    "atkdsp_cfar_events_max", "atkdsp_cfar_process", "atkdsp_cfar_flush", "atkdsp_cfar_last",
    "atkdsp_cfar_info", "atkdsp_cfar_describe", "atkdsp_cfar_stats",
}

_STUB_BODY = {
    "void": "",
    "neg": "return -1; /* ATKDSP_E_ARG */",
    "zero": "return 0;",
    "ptr": "return 0; /* NULL */",
    "double": "return __builtin_nan(\"\");",
    "float": "return __builtin_nanf(\"\");",
}
_STUB_TYPE = {"void": "void", "neg": "long", "zero": "unsigned long", "ptr": "void *",
              "double": "double", "float": "float"}


def abi_symbols() -> list[tuple[str, str]]:
    out = []
    for line in (CSRC / "abi" / "symbols.txt").read_text().splitlines():
        if line.strip():
            name, kind = line.split()
            out.append((name, kind))
    return out


def write_stubs(path: Path) -> None:
    """Stubs for the ABI names outside the wideband front end. A stub takes no
    arguments (the C calling convention lets a caller pass any) and returns the
    reference's failure value for its return type: NULL handles, ATKDSP_E_ARG
    (-1) for signed counts, 0 for unsigned counts, NaN for floating results."""
    lines = ["/* generated by atkdsp_amd/build.py from csrc/abi/symbols.txt */",
             "#define STUB __attribute__((visibility(\"default\")))"]
    for name, kind in abi_symbols():
        if name in IMPLEMENTED:
            continue
        body = _STUB_BODY[kind]
        lines.append(f"STUB {_STUB_TYPE[kind]} {name}(void) {{ {body} }}")
    path.write_text("\n".join(lines) + "\n")


def _run(cmd: list[str]) -> None:
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(" ".join(cmd) + "\n" + r.stdout + r.stderr)
        raise RuntimeError(f"compile failed: {cmd[-1]}")


def build(verbose: bool = False) -> Path:
    if not Path(HIPCC).exists() and not shutil.which("hipcc"):
        raise RuntimeError("hipcc not found: the HIP library cannot be built")
    BUILD.mkdir(exist_ok=True)
    stubs = BUILD / "stubs.c"
    write_stubs(stubs)
    jobs = []
    for src in sorted((CSRC / "dev").glob("*.hip")):
        obj = BUILD / (src.stem + ".hip.o")
        jobs.append([HIPCC, *HIP_FLAGS, "-c", "-o", str(obj), str(src)])
    for src in sorted((CSRC / "host").glob("*.c")) + [stubs]:
        obj = BUILD / (src.stem + ".c.o")
        jobs.append(["gcc", *C_FLAGS, "-c", "-o", str(obj), str(src)])
    with ThreadPoolExecutor(max_workers=min(8, os.cpu_count() or 4)) as ex:
        list(ex.map(_run, jobs))
    objs = [j[j.index("-o") + 1] for j in jobs]
    tmp = LIB.with_suffix(".so.tmp")
    _run([HIPCC, "-shared", f"--offload-arch={ARCH}", "-fPIC", "-o", str(tmp), *objs,
          "-lpthread", "-lm"])
    os.replace(tmp, LIB)
    # the serving-loop load generator, linked against the library beside it
    _run(["gcc", "-std=c11", "-O2", "-Wall", "-I", str(ROOT / "include"), "-o", str(PKG / "serve_load"),
          str(CSRC / "load" / "serve_load.c"), "-L", str(PKG), "-l:atkdsp.so", "-Wl,-rpath,$ORIGIN",
          "-lpthread", "-lm"])
    if verbose:
        print(f"built {LIB}")
    return LIB


def exported_names(lib: Path = LIB) -> set[str]:
    r = subprocess.run(["nm", "-D", "--defined-only", str(lib)], capture_output=True, text=True,
                       check=True)
    return {ln.split()[-1] for ln in r.stdout.splitlines() if re.search(r" T ", ln)}


if __name__ == "__main__":
    build(verbose=True)
