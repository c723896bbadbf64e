"""CPU check of the device FFT's index arithmetic: the team FFT stages
(atkdsp_amd/csrc/dev/fft_core.hpp, __host__ __device__) run on the host with
one loop iteration per thread and the LDS as an array, against an FP64 DFT."""
import shutil
import subprocess

import pytest

from conftest import ROOT


@pytest.mark.skipif(shutil.which("g++") is None, reason="needs g++")
def test_team_fft_emulation_matches_dft(tmp_path):
    exe = tmp_path / "emul"
    subprocess.run(["g++", "-std=c++17", "-O2", "-D__HIP_PLATFORM_AMD__", "-I/opt/rocm/include",
                    "-I", str(ROOT / "atkdsp_amd" / "csrc" / "dev"),
                    str(ROOT / "tests" / "cpu_emul" / "team_fft_emul.cpp"), "-o", str(exe)], check=True)
    out = subprocess.run([str(exe)], capture_output=True, text=True, check=True).stdout
    rows = [ln.split() for ln in out.splitlines()]
    sizes = [r for r in rows if not r[0].startswith("tw")]
    assert [int(r[0]) for r in sizes] == [64, 128, 256, 512, 1024, 2048, 4096]
    for n, err in sizes:
        assert float(err) < 3e-7, (n, err)
    # the register-twiddle stages equal the table-twiddle stages bit for bit
    assert [r for r in rows if r[0].startswith("tw")] == [["tw4096", "0"], ["tw256", "0"]]
