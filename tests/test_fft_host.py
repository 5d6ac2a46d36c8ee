"""FFT index math on the CPU: the __host__ __device__ building blocks of the
kernels (Stockham stages, register DFTs, the two-level line split) compiled
for the host with hipcc and checked against a direct double-precision DFT
(tests/native/fft_host_check.hip)."""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HIPCC = "/opt/rocm/bin/hipcc"


@pytest.mark.skipif(not os.path.exists(HIPCC), reason="hipcc not available")
def test_fft_building_blocks_on_host(tmp_path):
    exe = tmp_path / "fft_host_check"
    subprocess.run([HIPCC, "--offload-arch=gfx950", "-O2", "-std=c++17",
                    "-I", os.path.join(ROOT, "iqm_amd", "csrc"), "-I", os.path.join(ROOT, "include"),
                    "-o", str(exe), os.path.join(ROOT, "tests", "native", "fft_host_check.hip")],
                   check=True, capture_output=True, timeout=300)
    p = subprocess.run([str(exe)], capture_output=True, text=True, timeout=120)
    print(p.stdout)
    assert p.returncode == 0, p.stdout
