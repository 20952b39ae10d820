"""The register FFT of the pair-density kernels (respackdev_amd/csrc/rp_fft.h:
mixed radix 4, 2, 3, 5, roots read from the twiddle table) against a
long-double direct DFT on the CPU, for every line length the kernels support
(products of 2, 3 and 5 up to 32) and both signs.  The header is compiled for
the host by hipcc (tests/fft_host_check.hip); no GPU is used."""
import os
import shutil
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")


@pytest.mark.skipif(not os.path.exists(HIPCC) and shutil.which("hipcc") is None,
                    reason="hipcc not available")
def test_register_fft_all_lengths(tmp_path):
    exe = tmp_path / "fft_check"
    subprocess.run([HIPCC, "-O2", "-std=c++17", os.path.join(HERE, "fft_host_check.hip"), "-o",
                    str(exe)], check=True, capture_output=True)
    out = subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout
    rows = [line.split() for line in out.strip().splitlines()]
    sizes = [int(r[0]) for r in rows]
    assert sizes == [1, 2, 3, 4, 5, 6, 8, 9, 10, 12, 15, 16, 18, 20, 24, 25, 27, 30, 32]
    for n, em, ep in rows:
        print(f"N={n}: {em} {ep}")
        assert float(em) <= 1e-15 and float(ep) <= 1e-15, n
