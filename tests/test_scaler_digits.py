"""CPU check of the matrix-core scaler's digit tables: tests/native/
scaler_digits_check.cpp is built against the library's host code and the
oracle, and reproduces the kernel's exact integer arithmetic on the tables
for every scaler shape and both MFMA depths, output by output against the
oracle's RnsScaler::scale (math/src/rns/scaler.rs:249-352)."""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")


def test_scaler_digit_tables_match_oracle(tmp_path):
    lib = os.path.join(ROOT, "fhe.rs_amd", "lib")
    olib = os.path.join(ROOT, "oracle", "lib")
    if not os.path.exists(HIPCC) or not os.path.exists(os.path.join(lib, "libfhehip.so")):
        pytest.skip("needs hipcc and the built library")
    exe = str(tmp_path / "scaler_digits_check")
    subprocess.check_call([HIPCC, "-O2", "-std=c++17", "-I", os.path.join(ROOT, "fhe.rs_amd", "csrc"),
                           "-I", os.path.join(ROOT, "include"), "-o", exe,
                           os.path.join(ROOT, "tests", "native", "scaler_digits_check.cpp"),
                           "-L", lib, "-lfhehip", "-L", olib, "-loracle",
                           "-Wl,-rpath," + lib, "-Wl,-rpath," + olib])
    out = subprocess.run([exe], capture_output=True, text=True, timeout=600)
    assert out.returncode == 0 and out.stdout.strip().endswith("OK"), out.stdout[-3000:] + out.stderr[-2000:]
