"""The tile pass's row-interval inside bits (row_inside_bits in sdfibm_amd/csrc/shapes.h) equal
the exact per-vertex test on 400,000 random rows of quadric solids, surface vertices included:
the device code's header compiled for the host (hipcc, FMA contraction off as in the library)."""
from __future__ import annotations

import os
import shutil
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))


@pytest.mark.skipif(shutil.which("hipcc") is None, reason="hipcc not found")
def test_row_inside_bits_equal_exact_test(tmp_path):
    exe = str(tmp_path / "row_inside_check")
    src = os.path.join(HERE, "host", "row_inside_check.cpp")
    subprocess.run(["hipcc", "-O2", "-std=c++17", "-ffp-contract=off", "-w", "-o", exe, src], check=True,
                   capture_output=True, timeout=300)
    r = subprocess.run([exe, "400000"], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "mismatches 0" in r.stdout
