"""CPU check of the banded scan's shared-slot arithmetic (the functions in
slime_amd/csrc/slime_math.h that scan_band_kernel calls) against the oracle's
first-draw predicate over random seeds and full-range int32 coordinates: one
test per band must reproduce next(31) % 10 == 0 of the selected x-phase and
the other phase must never be a first-draw hit."""
import os
import subprocess

from conftest import ROOT


def test_band_slot_arithmetic_matches_oracle(oracle, tmp_path):
    exe = tmp_path / "band_math_check"
    subprocess.run(["g++", "-O2", "-std=c++17", "-I", os.path.join(ROOT, "slime_amd", "csrc"),
                    os.path.join(ROOT, "tests", "band_math_check.cpp"),
                    os.path.join(ROOT, "oracle", "liboracle.so"),
                    "-Wl,-rpath," + os.path.join(ROOT, "oracle"), "-o", str(exe)], check=True)
    out = subprocess.run([str(exe), "100000"], check=True, capture_output=True, text=True).stdout
    chunks, hits, bad = map(int, out.split())
    assert chunks == 100000 * 32
    assert bad == 0
    assert 0.09 < hits / chunks < 0.11
