"""CPU check of the Java-rejection correction used by the banded search: the
mismatch table and the per-row inversion (the code the GPU runs, compiled for
the host) against the oracle. See tests/mismatch_check.cpp."""
import os
import subprocess

from conftest import ROOT


def test_mismatch_table_and_row_search(oracle, tmp_path):
    exe = tmp_path / "mismatch_check"
    subprocess.run(["g++", "-O2", "-std=c++17", "-I", os.path.join(ROOT, "slime_amd", "csrc"),
                    os.path.join(ROOT, "tests", "mismatch_check.cpp"),
                    os.path.join(ROOT, "oracle", "liboracle.so"),
                    "-Wl,-rpath," + os.path.join(ROOT, "oracle"), "-o", str(exe)], check=True)
    out = subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout
    size, java_hits, windows, equal, planted = map(int, out.split())
    assert size == 209712
    # rejected-then-hit states, less the hits among next(31) = 2^31-8 (which the
    # first draw already counts as hits): the reference's 104,848 minus 13,104
    assert java_hits == 104848 - 13104
    assert windows == 400 and equal == 400 and planted == 400
