"""CPU check of band_tile_of, the map from a center to the tile of a band job
that decides it (slime_amd/csrc/search_util.h): after a cancel the engines
apply the exact recounts of centers near Java-rejected chunks to exactly the
centers their finished tiles decided. See tests/tile_geometry_check.cpp."""
import os
import shutil
import subprocess

import pytest

from conftest import ROOT

HIPCC = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"


@pytest.mark.skipif(not os.path.exists(HIPCC), reason="hipcc absent")
def test_band_tile_of_inverts_the_kernel_tiling(tmp_path):
    exe = tmp_path / "tile_geometry_check"
    subprocess.run([HIPCC, "-O2", "-std=c++17", "-I", os.path.join(ROOT, "slime_amd", "csrc"),
                    os.path.join(ROOT, "tests", "tile_geometry_check.cpp"), "-o", str(exe)],
                   check=True, capture_output=True)
    centers, tiles, bad = map(int, subprocess.run([str(exe)], check=True, capture_output=True,
                                                  text=True).stdout.split())
    assert centers == 400 * 300 and tiles > 40000
    assert bad == 0
