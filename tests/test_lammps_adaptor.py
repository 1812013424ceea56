"""The LAMMPS adaptor (pair_style tensormd/amd) compiles against the LAMMPS
headers of the reference tree, with LAMMPS' own MPI stubs.  Syntax check
only: building LAMMPS itself is out of scope.  Skipped where the reference
tree is absent."""
import os
import shutil
import subprocess

import pytest

from conftest import ROOT

LMP_SRC = "/root/reference/src"


@pytest.mark.skipif(not os.path.exists(os.path.join(LMP_SRC, "pair.h")) or
                    shutil.which("g++") is None, reason="LAMMPS headers not available")
@pytest.mark.parametrize("src", ["pair_tensormd_amd.cpp", "compute_tensormd_density_amd.cpp",
                                 "compute_tensormd_tdpe_amd.cpp", "fix_tensormd_etemp_amd.cpp",
                                 "fix_nve_tensormd_amd.cpp", "tensormd_amd_plugin.cpp"])
def test_adaptor_compiles(src):
    here = os.path.join(ROOT, "lammps-tensormd_amd", "lammps")
    cmd = ["g++", "-std=c++17", "-fsyntax-only", "-I" + LMP_SRC, "-I" + os.path.join(LMP_SRC, "STUBS"),
           "-I" + os.path.join(ROOT, "include"), "-I" + here, os.path.join(here, src)]
    r = subprocess.run(cmd, capture_output=True, text=True)
    assert r.returncode == 0, r.stderr[-2000:]
