"""The reference's own programs build unchanged against this repository.

mantiis.cpp:1-23 and test/{transport,adsorption}_test.cpp of the reference
include "include/mantiisApp.hpp" and use Settings / MantiisApp / init() /
run().  Each file is compiled where it lies (read from the reference tree, fed
to the compiler on stdin from an empty directory, so that its quoted include
resolves to this repository's include/ and not to the reference's) and linked
against the HIP library.  Nothing is run: the GPU runs of the same Settings
are tests/test_gpu_dropin.py.  Skipped where the reference tree is absent.
"""
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
REF = "/root/reference"
LIB = os.path.join(ROOT, "mantiis_amd", "libmantiis_hip.so")

PROGRAMS = ["mantiis.cpp", "test/transport_test.cpp", "test/adsorption_test.cpp"]


@pytest.mark.skipif(not os.path.isdir(REF), reason="reference tree absent")
@pytest.mark.parametrize("prog", PROGRAMS)
def test_reference_program_builds_unchanged(prog, tmp_path):
    src = os.path.join(REF, prog)
    assert os.path.isfile(src)
    cmd = ["g++", "-std=c++17", "-O1", "-x", "c++", "-I" + ROOT, "-I" + os.path.join(ROOT, "include"), "-",
           "-o", str(tmp_path / "prog")]
    if os.path.exists(LIB):
        cmd += ["-L" + os.path.dirname(LIB), "-lmantiis_hip", "-Wl,-rpath," + os.path.dirname(LIB)]
    else:  # library not built: compile only
        cmd += ["-c"]
    with open(src, "rb") as f:
        r = subprocess.run(cmd, stdin=f, cwd=str(tmp_path), capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr[-2000:]
    assert (tmp_path / "prog").exists()

