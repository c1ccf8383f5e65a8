"""The C-ABI library and the C++ drop-in headers, without a GPU: the library
loads, exports every function the header declares, the host-only entry points
agree with the oracle, and drivers written like the reference's compile and
link against include/ of this repository."""
import ctypes as C
import os
import re
import shutil
import subprocess

import numpy as np
import pytest

import _oracle as O

ROOT = O.ROOT
HEADER = os.path.join(ROOT, "include", "mantiis_hip.h")
LIB = os.path.join(ROOT, "mantiis_amd", "libmantiis_hip.so")


def declared():
    txt = open(HEADER).read()
    return sorted(set(re.findall(r"^(?:int|void|const char\*)\s+(mts_\w+)\s*\(", txt, re.M)))


def test_library_exports_every_declared_symbol():
    from mantiis_amd.hip import SYMBOLS, load_library
    lib = load_library()
    names = declared()
    assert len(names) >= 20
    for n in names:
        assert hasattr(lib, n), n
    out = subprocess.run(["nm", "-D", "--defined-only", LIB], capture_output=True, text=True).stdout
    exported = set(re.findall(r" T (mts_\w+)", out))
    assert set(names) <= exported
    assert set(SYMBOLS) <= exported


def test_host_globals_match_oracle():
    from mantiis_amd.hip import globals_for
    lib = O.load()
    for dx, T, P in [(1e-9, 300.0, 2e6), (12e-9, 300.0, 2e6), (5e-10, 320.0, 3e6)]:
        rc, g = globals_for(dx=dx, T=T, P=P)
        assert rc == 0
        out = np.zeros(10)
        lib.orc_globals_for(dx, T, P, out.ctypes.data_as(O._D))
        for i, k in enumerate(["rhoin", "mfp", "Cf", "T_lu", "a_alpha_lu", "ZG", "Crho", "Cu"]):
            assert g[k] == out[i], k


def test_host_globals_reject_liquid_like_state():
    """Three real roots (subcritical, high pressure): the reference's Cardano
    branch yields no root (utils.hpp:582-599); the library reports it."""
    from mantiis_amd.hip import globals_for
    rc, g = globals_for(T=150.0, P=4e6)
    assert rc == -9 and g["cubic_ok"] == 0.0


def test_load_existing_model_matches_fixture():
    from mantiis_amd.hip import load_existing_model
    pore, kn, lp = load_existing_model(O.FIXTURES + "/", 64)
    p2, k2, l2 = O.load_fixture_geometry()
    assert np.array_equal(pore, p2) and np.array_equal(kn, k2) and np.array_equal(lp, l2)


def test_load_existing_model_errors(tmp_path):
    from mantiis_amd.hip import MtsError, load_existing_model
    with pytest.raises(MtsError) as e:
        load_existing_model(str(tmp_path) + "/", 64)
    assert e.value.code == -3 and "File not found" in str(e.value)
    for name in ("pore.dat", "Kn.dat", "localporesize.dat"):
        shutil.copy(os.path.join(O.FIXTURES, name), tmp_path / name)
    with open(tmp_path / "Kn.dat", "w") as f:
        f.write("0\n1\n")
    with pytest.raises(MtsError) as e:
        load_existing_model(str(tmp_path) + "/", 64)
    assert e.value.code == -3 and "Insufficient data" in str(e.value)


@pytest.mark.parametrize("driver", ["app_driver", "phase_driver", "shape_driver"])
def test_drivers_compile_and_link(driver, tmp_path):
    src = os.path.join(ROOT, "tests", "drivers", driver + ".cpp")
    exe = tmp_path / driver
    r = subprocess.run(["g++", "-std=c++17", "-O1", "-Wall", "-I" + ROOT, "-I" + os.path.join(ROOT, "include"), src,
                        "-L" + os.path.join(ROOT, "mantiis_amd"), "-lmantiis_hip",
                        "-Wl,-rpath," + os.path.join(ROOT, "mantiis_amd"), "-o", str(exe)],
                       capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    # argc < 6 exits with the usage line (mantiisApp.hpp:131-135)
    if driver == "app_driver":
        r = subprocess.run([str(exe), "type=transport", "--", "1"], capture_output=True, text=True)
        assert r.returncode == 1 and "Usage:" in r.stderr
        r = subprocess.run([str(exe), "nonsense=1", "--"], capture_output=True, text=True)
        assert r.returncode == 2 and "bad setting" in r.stderr
