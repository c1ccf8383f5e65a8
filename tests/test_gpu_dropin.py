"""MantiisApp through this repository's include/ (tests/drivers/app_driver.cpp
takes the Settings from its command line), run the way test/run_tests.sh runs
the reference's regression cases (1 thread, single-block, 64, 100 steps, the
Settings of test/transport_test.cpp:9-16 and test/adsorption_test.cpp:10-17);
outputs compared with the gold files at the reference's criterion (max-abs
1e-10, scripts/compare_files.py:16-19)."""
import os
import shutil
import subprocess

import numpy as np
import pytest

import _oracle as O
from _gpu_util import banner

pytestmark = pytest.mark.gpu
ROOT = O.ROOT


def build(driver, tmp_path):
    exe = tmp_path / driver
    subprocess.run(["g++", "-std=c++17", "-O2", "-I" + ROOT, "-I" + os.path.join(ROOT, "include"),
                    os.path.join(ROOT, "tests", "drivers", driver + ".cpp"), "-L" + os.path.join(ROOT, "mantiis_amd"),
                    "-lmantiis_hip", "-Wl,-rpath," + os.path.join(ROOT, "mantiis_amd"), "-o", str(exe)], check=True)
    return exe


def fixture_copy(tmp_path):
    d = tmp_path / "io"
    d.mkdir()
    for name in ("pore.dat", "Kn.dat", "localporesize.dat"):
        shutil.copy(os.path.join(O.FIXTURES, name), d / name)
    return str(d) + "/"


# Settings of the reference's regression drivers (test/transport_test.cpp:9-16,
# test/adsorption_test.cpp:10-17)
GOLD_SETTINGS = {
    "transport": ["type=transport", "verbose=10", "T=300", "P=2e6", "dx=1e-9", "fbody=1e-8", "macro_every=1000000",
                  "dist_first=-1"],
    "adsorption": ["type=adsorption", "verbose=10", "T=300", "P=2e6", "dx=12e-9", "fbody=1e-8",
                   "macro_every=1000000", "dist_first=-1"],
}


@pytest.mark.parametrize("kind", ["transport", "adsorption"])
def test_app_reproduces_gold(kind, tmp_path):
    banner(f"drop-in MantiisApp {kind} 64x64 100 steps")
    exe = build("app_driver", tmp_path)
    folder = fixture_copy(tmp_path)
    r = subprocess.run([str(exe), *GOLD_SETTINGS[kind], "--", "1", "False", "64", folder, "100"],
                       capture_output=True, text=True, timeout=120)
    print(r.stdout[-600:])
    assert r.returncode == 0, r.stderr
    for name in ("ux", "uy", "rho"):
        a = np.loadtxt(folder + name + ".txt")
        g = np.loadtxt(os.path.join(O.FIXTURES, "gold", kind, name + ".txt"))
        err = np.abs(a - g).max()
        print(name, err)
        assert err <= 1e-10, name
    conv = np.loadtxt(folder + "convergence.txt")
    assert conv.size == 100
    assert os.path.exists(folder + "f.txt") and os.path.exists(folder + "SimulationDetails.txt")
    assert os.path.exists(folder + "ux_0.txt")  # dump at t = 0 (mantiisApp.hpp:328)
    if kind == "transport":
        # FP64 permeability report against the written fields and the oracle's
        # relaxation times: U_D = sum u_x / N^2, k = nu U_D / F
        rep = dict(line.split() for line in open(folder + "permeability.txt"))
        rep = {k: float(v) for k, v in rep.items()}
        ux = np.loadtxt(folder + "ux.txt")
        rho = np.loadtxt(folder + "rho.txt")
        pore, kn, lp = O.load_fixture_geometry()
        o = O.Oracle(pore, kn, lp, dx=1e-9, Fbody=1e-8)
        tau = o.get("taus").mean()
        o.close()
        UD = ux.sum() / ux.size
        print("permeability report", rep)
        # the report is text at 15 significant digits
        assert abs(rep["darcy_velocity"] - UD) <= 1e-12 * abs(UD)
        assert abs(rep["flux"] - (rho * ux).sum()) <= 1e-12 * abs((rho * ux).sum())
        assert abs(rep["porosity"] - (pore == 1).mean()) <= 1e-14
        assert abs(rep["tau"] - tau) <= 1e-13 * tau
        k = (tau - 0.5) / 3.0 * UD / 1e-8
        assert abs(rep["k_lattice"] - k) <= 1e-12 * abs(k)
        assert abs(rep["k_m2"] - k * 1e-18) <= 1e-12 * abs(k) * 1e-18


def test_method_by_method_driver(tmp_path):
    banner("drop-in LB2D method-by-method transport driver 64x64 100 steps")
    exe = build("phase_driver", tmp_path)
    folder = fixture_copy(tmp_path)
    r = subprocess.run([str(exe), "64", folder, "100", folder + "m_"], capture_output=True, text=True, timeout=120)
    print(r.stdout[-300:])
    assert r.returncode == 0, r.stderr
    for name in ("ux", "uy", "rho"):
        a = np.loadtxt(folder + "m_" + name + ".txt")
        g = np.loadtxt(os.path.join(O.FIXTURES, "gold", "transport", name + ".txt"))
        assert np.abs(a - g).max() <= 1e-10, name
    conv = np.loadtxt(folder + "m_convergence.txt")
    assert abs(conv[0] - 0.466825156246162) < 1e-13 and abs(conv[99] - 0.00151596085691759) < 1e-14


def test_reference_style_calculate_properties(tmp_path):
    """Shape::calculateProperties through include/shape.hpp on the fixture's
    pore image: the maps written by Shape::writeToText (15 digits) equal the
    oracle's."""
    banner("drop-in Shape::calculateProperties 64x64")
    exe = build("shape_driver", tmp_path)
    folder = fixture_copy(tmp_path)
    out = tmp_path / "out"
    out.mkdir()
    subprocess.run([str(exe), folder, str(out) + "/", "64"], check=True, timeout=120)
    pore, _, _ = O.load_fixture_geometry()
    lp, kn, _, _ = O.shape_properties(pore, 1e-9, O.mfp_of())
    glp = np.loadtxt(out / "localporesize.dat").reshape(64, 64)
    gkn = np.loadtxt(out / "Kn.dat").reshape(64, 64)
    assert np.allclose(glp, lp, rtol=1e-14, atol=0)
    assert np.allclose(gkn, kn, rtol=1e-14, atol=0)


def test_resume_from_dump_continues_exactly(tmp_path):
    """Resume (no reference counterpart): 100 steps, then 50 more started from
    the binary dump of f, equal the uninterrupted 150 steps bit for bit; from
    the 15-digit text dump they agree to the text's precision (absolute)."""
    banner("drop-in resume from f.bin / f.txt, fixture 64x64 transport")
    exe = build("app_driver", tmp_path)
    runs = {}
    for name in ("a", "b", "c", "d"):
        d = tmp_path / name
        d.mkdir()
        for f in ("pore.dat", "Kn.dat", "localporesize.dat"):
            shutil.copy(os.path.join(O.FIXTURES, f), d / f)
        runs[name] = str(d) + "/"
    env = dict(os.environ, OMP_NUM_THREADS="1")
    base = ["type=transport", "verbose=1000", "dx=1e-9", "fbody=1e-8", "macro_every=1000000", "dist_first=-1",
            "binary=1"]
    run = lambda name, nt, *extra: subprocess.run([str(exe), *base, *extra, "--", "1", "false", "64", runs[name],
                                                   str(nt)], check=True, timeout=300, env=env,
                                                  stdout=subprocess.DEVNULL)
    run("a", 100)
    run("b", 50, "init=" + runs["a"] + "f.bin")
    run("c", 150)
    run("d", 50, "init=" + runs["a"] + "f.txt")
    for field in ("f", "ux", "uy", "rho"):
        b = np.fromfile(runs["b"] + field + ".bin")
        c = np.fromfile(runs["c"] + field + ".bin")
        d = np.fromfile(runs["d"] + field + ".bin")
        assert b.size == c.size and np.array_equal(b, c), field
        # 15 significant digits of f (~0.03) carry ~1e-17 absolute error into
        # every moment; the velocities are differences of f of order 1e-6
        assert np.abs(d - c).max() <= 1e-13, field


def write_inputs(folder, pore, kn, lp):
    """Shape::loadExistingModel's column-major files (shape.hpp:253-285):
    the value of row i, column j on line j N + i."""
    np.savetxt(folder + "pore.dat", pore.T.reshape(-1), fmt="%d")
    np.savetxt(folder + "Kn.dat", kn.T.reshape(-1), fmt="%.17g")
    np.savetxt(folder + "localporesize.dat", lp.T.reshape(-1), fmt="%.17g")


@pytest.mark.parametrize("levels", [2, 3])
def test_app_multiblock_against_oracle(levels, tmp_path):
    """MantiisApp with is_multiblock = true (mantiisApp.hpp:138, 166: the
    quadtree Grid2D and the multi-block transport loop) through include/, on
    a sparse 128^2 packing closed on all four sides; MANTIIS_LEVELS=3 allows
    three quadtree levels.  The grid size printed at set-up and the binary
    output fields equal the oracle's multi-block run bit for bit."""
    from mantiis_amd import geometry as G
    banner(f"drop-in MantiisApp multi-block ({levels} levels) 128x128 60 steps vs oracle")
    pore = G.disk_packing(128, 0.9, 3, 5, seed=7)
    pore[0, :] = pore[-1, :] = pore[:, 0] = pore[:, -1] = 0
    pore = G.sanitize(pore)
    kn, lp = G.uniform_kn(pore, 0.1, 0.7730846760658538)
    d = tmp_path / "io"
    d.mkdir()
    folder = str(d) + "/"
    write_inputs(folder, pore, kn, lp)
    exe = build("app_driver", tmp_path)
    env = dict(os.environ, OMP_NUM_THREADS="1")
    env.pop("MANTIIS_LEVELS", None)
    env.pop("MANTIIS_TORUS", None)
    if levels == 3:
        env["MANTIIS_LEVELS"] = "3"
    settings = ["type=transport", "verbose=1000", "dx=1e-9", "fbody=1e-6", "macro_every=1000000", "dist_first=-1",
                "binary=1"]
    r = subprocess.run([str(exe), *settings, "--", "1", "true", "128", folder, "60"], capture_output=True, text=True,
                       timeout=300, env=env)
    print(r.stdout[-400:])
    assert r.returncode == 0, r.stderr
    o = O.Oracle(pore, kn, lp, dx=1e-9, Fbody=1e-6, multiblock=levels)
    assert int(o.grid()["level"].max()) == levels
    assert f"The number of active cells: {o.ncell}" in r.stdout
    o.run("transport", 60, 0.0)
    oux, ouy, orho = (np.zeros(pore.size) for _ in range(3))
    o.lib.orc_pixel_domain(o.h, oux.ctypes.data_as(O._D), ouy.ctypes.data_as(O._D), orho.ctypes.data_as(O._D))
    for name, want in (("ux", oux), ("uy", ouy), ("rho", orho)):
        got = np.fromfile(folder + name + ".bin")
        assert got.size == want.size and np.array_equal(got, want), (name, np.abs(got - want).max())
    o.close()


def test_method_by_method_driver_multiblock(tmp_path):
    """The method-by-method LB2D driver on a three-level quadtree
    (is_multiblock, MANTIIS_LEVELS=3): evolutionStepOfMultiBlock recurses
    over the levels and interpolates each pair of levels through the C ABI;
    the written fields (15 significant digits) equal the oracle's."""
    from mantiis_amd import geometry as G
    banner("drop-in LB2D method-by-method driver, three-level quadtree 128x128 30 steps")
    pore = G.disk_packing(128, 0.9, 3, 5, seed=7)
    pore[0, :] = pore[-1, :] = pore[:, 0] = pore[:, -1] = 0
    pore = G.sanitize(pore)
    kn, lp = G.uniform_kn(pore, 0.1, 0.7730846760658538)
    d = tmp_path / "io"
    d.mkdir()
    folder = str(d) + "/"
    write_inputs(folder, pore, kn, lp)
    exe = build("phase_driver", tmp_path)
    env = dict(os.environ, MANTIIS_LEVELS="3")
    env.pop("MANTIIS_TORUS", None)
    r = subprocess.run([str(exe), "128", folder, "30", folder + "m_", "true"], capture_output=True, text=True,
                       timeout=300, env=env)
    print(r.stdout[-300:])
    assert r.returncode == 0, r.stderr
    o = O.Oracle(pore, kn, lp, dx=1e-9, Fbody=1e-8, multiblock=3)
    assert int(o.grid()["level"].max()) == 3
    o.run("transport", 30, 0.0)
    oux, ouy, orho = (np.zeros(pore.size) for _ in range(3))
    o.lib.orc_pixel_domain(o.h, oux.ctypes.data_as(O._D), ouy.ctypes.data_as(O._D), orho.ctypes.data_as(O._D))
    for name, want in (("ux", oux), ("uy", ouy), ("rho", orho)):
        got = np.loadtxt(folder + "m_" + name + ".txt").reshape(-1)
        assert np.allclose(got, want, rtol=1e-13, atol=1e-300), name
    conv = np.loadtxt(folder + "m_convergence.txt")
    assert abs(conv[-1] - o.diff_history()[-1]) <= 1e-13 * abs(o.diff_history()[-1])
    o.close()
