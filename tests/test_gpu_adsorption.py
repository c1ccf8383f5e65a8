"""Shan-Chen adsorption and coupled steps of the HIP library against the CPU
oracle and the reference's adsorption gold."""
import numpy as np
import pytest

from _gpu_util import ADSORPTION, banner, pair, rel_l2

pytestmark = pytest.mark.gpu


def test_adsorption_gold_fp64(fixture_geometry, gold):
    banner("adsorption fixture FP64 100 x 1 step")
    pore, kn, lp = fixture_geometry
    o, g = pair(pore, kn, lp, collision=1, wall=0, **ADSORPTION)
    o.run("adsorption", 100, 1e-10)
    hist = []
    for _ in range(100):
        g.step("adsorption", 1)
        hist.append(g.residual())
    ux, uy, rho = g.complete_domain()
    oux, ouy, orho = o.complete_domain()
    for name, a, b in (("ux", ux, oux), ("uy", uy, ouy), ("rho", rho, orho)):
        e = np.abs(a - gold[("adsorption", name)]).max()
        print(name, "max|gpu-gold| =", e, "max|gpu-oracle| =", np.abs(a - b).max())
        assert e <= 1e-10, name
        assert rel_l2(a, b) <= 1e-12, name
    assert np.allclose(hist, o.diff_history(), rtol=1e-12, atol=0)


def test_adsorption_one_call(fixture_geometry):
    banner("adsorption fixture FP64 one call of 100 steps")
    pore, kn, lp = fixture_geometry
    o, g = pair(pore, kn, lp, collision=1, wall=0, **ADSORPTION)
    o.run("adsorption", 100, 0.0)
    g.step("adsorption", 100)
    for name in ("f", "rho", "ux", "uy", "rho_old", "psi", "Fx", "Fy"):
        a, b = g.get(name), o.get(name)
        print(name, np.abs(a - b).max())
        assert rel_l2(a, b) <= 1e-12, name
    assert abs(g.residual() - o.diff_history()[-1]) <= 1e-12 * o.diff_history()[-1]


def test_coupled(fixture_geometry):
    banner("coupled fixture FP64: 100 adsorption (F=0) + 30 coupled")
    pore, kn, lp = fixture_geometry
    o, g = pair(pore, kn, lp, collision=2, wall=2, **ADSORPTION)
    o.run("coupled", 30, 0.0)
    g.set_body_force(0.0)
    g.step("adsorption", 100)
    g.set_body_force(ADSORPTION["Fbody"])
    g.step("coupled", 30)
    for name in ("f", "rho", "ux", "uy"):
        a, b = g.get(name), o.get(name)
        print(name, np.abs(a - b).max())
        assert rel_l2(a, b) <= 1e-12, name


def test_adsorption_fp32_deviation(fixture_geometry):
    banner("adsorption fixture FP32 deviation storage 100 steps")
    pore, kn, lp = fixture_geometry
    o, g = pair(pore, kn, lp, precision="fp32dev", collision=1, wall=0, **ADSORPTION)
    o.run("adsorption", 100, 0.0)
    g.step("adsorption", 100)
    ux, uy, rho = g.complete_domain()
    oux, ouy, orho = o.complete_domain()
    e = dict(ux=rel_l2(ux, oux), uy=rel_l2(uy, ouy), rho=rel_l2(rho, orho))
    print("FP32-dev rel-l2 vs FP64 oracle", e)
    assert e["rho"] < 1e-5 and e["ux"] < 2e-3 and e["uy"] < 2e-3


def test_shanchen_fp32_large_tile_edges():
    """FP32 Shan-Chen on a lattice whose size is not a multiple of the 64x16
    tile (edge tiles, halo wrap) against the FP64 path of the library."""
    banner("adsorption 200x72 fp32dev vs fp64 50 steps")
    from mantiis_amd import geometry as G
    from mantiis_amd.hip import Lattice
    pore = G.disk_packing(200, 0.6, 3, 8, seed=4)[:72, :].copy()
    pore = G.sanitize(pore)
    kn, lp = G.uniform_kn(pore, 0.1, 0.7730846760658538)
    a = Lattice(pore, kn, lp, precision="fp64", collision=1, wall=0, periodic_y=1, **ADSORPTION)
    b = Lattice(pore, kn, lp, precision="fp32dev", collision=1, wall=0, periodic_y=1, **ADSORPTION)
    a.step("adsorption", 50)
    b.step("adsorption", 50)
    for name in ("rho", "ux", "uy", "psi"):
        e = rel_l2(b.get(name), a.get(name))
        print(name, e)
        assert e < 2e-3, name
