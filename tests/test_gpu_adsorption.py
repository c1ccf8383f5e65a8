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
    # measured: rho 1.6e-7, ux 1.1e-6, uy 1.0e-6
    assert e["rho"] < 2e-6 and e["ux"] < 2e-5 and e["uy"] < 2e-5


def test_shanchen_edge_tiles_fp64_and_fp32():
    """A 100x100 torus lattice (not a multiple of the 64x16 LDS tile: edge
    tiles and halo wrap): FP64 library vs oracle bit for bit, FP32 library vs
    FP64 within FP32 error."""
    banner("adsorption 100x100 torus fp64 vs oracle, fp32dev vs fp64, 50 steps")
    import _oracle as O
    from mantiis_amd import geometry as G
    from mantiis_amd.hip import Lattice
    pore = G.disk_packing(100, 0.6, 3, 8, seed=4)
    kn, lp = G.uniform_kn(pore, 0.1, 0.7730846760658538)
    o = O.Oracle(pore, kn, lp, periodic_y=True, **ADSORPTION)
    o.set_operators(1, 0)
    a = Lattice(pore, kn, lp, precision="fp64", collision=1, wall=0, periodic_y=1, **ADSORPTION)
    b = Lattice(pore, kn, lp, precision="fp32dev", collision=1, wall=0, periodic_y=1, **ADSORPTION)
    o.run("adsorption", 50, 0.0)
    a.step("adsorption", 50)
    b.step("adsorption", 50)
    for name in ("f", "rho", "ux", "uy", "psi"):
        d = np.abs(a.get(name) - o.get(name)).max()
        e = rel_l2(b.get(name), a.get(name))
        print(name, "fp64 vs oracle", d, "fp32 vs fp64", e)
        assert d <= 1e-15 * max(1.0, np.abs(o.get(name)).max()), name
        assert e < 1e-5, name  # measured: at most 5.8e-7 (ux)


def test_non_square_lattice():
    """nx != ny (the reference needs square lattices; the library does not):
    a 200x72 torus, FP64 and FP32 agree and conserve mass."""
    banner("transport 200x72 torus fp64 vs fp32dev 300 steps")
    from mantiis_amd import geometry as G
    from mantiis_amd.hip import Lattice
    pore = G.sanitize(G.disk_packing(200, 0.6, 3, 8, seed=4)[:72, :].copy())
    kn, lp = G.uniform_kn(pore, 0.2, 0.7730846760658538)
    a = Lattice(pore, kn, lp, precision="fp64", periodic_y=1, dx=1e-9, Fbody=1e-6)
    b = Lattice(pore, kn, lp, precision="fp32dev", periodic_y=1, dx=1e-9, Fbody=1e-6)
    m0 = a.diagnostics()[0]
    a.step("transport", 300)
    b.step("transport", 300)
    da, db = a.diagnostics(), b.diagnostics()
    print("mass", (da[0] - m0) / m0, (db[0] - m0) / m0, "rel ux", rel_l2(b.get("ux"), a.get("ux")))
    assert np.isfinite(da).all() and np.isfinite(db).all()
    assert abs(da[0] - m0) / m0 < 1e-12 and abs(db[0] - m0) / m0 < 1e-6
    assert rel_l2(b.get("ux"), a.get("ux")) < 1e-5  # measured: 6.6e-7


def test_shanchen_fp32_across_tile_columns():
    """The FP32 Shan-Chen pass runs one block per 128 x 8 site tile and keeps a
    cell's column as x mod 128 in its cell word.  Lattices wider than one tile
    column, with sides that are multiples of neither 128 nor 8, exercise tile
    columns tx >= 1, the ring of sites two tile columns share and partial
    edge tiles: FP64 library vs oracle on a 300 x 300 torus (3 tile columns),
    FP32 vs FP64 there and on a 300 x 90 torus, 50 steps each."""
    import _oracle as O
    from mantiis_amd import geometry as G
    from mantiis_amd.hip import Lattice
    for nx, ny in ((300, 300), (300, 90)):
        banner(f"adsorption {nx}x{ny} torus fp32dev vs fp64 (and fp64 vs oracle when square), 50 steps")
        pore = G.sanitize(G.disk_packing(300, 0.6, 3, 8, seed=8)[:ny, :nx].copy())
        kn, lp = G.uniform_kn(pore, 0.1, 0.7730846760658538)
        a = Lattice(pore, kn, lp, precision="fp64", collision=1, wall=0, periodic_y=1, **ADSORPTION)
        b = Lattice(pore, kn, lp, precision="fp32dev", collision=1, wall=0, periodic_y=1, **ADSORPTION)
        a.step("adsorption", 50)
        b.step("adsorption", 50)
        if nx == ny:
            o = O.Oracle(pore, kn, lp, periodic_y=True, **ADSORPTION)
            o.set_operators(1, 0)
            o.run("adsorption", 50, 0.0)
            for name in ("f", "rho", "ux", "uy", "psi"):
                d = np.abs(a.get(name) - o.get(name)).max()
                print(name, "fp64 vs oracle max-abs", d)
                assert d <= 1e-15 * max(1.0, np.abs(o.get(name)).max()), name
            o.close()
        for name, tol in (("rho", 1e-6), ("ux", 1e-4), ("uy", 1e-4), ("psi", 1e-6)):
            e = rel_l2(b.get(name), a.get(name))
            print(f"{nx}x{ny}", name, "fp32 vs fp64 rel-l2", e)
            assert e < tol, (name, e)
        a.close()
        b.close()
