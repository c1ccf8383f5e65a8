"""Transport step (RMRT + SRBB + periodic) of the HIP library against the CPU
oracle and the reference's gold (64x64, 100 steps, test/run_tests.sh:15-23)."""
import numpy as np
import pytest

import _oracle as O
from _gpu_util import TRANSPORT, banner, pair, rel_l2
from mantiis_amd import geometry as G
from mantiis_amd.hip import Lattice

pytestmark = pytest.mark.gpu


def test_grid_and_links_bit_exact(fixture_geometry):
    banner("grid/links fixture 64x64")
    pore, kn, lp = fixture_geometry
    o, g = pair(pore, kn, lp, **TRANSPORT)
    go, gg = o.grid(), g.grid()
    for k in ("gridID", "gridIJ", "level", "connect", "bnd"):
        assert np.array_equal(go[k], gg[k]), k
    for which in ("stream", "boundary", "periodic"):
        a, b = o.links(which), g.links(which)
        assert np.array_equal(a[0], b[0]) and np.array_equal(a[2], b[2]), which
        if which != "boundary":
            assert np.array_equal(a[1], b[1]), which
    assert np.array_equal(o.get("taus"), g.get("taus"))


def test_links_bit_exact_on_seamed_packing():
    """Fluid across both seams and at the corners: y-seam drop, x-seam periodic."""
    banner("links periodic disk packing 96x96")
    pore = G.disk_packing(96, 0.6, 3, 8, seed=3)
    kn, lp = G.uniform_kn(pore, 0.1, 0.773085)
    o, g = pair(pore, kn, lp, **TRANSPORT)
    for which in ("stream", "boundary", "periodic"):
        a, b = o.links(which), g.links(which)
        assert len(a[0]) == len(b[0]), which
        assert np.array_equal(a[0], b[0]) and np.array_equal(a[2], b[2]), which
    assert len(o.links("periodic")[0]) > 0


def test_transport_gold_fp64(fixture_geometry, gold):
    banner("transport fixture FP64 100 x 1 step")
    pore, kn, lp = fixture_geometry
    o, g = pair(pore, kn, lp, **TRANSPORT)
    o.run("transport", 100, 1e-10)
    hist = []
    for _ in range(100):
        g.step("transport", 1)
        hist.append(g.residual())
    ux, uy, rho = g.complete_domain()
    oux, ouy, orho = o.complete_domain()
    for name, a, b in (("ux", ux, oux), ("uy", uy, ouy), ("rho", rho, orho)):
        assert np.abs(a - gold[("transport", name)]).max() <= 1e-10, name
        d = np.abs(a - b).max()
        print(name, "max|gpu-oracle| =", d, "rel-l2 =", rel_l2(a, b))
        assert rel_l2(a, b) <= 5e-13, name
    assert np.allclose(hist, o.diff_history(), rtol=1e-11, atol=0)
    f_o, f_g = o.get("f"), g.get("f")
    print("f max abs diff", np.abs(f_o - f_g).max())
    assert np.abs(f_o - f_g).max() <= 1e-16


def test_multistep_call_equals_single_steps(fixture_geometry):
    banner("transport fixture FP64 one call of 100 steps")
    pore, kn, lp = fixture_geometry
    o, g1 = pair(pore, kn, lp, **TRANSPORT)
    from mantiis_amd.hip import Lattice
    g2 = Lattice(pore, kn, lp, precision="fp64", **TRANSPORT)
    for _ in range(37):
        g1.step("transport", 1)
    g2.step("transport", 37)
    assert np.array_equal(g1.get("f"), g2.get("f"))
    assert g1.residual() == g2.residual()
    o.run("transport", 37, 0.0)
    assert abs(g2.residual() - o.diff_history()[-1]) <= 1e-12 * o.diff_history()[-1]


def test_transport_fp32_deviation(fixture_geometry):
    banner("transport fixture FP32 deviation storage 100 steps")
    pore, kn, lp = fixture_geometry
    o, g = pair(pore, kn, lp, precision="fp32dev", **TRANSPORT)
    o.run("transport", 100, 0.0)
    g.step("transport", 100)
    ux, uy, rho = g.complete_domain()
    oux, ouy, orho = o.complete_domain()
    e = dict(ux=rel_l2(ux, oux), uy=rel_l2(uy, ouy), rho=rel_l2(rho, orho))
    print("FP32-dev rel-l2 vs FP64 oracle", e)
    assert e["rho"] < 1e-11 and e["ux"] < 3e-6 and e["uy"] < 4e-6  # measured: 1.4e-13, 2.5e-7, 3.5e-7


@pytest.mark.parametrize("wall", [0, 1])
@pytest.mark.parametrize("coll", [0, 1])
def test_transport_operators(fixture_geometry, coll, wall):
    banner(f"transport fixture FP64 collision={coll} wall={wall} 20 steps")
    pore, kn, lp = fixture_geometry
    o, g = pair(pore, kn, lp, collision=coll, wall=wall, **TRANSPORT)
    o.run("transport", 20, 0.0)
    g.step("transport", 20)
    d = np.abs(o.get("f") - g.get("f")).max()
    print("max |f| diff", d)
    assert d <= 1e-15


def test_transport_disk_packing_fp64():
    """A seeded 256^2 disk packing (r ~ U[4,12], porosity 0.6), solid rows at
    y = 0 and N-1 so that no link is lost at the y-seam; Kn = 0.1."""
    banner("transport disk packing 256x256 FP64 200 steps")
    pore = G.disk_packing(256, 0.6, 4, 12, seed=1, solid_rows=True)
    rc = O.load()
    kn, lp = G.uniform_kn(pore, 0.1, 0.7730846760658538)
    o, g = pair(pore, kn, lp, dx=1e-9, Fbody=1e-6, T=300.0, P=2e6)
    o.run("transport", 200, 0.0)
    g.step("transport", 200)
    ux, uy, rho = g.complete_domain()
    oux, ouy, orho = o.complete_domain()
    print("rel-l2 ux", rel_l2(ux, oux), "rho", rel_l2(rho, orho))
    assert rel_l2(ux, oux) < 5e-10 and rel_l2(rho, orho) < 1e-14


def test_physics_mode_y_periodic_fp64():
    """periodic_y=1 (torus): links across both seams are classified like
    interior links; the library matches the oracle bit for bit and conserves mass."""
    banner("transport torus 128x128 FP64 100 steps")
    pore = G.disk_packing(128, 0.6, 4, 10, seed=11)
    kn, lp = G.uniform_kn(pore, 0.1, 0.7730846760658538)
    o = O.Oracle(pore, kn, lp, dx=1e-9, Fbody=1e-6, periodic_y=True)
    from mantiis_amd.hip import Lattice
    g = Lattice(pore, kn, lp, precision="fp64", dx=1e-9, Fbody=1e-6, periodic_y=1)
    for which in ("stream", "boundary", "periodic"):
        a, b = o.links(which), g.links(which)
        assert np.array_equal(a[0], b[0]) and np.array_equal(a[2], b[2]), which
    assert len(g.links("periodic")[0]) == 0
    m0 = g.diagnostics()[0]
    o.run("transport", 100, 0.0)
    g.step("transport", 100)
    d = np.abs(o.get("f") - g.get("f")).max()
    print("max |f| diff", d)
    assert d <= 1e-15
    assert abs(g.diagnostics()[0] - m0) / m0 < 1e-12


@pytest.mark.parametrize("prec", ["fp64", "fp32dev"])
def test_mass_conservation_long_run(prec):
    """Bounce-back walls + periodic seams conserve mass; 2000 steps, 512^2."""
    banner(f"transport mass conservation 512x512 {prec} 2000 steps")
    from mantiis_amd.hip import Lattice
    pore = G.disk_packing(512, 0.5, 4, 12, seed=5)
    kn, lp = G.uniform_kn(pore, 0.5, 0.7730846760658538)
    g = Lattice(pore, kn, lp, precision=prec, dx=1e-9, Fbody=1e-6, periodic_y=1)
    m0 = g.diagnostics()[0]
    g.step("transport", 2000)
    d = g.diagnostics()
    print("mass drift", (d[0] - m0) / m0, "mean jx", d[1] / d[3])
    assert np.isfinite(d).all()
    assert abs(d[0] - m0) / m0 < (1e-12 if prec == "fp64" else 1e-6)
    assert d[1] > 0  # the body force drives a net +x flux


@pytest.mark.parametrize("kind,prec", [("transport", "fp64"), ("transport", "fp32dev"), ("adsorption", "fp64"),
                                       ("adsorption", "fp32dev")])
def test_nonfinite_state_is_reported(kind, prec):
    """A NaN planted in f (through mts_set_field) makes the fused step call
    return MTS_ERR_NONFINITE (-8) instead of running on: the reference's test
    diff < tol is never true for NaN, so a diverged run would use all its
    iterations."""
    banner(f"NaN injection {kind} {prec}")
    from mantiis_amd import geometry as G
    from mantiis_amd.hip import MtsError
    pore = G.disk_packing(64, 0.7, 3, 8, seed=2)
    kn, lp = G.uniform_kn(pore, 0.1, 0.7730846760658538)
    coll, wall = (2, 2) if kind == "transport" else (1, 0)
    g = Lattice(pore, kn, lp, precision=prec, collision=coll, wall=wall, periodic_y=1, dx=1e-9, Fbody=1e-6)
    g.step(kind, 2)
    f = g.get("f")
    f[9 * 100 + 3] = np.nan
    g.set("f", f)
    with pytest.raises(MtsError) as e:
        g.step(kind, 4)
    assert e.value.code == -8, e.value
    g.close()


def test_absurd_allocation_reports_oom():
    """A device allocation far beyond the GPU's memory (1 PiB) returns
    MTS_ERR_OOM (-5) with the runtime's message, and the library keeps
    working: a lattice created and stepped afterwards matches one stepped
    before (no failed status left behind for a later launch check)."""
    banner("OOM: 1 PiB device allocation, then a lattice")
    from mantiis_amd import geometry as G
    from mantiis_amd.hip import DeviceArray, MtsError
    pore = G.disk_packing(64, 0.7, 3, 8, seed=2)
    kn, lp = G.uniform_kn(pore, 0.1, 0.7730846760658538)
    a = Lattice(pore, kn, lp, precision="fp32dev", periodic_y=1, dx=1e-9, Fbody=1e-6)
    a.step("transport", 20)
    with pytest.raises(MtsError) as e:
        DeviceArray((1 << 48,), np.float32)
    print("error", e.value)
    assert e.value.code == -5 and "hipMalloc" in str(e.value)
    b = Lattice(pore, kn, lp, precision="fp32dev", periodic_y=1, dx=1e-9, Fbody=1e-6)
    b.step("transport", 20)
    assert np.array_equal(a.get("f"), b.get("f"))
    a.close()
    b.close()


@pytest.mark.parametrize("coll", [2, 1, 0], ids=["RMRT", "MRT", "SRT"])
def test_fp32_mdbb_against_fp64_oracle(coll):
    """MDBB walls (coreLB.hpp:936-967) in the fused FP32 step: each wall cell
    keeps Kno / Kdeno and the moments of the state it collided for the next
    pass's wall rule.  Against the FP64 oracle with the same operators; a call
    of 2 x 50 steps equals one of 100 bit for bit."""
    banner(f"FP32 MDBB collision {coll} vs FP64 oracle, 128^2 torus, 100 steps")
    pore = G.disk_packing(128, 0.65, 3, 9, seed=6)
    kn, lp = G.uniform_kn(pore, 0.2, 0.7730846760658538)
    steps = 20 if coll == 1 else 100
    o = O.Oracle(pore, kn, lp, dx=1e-9, Fbody=1e-6, periodic_y=True)
    o.set_operators(coll, 1)
    o.run("transport", steps, 0.0)
    g = Lattice(pore, kn, lp, precision="fp32dev", collision=coll, wall=1, periodic_y=1, dx=1e-9, Fbody=1e-6)
    g.step("transport", steps // 2)
    g.step("transport", steps - steps // 2)
    h = Lattice(pore, kn, lp, precision="fp32dev", collision=coll, wall=1, periodic_y=1, dx=1e-9, Fbody=1e-6)
    h.step("transport", steps)
    # a call re-derives the wall cells' MDBB state from the canonical FP64
    # populations; a deviation far below w rho0 does not survive that round
    # trip in every bit, so the split run agrees to FP32 rounding
    split = float(np.linalg.norm(g.get("f") - h.get("f")) / np.linalg.norm(h.get("f")))
    print("split vs continuous: rel |df|", split)
    assert split < 1e-9
    e = {k: rel_l2(g.get(k), o.get(k)) for k in ("rho", "ux", "uy")}
    print("FP32 MDBB rel-l2 vs oracle", e)
    assert e["rho"] < 1e-9 and e["ux"] < 1e-5 and e["uy"] < 1e-5, e  # measured: at most 8.9e-11, 7.9e-7, 8.8e-7
    g.close()
    h.close()


@pytest.mark.parametrize("torus", [False, True])
def test_method_by_method_single_block_against_oracle(torus):
    """The transport loop called method by method (mantiisApp.hpp:239-246)
    on a single-block disk packing, under the reference's seam rule and on a
    torus (where links across the x-seam stream like interior links and the
    periodic step has nothing to do), FP64 against the oracle, bit for bit."""
    banner(f"single-block disks96 method by method, torus={torus}")
    pore = G.disk_packing(96, 0.7, 3, 8, seed=12)
    kn, lp = G.uniform_kn(pore, 0.1, 0.7730846760658538)
    o = O.Oracle(pore, kn, lp, periodic_y=torus, Fbody=1e-6)
    g = Lattice(pore, kn, lp, precision="fp64", periodic_y=int(torus), Fbody=1e-6)
    for _ in range(40):
        g.method("copy_usq_to_usq_old")
        g.method("equilibrium")
        g.method("evolutionStepOfMultiBlock", 1)
        g.method("calculateMacroscopicPorperties")
        g.method("calculateVelocityDifference")
    o.run("transport", 40, 0.0)
    d = np.abs(o.get("f") - g.get("f")).max()
    print("max |f| diff", d, "residual", g.residual(), o.diff_history()[-1])
    assert d <= 1e-15
    assert abs(g.residual() - o.diff_history()[-1]) <= 1e-12 * o.diff_history()[-1]
    o.close()
    g.close()
