"""Quadtree multi-block transport (coarse step: level-2 collide/stream,
two level-1 substeps with the step's stored equilibrium, interface
interpolation, moments) of the HIP library against the oracle."""
import numpy as np
import pytest

import _oracle as O
from _gpu_util import banner, rel_l2
from mantiis_amd import geometry as G
from mantiis_amd.hip import COLL, WALL, Lattice

pytestmark = pytest.mark.gpu
MFP_DX = 0.7730846760658538


def box64():
    N = 64
    p = np.ones((N, N), np.int32)
    p[0, :] = p[-1, :] = p[:, 0] = p[:, -1] = 0
    yy, xx = np.mgrid[0:N, 0:N]
    p[(xx - 32) ** 2 + (yy - 30) ** 2 <= 36] = 0
    return G.sanitize(p)


CASES = [("box64_reference", box64, False), ("disks128_torus", lambda: G.disk_packing(128, 0.7, 3, 8, seed=3), True)]


@pytest.mark.parametrize("name,geom,torus", CASES, ids=[c[0] for c in CASES])
def test_multiblock_links_and_fp64_steps(name, geom, torus):
    banner(f"multi-block {name} FP64 links + 60 steps")
    pore = geom()
    kn, lp = G.uniform_kn(pore, 0.1, MFP_DX)
    o = O.Oracle(pore, kn, lp, multiblock=True, periodic_y=torus, Fbody=1e-6)
    g = Lattice(pore, kn, lp, precision="fp64", multiblock=1, periodic_y=int(torus), Fbody=1e-6)
    assert g.ncell == o.ncell
    for which in ("stream", "boundary", "periodic"):
        a, b = o.links(which), g.links(which)
        assert np.array_equal(a[0], b[0]) and np.array_equal(a[2], b[2]), which
        if which != "boundary":
            assert np.array_equal(a[1], b[1]), which
    o.run("transport", 60, 0.0)
    g.step("transport", 60)
    d = np.abs(o.get("f") - g.get("f")).max()
    print("max |f| diff", d, "residual", g.residual(), o.diff_history()[-1])
    assert d <= 1e-15
    assert abs(g.residual() - o.diff_history()[-1]) <= 1e-12 * o.diff_history()[-1]
    ux, uy, rho = g.complete_domain()
    oux, ouy, orho = (np.zeros(pore.size) for _ in range(3))
    o.lib.orc_pixel_domain(o.h, oux.ctypes.data_as(O._D), ouy.ctypes.data_as(O._D), orho.ctypes.data_as(O._D))
    assert np.array_equal(ux, oux) and np.array_equal(rho, orho)


def test_multiblock_fp32_against_fp64_oracle():
    banner("multi-block disks128 torus FP32 vs FP64 oracle 200 steps")
    pore = G.disk_packing(128, 0.7, 3, 8, seed=3)
    kn, lp = G.uniform_kn(pore, 0.1, MFP_DX)
    o = O.Oracle(pore, kn, lp, multiblock=True, periodic_y=True, Fbody=1e-6)
    g = Lattice(pore, kn, lp, precision="fp32dev", multiblock=1, periodic_y=1, Fbody=1e-6)
    o.run("transport", 200, 0.0)
    g.step("transport", 200)
    e = {k: rel_l2(g.get(k), o.get(k)) for k in ("rho", "ux", "uy")}
    print("FP32 multi-block rel-l2", e)
    assert e["rho"] < 1e-9 and e["ux"] < 1e-5 and e["uy"] < 1.5e-5  # measured: 7e-11, 1.1e-6, 1.3e-6


OPS = [("SRT", "BB"), ("SRT", "SRBB"), ("MRT", "BB"), ("MRT", "SRBB"), ("RMRT", "BB")]


@pytest.mark.parametrize("coll,wall", OPS, ids=[f"{c}-{w}" for c, w in OPS])
def test_multiblock_fp32_operators_against_fp64_oracle(coll, wall):
    """Every collision x wall pair of the FP32 quadtree step (each has its own
    instantiation of the fused interface + first-substep launch) against the
    FP64 oracle with the same operators.  The reference's non-regularised MRT
    diverges on this multi-block lattice (the oracle reaches max |u| ~ 740
    after 200 steps, the single-block lattice stays at 3e-4), so the MRT pairs
    are compared over the first 20 steps."""
    steps = 20 if coll == "MRT" else 200
    banner(f"multi-block disks128 torus FP32 {coll}+{wall} vs FP64 oracle {steps} steps")
    pore = G.disk_packing(128, 0.7, 3, 8, seed=3)
    kn, lp = G.uniform_kn(pore, 0.1, MFP_DX)
    o = O.Oracle(pore, kn, lp, multiblock=True, periodic_y=True, Fbody=1e-6)
    o.set_operators(COLL[coll], WALL[wall])
    g = Lattice(pore, kn, lp, precision="fp32dev", multiblock=1, periodic_y=1, Fbody=1e-6, collision=coll, wall=wall)
    o.run("transport", steps, 0.0)
    g.step("transport", steps)
    e = {k: rel_l2(g.get(k), o.get(k)) for k in ("rho", "ux", "uy")}
    print(f"FP32 multi-block {coll}+{wall} rel-l2", e)
    assert e["rho"] < 1e-9 and e["ux"] < 1e-5 and e["uy"] < 1.5e-5  # measured: 7e-11, 1.1e-6, 1.3e-6


def test_multiblock_large_torus_runs():
    banner("multi-block disks512 torus FP32 1000 steps")
    pore = G.disk_packing(512, 0.6, 4, 12, seed=7)
    kn, lp = G.uniform_kn(pore, 0.5, MFP_DX)
    g = Lattice(pore, kn, lp, precision="fp32dev", multiblock=1, periodic_y=1, Fbody=1e-6)
    d0 = g.diagnostics()
    g.step("transport", 1000)
    d1 = g.diagnostics()
    print("mass drift", (d1[0] - d0[0]) / d0[0], "flux", d1[1], "residual", g.residual())
    assert np.isfinite(d1).all() and abs(d1[0] - d0[0]) / d0[0] < 1e-4 and d1[1] > 0


def test_multiblock_fp32_reference_seams_against_oracle():
    banner("multi-block box64 reference seams FP32 vs FP64 oracle 200 steps")
    pore = box64()
    kn, lp = G.uniform_kn(pore, 0.1, MFP_DX)
    o = O.Oracle(pore, kn, lp, multiblock=True, periodic_y=False, Fbody=1e-6)
    g = Lattice(pore, kn, lp, precision="fp32dev", multiblock=1, periodic_y=0, Fbody=1e-6)
    o.run("transport", 200, 0.0)
    g.step("transport", 200)
    e = {k: rel_l2(g.get(k), o.get(k)) for k in ("rho", "ux", "uy")}
    print("FP32 multi-block (reference seams) rel-l2", e)
    assert e["rho"] < 1e-9 and e["ux"] < 1e-5 and e["uy"] < 1.5e-5  # measured: 7e-11, 1.1e-6, 1.3e-6


def test_multiblock_fp32_call_chunking_is_exact():
    """A call starts from and ends in the canonical FP64 state (with the
    interface overrides applied), so 2 x 50 steps equal 100 steps."""
    banner("multi-block disks128 FP32 100 steps in one call vs two calls")
    pore = G.disk_packing(128, 0.7, 3, 8, seed=3)
    kn, lp = G.uniform_kn(pore, 0.1, MFP_DX)
    a = Lattice(pore, kn, lp, precision="fp32dev", multiblock=1, periodic_y=1, Fbody=1e-6)
    b = Lattice(pore, kn, lp, precision="fp32dev", multiblock=1, periodic_y=1, Fbody=1e-6)
    a.step("transport", 100)
    b.step("transport", 50)
    b.step("transport", 50)
    d = np.abs(a.get("f") - b.get("f")).max()
    print("max |f| diff", d)
    assert d <= 1e-14


def test_multiblock_taus_per_level():
    """LB2D::calculateMultigridTauS (coreLB.hpp:432-480): every cell's tau
    from its level's resized image and calculateProperties, through the
    device lattice and the host grid, equals the oracle's; a straight
    channel gives the closed form on both levels."""
    banner("multi-block per-level taus vs oracle (box64, disks128, channel64)")
    from mantiis_amd.hip import HostGrid
    N = 64
    chan = np.ones((N, N), np.int32)
    chan[0] = chan[-1] = 0
    for name, pore, torus in (("box64", box64(), False), ("disks128", G.disk_packing(128, 0.7, 3, 8, seed=3), True),
                              ("channel64", chan, False)):
        kn, lp = G.uniform_kn(pore, 0.1, MFP_DX)
        o = O.Oracle(pore, kn, lp, multiblock=True, periodic_y=torus)
        g = Lattice(pore, kn, lp, precision="fp32dev", multiblock=1, periodic_y=int(torus))
        h = HostGrid(pore, kn, lp, multiblock=1, periodic_y=int(torus))
        t = o.get("taus")
        assert np.array_equal(g.get("taus"), t), name
        assert np.array_equal(h.arrays()["taus"], t), name
        if name == "channel64":
            lvl = h.arrays()["level"]
            mfp = O.mfp_of()
            for L, side in ((1, N), (2, N // 2)):
                lpL = 2 * ((side - 1) // 2)  # middle row of the side-row channel, doubled
                knL = mfp / lpL / (1e-9 * 2 ** (L - 1))
                want = 0.5 + np.sqrt(6 / np.pi) * lpL * knL / (1 + 2 * knL)
                assert (lvl == L).any() and np.allclose(t[lvl == L], want, rtol=1e-14), L
        g.close()


TABLE_CASES = [
    ("box64_reference", box64, False),
    ("disks128_torus", lambda: G.disk_packing(128, 0.7, 3, 8, seed=3), True),
    ("disks256_reference", lambda: G.disk_packing(256, 0.6, 4, 12, seed=8), False),
    ("disks512_torus", lambda: G.disk_packing(512, 0.55, 3, 16, seed=9), True),
    ("all_fluid64_torus", lambda: np.ones((64, 64), np.int32), True),
    ("narrow_pores64", lambda: G.disk_packing(64, 0.4, 2, 3, seed=2), True),
]


@pytest.mark.parametrize("name,geom,torus", TABLE_CASES, ids=[c[0] for c in TABLE_CASES])
def test_device_quadtree_tables_equal_host(name, geom, torus):
    """The quadtree tables built on the device (level decision by a window
    test, connectivity, boundary types, parent/child and buffer rows, level
    image link words) equal the host builder's entry for entry; the
    lattices of this file's other tests are built by the device builder and
    match the oracle."""
    from mantiis_amd.hip import quadtree_check
    banner(f"quadtree tables device vs host: {name}")
    pore = geom()
    kn, _ = G.uniform_kn(pore, 0.1, MFP_DX)
    assert quadtree_check(pore, kn, periodic_y=int(torus)) == 0


L3_TABLE_CASES = [
    ("sparse128_reference", lambda: G.disk_packing(128, 0.85, 3, 6, seed=4), False),
    ("sparse256_torus", lambda: G.disk_packing(256, 0.8, 4, 9, seed=3), True),
    ("disks512_torus", lambda: G.disk_packing(512, 0.8, 3, 16, seed=3), True),
]


@pytest.mark.parametrize("name,geom,torus", L3_TABLE_CASES, ids=[c[0] for c in L3_TABLE_CASES])
def test_device_quadtree_tables_equal_host_three_levels(name, geom, torus):
    """The same comparison on quadtrees of at most three levels (level
    decision launched per level, index images per level, rows per pair of
    levels)."""
    from mantiis_amd.hip import quadtree_check
    banner(f"quadtree tables device vs host, three levels: {name}")
    pore = geom()
    kn, _ = G.uniform_kn(pore, 0.1, MFP_DX)
    assert quadtree_check(pore, kn, periodic_y=int(torus), multiblock=3) == 0


def test_device_quadtree_tables_at_config_size():
    """The 1024^2 lattice of the quadtree benchmark configuration."""
    import bench
    from mantiis_amd.hip import quadtree_check
    banner("quadtree tables device vs host: 1024^2 benchmark lattice")
    pore, kn, lp, phi = bench.quadtree_inputs(1024)
    assert quadtree_check(pore, kn, periodic_y=1) == 0
    assert quadtree_check(pore, kn, periodic_y=1, multiblock=3) == 0


# ---------------------------------------------------------------- three levels
def boxed128():
    """A sparse packing inside a solid frame.  The reference's seam rule
    (x-seam links overwritten periodically, y-seam links dropped) diverges
    on quadtree lattices whose seams are open (the oracle too, on two levels
    as on three), so the reference-seam case is closed on all four sides."""
    p = G.disk_packing(128, 0.9, 3, 5, seed=7)
    p[0, :] = p[-1, :] = p[:, 0] = p[:, -1] = 0
    return G.sanitize(p)


L3_CASES = [("boxed128_reference", boxed128, False),
            ("sparse256_torus", lambda: G.disk_packing(256, 0.8, 4, 9, seed=3), True)]


@pytest.mark.parametrize("name,geom,torus", L3_CASES, ids=[c[0] for c in L3_CASES])
def test_three_levels_links_and_fp64_steps(name, geom, torus):
    """A quadtree of three levels (maxAllowedLevel = 3): per coarse step one
    level-3 step, two level-2 and four level-1 substeps, then the interfaces
    of levels (1, 2) and (2, 3) in that order (coreLB.hpp:1310-1331), in
    FP64 against the oracle: link lists equal, populations to 1e-15."""
    steps = 40 if torus else 60
    banner(f"three-level {name} FP64 links + {steps} steps")
    pore = geom()
    kn, lp = G.uniform_kn(pore, 0.1, MFP_DX)
    o = O.Oracle(pore, kn, lp, multiblock=3, periodic_y=torus, Fbody=1e-6)
    g = Lattice(pore, kn, lp, precision="fp64", multiblock=3, periodic_y=int(torus), Fbody=1e-6)
    assert g.ncell == o.ncell and int(np.max(g.grid()["level"])) == 3
    for which in ("stream", "boundary", "periodic"):
        a, b = o.links(which), g.links(which)
        assert np.array_equal(a[0], b[0]) and np.array_equal(a[2], b[2]), which
        if which != "boundary":
            assert np.array_equal(a[1], b[1]), which
    o.run("transport", steps, 0.0)
    g.step("transport", steps)
    d = np.abs(o.get("f") - g.get("f")).max()
    print("max |f| diff", d, "residual", g.residual(), o.diff_history()[-1])
    assert d <= 1e-15
    assert abs(g.residual() - o.diff_history()[-1]) <= 1e-12 * o.diff_history()[-1]
    ux, uy, rho = g.complete_domain()
    oux, ouy, orho = (np.zeros(pore.size) for _ in range(3))
    o.lib.orc_pixel_domain(o.h, oux.ctypes.data_as(O._D), ouy.ctypes.data_as(O._D), orho.ctypes.data_as(O._D))
    assert np.array_equal(ux, oux) and np.array_equal(rho, orho)
    o.close()
    g.close()


@pytest.mark.parametrize("name,geom,torus", L3_CASES, ids=[c[0] for c in L3_CASES])
def test_three_levels_fp32_against_fp64_oracle(name, geom, torus):
    """The FP32 three-level step (each level's fused substeps, then each
    pair's interface as overrides; above the finest pair a child that is a
    coarse buffer of the pair below takes the value just formed for it)
    against the FP64 oracle, to FP32 rounding."""
    steps = 100
    banner(f"three-level {name} FP32 vs FP64 oracle {steps} steps")
    pore = geom()
    kn, lp = G.uniform_kn(pore, 0.1, MFP_DX)
    o = O.Oracle(pore, kn, lp, multiblock=3, periodic_y=torus, Fbody=1e-6)
    g = Lattice(pore, kn, lp, precision="fp32dev", multiblock=3, periodic_y=int(torus), Fbody=1e-6)
    o.run("transport", steps, 0.0)
    g.step("transport", steps)
    e = {k: rel_l2(g.get(k), o.get(k)) for k in ("rho", "ux", "uy")}
    print("FP32 three-level rel-l2", e)
    assert e["rho"] < 1e-9 and e["ux"] < 6e-6 and e["uy"] < 7e-6  # measured: 1e-10, 5.9e-7, 7.1e-7
    o.close()
    g.close()


def test_three_levels_fp32_call_chunking_is_exact():
    """2 x 25 coarse steps equal 50 in one call on three levels."""
    banner("three-level sparse256 FP32 50 steps in one call vs two calls")
    pore = G.disk_packing(256, 0.8, 4, 9, seed=3)
    kn, lp = G.uniform_kn(pore, 0.1, MFP_DX)
    a = Lattice(pore, kn, lp, precision="fp32dev", multiblock=3, periodic_y=1, Fbody=1e-6)
    b = Lattice(pore, kn, lp, precision="fp32dev", multiblock=3, periodic_y=1, Fbody=1e-6)
    a.step("transport", 50)
    b.step("transport", 25)
    b.step("transport", 25)
    d = np.abs(a.get("f") - b.get("f")).max()
    print("max |f| diff", d)
    assert d <= 1e-14


# ------------------------------------------------- per-method calls, MDBB
MB_METHOD_CASES = [("boxed128_reference_2", boxed128, False, 2), ("boxed128_reference_3", boxed128, False, 3),
                   ("sparse128_torus_3", lambda: G.disk_packing(128, 0.85, 3, 6, seed=4), True, 3)]


@pytest.mark.parametrize("name,geom,torus,levels", MB_METHOD_CASES, ids=[c[0] for c in MB_METHOD_CASES])
def test_multiblock_method_by_method_against_oracle(name, geom, torus, levels):
    """The transport loop of mantiisApp.hpp:239-246 called method by method
    on a multi-block lattice (equilibrium, evolutionStepOfMultiBlock(maxLevel)
    -- its level substeps and interfaces -- moments, residual), FP64, against
    the oracle's loop: populations and residual bit for bit.  Then one
    regularised collision of a single level: f of every level and fneq
    (formed over all cells, coreLB.hpp:744-792) equal the oracle's."""
    steps = 30
    banner(f"multi-block {name}: {steps} transport steps method by method vs oracle")
    pore = geom()
    kn, lp = G.uniform_kn(pore, 0.1, MFP_DX)
    o = O.Oracle(pore, kn, lp, multiblock=levels, periodic_y=torus, Fbody=1e-6)
    g = Lattice(pore, kn, lp, precision="fp64", multiblock=levels, periodic_y=int(torus), Fbody=1e-6)
    L = int(np.max(g.grid()["level"]))
    assert L == levels
    for _ in range(steps):
        g.method("copy_usq_to_usq_old")
        g.method("equilibrium")
        g.method("evolutionStepOfMultiBlock", L)
        g.method("calculateMacroscopicPorperties")
        g.method("calculateVelocityDifference")
    o.run("transport", steps, 0.0)
    d = np.abs(o.get("f") - g.get("f")).max()
    print("max |f| diff", d, "residual", g.residual(), o.diff_history()[-1])
    assert d <= 1e-15
    # the residual's sums run per level image, in another order than the oracle's
    assert abs(g.residual() - o.diff_history()[-1]) <= 1e-12 * o.diff_history()[-1]
    for lib, orc in ((g, o),):
        lib.method("equilibrium")
        orc.method("equilibrium")
        lib.method("MRTCollisionRegularized", 2)
        orc.method("MRTCollisionRegularized", 2)
    for field in ("f", "fneq", "feq"):
        assert np.array_equal(g.get(field), o.get(field)), field
    o.close()
    g.close()


@pytest.mark.parametrize("levels", [2, 3])
def test_multiblock_mdbb_against_oracle(levels):
    """MDBB walls (coreLB.hpp:936-967) on a quadtree lattice: the transport
    step runs method by method (no fused multi-block kernel takes MDBB), FP64
    against the oracle with the same operators, bit for bit."""
    banner(f"multi-block boxed128 {levels} levels MDBB FP64 vs oracle 30 steps")
    pore = boxed128()
    kn, lp = G.uniform_kn(pore, 0.1, MFP_DX)
    o = O.Oracle(pore, kn, lp, multiblock=levels, Fbody=1e-6)
    o.set_operators(COLL["RMRT"], WALL["MDBB"])
    g = Lattice(pore, kn, lp, precision="fp64", multiblock=levels, Fbody=1e-6, collision="RMRT", wall="MDBB")
    o.run("transport", 30, 0.0)
    g.step("transport", 30)
    d = np.abs(o.get("f") - g.get("f")).max()
    print("max |f| diff", d, "residual", g.residual(), o.diff_history()[-1])
    assert d <= 1e-15 and abs(g.residual() - o.diff_history()[-1]) <= 1e-12 * o.diff_history()[-1]
    o.close()
    g.close()


def test_multiblock_fp32_mdbb_is_refused():
    from mantiis_amd.hip import MtsError
    pore = boxed128()
    kn, lp = G.uniform_kn(pore, 0.1, MFP_DX)
    with pytest.raises(MtsError) as e:
        Lattice(pore, kn, lp, precision="fp32dev", multiblock=3, wall="MDBB")
    assert "FP64" in str(e.value)


@pytest.mark.parametrize("levels", [1, 3])
def test_multiblock_wall_free_torus_is_refused(levels):
    """A torus with no solid has no wall to refine at: every cell lands on
    the coarsest level and the finer ones are empty.  The lattice refuses it
    with a clear error (the single-block lattice covers such a domain)."""
    from mantiis_amd.hip import MtsError
    pore = np.ones((64, 64), np.int32)
    kn, lp = G.uniform_kn(pore, 0.1, MFP_DX)
    with pytest.raises(MtsError) as e:
        Lattice(pore, kn, lp, precision="fp32dev", multiblock=levels, periodic_y=1)
    assert "holds no cells" in str(e.value)
