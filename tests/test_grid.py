"""Grid2D built by the library on the host (mts_grid_*) against the oracle's
restatement of grid.hpp: cell lists, levels, connectivity, boundary types,
buffer cells, parent/child rows and relaxation times, bit for bit."""
import numpy as np
import pytest

import _oracle as O
from mantiis_amd import geometry as G
from mantiis_amd.hip import HostGrid, MtsError

MFP_DX = 0.7730846760658538


def geometries():
    out = []
    pore, kn, lp = O.load_fixture_geometry()
    out.append(("fixture64", pore, kn, lp))
    for n, phi, seed in ((64, 0.8, 2), (128, 0.7, 3), (128, 0.55, 9)):
        p = G.disk_packing(n, phi, 3, 8, seed=seed)
        k, l = G.uniform_kn(p, 0.1, MFP_DX)
        out.append((f"disks{n}_{seed}", p, k, l))
    # closed box with an obstacle: interfaces away from the seams
    N = 64
    p = np.ones((N, N), np.int32)
    p[0, :] = p[-1, :] = p[:, 0] = p[:, -1] = 0
    yy, xx = np.mgrid[0:N, 0:N]
    p[(xx - 32) ** 2 + (yy - 30) ** 2 <= 36] = 0
    p = G.sanitize(p)
    k, l = G.uniform_kn(p, 0.2, MFP_DX)
    out.append(("box64", p, k, l))
    return out


GEOMS = geometries()


@pytest.mark.parametrize("torus", [False, True])
@pytest.mark.parametrize("multiblock", [False, True])
@pytest.mark.parametrize("geom", GEOMS, ids=[g[0] for g in GEOMS])
def test_library_grid_equals_oracle(geom, multiblock, torus):
    name, pore, kn, lp = geom
    o = O.Oracle(pore, kn, lp, multiblock=multiblock, periodic_y=torus)
    g = HostGrid(pore, kn, lp, multiblock=int(multiblock), periodic_y=int(torus))
    n, L = g.size()
    assert n == o.ncell and L == o.lib.orc_max_level(o.h)
    # a multi-block grid's relaxation times come from per-level pore sizes
    # computed on the device: tests/test_gpu_multiblock.py compares them
    a, b = o.grid(), g.arrays(taus=not multiblock)
    for k in ("gridID", "gridIJ", "level", "connect", "bnd", "isBuffer", "type"):
        assert np.array_equal(a[k], b[k]), k
    if not multiblock:
        assert np.array_equal(o.get("taus"), b["taus"])
    if multiblock:
        for which in ("parent_child", "coarse", "fine"):
            assert np.array_equal(O.oracle_rows(o, which), g.rows(which)), which
        if name != "fixture64":
            assert L == 2 and (b["type"] == 12).sum() > 0 and (b["type"] == 21).sum() > 0


def test_multiblock_structure_invariants():
    """Coarse cells are aligned 2x2 all-fluid blocks at distance >= 6 from every
    wall pixel; fine buffers come in fours under their parent; every coarse
    buffer row lists its four level-1 children."""
    pore = G.disk_packing(128, 0.7, 3, 8, seed=3)
    kn, lp = G.uniform_kn(pore, 0.1, MFP_DX)
    g = HostGrid(pore, kn, lp, multiblock=1, periodic_y=1)
    a = g.arrays(taus=False)
    walls = np.argwhere(pore == 0)
    N = pore.shape[0]
    for c in np.nonzero((a["level"] == 2) & (a["isBuffer"] == 0))[0][:200]:
        x, y = a["gridIJ"][c]
        assert x % 2 == 0 and y % 2 == 0
        blk = pore[y:y + 2, x:x + 2]
        assert blk.all()
        dy = np.abs(walls[:, 0][:, None] - np.array([y, y, y + 1, y + 1])[None, :])
        dx = np.abs(walls[:, 1][:, None] - np.array([x, x + 1, x, x + 1])[None, :])
        dy = np.minimum(dy, N - dy)
        dx = np.minimum(dx, N - dx)
        assert (dx * dx + dy * dy).min() >= 36
    fine = g.rows("fine")
    for row in fine:
        px, py = a["gridIJ"][row[0]]
        kids = {tuple(a["gridIJ"][k]) for k in row[1:]}
        assert kids == {(px, py), (px + 1, py), (px, py + 1), (px + 1, py + 1)}
        assert (a["type"][row[1:]] == 12).all()
    for row in g.rows("coarse"):
        assert a["type"][row[0]] == 21
        assert (row[1:] >= 0).all() and (a["level"][row[1:]] == 1).all()


def test_unsupported_wall_pattern_raises():
    """A lone fluid pixel between walls (grid.hpp:1143-1150)."""
    pore = np.ones((16, 16), np.int32)
    pore[5, :] = 0
    pore[7, :] = 0
    pore[6, :] = 0
    pore[6, 8] = 1
    kn = np.where(pore == 1, 0.1, 0.0)
    with pytest.raises(MtsError) as e:
        HostGrid(pore, kn, kn * 10)
    assert e.value.code == -2 and "Boundary type is not contained" in str(e.value)
    with pytest.raises(RuntimeError):
        O.Oracle(pore, kn, kn * 10)
