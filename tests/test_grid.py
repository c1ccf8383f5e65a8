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
@pytest.mark.parametrize("multiblock", [0, 1, 3], ids=["single", "quadtree2", "quadtree3"])
@pytest.mark.parametrize("geom", GEOMS, ids=[g[0] for g in GEOMS])
def test_library_grid_equals_oracle(geom, multiblock, torus):
    name, pore, kn, lp = geom
    o = O.Oracle(pore, kn, lp, multiblock=multiblock, periodic_y=torus)
    g = HostGrid(pore, kn, lp, multiblock=multiblock, periodic_y=int(torus))
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
        assert np.array_equal(O.oracle_rows(o, "parent_child"), g.rows("parent_child"))
        for key, w in ((21, 0), (32, 0), (12, 1), (23, 1)):
            assert np.array_equal(O.oracle_rows_key(o, w, key).reshape(-1, 5), g.rows(key)), key
        if name != "fixture64" and multiblock == 1:
            assert L == 2 and (b["type"] == 12).sum() > 0 and (b["type"] == 21).sum() > 0


@pytest.mark.parametrize("torus", [False, True])
def test_three_level_grid_equals_oracle(torus):
    """A packing sparse enough for level-3 cells (4 x 4 blocks at distance
    >= 24 from every wall pixel; level 2 then needs distance >= 12): every
    table and the rows of each buffer type (21, 32, 12, 23), bit for bit."""
    pore = G.disk_packing(256, 0.8, 4, 9, seed=3)
    kn, lp = G.uniform_kn(pore, 0.1, MFP_DX)
    o = O.Oracle(pore, kn, lp, multiblock=3, periodic_y=torus)
    g = HostGrid(pore, kn, lp, multiblock=3, periodic_y=int(torus))
    n, L = g.size()
    assert n == o.ncell and L == o.lib.orc_max_level(o.h) == 3
    a, b = o.grid(), g.arrays(taus=False)
    for k in ("gridID", "gridIJ", "level", "connect", "bnd", "isBuffer", "type"):
        assert np.array_equal(a[k], b[k]), k
    assert np.array_equal(O.oracle_rows(o, "parent_child"), g.rows("parent_child"))
    for key, w in ((21, 0), (32, 0), (12, 1), (23, 1)):
        rows = g.rows(key)
        assert len(rows) > 0, key
        assert np.array_equal(O.oracle_rows_key(o, w, key), rows), key
    # the flat lists are the pairs one after the other
    assert np.array_equal(g.rows("coarse"), np.concatenate([g.rows(21), g.rows(32)]))
    assert np.array_equal(g.rows("fine"), np.concatenate([g.rows(12), g.rows(23)]))
    o.close()


def test_three_level_structure_invariants():
    """Cells of level l >= 2 are aligned all-fluid blocks of side 2^(l-1) at
    distance >= ceil(4 sqrt 2) 2^(l-1) = 6 * 2^(l-1) from every wall pixel
    (minimum image on the torus), and the block that holds a level-2 cell is
    not itself at distance >= 24; each fine row lists the four quadrants of
    its parent."""
    pore = G.disk_packing(256, 0.8, 4, 9, seed=3)
    kn, lp = G.uniform_kn(pore, 0.1, MFP_DX)
    g = HostGrid(pore, kn, lp, multiblock=3, periodic_y=1)
    a = g.arrays(taus=False)
    walls = np.argwhere(pore == 0)
    N = pore.shape[0]

    def mind2(x, y, h):
        px = np.arange(x, x + h).repeat(h)
        py = np.tile(np.arange(y, y + h), h)
        dy = np.abs(walls[:, 0][:, None] - py[None, :])
        dx = np.abs(walls[:, 1][:, None] - px[None, :])
        dy = np.minimum(dy, N - dy)
        dx = np.minimum(dx, N - dx)
        return (dx * dx + dy * dy).min()

    orig = a["isBuffer"] == 0
    for lvl in (2, 3):
        h = 1 << (lvl - 1)
        cells = np.nonzero((a["level"] == lvl) & orig)[0]
        assert len(cells) > 0
        for c in cells[:: max(1, len(cells) // 60)]:
            x, y = a["gridIJ"][c]
            assert x % h == 0 and y % h == 0 and pore[y:y + h, x:x + h].all()
            assert mind2(x, y, h) >= (6 * h) ** 2
            if lvl == 2:
                bx, by = x - x % 4, y - y % 4
                assert mind2(bx, by, 4) < 24 ** 2
    for key in (12, 23):
        for row in g.rows(key)[:200]:
            px, py = a["gridIJ"][row[0]]
            hj = 1 << (a["level"][row[0]] - 2)
            kids = {tuple(a["gridIJ"][k]) for k in row[1:]}
            assert kids == {(px, py), (px + hj, py), (px, py + hj), (px + hj, py + hj)}, key


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
