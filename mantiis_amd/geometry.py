"""Seeded synthetic porous geometries in the reference's input conventions.

Disk packings are sanitised so that every fluid site's 3x3 wall pattern is one
of the 52 configurations the reference supports (an unsupported pattern makes
Grid2D::getBoundaryTypes throw, grid.hpp:1143-1150): offending fluid sites are
turned solid until none is left.  Kn / local pore size follow the reference's
validation mode (shape.hpp:247-249): a uniform Kn with LocPore = (mfp/dx)/Kn.
"""
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
_EX = [0, 1, 0, -1, 0, 1, -1, -1, 1]
_EY = [0, 0, 1, 0, -1, 1, 1, -1, -1]
_ORDER = [7, 3, 6, 4, 0, 2, 8, 1, 5]


def supported_codes():
    bt = np.loadtxt(os.path.join(HERE, "data", "system_files", "bnd_types.txt")).astype(int)
    ok = np.zeros(512, bool)
    ok[bt] = True
    ok[511] = True
    return ok


def wall_codes(pore):
    """9-bit neighbourhood code of every site (grid.hpp:1111-1141), x/y periodic."""
    code = np.zeros(pore.shape, np.int32)
    for k in _ORDER:
        nb = np.roll(pore, shift=(-_EY[k], -_EX[k]), axis=(0, 1))
        code = (code << 1) | (nb != 0).astype(np.int32)
    return code


def sanitize(pore, max_iter=100000):
    """Turn fluid sites with an unsupported wall pattern solid, to a fixed point.
    After the first sweep only the 3x3 neighbourhoods of changed sites are
    re-examined."""
    pore = pore.copy().astype(np.int32)
    ok = supported_codes()
    n0, n1 = pore.shape
    ys, xs = np.nonzero(pore == 1)
    for _ in range(max_iter):
        code = np.zeros(ys.shape, np.int32)
        for k in _ORDER:
            code = (code << 1) | (pore[(ys + _EY[k]) % n0, (xs + _EX[k]) % n1] != 0)
        bad = (pore[ys, xs] == 1) & ~ok[code]
        if not bad.any():
            return pore
        by, bx = ys[bad], xs[bad]
        pore[by, bx] = 0
        cy = np.concatenate([(by + dy) % n0 for dy in (-1, 0, 1) for dx in (-1, 0, 1)])
        cx = np.concatenate([(bx + dx) % n1 for dy in (-1, 0, 1) for dx in (-1, 0, 1)])
        lin = np.unique(cy.astype(np.int64) * n1 + cx)
        ys, xs = (lin // n1).astype(np.int64), (lin % n1).astype(np.int64)
    raise RuntimeError("sanitize did not converge")


def disk_packing(n, porosity, rmin, rmax, seed, solid_rows=False):
    """Overlapping disks (centres U[0,n)^2, radius U[rmin,rmax]) added until the
    fluid fraction drops to `porosity`; periodic in x and y.  Each disk is
    stamped in its own bounding window, so the cost is O(disks * r^2)."""
    rng = np.random.default_rng(seed)
    pore = np.ones((n, n), np.int32)
    fluid = n * n
    target = porosity * n * n
    R = int(np.ceil(rmax)) + 1
    off = np.arange(-R, R + 1)
    while fluid > target:
        cx, cy = rng.uniform(0, n, 2)
        r = rng.uniform(rmin, rmax)
        xs = np.floor(cx).astype(int) + off
        ys = np.floor(cy).astype(int) + off
        dx = xs - cx
        dy = ys - cy
        inside = (dy[:, None] ** 2 + dx[None, :] ** 2) <= r * r
        sub = pore[np.ix_(ys % n, xs % n)]
        newly = int((sub[inside] == 1).sum())
        sub[inside] = 0
        pore[np.ix_(ys % n, xs % n)] = sub
        fluid -= newly
    if solid_rows:
        pore[0, :] = 0
        pore[-1, :] = 0
    return sanitize(pore)


def uniform_kn(pore, kn, mfp_over_dx):
    knf = np.where(pore == 1, kn, 0.0)
    lp = np.where(pore == 1, mfp_over_dx / kn, 0.0)
    return knf, lp


def write_reference_format(folder, pore, kn, locpore):
    """Column-major text files read by Shape::loadExistingModel (shape.hpp:253-285)."""
    os.makedirs(folder, exist_ok=True)
    for name, a, fmt in (("pore.dat", pore, "%d"), ("Kn.dat", kn, "%.6e"), ("localporesize.dat", locpore, "%.6e")):
        np.savetxt(os.path.join(folder, name), np.asarray(a).T.reshape(-1), fmt=fmt)
