"""Pins of the oracle's geometry preprocessing (Shape::calculateProperties,
shape.hpp:127-251) -- CPU only.

The reference's fixture test/test_input_output/localporesize.dat is not an
output of its current calculateProperties (no thinning bound, padding or
search variant reproduces it), so it pins nothing here; the checks below
come from first principles instead:
  * thinning against a numpy Zhang-Suen written from the paper's P2..P9
    notation (with the reference's bound B <= 5, which the textbook B <= 6
    contradicts on the same images), and known answers;
  * distances against a brute-force minimum over all zero pixels;
  * the nearest-skeleton search against a brute-force first-in-row-major
    search of the skeleton pixels at the minimum distance;
  * a straight channel's LocPore against its closed form;
  * invariants of the whole map on disk packings.
"""
import numpy as np
import pytest

import _oracle as O


def zhang_suen(img, bmax=5):
    """Zhang & Suen (1984): P2..P9 = N, NE, E, SE, S, SW, W, NW of P1; a set
    P1 is deleted when 2 <= B(P1) <= bmax, A(P1) = 1 and, in the first
    subiteration, P2 P4 P6 = P4 P6 P8 = 0, in the second P2 P4 P8 =
    P2 P6 P8 = 0; border pixels are kept."""
    img = np.array(img, np.uint8)
    H, W = img.shape
    while True:
        changed = False
        for first in (True, False):
            P = np.pad(img, 1).astype(int)
            p2, p3, p4 = P[:-2, 1:-1], P[:-2, 2:], P[1:-1, 2:]
            p5, p6, p7 = P[2:, 2:], P[2:, 1:-1], P[2:, :-2]
            p8, p9 = P[1:-1, :-2], P[:-2, :-2]
            seq = [p2, p3, p4, p5, p6, p7, p8, p9, p2]
            A = sum((seq[k] == 0) & (seq[k + 1] == 1) for k in range(8))
            B = p2 + p3 + p4 + p5 + p6 + p7 + p8 + p9
            if first:
                c = (p2 * p4 * p6 == 0) & (p4 * p6 * p8 == 0)
            else:
                c = (p2 * p4 * p8 == 0) & (p2 * p6 * p8 == 0)
            d = (img == 1) & (A == 1) & (B >= 2) & (B <= bmax) & c
            d[0, :] = d[-1, :] = d[:, 0] = d[:, -1] = False
            if d.any():
                img[d] = 0
                changed = True
        if not changed:
            return img


def random_images(n, seed=0):
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(n):
        H, W = int(rng.integers(8, 40)), int(rng.integers(8, 40))
        img = (rng.random((H, W)) < rng.uniform(0.5, 0.9)).astype(np.uint8)
        yy, xx = np.mgrid[:H, :W]
        for _ in range(int(rng.integers(0, 5))):
            cy, cx, r = rng.integers(0, H), rng.integers(0, W), rng.integers(1, 6)
            img[(yy - cy) ** 2 + (xx - cx) ** 2 <= r * r] = int(rng.integers(0, 2))
        out.append(img)
    return out


def test_thinning_matches_paper_rule():
    differs_from_textbook = 0
    for img in random_images(40):
        got, _ = O.thinning(img)
        assert np.array_equal(got, zhang_suen(img, 5))
        differs_from_textbook += not np.array_equal(got, zhang_suen(img, 6))
    assert differs_from_textbook > 0


def test_thinning_known_answers():
    img = np.zeros((8, 8), np.uint8)
    img[3:5, 3:5] = 1  # a 2x2 block: every pixel deletable on the first pass
    assert O.thinning(img)[0].sum() == 0
    img = np.zeros((7, 9), np.uint8)
    img[3, 2:7] = 1  # a one-pixel line and an isolated pixel stay
    img[1, 1] = 1
    got, it = O.thinning(img)
    assert np.array_equal(got, img) and it == 1
    img = np.ones((6, 6), np.uint8)  # the border is never touched
    got, _ = O.thinning(img)
    assert got[0].all() and got[-1].all() and got[:, 0].all() and got[:, -1].all()
    sub = zhang_suen(img, 5)
    assert np.array_equal(got, sub)


def brute_distance(fg):
    H, W = fg.shape
    zr, zc = np.nonzero(fg == 0)
    if zr.size == 0:
        return np.full((H, W), np.sqrt(np.float32(1e15)), np.float32), np.full((H, W), -1)
    yy, xx = np.mgrid[:H, :W]
    k2 = ((yy[..., None] - zr) ** 2 + (xx[..., None] - zc) ** 2).min(axis=2)
    return np.sqrt(k2.astype(np.float32)), k2


def test_distance_brute_force():
    for img in random_images(25, seed=3) + [np.ones((5, 7), np.uint8)]:
        v, k = O.distance(img)
        bv, bk = brute_distance(img)
        assert np.array_equal(k, bk)
        assert np.array_equal(v, bv)


def test_nearest_skeleton_brute_force():
    rng = np.random.default_rng(5)
    for _ in range(20):
        H, W = int(rng.integers(4, 30)), int(rng.integers(4, 30))
        sk = (rng.random((H, W)) < rng.choice([0.02, 0.1, 0.3])).astype(np.uint8)
        sk[rng.integers(0, H), rng.integers(0, W)] = 1
        v, k = O.distance((sk == 0).astype(np.uint8))
        sr, sc = np.nonzero(sk)
        for i in range(H):
            for j in range(W):
                d2 = (sr - i) ** 2 + (sc - j) ** 2
                at_min = np.nonzero(d2 == d2.min())[0]
                want = int(sr[at_min[0]] * W + sc[at_min[0]])  # np.nonzero is row-major
                assert O.nearest_set(sk, i, j, float(v[i, j])) == want


@pytest.mark.parametrize("N", [6, 7, 16, 17, 32, 33])
def test_channel_closed_form(N):
    """Walls on rows 0 and N-1: the skeleton is the middle row and LocPore
    is twice its distance to the nearer wall, 2 floor((N-1)/2)."""
    mfp = O.mfp_of()
    d = np.ones((N, N), np.int32)
    d[0] = d[-1] = 0
    lp, kn, sk, _ = O.shape_properties(d, 1e-9, mfp)
    assert np.all(lp[1:-1] == 2 * ((N - 1) // 2))
    assert np.all(lp[[0, -1]] == 0) and np.all(kn[[0, -1]] == 0)
    assert np.allclose(kn[1:-1], mfp / lp[1:-1] / 1e-9, rtol=1e-15)
    assert set(np.nonzero(sk)[0]) == {(N - 1) // 2}


def test_map_invariants_on_disk_packings():
    import os
    import sys
    sys.path.insert(0, os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
    from mantiis_amd import geometry as G
    mfp = O.mfp_of()
    for N, seed in ((48, 1), (64, 2), (80, 3)):
        pore = G.disk_packing(N, 0.6, 3, 10, seed=seed)
        lp, kn, sk, _ = O.shape_properties(pore, 1e-9, mfp)
        fl = pore == 1
        assert np.all(sk[~fl] == 0)
        dw, _ = O.distance((pore != 0).astype(np.uint8))
        allowed = set(np.round(2 * dw[sk == 1].astype(np.float64), 9))
        assert set(np.round(lp[fl], 9)) <= allowed
        assert np.all(lp[fl] > 0) and np.all(lp[~fl] == 0)
        assert np.allclose(kn[fl] * lp[fl] * 1e-9, mfp, rtol=1e-14)
        # skeleton pixels are their own nearest skeleton pixel
        assert np.all(lp[sk == 1] == 2 * dw[sk == 1].astype(np.float64))


def test_rejects_non_binary():
    d = np.ones((8, 8), np.int32)
    d[3, 3] = 2
    with pytest.raises(RuntimeError):
        O.shape_properties(d, 1e-9, O.mfp_of())


def test_multiblock_taus_channel_closed_form():
    """LB2D::calculateMultigridTauS (coreLB.hpp:432-480) in the oracle: in a
    straight channel (walls on rows 0 and N-1) level L sees a channel of
    N / 2^(L-1) rows with forced wall rows, so LocPore_L = 2 floor((side-1)/2)
    at resolution 1e-9 2^(L-1), and every level-L cell has
    tau = 0.5 + sqrt(6/pi) LocPore Kn / (1 + 2 Kn)."""
    import os
    import sys
    sys.path.insert(0, os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
    from mantiis_amd import geometry as G
    N = 64
    chan = np.ones((N, N), np.int32)
    chan[0] = chan[-1] = 0
    kn, lp = G.uniform_kn(chan, 0.1, 0.7730846760658538)
    o = O.Oracle(chan, kn, lp, multiblock=True)
    t, lvl = o.get("taus"), o.grid()["level"]
    mfp = O.mfp_of()
    for L, side in ((1, N), (2, N // 2)):
        lpL = 2 * ((side - 1) // 2)
        knL = mfp / lpL / (1e-9 * 2 ** (L - 1))
        assert (lvl == L).any()
        assert np.allclose(t[lvl == L], 0.5 + np.sqrt(6 / np.pi) * lpL * knL / (1 + 2 * knL), rtol=1e-14)
