"""Geometry preprocessing on the device (mts_shape_*): Shape::calculateProperties
(shape.hpp:127-251).  The skeleton, LocPore and Kn must equal the oracle's
bit for bit -- integer and byte work up to the final double division, which
both sides do in IEEE double with the same operands."""
import numpy as np
import pytest

import _oracle as O
from _gpu_util import banner
from mantiis_amd import geometry as G
from mantiis_amd.hip import DeviceArray, MtsError, ShapeProperties, shape_properties

pytestmark = pytest.mark.gpu
RES = 1e-9


def mfp():
    return O.mfp_of()


def check(pore, res=RES):
    lp, kn, sk, _ = O.shape_properties(pore, res, mfp())
    glp, gkn, gsk = shape_properties(pore, res, mfp(), skeleton=True)
    assert np.array_equal(gsk, sk), "skeleton differs"
    assert np.array_equal(glp, lp), "LocPore differs"
    assert np.array_equal(gkn, kn), "Kn differs"
    return glp, gkn, gsk


def test_fixture_pore_and_channels():
    banner("shape: reference fixture pore.dat 64x64, channels 6..33")
    pore, _, _ = O.load_fixture_geometry()
    check(pore)
    for N in (6, 7, 16, 33):
        d = np.ones((N, N), np.int32)
        d[0] = d[-1] = 0
        lp, _, _ = check(d)
        assert np.all(lp[1:-1] == 2 * ((N - 1) // 2))


@pytest.mark.parametrize("n,phi,seed", [(48, 0.5, 1), (64, 0.7, 2), (96, 0.6, 3), (128, 0.45, 4), (200, 0.65, 5),
                                        (256, 0.6, 6)])
def test_disk_packings(n, phi, seed):
    banner(f"shape: disk packing n={n} phi={phi} seed={seed}")
    check(G.disk_packing(n, phi, 3, 16, seed=seed))


def test_non_square_and_tiny():
    banner("shape: non-square and tiny images")
    rng = np.random.default_rng(7)
    for ny, nx in ((2, 2), (3, 5), (5, 3), (8, 40), (40, 8), (70, 130), (130, 70)):
        d = (rng.random((ny, nx)) < 0.8).astype(np.int32)
        d[0] = d[-1] = 0
        try:
            ref = O.shape_properties(d, RES, mfp())
        except RuntimeError:
            with pytest.raises(MtsError):
                shape_properties(d, RES, mfp())
            continue
        glp, gkn, gsk = shape_properties(d, RES, mfp(), skeleton=True)
        assert np.array_equal(gsk, ref[2]) and np.array_equal(glp, ref[0]) and np.array_equal(gkn, ref[1])


def test_no_wall_pixel():
    banner("shape: all-fluid domain (wall distance at OpenCV's infinity)")
    check(np.ones((12, 12), np.int32))


def test_device_resident_and_reuse():
    banner("shape: device-resident handle, two images of one size")
    n = 160
    h = ShapeProperties(n, n)
    for seed in (11, 12):
        pore = G.disk_packing(n, 0.6, 3, 12, seed=seed)
        d = DeviceArray.from_numpy(pore.astype(np.uint8))
        lp, kn = DeviceArray((n, n), np.float64), DeviceArray((n, n), np.float64)
        sk = DeviceArray((n, n), np.uint8)
        it = h.run(d, RES, mfp(), lp, kn, sk)
        rlp, rkn, rsk, rit = O.shape_properties(pore, RES, mfp())
        assert np.array_equal(sk.numpy(), rsk) and np.array_equal(lp.numpy(), rlp)
        assert np.array_equal(kn.numpy(), rkn)
        assert it >= rit and it % 8 == 0
    h.close()


def test_rejects_non_binary():
    banner("shape: domain value 2 rejected")
    d = np.ones((16, 16), np.int32)
    d[0] = d[-1] = 0
    d[5, 5] = 2
    with pytest.raises(MtsError):
        shape_properties(d, RES, mfp())


def test_512_against_oracle():
    banner("shape: disk packing 512x512 against the oracle")
    check(G.disk_packing(512, 0.6, 3, 16, seed=21))


def test_full_size_properties():
    """1024 x 1024 (the largest size the reference recommends): invariants
    that hold at any size -- the skeleton lies in the fluid, every LocPore
    is twice the wall distance of a skeleton pixel, skeleton pixels are their
    own nearest skeleton pixel, and Kn LocPore resolution = mfp."""
    banner("shape: disk packing 1024x1024 invariants")
    pore = G.disk_packing(1024, 0.6, 3, 16, seed=31)
    lp, kn, sk = shape_properties(pore, RES, mfp(), skeleton=True)
    fl = pore == 1
    assert np.all(sk[~fl] == 0) and sk.sum() > 0
    dw, _ = O.distance((pore != 0).astype(np.uint8))
    assert np.all(lp[sk == 1] == 2 * dw[sk == 1].astype(np.float64))
    assert set(np.unique(lp[fl])) <= set(np.unique(2 * dw[sk == 1].astype(np.float64)))
    assert np.all(lp[~fl] == 0) and np.all(kn[~fl] == 0)
    assert np.allclose(kn[fl] * lp[fl] * RES, mfp(), rtol=1e-14)
