"""mts_shape_run_static (the allocation-free form of mts_shape_run, whose
work buffers live in the library's static device array) against
mts_shape_run with a handle: skeleton, LocPore, Kn and the iteration count
bit for bit, on two calls in a row -- the second on a smaller image, so that
what the first call left in the shared array must not leak into it -- and a
too-large image rejected with MTS_ERR_ARG."""
import numpy as np
import pytest

import _oracle as O
from _gpu_util import banner
from mantiis_amd import geometry as G
from mantiis_amd.hip import DeviceArray, MtsError, ShapeProperties, shape_run_static

pytestmark = pytest.mark.gpu
RES = 1e-9


def run_handle(pore):
    ny, nx = pore.shape
    h = ShapeProperties(nx, ny)
    d = DeviceArray.from_numpy(pore.astype(np.uint8))
    lp, kn = DeviceArray((ny, nx), np.float64), DeviceArray((ny, nx), np.float64)
    sk = DeviceArray((ny, nx), np.uint8)
    it = h.run(d, RES, O.mfp_of(), lp, kn, sk)
    out = (lp.numpy(), kn.numpy(), sk.numpy(), it)
    for a in (d, lp, kn, sk):
        a.close()
    h.close()
    return out


def run_static(pore):
    ny, nx = pore.shape
    d = DeviceArray.from_numpy(pore.astype(np.uint8))
    lp, kn = DeviceArray((ny, nx), np.float64), DeviceArray((ny, nx), np.float64)
    sk = DeviceArray((ny, nx), np.uint8)
    it = shape_run_static(d, RES, O.mfp_of(), lp, kn, sk)
    out = (lp.numpy(), kn.numpy(), sk.numpy(), it)
    for a in (d, lp, kn, sk):
        a.close()
    return out


def test_static_shape_run_equals_handle_run():
    first = G.disk_packing(512, 0.6, 4, 12, seed=21)[:, :480].copy()    # 480 x 512, sides up to the limit
    second = G.disk_packing(160, 0.7, 3, 9, seed=22)[:96, :].copy()      # 160 x 96, smaller
    for name, pore in (("480x512", first), ("160x96", second)):
        banner(f"mts_shape_run_static vs mts_shape_run, {name}")
        a = run_handle(pore)
        b = run_static(pore)
        assert np.array_equal(a[2], b[2]), f"{name}: skeleton differs"
        assert np.array_equal(a[0], b[0]), f"{name}: LocPore differs"
        assert np.array_equal(a[1], b[1]), f"{name}: Kn differs"
        assert a[3] == b[3], f"{name}: iterations {a[3]} vs {b[3]}"
        print(name, "equal; iterations", b[3], flush=True)
    # the oracle agrees on the smaller image as well
    lp, kn, sk, _ = O.shape_properties(second, RES, O.mfp_of())
    b = run_static(second)
    assert np.array_equal(b[2], sk) and np.array_equal(b[0], lp) and np.array_equal(b[1], kn)


def test_static_shape_run_rejects_sides_beyond_its_memory():
    banner("mts_shape_run_static rejects 600 x 600")
    d = DeviceArray((600, 600), np.uint8)
    lp, kn = DeviceArray((600, 600), np.float64), DeviceArray((600, 600), np.float64)
    with pytest.raises(MtsError) as e:
        shape_run_static(d, RES, O.mfp_of(), lp, kn)
    assert e.value.code == -1 and "512" in str(e.value)
    for a in (d, lp, kn):
        a.close()
