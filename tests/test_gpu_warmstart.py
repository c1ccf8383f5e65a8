"""What the learned warm start gains: a small network trained on steady
states of the batched solver seeds unseen lattices, which then reach the
same tolerance in fewer steps than from rest.  Both runs stop at the same
per-step tolerance; their permeabilities are compared with a tightly
converged reference (a warm start approaches from either side, so its
stopping error can differ in sign from the cold start's)."""
import numpy as np
import pytest

from _gpu_util import banner

pytestmark = pytest.mark.gpu
F = 1e-6


def test_warm_start_saves_steps_on_unseen_lattices():
    from mantiis_amd import warmstart as WS
    from mantiis_amd.hip import globals_for
    banner("warm start: train on 384 steady 128^2 lattices, seed 16 unseen ones")
    rho_in = globals_for(dx=1e-9)[1]["rhoin"]
    train_s = WS.samples(range(100, 484), 128, workers=8)
    r, fields = WS.solve(train_s, F)
    assert (r["steps"] > 0).all()
    ds = WS.Dataset([WS.features(s, F) for s in train_s], [WS.targets(s, f, F) for s, f in zip(train_s, fields)])
    net, hist = WS.train(ds, c0=8, batch=8, epochs=30, lr=2e-3)
    print("training loss", hist[0], "->", hist[-1])
    assert hist[-1] < 0.5 * hist[0]
    test_s = WS.samples(range(900, 916), 128, workers=8)
    rc, fc = WS.solve(test_s, F)
    pred = WS.predict(net.params(), 8, test_s, F)
    warm = [WS.seed_populations(s, ux, uy, rho_in) for s, (ux, uy) in zip(test_s, pred)]
    rw, fw = WS.solve(test_s, F, warm=warm)
    kc = np.array([WS.permeability(s, f[1], F) for s, f in zip(test_s, fc)])
    kw = np.array([WS.permeability(s, f[1], F) for s, f in zip(test_s, fw)])
    kp = np.array([WS.permeability(s, p[0], F) for s, p in zip(test_s, pred)])
    rr, fr = WS.solve(test_s, F, tol=1e-9, max_steps=300000)
    kr = np.array([WS.permeability(s, f[1], F) for s, f in zip(test_s, fr)])
    print("cold steps", rc["steps"], "warm steps", rw["steps"])
    print("k error per lattice: network", np.round(kp / kr - 1, 3))
    print("k error vs the 1e-9 reference: network", np.abs(kp / kr - 1).max(), "cold", np.abs(kc / kr - 1).max(),
          "warm", np.abs(kw / kr - 1).max())
    assert (rc["steps"] > 0).all() and (rw["steps"] > 0).all()
    assert rw["steps"].sum() < rc["steps"].sum()
    # the per-step stopping test is met early by a run that approaches from a
    # flux far above the steady one (the slowest lattices relax over ~1e5
    # steps): a seed that over-predicts several-fold can stop a few percent
    # away, so the warm runs are held to the median and a wider maximum
    assert np.abs(kc / kr - 1).max() < 0.05
    assert np.median(np.abs(kw / kr - 1)) < 0.01 and np.abs(kw / kr - 1).max() < 0.1


def test_warm_start_at_config_sizes():
    """The warm start at the sizes of its configuration: trained on steady
    states of 256^2 lattices of the batched distribution, evaluated on unseen
    512^2 lattices.  Warm-started and cold runs stop at the same tolerance
    with the same permeability (within the stopping error), and the steps of
    both are reported; the network's own permeability error is printed."""
    from mantiis_amd import warmstart as WS
    from mantiis_amd.hip import globals_for
    banner("warm start: train on 256 steady 256^2 lattices, seed 8 unseen 512^2 ones")
    rho_in = globals_for(dx=1e-9)[1]["rhoin"]
    train_s = WS.samples(range(20000, 20256), 256, workers=8)
    r, fields = WS.solve(train_s, F)
    assert (r["steps"] > 0).all()
    ds = WS.Dataset([WS.features(s, F) for s in train_s], [WS.targets(s, f, F) for s, f in zip(train_s, fields)])
    net, hist = WS.train(ds, c0=16, batch=8, epochs=12, lr=2e-3)
    print("training loss", hist[0], "->", hist[-1], flush=True)
    assert hist[-1] < hist[0]
    test_s = WS.samples(range(30000, 30008), 512, workers=8)
    rc, fc = WS.solve(test_s, F)
    pred = WS.predict(net.params(), 16, test_s, F)
    warm = [WS.seed_populations(s, ux, uy, rho_in) for s, (ux, uy) in zip(test_s, pred)]
    rw, fw = WS.solve(test_s, F, warm=warm)
    kc = np.array([WS.permeability(s, f[1], F) for s, f in zip(test_s, fc)])
    kw = np.array([WS.permeability(s, f[1], F) for s, f in zip(test_s, fw)])
    kp = np.array([WS.permeability(s, p[0], F) for s, p in zip(test_s, pred)])
    print("512^2 cold steps", rc["steps"], "warm steps", rw["steps"], flush=True)
    print("network k error vs the cold steady state", np.round(kp / kc - 1, 3), flush=True)
    print("warm vs cold steady k", np.abs(kw / kc - 1).max(), flush=True)
    assert (rc["steps"] > 0).all() and (rw["steps"] > 0).all()
    assert np.abs(kw / kc - 1).max() < 0.02


def test_device_seed_equals_host_seed():
    """The warm-start populations formed on the device from a velocity field
    (mts_batch_seed) equal seed_populations' host formulas: the same FP64
    operations in the same order, so the runs that start from them agree bit
    for bit."""
    from mantiis_amd import warmstart as WS
    from mantiis_amd.hip import globals_for
    banner("warm start seeding on the device vs on the host, 3 lattices of 128^2")
    rho_in = globals_for(dx=1e-9)[1]["rhoin"]
    s_list = WS.samples(range(900, 903), 128, workers=2)
    _, fields = WS.solve(s_list, F, tol=1e-3)
    us = [(f[1] * 1.3, f[2] * 0.7) for f in fields]  # a velocity that is not the steady one
    host = [WS.seed_populations(s, ux, uy, rho_in) for s, (ux, uy) in zip(s_list, us)]
    ra, fa = WS.solve(s_list, F, warm=host, tol=1e-300, max_steps=200)
    rb, fb = WS.solve(s_list, F, warm_u=us, rho_in=rho_in, tol=1e-300, max_steps=200)
    for a, b in zip(fa, fb):
        for x, y in zip(a, b):
            assert np.array_equal(x, y)
