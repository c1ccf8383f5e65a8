"""Batched solver (mts_batch_*): independent single-block lattices stepped by
one launch per step over a work queue of unconverged lattices.  Each lattice
of a batch must follow exactly the single-block FP32 path of the same
lattice; graph replays must equal eager launches; converged lattices leave
the queue and stop changing."""
import numpy as np
import pytest

from _gpu_util import banner, rel_l2
from mantiis_amd import geometry as G
from mantiis_amd.hip import Batch, Lattice

pytestmark = pytest.mark.gpu
MFP_DX = 0.7730846760658538


def geoms():
    out = []
    for n, phi, kn, seed in ((64, 0.7, 0.1, 1), (96, 0.55, 0.03, 2), (128, 0.75, 0.8, 3)):
        pore = G.disk_packing(n, phi, 3, 9, seed=seed)
        k, lp = G.uniform_kn(pore, kn, MFP_DX)
        out.append((pore, k, lp))
    return out


PARAMS = dict(collision=2, wall=2, dx=1e-9, Fbody=1e-6)


@pytest.mark.parametrize("batch_steps", [1, 2])
def test_batch_lattices_equal_single_block(batch_steps):
    """batch_steps 1: the one-step kernel equals the single-block path bit for
    bit (the same cell-order gather); 2: the tile kernels agree to FP32
    rounding (the same functions in another inlining context may contract
    multiply-adds differently)."""
    banner(f"batch of 3 lattices (batch_steps {batch_steps}) vs single-block FP32 runs, 300 steps")
    gs = geoms()
    b = Batch(gs, batch_steps=batch_steps, **PARAMS)
    r = b.run(300, check_every=300, tol=1e-300, graph=False)
    assert (r["steps"] == -300).all()
    for g, (pore, kn, lp) in enumerate(gs):
        lat = Lattice(pore, kn, lp, precision="fp32dev", periodic_y=1, **PARAMS)
        lat.step("transport", 300)
        ux, uy, rho = lat.complete_domain()
        brho, bux, buy = b.fields(g)
        fl = pore.reshape(-1) == 1
        d = max(np.abs(bux.reshape(-1) - ux)[fl].max(), np.abs(buy.reshape(-1) - uy)[fl].max(),
                np.abs(brho.reshape(-1) - rho)[fl].max())
        print("lattice", g, "max |diff|", d, "rel-l2 ux", rel_l2(bux.reshape(-1)[fl], ux[fl]))
        if batch_steps == 1:
            assert d == 0.0
        else:
            assert d <= 1e-8 and rel_l2(bux.reshape(-1)[fl], ux[fl]) < 1e-6


def test_batch_beyond_4gib_of_populations():
    """A batch whose population planes span more than 4 GiB (240 lattices of
    1024^2, about 150 M cells, so plane offsets need 64 bits): its last
    lattice equals the same lattice run alone, bit for bit."""
    banner("batch of 240 x 1024^2 lattices (> 4 GiB of populations) vs the last lattice alone, 50 steps")
    pore = G.disk_packing(1024, 0.6, 3, 16, seed=11)
    k, lp = G.uniform_kn(pore, 0.2, MFP_DX)
    n = 240
    b = Batch([(pore, k, lp)] * n, **PARAMS)
    total = int(b.cells().sum())
    print("cells", total, "population bytes per buffer", 36 * total, flush=True)
    assert 36 * total > 2 ** 32
    b.run(50, check_every=50, tol=1e-300, graph=False)
    last = b.fields(n - 1)
    b.close()
    c = Batch([(pore, k, lp)], **PARAMS)
    c.run(50, check_every=50, tol=1e-300, graph=False)
    one = c.fields(0)
    c.close()
    for x, y in zip(last, one):
        assert np.array_equal(x, y)


def test_batch_graph_replays_equal_eager():
    banner("batch graph replays vs eager launches, 6 chunks of 50 steps")
    gs = geoms()
    a = Batch(gs, **PARAMS)
    b = Batch(gs, **PARAMS)
    for _ in range(6):
        ra = a.run(50, check_every=50, tol=1e-300, graph=True)
        rb = b.run(50, check_every=50, tol=1e-300, graph=False)
        assert np.array_equal(ra["ux"], rb["ux"])
        for g in range(len(gs)):
            fa, fb = a.fields(g), b.fields(g)
            assert all(np.array_equal(x, y) for x, y in zip(fa, fb))


def test_batch_converges_and_freezes_converged_lattices():
    banner("batch run to 1e-6, converged lattices leave the queue")
    gs = geoms()
    b = Batch(gs, **PARAMS)
    r = b.run(200000, check_every=100, tol=1e-6, graph=True)
    print("steps", r["steps"], "seconds", r["seconds"], "ux", r["ux"], "updates", r["updates"])
    assert (r["steps"] > 0).all() and (r["ux"] > 0).all()
    assert r["updates"] == sum(int(s) * int(n) for s, n in zip(r["steps"], b.cells()))
    before = [b.fields(g) for g in range(len(gs))]
    r2 = b.run(400, check_every=100, tol=1e-6)  # nothing is active any more
    assert r2["updates"] == 0
    after = [b.fields(g) for g in range(len(gs))]
    assert all(np.array_equal(x, y) for fa, fb in zip(before, after) for x, y in zip(fa, fb))


def test_batch_warm_start_from_converged_state():
    banner("batch warm start: a converged state converges again at the first check")
    gs = geoms()[:2]
    b = Batch(gs, **PARAMS)
    r = b.run(200000, check_every=100, tol=1e-6)
    cold = r["steps"].copy()
    w = np.array([4 / 9] + [1 / 9] * 4 + [1 / 36] * 4)
    ex = np.array([0, 1, 0, -1, 0, 1, -1, -1, 1])
    ey = np.array([0, 0, 1, 0, -1, 1, 1, -1, -1])
    c = Batch(gs, **PARAMS)
    for g in range(len(gs)):
        rho, ux, uy = b.fields(g)
        cu = ex[:, None, None] * ux[None] + ey[:, None, None] * uy[None]
        f = w[:, None, None] * rho[None] * (1 + 3 * cu + 4.5 * cu ** 2 - 1.5 * (ux ** 2 + uy ** 2)[None])
        c.set_state(g, f)
    r = c.run(200000, check_every=100, tol=1e-6)
    print("cold", cold, "warm", r["steps"])
    assert (r["steps"] > 0).all() and (r["steps"] < cold).all()
    # seeding several lattices in one launch equals seeding them one by one
    d = Batch(gs, **PARAMS)
    e = Batch(gs, **PARAMS)
    seeds = []
    for g in range(len(gs)):
        rho, ux, uy = b.fields(g)
        cu = ex[:, None, None] * ux[None] + ey[:, None, None] * uy[None]
        seeds.append(w[:, None, None] * rho[None] * (1 + 3 * cu + 4.5 * cu ** 2 - 1.5 * (ux ** 2 + uy ** 2)[None]))
    for g, f in enumerate(seeds):
        d.set_state(g, f)
    e.set_states(list(enumerate(seeds)))
    d.run(300, check_every=100, tol=1e-300, graph=False)
    e.run(300, check_every=100, tol=1e-300, graph=False)
    for g in range(len(gs)):
        for x, y in zip(d.fields(g), e.fields(g)):
            assert np.array_equal(x, y)


@pytest.mark.parametrize("odd", [False, True])
def test_tiled_batch_equals_one_step_batch(odd):
    """The tile kernels (two steps per launch, staged in LDS; batch_steps = 2,
    sides multiples of 32 x 16) against the one-step cell-order kernel:
    the same fields after the same steps to FP32 rounding (the same wall rule
    and collision functions, contracted independently by the compiler) -- an
    even count (two-step launches only) and an odd one (a one-step tile
    launch first)."""
    banner(f"batch tile kernels vs one-step kernel, {'301' if odd else '300'} steps")
    gs = geoms()
    pore = G.disk_packing(256, 0.6, 3, 10, seed=5)[:96, :].copy()  # 256 x 96: 8 x 3 tiles
    pore = G.sanitize(pore)
    k, lp = G.uniform_kn(pore, 0.2, MFP_DX)
    gs.append((pore, k, lp))
    a = Batch(gs, batch_steps=2, **PARAMS)
    b = Batch(gs, batch_steps=1, **PARAMS)
    if odd:
        a.step(1)
        b.step(1)
    ra = a.run(300, check_every=300, tol=1e-300, graph=False)
    rb = b.run(300, check_every=300, tol=1e-300, graph=True)
    worst, worst_u = 0.0, 0.0
    for g in range(len(gs)):
        fa, fb = a.fields(g), b.fields(g)
        for x, y in zip(fa, fb):
            worst = max(worst, float(np.abs(x - y).max()))
        worst_u = max(worst_u, rel_l2(fa[1], fb[1]))
    print("max |tile - one-step|", worst, "ux rel-l2", worst_u, "ux", ra["ux"], rb["ux"])
    assert worst <= 1e-8 and worst_u < 1e-6
    assert np.allclose(ra["ux"], rb["ux"], rtol=1e-6, atol=0)


def test_batch_launch_configurations_agree_and_tuning_is_invisible():
    """Every launch configuration of the autotuner (threads per block x XCD
    run length) gives the same fields bit for bit, and tuning puts the state
    back: a tuned batch and an untuned one agree after the same run."""
    banner("batch autotuner: 9 configurations agree; tuning leaves the state")
    gs = geoms()
    ref = Batch(gs, **PARAMS)
    ref.set_config(256, 32)
    ref.run(200, check_every=100, tol=1e-300, graph=False)
    want = [ref.fields(g) for g in range(len(gs))]
    for t in (64, 128, 256):
        for k in (0, 8, 32):
            b = Batch(gs, **PARAMS)
            b.set_config(t, k)
            b.run(200, check_every=100, tol=1e-300, graph=(t == 128))
            got = [b.fields(g) for g in range(len(gs))]
            assert all(np.array_equal(x, y) for fa, fb in zip(got, want) for x, y in zip(fa, fb)), (t, k)
            b.close()
    tb = Batch(gs, **PARAMS)
    t, k, ms = tb.tune(6)
    print("tuned", t, k, "\n", ms)
    # the tile kernels run 256 threads per tile: only that row is timed
    timed = ms[2] if (ms[0] < 0).all() else ms
    assert (timed > 0).all() and ms[[64, 128, 256].index(t), [0, 8, 32].index(k)] == timed.min()
    tb.run(200, check_every=100, tol=1e-300)
    got = [tb.fields(g) for g in range(len(gs))]
    assert all(np.array_equal(x, y) for fa, fb in zip(got, want) for x, y in zip(fa, fb))


@pytest.mark.parametrize("batch_steps", [1, 2])
def test_batch_grid_follows_the_queue(batch_steps):
    """The step grid is the active block count rounded up to a power of two
    (at least 1024); the captured chunk is replaced when that bucket changes.
    Four 512^2 lattices (~2400 blocks of 256 cells) are deactivated one by one,
    which crosses the 4096 / 2048 / 1024 buckets; every stage's graph replays
    must equal eager launches bit for bit, and inactive lattices must not
    change."""
    banner("batch grid buckets: 4 x 512^2, lattices deactivated one by one, graph vs eager")
    gs = []
    for phi, kn, seed in ((0.6, 0.1, 11), (0.65, 0.05, 12), (0.55, 0.3, 13), (0.7, 0.02, 14)):
        pore = G.disk_packing(512, phi, 3, 12, seed=seed)
        k, lp = G.uniform_kn(pore, kn, MFP_DX)
        gs.append((pore, k, lp))
    a = Batch(gs, batch_steps=batch_steps, **PARAMS)
    b = Batch(gs, batch_steps=batch_steps, **PARAMS)
    blocks = sum((int(n) + 255) // 256 for n in a.cells())
    assert blocks > 2048
    for stage in range(4):
        mask = np.array([g >= stage for g in range(4)])
        a.activate(mask)
        b.activate(mask)
        frozen = [a.fields(g) for g in range(stage)]
        ra = a.run(40, check_every=40, tol=1e-300, graph=True)
        rb = b.run(40, check_every=40, tol=1e-300, graph=False)
        assert ra["updates"] == rb["updates"] == 40 * sum(int(n) for n in a.cells()[stage:])
        for g in range(4):
            fa, fb = a.fields(g), b.fields(g)
            assert all(np.array_equal(x, y) for x, y in zip(fa, fb)), (stage, g)
        for g in range(stage):
            assert all(np.array_equal(x, y) for x, y in zip(frozen[g], a.fields(g))), (stage, g)


def test_batch_mdbb_equals_single_block():
    """MDBB walls in the batched solver (the one-step kernel; the tile
    kernels have no MDBB): each lattice equals the single-block FP32 MDBB
    path bit for bit, also across a tuning run (the MDBB state is saved and
    restored with the populations)."""
    banner("batch MDBB vs single-block FP32 MDBB, 200 steps")
    gs = geoms()
    prm = dict(PARAMS, wall=1)
    b = Batch(gs, **prm)
    b.tune(4)
    b.run(200, check_every=200, tol=1e-300, graph=False)
    for g, (pore, kn, lp) in enumerate(gs):
        lat = Lattice(pore, kn, lp, precision="fp32dev", periodic_y=1, **prm)
        lat.step("transport", 200)
        ux, uy, rho = lat.complete_domain()
        brho, bux, buy = b.fields(g)
        fl = pore.reshape(-1) == 1
        d = max(np.abs(bux.reshape(-1) - ux)[fl].max(), np.abs(brho.reshape(-1) - rho)[fl].max())
        print("lattice", g, "max |diff|", d)
        assert d == 0.0


@pytest.mark.parametrize("coll,wall", [(2, 0), (1, 2), (1, 0), (0, 2), (0, 0)],
                         ids=["RMRT-BB", "MRT-SRBB", "MRT-BB", "SRT-SRBB", "SRT-BB"])
def test_batch_operators_against_oracle(coll, wall):
    """Every other collision / wall pair of the batched kernel: two lattices
    of a batch against the FP64 oracle with the same operators after 100
    steps (MRT without regularisation drifts on these lattices, so the pairs
    are compared early), and against the single-block FP32 path bit for bit."""
    import _oracle as O
    banner(f"batch collision {coll} wall {wall} vs the oracle and the single-block path, 100 steps")
    gs = geoms()[:2]
    prm = dict(PARAMS, collision=coll, wall=wall)
    b = Batch(gs, **prm)
    b.run(100, check_every=100, tol=1e-300, graph=False)
    for g, (pore, kn, lp) in enumerate(gs):
        o = O.Oracle(pore, kn, lp, dx=1e-9, Fbody=1e-6, periodic_y=True)
        o.set_operators(coll, wall)
        o.run("transport", 100, 0.0)
        oux, ouy, orho = o.complete_domain()
        rho, ux, uy = b.fields(g)
        e = dict(rho=rel_l2(rho.reshape(-1), orho), ux=rel_l2(ux.reshape(-1), oux), uy=rel_l2(uy.reshape(-1), ouy))
        print("lattice", g, "batch FP32 vs oracle FP64 rel-l2", e, flush=True)
        assert e["rho"] < 1e-8 and e["ux"] < 1e-5 and e["uy"] < 1e-5, e
        o.close()
        lat = Lattice(pore, kn, lp, precision="fp32dev", periodic_y=1, **prm)
        lat.step("transport", 100)
        lux, luy, lrho = lat.complete_domain()
        fl = pore.reshape(-1) == 1
        assert np.array_equal(ux.reshape(-1)[fl], lux[fl]) and np.array_equal(rho.reshape(-1)[fl], lrho[fl])
        lat.close()
    b.close()
