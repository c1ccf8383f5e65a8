"""Learned warm start of the batched solver.

A convolutional encoder-decoder (mantiis_amd.hip.Net, kernels in
csrc/mts_net.hip) maps (fluid mask, log10 Kn, forcing) to the steady
velocity of a periodic porous lattice.  It is trained on steady states of
the batched solver.  Its prediction seeds the populations -- equilibrium at
rho_in plus the regularised non-equilibrium part of the predicted velocity
gradient -- and plain LBM then runs to the same tolerance, so the plain
method is both fallback and judge.

Scaling: the network predicts u nu / (F k_KC), a local permeability in
units of the lattice's Kozeny-Carman estimate k_KC = phi^3 / (5 S^2) (phi the
porosity, S the fluid-solid interface length per lattice site, both from the
mask), with nu = (tau - 1/2)/3 the lattice viscosity of the (uniform) Kn and F
the body force.  Permeabilities of the training distribution span two
decades; the estimate takes out most of that spread before the loss sees it.
"""
import numpy as np

from .hip import Batch, DeviceArray, Net

MFP_DX = 0.7730846760658538  # mfp / dx at the default T, P and dx = 1 nm
W9 = np.array([4 / 9] + [1 / 9] * 4 + [1 / 36] * 4)
EX = np.array([0, 1, 0, -1, 0, 1, -1, -1, 1])
EY = np.array([0, 0, 1, 0, -1, 1, 1, -1, -1])


def tau_of(kn):
    """tau_s of a cell with LocPore = (mfp/dx)/Kn (coreLB.hpp:482-502)."""
    return 0.5 + np.sqrt(6.0 / np.pi) * (MFP_DX / kn) * kn / (1.0 + 2.0 * kn)


def nu_of(kn):
    return (tau_of(kn) - 0.5) / 3.0


def sample(seed, L=None, sides=(128, 256, 512, 1024)):
    """One lattice of the batched distribution (porosity U[0.4, 0.8], disk
    radius U[3, 16], Kn log-uniform in [0.01, 1]); side L or drawn."""
    from . import geometry as G
    rng = np.random.default_rng(seed)
    side = int(L) if L else int(rng.choice(sides))
    phi = float(rng.uniform(0.4, 0.8))
    kn = float(np.exp(rng.uniform(np.log(0.01), np.log(1.0))))
    pore = G.disk_packing(side, phi, 3, 16, seed=seed)
    knf, lp = G.uniform_kn(pore, kn, MFP_DX)
    return dict(pore=pore, kn=kn, knf=knf, lp=lp)


def samples(seeds, L, workers=8):
    import multiprocessing as mp
    if workers <= 1 or len(seeds) < 4:
        return [sample(s, L) for s in seeds]
    with mp.get_context("fork").Pool(workers) as pool:
        return pool.starmap(sample, [(s, L) for s in seeds])


def k_scale(pore):
    """Kozeny-Carman permeability estimate phi^3 / (5 S^2) in lattice units."""
    fl = pore == 1
    phi = fl.mean()
    edges = (fl != np.roll(fl, 1, 0)).sum() + (fl != np.roll(fl, 1, 1)).sum()
    S = max(edges / fl.size, 1e-6)
    return max(phi ** 3 / (5.0 * S * S), 1e-3)


def features(s, F):
    """Network input [H, W, 3]: fluid mask, log10(Kn) + 1, F / 1e-6."""
    pore = s["pore"]
    x = np.zeros(pore.shape + (3,), np.float32)
    x[..., 0] = pore
    x[..., 1] = np.log10(s["kn"]) + 1.0
    x[..., 2] = F / 1e-6
    return x


def solve(batch_samples, F, tol=1e-6, check=100, max_steps=400000, warm=None, warm_u=None, rho_in=None, graph=True):
    """Runs the batched solver.  warm: list of populations (9, H, W) to start
    from; warm_u: list of velocity fields (ux, uy) that the device turns into
    the same populations as seed_populations (needs rho_in).  Returns the run
    record and the velocity fields."""
    b = Batch([(s["pore"], s["knf"], s["lp"]) for s in batch_samples], collision=2, wall=2, dx=1e-9, Fbody=F)
    if warm is not None:
        b.set_states(list(enumerate(warm)))
    if warm_u is not None:
        items = []
        for g, (s, (ux, uy)) in enumerate(zip(batch_samples, warm_u)):
            fl = s["pore"] == 1
            items.append((g, np.where(fl, ux, 0.0), np.where(fl, uy, 0.0), tau_of(s["kn"])))
        b.seed(items, rho_in)
    r = b.run(max_steps, check_every=check, tol=tol, graph=graph)
    fields = [b.fields(g) for g in range(len(batch_samples))]
    b.close()
    return r, fields


def targets(s, fields, F):
    rho, ux, uy = fields
    y = np.zeros(ux.shape + (2,), np.float32)
    sc = nu_of(s["kn"]) / (F * k_scale(s["pore"]))
    y[..., 0] = ux * sc
    y[..., 1] = uy * sc
    return y


def permeability(s, ux, F):
    """Darcy permeability in lattice units: nu <u_x> / F with <u_x> the
    superficial mean velocity over the whole lattice."""
    return nu_of(s["kn"]) * ux.sum() / ux.size / F


def seed_populations(s, ux, uy, rho_in):
    """Equilibrium at rho_in plus the regularised non-equilibrium part
    -3 tau w_i rho (Q_i : grad u) of the given velocity (central periodic
    differences with u = 0 in the solid)."""
    tau = tau_of(s["kn"])
    fl = s["pore"] == 1
    ux = np.where(fl, ux, 0.0)
    uy = np.where(fl, uy, 0.0)
    dxux = (np.roll(ux, -1, 1) - np.roll(ux, 1, 1)) / 2
    dyuy = (np.roll(uy, -1, 0) - np.roll(uy, 1, 0)) / 2
    dxuy = (np.roll(uy, -1, 1) - np.roll(uy, 1, 1)) / 2
    dyux = (np.roll(ux, -1, 0) - np.roll(ux, 1, 0)) / 2
    f = np.empty((9,) + ux.shape)
    usq = ux * ux + uy * uy
    for j in range(9):
        cu = EX[j] * ux + EY[j] * uy
        feq = W9[j] * rho_in * (1 + 3 * cu + 4.5 * cu * cu - 1.5 * usq)
        q = (EX[j] ** 2 - 1 / 3) * dxux + (EY[j] ** 2 - 1 / 3) * dyuy + EX[j] * EY[j] * (dxuy + dyux)
        f[j] = feq - 3 * tau * W9[j] * rho_in * q
    return f


class Dataset:
    """Inputs, targets and masks of n lattices of one size, on the device."""

    def __init__(self, xs, ys):
        self.n = len(xs)
        self.shape = xs[0].shape[:2]
        self.x = DeviceArray.from_numpy(np.stack(xs).astype(np.float32))
        self.y = DeviceArray.from_numpy(np.stack(ys).astype(np.float32))
        self.m = DeviceArray.from_numpy(np.stack([x[..., 0] for x in xs]).astype(np.float32))

    def view(self, i, B):
        H, W = self.shape
        return (_View(self.x.ptr + i * H * W * 3 * 4, (B, H, W, 3)), _View(self.y.ptr + i * H * W * 2 * 4, (B, H, W, 2)),
                _View(self.m.ptr + i * H * W * 4, (B, H, W)))


class _View:
    def __init__(self, ptr, shape):
        self.ptr = ptr
        self.shape = shape

    def data_ptr(self):
        return self.ptr


def train(ds, c0=24, batch=8, epochs=10, lr=1e-3, seed=1, log=None):
    H, W = ds.shape
    net = Net(c0=c0, max_batch=batch, max_h=H, max_w=W, seed=seed)
    rng = np.random.default_rng(seed)
    nb = ds.n // batch
    hist = []
    for ep in range(epochs):
        # batches are contiguous runs of a shuffled start order
        order = rng.permutation(nb)
        tot = 0.0
        for k in order:
            x, y, m = ds.view(int(k) * batch, batch)
            net.forward(x)
            tot += net.backward(y, m)
            net.adam(lr=lr * (0.5 ** (ep // max(1, epochs // 3))))
        hist.append(tot / max(nb, 1))
        if log:
            log(f"epoch {ep} loss {hist[-1]:.5f}")
    return net, hist


def predict(params, c0, s_list, F, batch=4):
    """Velocity fields predicted for lattices of one size."""
    H, W = s_list[0]["pore"].shape
    net = Net(c0=c0, max_batch=batch, max_h=H, max_w=W)
    net.set_params(params)
    out = []
    for i in range(0, len(s_list), batch):
        chunk = s_list[i:i + batch]
        x = DeviceArray.from_numpy(np.stack([features(s, F) for s in chunk]))
        o = DeviceArray((len(chunk), H, W, 2))
        net.forward(x, o)
        o = o.numpy()
        for s, oo in zip(chunk, o):
            sc = F * k_scale(s["pore"]) / nu_of(s["kn"])
            fl = s["pore"] == 1
            out.append((np.where(fl, oo[..., 0] * sc, 0.0), np.where(fl, oo[..., 1] * sc, 0.0)))
    net.close()
    return out
