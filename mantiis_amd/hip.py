"""ctypes binding of libmantiis_hip.so (the C ABI of include/mantiis_hip.h).

There is no fallback: if the library is missing or fails to load, every entry
point raises HipLibraryMissing.
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libmantiis_hip.so")

COLL = {"SRT": 0, "MRT": 1, "RMRT": 2}
WALL = {"BB": 0, "MDBB": 1, "SRBB": 2}
PREC = {"fp64": 0, "fp32dev": 1}
METHODS = {
    "equilibrium": 1, "regularize": 2, "SRTCollision": 3, "MRTCollisionNonRegularized": 4,
    "MRTCollisionRegularized": 5, "collide": 6, "zeroDown": 7, "stream": 8, "BBWall": 9, "SRBBWall": 10,
    "MDBBWall": 11, "periodicBoundary": 12, "generalizedPeriodicBoundaryAdsorption": 13,
    "calculateMacroscopicPorperties": 14, "calculateDensity": 15, "calculateVelocity": 16,
    "getPseudoPotential": 17, "calculateForces": 18, "applySourceTerm": 19,
    "calculateVelocityDifference": 20, "calculateDensityDifference": 21, "applyBoundaryConditions": 22,
    "evolutionStepOfMultiBlock": 23, "copy_f_to_ftemp": 24, "copy_f_to_fb": 25, "copy_usq_to_usq_old": 26,
    "copy_rho_to_rho_old": 27, "interpolateBlockInterface": 28,
}
FIELDS = {"f": 0, "ftemp": 1, "feq": 2, "fneq": 3, "fb": 4, "rho": 5, "rho_old": 6, "ux": 7, "uy": 8,
          "usq": 9, "usq_old": 10, "psi": 11, "Fx": 12, "Fy": 13, "taus": 14, "kn": 15}
SYMBOLS = ["mts_default_params", "mts_last_error", "mts_globals", "mts_load_existing_model", "mts_create",
           "mts_destroy", "mts_cell_count", "mts_grid_info", "mts_links", "mts_set_operators",
           "mts_set_body_force", "mts_initialize", "mts_step_transport", "mts_step_adsorption",
           "mts_step_coupled", "mts_residual", "mts_method", "mts_field_size", "mts_get_field",
           "mts_set_field", "mts_complete_domain", "mts_diagnostics", "mts_permeability", "mts_get_stream", "mts_synchronize",
           "mts_grid_create", "mts_grid_destroy", "mts_grid_size", "mts_grid_arrays", "mts_grid_rows", "mts_quadtree_check",
           "mts_kernel_timer", "mts_kernel_timer_read", "mts_probe_copy", "mts_probe_copy16", "mts_batch_create", "mts_batch_destroy",
           "mts_batch_cells", "mts_batch_run", "mts_batch_step", "mts_batch_kernel", "mts_batch_activate", "mts_batch_fields", "mts_batch_set_state", "mts_batch_set_states", "mts_batch_seed",
           "mts_batch_stream", "mts_batch_tune", "mts_batch_set_config", "mts_dev_alloc", "mts_dev_free", "mts_dev_upload", "mts_dev_download",
           "mts_net_create", "mts_net_destroy", "mts_net_param_count", "mts_net_flops", "mts_net_get_params", "mts_net_set_params",
           "mts_net_get_grads", "mts_net_forward", "mts_net_backward", "mts_net_adam", "mts_net_gemm_check",
           "mts_net_stream", "mts_shape_create", "mts_shape_destroy", "mts_shape_stream", "mts_shape_run", "mts_shape_run_static",
           "mts_shape_properties"]


class HipLibraryMissing(RuntimeError):
    pass


class MtsError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"mts error {code}: {msg}")
        self.code = code


class Params(C.Structure):
    _fields_ = [("dx", C.c_double), ("Fbody", C.c_double), ("T", C.c_double), ("P", C.c_double),
                ("Gff", C.c_double), ("Gfs", C.c_double), ("collision", C.c_int), ("wall", C.c_int),
                ("precision", C.c_int), ("multiblock", C.c_int), ("device", C.c_int), ("periodic_y", C.c_int),
                ("batch_steps", C.c_int)]


_LIB = None
_D = C.POINTER(C.c_double)
_I32 = C.POINTER(C.c_int32)
_I64 = C.POINTER(C.c_int64)


def load_library(path=LIB_PATH):
    global _LIB
    if _LIB is not None:
        return _LIB
    if not os.path.exists(path):
        raise HipLibraryMissing(f"{path} is not built (run __graft_entry__.build())")
    try:
        lib = C.CDLL(path)
    except OSError as e:
        raise HipLibraryMissing(f"cannot load {path}: {e}") from e
    lib.mts_last_error.restype = C.c_char_p
    lib.mts_default_params.argtypes = [C.POINTER(Params)]
    lib.mts_default_params.restype = None
    lib.mts_globals.argtypes = [C.POINTER(Params), _D]
    lib.mts_load_existing_model.argtypes = [C.c_char_p, C.c_int, _I32, _D, _D]
    lib.mts_create.argtypes = [C.c_int, C.c_int, _I32, _D, _D, C.POINTER(Params), C.POINTER(C.c_void_p)]
    lib.mts_destroy.argtypes = [C.c_void_p]
    lib.mts_cell_count.argtypes = [C.c_void_p, _I64]
    lib.mts_grid_info.argtypes = [C.c_void_p, _I64, _I32, _I32, _I64, _I32]
    lib.mts_links.argtypes = [C.c_void_p, C.c_int, _I64, _I64, _I32, _I64]
    lib.mts_set_operators.argtypes = [C.c_void_p, C.c_int, C.c_int]
    lib.mts_set_body_force.argtypes = [C.c_void_p, C.c_double]
    lib.mts_initialize.argtypes = [C.c_void_p]
    lib.mts_step_transport.argtypes = [C.c_void_p, C.c_int]
    lib.mts_step_adsorption.argtypes = [C.c_void_p, C.c_int]
    lib.mts_step_coupled.argtypes = [C.c_void_p, C.c_int]
    lib.mts_residual.argtypes = [C.c_void_p, _D]
    lib.mts_method.argtypes = [C.c_void_p, C.c_int, C.c_int]
    lib.mts_field_size.argtypes = [C.c_void_p, C.c_int, _I64]
    lib.mts_get_field.argtypes = [C.c_void_p, C.c_int, _D]
    lib.mts_set_field.argtypes = [C.c_void_p, C.c_int, _D]
    lib.mts_complete_domain.argtypes = [C.c_void_p, _D, _D, _D]
    lib.mts_diagnostics.argtypes = [C.c_void_p, _D]
    lib.mts_permeability.argtypes = [C.c_void_p, _D]
    lib.mts_get_stream.argtypes = [C.c_void_p, C.POINTER(C.c_void_p)]
    lib.mts_synchronize.argtypes = [C.c_void_p]
    lib.mts_grid_create.argtypes = [C.c_int, C.c_int, _I32, _D, _D, C.POINTER(Params), C.POINTER(C.c_void_p)]
    lib.mts_grid_destroy.argtypes = [C.c_void_p]
    lib.mts_grid_size.argtypes = [C.c_void_p, _I64, _I32]
    lib.mts_grid_arrays.argtypes = [C.c_void_p, _I64, _I32, _I32, _I64, _I32, _I32, _I32, _D]
    lib.mts_grid_rows.argtypes = [C.c_void_p, C.c_int, _I64, _I64]
    lib.mts_kernel_timer.argtypes = [C.c_void_p, C.c_int]
    lib.mts_kernel_timer_read.argtypes = [C.c_void_p, _D, _I64]
    lib.mts_probe_copy.argtypes = [C.c_void_p, C.c_void_p, C.c_int64, C.c_int]
    lib.mts_probe_copy16.argtypes = [C.c_void_p, C.c_void_p, C.c_int64, C.c_int]
    _PP = C.POINTER(C.c_void_p)
    lib.mts_batch_create.argtypes = [C.c_int, _I32, _I32, _PP, _PP, _PP, C.POINTER(Params), _PP]
    lib.mts_batch_destroy.argtypes = [C.c_void_p]
    lib.mts_batch_cells.argtypes = [C.c_void_p, _I64]
    lib.mts_batch_run.argtypes = [C.c_void_p, C.c_int64, C.c_int, C.c_double, C.c_int, _I64, _D, _D, _I64]
    lib.mts_batch_step.argtypes = [C.c_void_p, C.c_int]
    lib.mts_batch_kernel.argtypes = [C.c_void_p, C.POINTER(C.c_int), _I64]
    lib.mts_batch_activate.argtypes = [C.c_void_p, _I32]
    lib.mts_batch_fields.argtypes = [C.c_void_p, C.c_int, _D, _D, _D]
    lib.mts_batch_set_state.argtypes = [C.c_void_p, C.c_int, _D]
    lib.mts_batch_stream.argtypes = [C.c_void_p, _PP]
    lib.mts_batch_tune.argtypes = [C.c_void_p, C.c_int, _I32, _I32, _D]
    lib.mts_batch_set_config.argtypes = [C.c_void_p, C.c_int, C.c_int]
    lib.mts_dev_alloc.argtypes = [C.c_size_t, _PP]
    lib.mts_dev_free.argtypes = [C.c_void_p]
    lib.mts_dev_upload.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t]
    lib.mts_dev_download.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t]
    _LIB = lib
    return lib


def _p(a, t):
    return a.ctypes.data_as(t) if a is not None else None


def _check(lib, rc):
    if rc != 0:
        raise MtsError(rc, lib.mts_last_error().decode())


def default_params(**kw):
    lib = load_library()
    p = Params()
    lib.mts_default_params(C.byref(p))
    for k, v in kw.items():
        if k == "collision" and isinstance(v, str):
            v = COLL[v]
        if k == "wall" and isinstance(v, str):
            v = WALL[v]
        if k == "precision" and isinstance(v, str):
            v = PREC[v]
        setattr(p, k, v)
    return p


def globals_for(**kw):
    lib = load_library()
    p = default_params(**kw)
    out = np.zeros(10)
    rc = lib.mts_globals(C.byref(p), _p(out, _D))
    keys = ["rhoin", "mfp", "Cf", "T_lu", "a_alpha_lu", "ZG", "Crho", "Cu", "Cl", "cubic_ok"]
    return rc, dict(zip(keys, out.tolist()))


def load_existing_model(folder, n):
    lib = load_library()
    pore = np.zeros((n, n), np.int32)
    kn = np.zeros((n, n))
    lp = np.zeros((n, n))
    folder = folder if folder.endswith("/") else folder + "/"
    _check(lib, lib.mts_load_existing_model(folder.encode(), n, _p(pore, _I32), _p(kn, _D), _p(lp, _D)))
    return pore, kn, lp


class Lattice:
    """One LB2D lattice on the GPU (handle of the C ABI)."""

    def __init__(self, pore, kn, locpore, **params):
        self.lib = load_library()
        pore = np.ascontiguousarray(pore, dtype=np.int32)
        kn = np.ascontiguousarray(kn, dtype=np.float64)
        locpore = np.ascontiguousarray(locpore, dtype=np.float64)
        self.ny, self.nx = pore.shape
        self.params = default_params(**params)
        h = C.c_void_p()
        _check(self.lib, self.lib.mts_create(self.nx, self.ny, _p(pore, _I32), _p(kn, _D), _p(locpore, _D),
                                             C.byref(self.params), C.byref(h)))
        self.h = h

    def close(self):
        if getattr(self, "h", None) and self.h.value:
            self.lib.mts_destroy(self.h)
            self.h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def ncell(self):
        n = C.c_int64()
        _check(self.lib, self.lib.mts_cell_count(self.h, C.byref(n)))
        return n.value

    def set_operators(self, collision, wall):
        c = COLL[collision] if isinstance(collision, str) else collision
        w = WALL[wall] if isinstance(wall, str) else wall
        _check(self.lib, self.lib.mts_set_operators(self.h, c, w))

    def set_body_force(self, F):
        _check(self.lib, self.lib.mts_set_body_force(self.h, F))

    def initialize(self):
        _check(self.lib, self.lib.mts_initialize(self.h))

    def step(self, kind, nsteps):
        fn = {"transport": self.lib.mts_step_transport, "adsorption": self.lib.mts_step_adsorption,
              "coupled": self.lib.mts_step_coupled}[kind]
        _check(self.lib, fn(self.h, int(nsteps)))

    def residual(self):
        d = C.c_double()
        _check(self.lib, self.lib.mts_residual(self.h, C.byref(d)))
        return d.value

    def method(self, name, lvl=1):
        _check(self.lib, self.lib.mts_method(self.h, METHODS[name], lvl))

    def get(self, name):
        n = C.c_int64()
        _check(self.lib, self.lib.mts_field_size(self.h, FIELDS[name], C.byref(n)))
        out = np.zeros(n.value)
        _check(self.lib, self.lib.mts_get_field(self.h, FIELDS[name], _p(out, _D)))
        return out

    def set(self, name, arr):
        arr = np.ascontiguousarray(arr, dtype=np.float64)
        _check(self.lib, self.lib.mts_set_field(self.h, FIELDS[name], _p(arr, _D)))

    def complete_domain(self):
        n = self.nx * self.ny
        ux, uy, rho = np.zeros(n), np.zeros(n), np.zeros(n)
        _check(self.lib, self.lib.mts_complete_domain(self.h, _p(ux, _D), _p(uy, _D), _p(rho, _D)))
        return ux, uy, rho

    def diagnostics(self):
        out = np.zeros(4)
        _check(self.lib, self.lib.mts_diagnostics(self.h, _p(out, _D)))
        return out

    def permeability(self):
        """FP64 permeability report (mts_permeability) as a dict."""
        out = np.zeros(8)
        _check(self.lib, self.lib.mts_permeability(self.h, _p(out, _D)))
        keys = ("darcy_velocity", "mean_ux", "flux", "mass", "porosity", "tau", "k_lattice", "k_m2")
        return dict(zip(keys, out.tolist()))

    def grid(self):
        n = self.ncell
        gid = np.zeros(n, np.int64)
        ij = np.zeros((n, 2), np.int32)
        lvl = np.zeros(n, np.int32)
        con = np.zeros((n, 9), np.int64)
        bnd = np.zeros(n, np.int32)
        _check(self.lib, self.lib.mts_grid_info(self.h, _p(gid, _I64), _p(ij, _I32), _p(lvl, _I32), _p(con, _I64),
                                                _p(bnd, _I32)))
        return dict(gridID=gid, gridIJ=ij, level=lvl, connect=con, bnd=bnd)

    def links(self, which):
        w = {"stream": 0, "boundary": 1, "periodic": 2}[which]
        cnt = C.c_int64()
        _check(self.lib, self.lib.mts_links(self.h, w, None, None, None, C.byref(cnt)))
        a = np.zeros(cnt.value, np.int64)
        b = np.zeros(cnt.value, np.int64)
        c = np.zeros(cnt.value, np.int32)
        _check(self.lib, self.lib.mts_links(self.h, w, _p(a, _I64), _p(b, _I64), _p(c, _I32), C.byref(cnt)))
        return a, b, c

    def stream_handle(self):
        s = C.c_void_p()
        _check(self.lib, self.lib.mts_get_stream(self.h, C.byref(s)))
        return s.value

    def synchronize(self):
        _check(self.lib, self.lib.mts_synchronize(self.h))

    def kernel_timer(self, enable):
        _check(self.lib, self.lib.mts_kernel_timer(self.h, int(bool(enable))))

    def kernel_timer_read(self):
        ms = C.c_double()
        n = C.c_int64()
        _check(self.lib, self.lib.mts_kernel_timer_read(self.h, C.byref(ms), C.byref(n)))
        return ms.value, n.value


class HostGrid:
    """Grid2D built on the host by the library (no device)."""

    def __init__(self, pore, kn, locpore, **params):
        self.lib = load_library()
        pore = np.ascontiguousarray(pore, dtype=np.int32)
        kn = np.ascontiguousarray(kn, dtype=np.float64)
        locpore = np.ascontiguousarray(locpore, dtype=np.float64)
        ny, nx = pore.shape
        p = default_params(**params)
        h = C.c_void_p()
        _check(self.lib, self.lib.mts_grid_create(nx, ny, _p(pore, _I32), _p(kn, _D), _p(locpore, _D), C.byref(p),
                                                  C.byref(h)))
        self.h = h

    def __del__(self):
        try:
            if self.h.value:
                self.lib.mts_grid_destroy(self.h)
        except Exception:
            pass

    def size(self):
        n, L = C.c_int64(), C.c_int32()
        _check(self.lib, self.lib.mts_grid_size(self.h, C.byref(n), C.byref(L)))
        return n.value, L.value

    def arrays(self, taus=True):
        """Grid2D arrays; taus=False leaves out the relaxation times, which a
        multi-block grid computes on the device."""
        n, _ = self.size()
        gid = np.zeros(n, np.int64)
        ij = np.zeros((n, 2), np.int32)
        lvl = np.zeros(n, np.int32)
        con = np.zeros((n, 9), np.int64)
        bnd = np.zeros(n, np.int32)
        typ = np.zeros(n, np.int32)
        isb = np.zeros(n, np.int32)
        t = np.zeros(n) if taus else None
        _check(self.lib, self.lib.mts_grid_arrays(self.h, _p(gid, _I64), _p(ij, _I32), _p(lvl, _I32), _p(con, _I64),
                                                  _p(bnd, _I32), _p(typ, _I32), _p(isb, _I32),
                                                  _p(t, _D) if taus else None))
        out = dict(gridID=gid, gridIJ=ij, level=lvl, connect=con, bnd=bnd, type=typ, isBuffer=isb)
        if taus:
            out["taus"] = t
        return out

    def rows(self, which):
        """which: "parent_child", "coarse" (every pair's coarse buffer rows),
        "fine", or a buffer type (21, 32: coarse; 12, 23: fine)."""
        w = which if isinstance(which, int) else {"parent_child": 0, "coarse": 1, "fine": 2}[which]
        cnt = C.c_int64()
        _check(self.lib, self.lib.mts_grid_rows(self.h, w, None, C.byref(cnt)))
        out = np.zeros((cnt.value, 5), np.int64)
        _check(self.lib, self.lib.mts_grid_rows(self.h, w, _p(out, _I64), C.byref(cnt)))
        return out


class Batch:
    """Independent single-block lattices stepped together (mts_batch_*).
    geoms: list of (pore, kn, locpore) 2-D arrays; params as for Lattice
    (FP32 deviations, torus seams)."""

    def __init__(self, geoms, **params):
        self.lib = load_library()
        params.setdefault("precision", "fp32dev")
        p = default_params(**params)
        self.shapes = [np.asarray(g[0]).shape for g in geoms]
        self._keep = []
        arrs = []
        for pore, kn, lp in geoms:
            a = (np.ascontiguousarray(pore, np.int32), np.ascontiguousarray(kn, np.float64),
                 np.ascontiguousarray(lp, np.float64))
            self._keep.append(a)
            arrs.append(a)
        nb = len(arrs)
        nx = np.array([a[0].shape[1] for a in arrs], np.int32)
        ny = np.array([a[0].shape[0] for a in arrs], np.int32)
        PP = C.c_void_p * nb
        pp = PP(*[a[0].ctypes.data for a in arrs])
        kp = PP(*[a[1].ctypes.data for a in arrs])
        lpp = PP(*[a[2].ctypes.data for a in arrs])
        h = C.c_void_p()
        _check(self.lib, self.lib.mts_batch_create(nb, _p(nx, _I32), _p(ny, _I32), pp, kp, lpp, C.byref(p),
                                                    C.byref(h)))
        self._keep = None
        self.h = h
        self.nb = nb

    def close(self):
        if getattr(self, "h", None):
            self.lib.mts_batch_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def cells(self):
        out = np.zeros(self.nb, np.int64)
        _check(self.lib, self.lib.mts_batch_cells(self.h, _p(out, _I64)))
        return out

    def run(self, max_steps, check_every=100, tol=1e-6, graph=True):
        steps = np.zeros(self.nb, np.int64)
        sec = np.zeros(self.nb)
        ux = np.zeros(self.nb)
        upd = C.c_int64()
        _check(self.lib, self.lib.mts_batch_run(self.h, int(max_steps), int(check_every), float(tol), int(bool(graph)),
                                                _p(steps, _I64), _p(sec, _D), _p(ux, _D), C.byref(upd)))
        return dict(steps=steps, seconds=sec, ux=ux, updates=upd.value)

    def kernel(self):
        """(steps per launch, tiles or 256-cell blocks) of the batch's step kernel."""
        k = C.c_int(0)
        u = C.c_int64(0)
        _check(self.lib, self.lib.mts_batch_kernel(self.h, C.byref(k), C.byref(u)))
        return k.value, u.value

    def step(self, nsteps):
        """nsteps steps of the active lattices (mts_batch_step): no statistics,
        no convergence test; asynchronous on the batch's stream."""
        _check(self.lib, self.lib.mts_batch_step(self.h, int(nsteps)))

    def activate(self, mask):
        m = np.ascontiguousarray(mask, np.int32)
        _check(self.lib, self.lib.mts_batch_activate(self.h, _p(m, _I32)))

    def fields(self, g):
        shp = self.shapes[g]
        rho, ux, uy = (np.zeros(shp) for _ in range(3))
        _check(self.lib, self.lib.mts_batch_fields(self.h, int(g), _p(rho, _D), _p(ux, _D), _p(uy, _D)))
        return rho, ux, uy

    def set_state(self, g, f):
        f = np.ascontiguousarray(f, np.float64)
        _check(self.lib, self.lib.mts_batch_set_state(self.h, int(g), _p(f, _D)))

    def set_states(self, items):
        """Warm start of several lattices, collided in one launch: items is a
        list of (g, populations)."""
        fs = [np.ascontiguousarray(f, np.float64) for _, f in items]
        gs = np.array([g for g, _ in items], np.int32)
        ptrs = (_D * len(fs))(*[_p(f, _D) for f in fs])
        self.lib.mts_batch_set_states.argtypes = [C.c_void_p, C.c_int, _I32, C.POINTER(_D)]
        _check(self.lib, self.lib.mts_batch_set_states(self.h, len(fs), _p(gs, _I32), ptrs))

    def seed(self, items, rho):
        """Warm start from velocity fields on the device (mts_batch_seed):
        items is a list of (g, ux, uy, tau) with ux, uy the lattice's
        row-major velocity (zero in the solids)."""
        us = [np.ascontiguousarray(u, np.float64) for _, u, _, _ in items]
        vs = [np.ascontiguousarray(v, np.float64) for _, _, v, _ in items]
        gs = np.array([it[0] for it in items], np.int32)
        taus = np.array([it[3] for it in items], np.float64)
        pu = (_D * len(us))(*[_p(u, _D) for u in us])
        pv = (_D * len(vs))(*[_p(v, _D) for v in vs])
        self.lib.mts_batch_seed.argtypes = [C.c_void_p, C.c_int, _I32, C.POINTER(_D), C.POINTER(_D), _D, C.c_double]
        _check(self.lib, self.lib.mts_batch_seed(self.h, len(us), _p(gs, _I32), pu, pv, _p(taus, _D), float(rho)))

    def stream_handle(self):
        s = C.c_void_p()
        _check(self.lib, self.lib.mts_batch_stream(self.h, C.byref(s)))
        return s.value

    def tune(self, steps_each=8):
        """Autotuner (mts_batch_tune): steps_each > 0 times the launch
        configurations and keeps the fastest, 0 takes the default, < 0 only
        reports.  Returns (threads per block, XCD run length, ms per step of
        each candidate as a 3 x 3 array: threads {64,128,256} x run {0,8,32})."""
        t, k = C.c_int32(0), C.c_int32(0)
        ms = np.full(9, -1.0)
        _check(self.lib, self.lib.mts_batch_tune(self.h, int(steps_each), C.byref(t), C.byref(k), _p(ms, _D)))
        return t.value, k.value, ms.reshape(3, 3)

    def set_config(self, threads, xcd_run):
        _check(self.lib, self.lib.mts_batch_set_config(self.h, int(threads), int(xcd_run)))


class DeviceArray:
    """A device buffer of the library (mts_dev_*) with numpy up/download."""

    def __init__(self, shape, dtype=np.float32):
        self.lib = load_library()
        self.shape = tuple(int(s) for s in shape)
        self.dtype = np.dtype(dtype)
        self.nbytes = int(np.prod(self.shape)) * self.dtype.itemsize
        p = C.c_void_p()
        _check(self.lib, self.lib.mts_dev_alloc(max(self.nbytes, 1), C.byref(p)))
        self.ptr = p.value

    @classmethod
    def from_numpy(cls, a):
        a = np.ascontiguousarray(a)
        d = cls(a.shape, a.dtype)
        d.upload(a)
        return d

    def upload(self, a):
        a = np.ascontiguousarray(a, self.dtype)
        assert a.nbytes == self.nbytes
        _check(self.lib, self.lib.mts_dev_upload(self.ptr, a.ctypes.data, self.nbytes))

    def numpy(self):
        out = np.empty(self.shape, self.dtype)
        _check(self.lib, self.lib.mts_dev_download(out.ctypes.data, self.ptr, self.nbytes))
        return out

    def data_ptr(self):
        return self.ptr

    def close(self):
        if getattr(self, "ptr", None):
            self.lib.mts_dev_free(self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Net:
    """The warm-start encoder-decoder (mts_net_*).  Inputs, targets, masks
    and outputs are DeviceArray buffers; parameters are host numpy arrays."""

    def __init__(self, c0=24, max_batch=8, max_h=256, max_w=256, seed=1):
        self.lib = load_library()
        for name, args in (("mts_net_create", [C.c_int, C.c_int, C.c_int, C.c_int, C.c_uint64, C.POINTER(C.c_void_p)]),
                           ("mts_net_destroy", [C.c_void_p]), ("mts_net_param_count", [C.c_void_p, _I64]),
                           ("mts_net_flops", [C.c_void_p, C.c_int, C.c_int, C.c_int, C.POINTER(C.c_double)]),
                           ("mts_net_get_params", [C.c_void_p, C.POINTER(C.c_float)]),
                           ("mts_net_set_params", [C.c_void_p, C.POINTER(C.c_float)]),
                           ("mts_net_get_grads", [C.c_void_p, C.POINTER(C.c_float)]),
                           ("mts_net_forward", [C.c_void_p, C.c_void_p, C.c_int, C.c_int, C.c_int, C.c_void_p]),
                           ("mts_net_backward", [C.c_void_p, C.c_void_p, C.c_void_p, C.POINTER(C.c_float)]),
                           ("mts_net_adam", [C.c_void_p, C.c_float, C.c_float, C.c_float, C.c_float, C.c_int]),
                           ("mts_net_stream", [C.c_void_p, C.POINTER(C.c_void_p)])):
            getattr(self.lib, name).argtypes = args
        h = C.c_void_p()
        _check(self.lib, self.lib.mts_net_create(c0, max_batch, max_h, max_w, seed, C.byref(h)))
        self.h = h
        self.c0 = c0
        n = C.c_int64()
        _check(self.lib, self.lib.mts_net_param_count(self.h, C.byref(n)))
        self.nparam = n.value
        self.t = 0

    def close(self):
        if getattr(self, "h", None):
            self.lib.mts_net_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def flops(self, batch, h, w):
        """Floating-point operations of one forward pass over `batch` images
        of h x w (2 per multiply-add of every convolution)."""
        f = C.c_double()
        _check(self.lib, self.lib.mts_net_flops(self.h, batch, h, w, C.byref(f)))
        return f.value

    def params(self):
        out = np.zeros(self.nparam, np.float32)
        _check(self.lib, self.lib.mts_net_get_params(self.h, out.ctypes.data_as(C.POINTER(C.c_float))))
        return out

    def set_params(self, p):
        p = np.ascontiguousarray(p, np.float32)
        _check(self.lib, self.lib.mts_net_set_params(self.h, p.ctypes.data_as(C.POINTER(C.c_float))))

    def grads(self):
        out = np.zeros(self.nparam, np.float32)
        _check(self.lib, self.lib.mts_net_get_grads(self.h, out.ctypes.data_as(C.POINTER(C.c_float))))
        return out

    def forward(self, x, out=None):
        """x: DeviceArray float32 [B, H, W, 3]; out: [B, H, W, 2] or None."""
        B, H, W, _ = x.shape
        _check(self.lib, self.lib.mts_net_forward(self.h, x.data_ptr(), B, H, W,
                                                  out.data_ptr() if out is not None else None))
        return out

    def backward(self, target, mask):
        loss = C.c_float()
        _check(self.lib, self.lib.mts_net_backward(self.h, target.data_ptr(), mask.data_ptr(), C.byref(loss)))
        return loss.value

    def adam(self, lr=1e-3, beta1=0.9, beta2=0.999, eps=1e-8):
        self.t += 1
        _check(self.lib, self.lib.mts_net_adam(self.h, lr, beta1, beta2, eps, self.t))


def gemm_check(A, B):
    """C = A B through the network's matrix-core GEMM kernel (BF16 operands)."""
    lib = load_library()
    f = C.POINTER(C.c_float)
    lib.mts_net_gemm_check.argtypes = [f, f, C.c_int, C.c_int, C.c_int, f]
    A = np.ascontiguousarray(A, np.float32)
    B = np.ascontiguousarray(B, np.float32)
    M, K = A.shape
    N = B.shape[1]
    Cm = np.zeros((M, N), np.float32)
    _check(lib, lib.mts_net_gemm_check(A.ctypes.data_as(f), B.ctypes.data_as(f), M, N, K, Cm.ctypes.data_as(f)))
    return Cm


def quadtree_check(pore, kn, **kw):
    """Builds the quadtree of `pore` with the host builder and with the
    device builder (mts_quadtree_check) and returns the number of table
    entries that differ (level decision, cells, connectivity, boundary
    types, parent/child and buffer rows, level-image link words)."""
    lib = load_library()
    lib.mts_quadtree_check.argtypes = [C.c_int, C.c_int, _I32, _D, C.POINTER(Params), _I64]
    pore = np.ascontiguousarray(pore, np.int32)
    kn = np.ascontiguousarray(kn, np.float64)
    prm = default_params(**kw)
    n = C.c_int64(-1)
    _check(lib, lib.mts_quadtree_check(pore.shape[1], pore.shape[0], _p(pore, _I32), _p(kn, _D), C.byref(prm),
                                       C.byref(n)))
    return n.value


def shape_properties(domain, resolution, mfp, skeleton=False):
    """Shape::calculateProperties (shape.hpp:127-251) on the device, one
    shot: domain is the row-major [ny][nx] 0/1 image (Shape::domain).
    Returns (LocPore, Kn) -- and the cropped skeleton when skeleton=True."""
    lib = load_library()
    u8 = C.POINTER(C.c_uint8)
    lib.mts_shape_properties.argtypes = [_I32, C.c_int, C.c_int, C.c_double, C.c_double, _D, _D, u8]
    d = np.ascontiguousarray(domain, np.int32)
    ny, nx = d.shape
    lp = np.zeros((ny, nx))
    kn = np.zeros((ny, nx))
    sk = np.zeros((ny, nx), np.uint8) if skeleton else None
    _check(lib, lib.mts_shape_properties(_p(d, _I32), nx, ny, float(resolution), float(mfp), _p(lp, _D), _p(kn, _D),
                                         sk.ctypes.data_as(u8) if skeleton else None))
    return (lp, kn, sk) if skeleton else (lp, kn)


def shape_run_static(domain, resolution, mfp, locpore, kn, skeleton=None, stream=None):
    """mts_shape_run_static: the device-resident calculateProperties with no
    handle and no device allocation (the library's static work memory, sides
    up to 512).  domain, locpore, kn (and skeleton) are DeviceArrays of one
    [ny][nx] shape; stream a hipStream_t (None: the null stream).  Returns
    the thinning iterations run."""
    lib = load_library()
    lib.mts_shape_run_static.argtypes = [C.c_int, C.c_int, C.c_void_p, C.c_double, C.c_double, C.c_void_p,
                                         C.c_void_p, C.c_void_p, C.POINTER(C.c_int), C.c_void_p]
    ny, nx = domain.shape
    it = C.c_int(0)
    _check(lib, lib.mts_shape_run_static(int(nx), int(ny), domain.data_ptr(), float(resolution), float(mfp),
                                         locpore.data_ptr(), kn.data_ptr(),
                                         skeleton.data_ptr() if skeleton is not None else None, C.byref(it),
                                         stream))
    return it.value


class ShapeProperties:
    """Device-resident calculateProperties for one image size (mts_shape_*):
    run() reads a DeviceArray domain (uint8 0/1) and writes DeviceArray
    LocPore / Kn (float64) and optionally the skeleton (uint8)."""

    def __init__(self, nx, ny):
        self.lib = load_library()
        for name, args in (("mts_shape_create", [C.c_int, C.c_int, C.POINTER(C.c_void_p)]),
                           ("mts_shape_destroy", [C.c_void_p]),
                           ("mts_shape_run", [C.c_void_p, C.c_void_p, C.c_double, C.c_double, C.c_void_p, C.c_void_p,
                                              C.c_void_p, C.POINTER(C.c_int)]),
                           ("mts_shape_stream", [C.c_void_p, C.POINTER(C.c_void_p)])):
            getattr(self.lib, name).argtypes = args
        h = C.c_void_p()
        _check(self.lib, self.lib.mts_shape_create(int(nx), int(ny), C.byref(h)))
        self.h = h.value
        self.nx, self.ny = int(nx), int(ny)

    def run(self, domain, resolution, mfp, locpore, kn, skeleton=None):
        it = C.c_int(0)
        _check(self.lib, self.lib.mts_shape_run(self.h, domain.data_ptr(), float(resolution), float(mfp),
                                                locpore.data_ptr(), kn.data_ptr(),
                                                skeleton.data_ptr() if skeleton is not None else None, C.byref(it)))
        return it.value

    def stream(self):
        s = C.c_void_p()
        _check(self.lib, self.lib.mts_shape_stream(self.h, C.byref(s)))
        return s.value

    def close(self):
        if getattr(self, "h", None):
            self.lib.mts_shape_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
