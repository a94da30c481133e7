"""Host-side Python API of the MI355X moment-representation LBM library.

A thin layer over the C ABI (include/mrlbm.h): `Lattice` owns one
device-resident lattice; node-type maps use the reference's encoding
(include/mrlbm_nodetype.h). Arrays crossing the boundary are numpy arrays
of shape (nx, ny, nz) (x-fastest in memory when flattened in Fortran
order) or flat x-fastest vectors.
"""
import ctypes

import numpy as np

from . import _lib
from ._lib import (INIT_EQUILIBRIUM, INIT_REFERENCE, KERNEL_AUTO, KERNEL_COLUMN, KERNEL_TILE, STORE_BF16,
                   STORE_FP16, STORE_FP32, MrlbmError)

__all__ = ["Lattice", "MrlbmError", "STORE_FP32", "STORE_FP16", "STORE_BF16", "INIT_EQUILIBRIUM",
           "INIT_REFERENCE", "KERNEL_AUTO", "KERNEL_TILE", "KERNEL_COLUMN", "MOMENT_NAMES"]

MOMENT_NAMES = ("rho-rho0", "3ux", "3uy", "3uz", "4.5mxx", "9mxy", "9mxz", "4.5myy", "9myz", "4.5mzz")


def _flat(a, n, dtype):
    a = np.asarray(a)
    if a.ndim == 3:
        a = a.reshape(-1, order="F")
    a = np.ascontiguousarray(a, dtype=dtype)
    if a.size != n:
        raise ValueError(f"expected {n} values, got {a.size}")
    return a


class Lattice:
    """One lattice resident in HBM.

    shape: (nx, ny, nz); q: 19 or 27; tau: relaxation time; walls: per-axis
    0 periodic / 1 wall; force: body force (unscaled); tile: (bx, by, bz) or
    None for the library default; storage: STORE_*.
    """

    def __init__(self, shape, q=19, tau=0.6, rho0=1.0, walls=(0, 0, 0), force=(0.0, 0.0, 0.0), tile=None,
                 storage=STORE_FP32, kernel=KERNEL_AUTO, device=0, seed=0):
        L = _lib.load()
        nx, ny, nz = shape
        bx, by, bz = tile if tile else (0, 0, 0)
        p = _lib.Params(nx, ny, nz, q, bx, by, bz, storage, walls[0], walls[1], walls[2], kernel, tau, rho0,
                        force[0], force[1], force[2], device, seed)
        h = ctypes.c_void_p()
        _lib.check(L.mrlbm_create(ctypes.byref(p), ctypes.byref(h)))
        self._h = h
        self.shape = (nx, ny, nz)
        self.n = nx * ny * nz
        self.q = q

    def close(self):
        if getattr(self, "_h", None):
            _lib.load().mrlbm_destroy(self._h)
            self._h = None

    def __del__(self):
        self.close()

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    # -- set-up ---------------------------------------------------------
    def set_bc_velocity(self, k, u):
        _lib.check(_lib.load().mrlbm_set_bc_velocity(self._h, k, *map(float, u)))

    def set_node_types(self, types):
        L = _lib.load()
        if types is None:
            _lib.check(L.mrlbm_set_node_types(self._h, None))
            return
        t = _flat(types, self.n, np.uint32)
        _lib.check(L.mrlbm_set_node_types(self._h, t.ctypes.data_as(_lib._u32p)))

    def set_geometry(self, solid=None, centres=(), radii=()):
        """No-slip geometry classified on the device: the lattice's walls, a
        solid mask (flat x-fastest, or an (nx, ny, nz) array) and spheres."""
        L = _lib.load()
        sp = np.ascontiguousarray([[*c, r] for c, r in zip(centres, radii)], dtype=np.float64).reshape(-1, 4)
        m = None
        if solid is not None:
            m = _flat(solid, self.n, np.uint8)
        _lib.check(L.mrlbm_set_geometry(self._h, None if m is None else m.ctypes.data_as(ctypes.POINTER(ctypes.c_uint8)),
                                        len(sp), sp.ctypes.data_as(_lib._dp) if len(sp) else None))

    def init_fields(self, rho, ux, uy, uz, mode=INIT_EQUILIBRIUM):
        a = [_flat(v, self.n, np.float64) for v in (rho, ux, uy, uz)]
        _lib.check(_lib.load().mrlbm_init_fields(self._h, *[v.ctypes.data_as(_lib._dp) for v in a], mode))

    def init_uniform(self, rho=1.0, u=(0.0, 0.0, 0.0), mode=INIT_EQUILIBRIUM):
        _lib.check(_lib.load().mrlbm_init_uniform(self._h, rho, *map(float, u), mode))

    def init_taylor_green(self, u0, mode=INIT_EQUILIBRIUM):
        _lib.check(_lib.load().mrlbm_init_taylor_green(self._h, u0, mode))

    # -- stepping -------------------------------------------------------
    def step(self, n=1):
        _lib.check(_lib.load().mrlbm_step(self._h, n))

    def time_steps(self, n):
        """Run n steps between HIP events on the lattice's stream; returns ms."""
        ms = ctypes.c_double()
        _lib.check(_lib.load().mrlbm_time_steps(self._h, n, ctypes.byref(ms)))
        return ms.value

    def synchronize(self):
        _lib.check(_lib.load().mrlbm_synchronize(self._h))

    # -- read-back ------------------------------------------------------
    def moments(self):
        """(10, N) float64, x-fastest, stored convention (MOMENT_NAMES)."""
        out = np.empty(10 * self.n)
        _lib.check(_lib.load().mrlbm_read_moments(self._h, out.ctypes.data_as(_lib._dp)))
        return out.reshape(10, self.n)

    def write_moments(self, m):
        m = np.ascontiguousarray(m, dtype=np.float64).reshape(-1)
        if m.size != 10 * self.n:
            raise ValueError("expected 10*N moments")
        _lib.check(_lib.load().mrlbm_write_moments(self._h, m.ctypes.data_as(_lib._dp)))

    def macroscopic(self, rho0=1.0):
        """rho, ux, uy, uz as (nx, ny, nz) arrays (unscaled)."""
        m = self.moments()
        f = lambda v: v.reshape(self.shape, order="F")
        return f(m[0] + rho0), f(m[1] / 3), f(m[2] / 3), f(m[3] / 3)

    def reduce(self):
        """(mean(rho - rho0), mean(rho |u|^2 / 2)) in FP64."""
        a, b = ctypes.c_double(), ctypes.c_double()
        _lib.check(_lib.load().mrlbm_reduce(self._h, ctypes.byref(a), ctypes.byref(b)))
        return a.value, b.value

    def reduce_checked(self):
        """(mean mass, mean KE, number of non-finite nodes), FP64."""
        a, b, n = ctypes.c_double(), ctypes.c_double(), ctypes.c_int64()
        _lib.check(_lib.load().mrlbm_reduce_checked(self._h, ctypes.byref(a), ctypes.byref(b), ctypes.byref(n)))
        return a.value, b.value, n.value

    def velocity_residual(self):
        """||u - u_prev|| / ||u|| since the previous call (-1 on the first)."""
        r = ctypes.c_double()
        _lib.check(_lib.load().mrlbm_velocity_residual(self._h, ctypes.byref(r)))
        return r.value

    # -- diagnostics ----------------------------------------------------
    def sample(self, points):
        """Stored-convention moments (n, 10) at integer nodes points (n, 3)."""
        p = np.ascontiguousarray(points, dtype=np.int32).reshape(-1, 3)
        out = np.empty((len(p), 10))
        if len(p):
            _lib.check(_lib.load().mrlbm_sample(self._h, len(p), p.ctypes.data_as(ctypes.POINTER(ctypes.c_int32)),
                                                out.ctypes.data_as(_lib._dp)))
        return out

    def moment_means(self):
        out = np.empty(10)
        _lib.check(_lib.load().mrlbm_moment_means(self._h, out.ctypes.data_as(_lib._dp)))
        return out

    def mean_flow_accumulate(self):
        _lib.check(_lib.load().mrlbm_mean_flow_accumulate(self._h))

    def mean_flow_reset(self):
        _lib.check(_lib.load().mrlbm_mean_flow_reset(self._h))

    def mean_flow(self):
        """(rho, ux, uy, uz) time means as (nx, ny, nz) arrays, and the count."""
        out = np.empty(4 * self.n)
        cnt = ctypes.c_int64()
        _lib.check(_lib.load().mrlbm_mean_flow_read(self._h, out.ctypes.data_as(_lib._dp), ctypes.byref(cnt)))
        f = out.reshape(4, self.n)
        return tuple(v.reshape(self.shape, order="F") for v in f), cnt.value

    def turbulent_ke(self):
        v = ctypes.c_double()
        _lib.check(_lib.load().mrlbm_turbulent_ke(self._h, ctypes.byref(v)))
        return v.value

    def checkpoint_save(self, path):
        _lib.check(_lib.load().mrlbm_checkpoint_save(self._h, str(path).encode()))

    def checkpoint_load(self, path):
        _lib.check(_lib.load().mrlbm_checkpoint_load(self._h, str(path).encode()))

    def save_vtk(self, path):
        _lib.check(_lib.load().mrlbm_save_vtk(self._h, str(path).encode()))

    def save_bin(self, directory, sim_id):
        """directory/ID_{rho,ux,uy,uz}<step>.bin as the reference's saveMacr writes them."""
        _lib.check(_lib.load().mrlbm_save_bin(self._h, str(directory).encode(), str(sim_id).encode()))

    def save_info(self, directory, sim_id, case_name="", u_max=0.0, save_steps=0, mlups=0.0):
        """directory/ID_info.txt in the reference's layout."""
        _lib.check(_lib.load().mrlbm_save_info(self._h, str(directory).encode(), str(sim_id).encode(),
                                               str(case_name).encode(), float(u_max), int(save_steps),
                                               float(mlups)))

    def info(self):
        i = _lib.Info()
        _lib.check(_lib.load().mrlbm_get_info(self._h, ctypes.byref(i)))
        return {k: getattr(i, k) for k, _ in _lib.Info._fields_}


def checkpoint_info(path):
    """(parameters dict, step) a checkpoint file was saved with; validates
    the file's structure on the host (no device needed)."""
    p = _lib.Params()
    step = ctypes.c_int64()
    _lib.check(_lib.load().mrlbm_checkpoint_info(str(path).encode(), ctypes.byref(p), ctypes.byref(step)))
    return {k: getattr(p, k) for k, _ in _lib.Params._fields_}, step.value


def copy_bandwidth(nbytes=1 << 30, iters=20, device=0):
    """Measured device copy bandwidth in GB/s (read + write bytes / time)."""
    ms = ctypes.c_double()
    _lib.check(_lib.load().mrlbm_copy_bench(device, nbytes, iters, ctypes.byref(ms)))
    return 2.0 * nbytes / (ms.value * 1e-3) / 1e9, ms.value


def bc_class_info(q, node_type, omega, u=(0.0, 0.0, 0.0)):
    """(known mask, 7x7 inverse) the kernels use for a boundary node type (host only)."""
    known = ctypes.c_uint32()
    a = np.zeros(49)
    uu = np.asarray(u, dtype=np.float64)
    _lib.check(_lib.load().mrlbm_bc_class_info(q, int(node_type), float(omega), uu.ctypes.data_as(_lib._dp),
                                               ctypes.byref(known), a.ctypes.data_as(_lib._dp)))
    return known.value, a.reshape(7, 7)


def bc_solve_host(q, node_type, omega, u, pop, rho0=1.0):
    """FP64 evaluation of the kernels' boundary rule for one node (host, for
    checks), in the kernels' formulation: populations as deviations
    g = f - w rho0 and the solve split around the rest state. Returns
    unscaled rho, u, m as a bc_definition.inc fragment leaves them."""
    known, ainv = bc_class_info(q, node_type, omega, u)
    c = np.array([[0, 1, -1, 0, 0, 0, 0, 1, -1, 1, -1, 0, 0, 1, -1, 1, -1, 0, 0, 1, -1, 1, -1, 1, -1, -1, 1],
                  [0, 0, 0, 1, -1, 0, 0, 1, -1, 0, 0, 1, -1, -1, 1, 0, 0, 1, -1, 1, -1, 1, -1, -1, 1, 1, -1],
                  [0, 0, 0, 0, 0, 1, -1, 0, 0, 1, -1, 1, -1, 0, 0, -1, 1, -1, 1, 1, -1, -1, 1, 1, -1, 1, -1]])[:, :q]
    kcls = np.array([0 if i == 0 else (1 if i < 7 else (2 if i < 19 else 3)) for i in range(q)])
    w = (np.array([1 / 3, 1 / 18, 1 / 36]) if q == 19 else np.array([8 / 27, 2 / 27, 1 / 54, 1 / 216]))[kcls]
    k = np.array([(known >> i) & 1 for i in range(q)], bool)
    t = 1.0 / 3.0
    H = np.stack([c[0] * c[0] - t, c[0] * c[1], c[0] * c[2], c[1] * c[1] - t, c[1] * c[2], c[2] * c[2] - t])
    g = np.where(k, np.asarray(pop, float) - w * rho0, 0.0)
    wk = w[k].sum()
    cvec = (H[:, k] * w[k]).sum(axis=1) / wk
    yrest = ainv @ np.concatenate([[1.0], cvec])
    r = g.sum()
    T = H @ g
    rI = rho0 * wk + r
    db = (T - cvec * r) / rI
    dy = ainv[:, 1:] @ db
    y = yrest + dy
    drho = rho0 * (wk * yrest[0] - 1.0) + r * yrest[0] + rI * dy[0]
    return np.concatenate([[rho0 + drho], np.asarray(u, float), y[1:] / y[0]])


class Batch:
    """Lattices of one type (q, tile, storage) advanced together, one launch
    per step. Members are `Lattice` objects created with the same q, tile,
    storage and the tile kernel."""

    def __init__(self, q=19, tile=(8, 8, 8), storage=STORE_FP32, device=0):
        L = _lib.load()
        h = ctypes.c_void_p()
        _lib.check(L.mrlbm_batch_create(device, q, tile[0], tile[1], tile[2], storage, ctypes.byref(h)))
        self._h = h
        self.members = []

    def add(self, lat):
        _lib.check(_lib.load().mrlbm_batch_add(self._h, lat._h))
        self.members.append(lat)

    def remove(self, lat):
        _lib.check(_lib.load().mrlbm_batch_remove(self._h, lat._h))
        self.members.remove(lat)

    def set_grid(self, max_blocks):
        _lib.check(_lib.load().mrlbm_batch_set_grid(self._h, max_blocks))

    def step(self, n, graph_steps=0):
        _lib.check(_lib.load().mrlbm_batch_step(self._h, n, graph_steps))

    def time_steps(self, n, graph_steps=0):
        ms = ctypes.c_double()
        _lib.check(_lib.load().mrlbm_batch_time_steps(self._h, n, graph_steps, ctypes.byref(ms)))
        return ms.value

    def synchronize(self):
        _lib.check(_lib.load().mrlbm_batch_synchronize(self._h))

    def size(self):
        m, n = ctypes.c_int(), ctypes.c_int64()
        _lib.check(_lib.load().mrlbm_batch_size(self._h, ctypes.byref(m), ctypes.byref(n)))
        return m.value, n.value

    def close(self):
        if getattr(self, "_h", None):
            _lib.load().mrlbm_batch_destroy(self._h)
            self._h = None

    def __del__(self):
        self.close()

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()


def autotune(dims, shapes=((8, 8, 8), (16, 8, 4), (32, 4, 4), (64, 4, 2)), q=19, storage=STORE_FP32, steps=20,
             device=0, walls=None):
    """Fastest tile shape for batching lattices of extents `dims` (and wall
    axes `walls`, one (wx, wy, wz) per lattice, default all periodic);
    returns (shape, {shape: ms or None})."""
    sh = (ctypes.c_int * (3 * len(shapes)))(*[v for s in shapes for v in s])
    dm = (ctypes.c_int * (3 * len(dims)))(*[v for d in dims for v in d])
    best = ctypes.c_int()
    ms = np.zeros(len(shapes))
    if walls is None:
        _lib.check(_lib.load().mrlbm_batch_autotune(device, q, storage, len(shapes), sh, len(dims), dm, steps,
                                                    ctypes.byref(best), ms.ctypes.data_as(_lib._dp)))
    else:
        if len(walls) != len(dims):
            raise ValueError("one wall triple per lattice")
        wl = (ctypes.c_int * (3 * len(walls)))(*[int(bool(v)) for w in walls for v in w])
        _lib.check(_lib.load().mrlbm_batch_autotune_walls(device, q, storage, len(shapes), sh, len(dims), dm, wl,
                                                          steps, ctypes.byref(best), ms.ctypes.data_as(_lib._dp)))
    return tuple(shapes[best.value]), {tuple(s): (None if m < 0 else float(m)) for s, m in zip(shapes, ms)}
