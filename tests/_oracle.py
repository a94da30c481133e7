"""ctypes binding of the FP64 CPU oracle (oracle/liboracle.so).

Test infrastructure: imported by tests, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py only, as the checker / CPU baseline.
"""
import ctypes
import os

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB_PATH = os.path.join(ROOT, "oracle", "liboracle.so")

_dp = ctypes.POINTER(ctypes.c_double)
_u32p = ctypes.POINTER(ctypes.c_uint32)
_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} missing: run make -C oracle")
        L = ctypes.CDLL(LIB_PATH)
        L.orc_create.restype = ctypes.c_void_p
        L.orc_create.argtypes = [ctypes.c_int] * 4 + [ctypes.c_double] * 5 + [ctypes.c_int] * 3
        L.orc_destroy.argtypes = [ctypes.c_void_p]
        L.orc_set_bc_velocity.argtypes = [ctypes.c_void_p, ctypes.c_int] + [ctypes.c_double] * 3
        L.orc_set_node_types.argtypes = [ctypes.c_void_p, _u32p]
        L.orc_init_fields.argtypes = [ctypes.c_void_p, _dp, _dp, _dp, _dp, ctypes.c_int]
        L.orc_init_taylor_green.argtypes = [ctypes.c_void_p, ctypes.c_double, ctypes.c_int]
        L.orc_init_uniform.argtypes = [ctypes.c_void_p] + [ctypes.c_double] * 4 + [ctypes.c_int]
        L.orc_step.argtypes = [ctypes.c_void_p, ctypes.c_int]
        L.orc_step_timed.argtypes = [ctypes.c_void_p, ctypes.c_int]
        L.orc_step_timed.restype = ctypes.c_double
        L.orc_get_moments.argtypes = [ctypes.c_void_p, _dp]
        L.orc_set_moments.argtypes = [ctypes.c_void_p, _dp]
        L.orc_set_ghost_tiles.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_int, ctypes.c_int]
        L.orc_reduce.argtypes = [ctypes.c_void_p, _dp, _dp]
        L.orc_count_nan.argtypes = [ctypes.c_void_p]
        L.orc_count_nan.restype = ctypes.c_int64
        L.orc_set_threads.argtypes = [ctypes.c_int]
        L.orc_get_threads.restype = ctypes.c_int
        L.orc_bc_node.argtypes = [ctypes.c_int, ctypes.c_uint32, ctypes.c_double, _dp, _dp, _dp]
        L.orc_known_mask.argtypes = [ctypes.c_int, ctypes.c_uint32]
        L.orc_known_mask.restype = ctypes.c_uint32
        L.orc_reconstruct.argtypes = [ctypes.c_int, _dp, _dp]
        L.orc_collide.argtypes = [_dp] + [ctypes.c_double] * 4
        _lib = L
    return _lib


def _p(a):
    return a.ctypes.data_as(_dp)


class Oracle:
    """One FP64 lattice; moments in the reference's stored convention."""

    def __init__(self, nx, ny, nz, q=19, tau=0.6, rho0=1.0, force=(0.0, 0.0, 0.0), walls=(0, 0, 0)):
        L = lib()
        self.shape = (nx, ny, nz)
        self.n = nx * ny * nz
        self.h = L.orc_create(nx, ny, nz, q, tau, rho0, *force, *walls)
        if not self.h:
            raise ValueError("orc_create rejected the parameters")

    def __del__(self):
        if getattr(self, "h", None):
            lib().orc_destroy(self.h)
            self.h = None

    def set_bc_velocity(self, k, u):
        assert lib().orc_set_bc_velocity(self.h, k, *u) == 0

    def set_node_types(self, types):
        if types is None:
            assert lib().orc_set_node_types(self.h, None) == 0
            return
        t = np.ascontiguousarray(types.reshape(-1, order="F") if types.ndim == 3 else types, dtype=np.uint32)
        assert lib().orc_set_node_types(self.h, t.ctypes.data_as(_u32p)) == 0

    def set_ghost_tiles(self, tile):
        """Reference ghost semantics for this tile shape (None: off)."""
        assert lib().orc_set_ghost_tiles(self.h, *(tile or (0, 0, 0))) == 0

    def init_taylor_green(self, u0, mode=0):
        assert lib().orc_init_taylor_green(self.h, u0, mode) == 0

    def init_uniform(self, rho=1.0, u=(0.0, 0.0, 0.0), mode=0):
        assert lib().orc_init_uniform(self.h, rho, *u, mode) == 0

    def init_fields(self, rho, ux, uy, uz, mode=0):
        a = [np.ascontiguousarray(v.reshape(-1, order="F"), dtype=np.float64) for v in (rho, ux, uy, uz)]
        assert lib().orc_init_fields(self.h, *[_p(v) for v in a], mode) == 0

    def step(self, n):
        lib().orc_step(self.h, n)

    def step_timed(self, n):
        return lib().orc_step_timed(self.h, n)

    def moments(self):
        out = np.empty(10 * self.n)
        lib().orc_get_moments(self.h, _p(out))
        return out.reshape(10, self.n)

    def set_moments(self, m):
        m = np.ascontiguousarray(m, dtype=np.float64).reshape(-1)
        lib().orc_set_moments(self.h, _p(m))

    def reduce(self):
        a, b = ctypes.c_double(), ctypes.c_double()
        lib().orc_reduce(self.h, ctypes.byref(a), ctypes.byref(b))
        return a.value, b.value

    def count_nan(self):
        return lib().orc_count_nan(self.h)


def bc_node(q, node_type, omega, u, pop):
    out = np.zeros(10)
    u = np.asarray(u, dtype=np.float64)
    pop = np.ascontiguousarray(pop, dtype=np.float64)
    rc = lib().orc_bc_node(q, int(node_type), omega, _p(u), _p(pop), _p(out))
    return out if rc == 0 else None


def known_mask(q, node_type):
    return lib().orc_known_mask(q, int(node_type))


def reconstruct(q, m):
    m = np.ascontiguousarray(m, dtype=np.float64)
    f = np.zeros(q)
    lib().orc_reconstruct(q, _p(m), _p(f))
    return f


def collide(m, omega, force=(0.0, 0.0, 0.0)):
    m = np.array(m, dtype=np.float64)
    lib().orc_collide(_p(m), omega, *force)
    return m
