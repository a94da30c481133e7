"""ctypes binding of libmrlbm.so (include/mrlbm.h).

The shared library is the product: every compute call below goes to the HIP
kernels inside it. There is no fallback; if the library is missing, loading
raises.
"""
import ctypes
import os

PKG_ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))  # mr-lbm_amd/
LIB_PATH = os.path.join(PKG_ROOT, "libmrlbm.so")

STORE_FP32, STORE_FP16, STORE_BF16 = 0, 1, 2
INIT_EQUILIBRIUM, INIT_REFERENCE = 0, 1
KERNEL_AUTO, KERNEL_TILE, KERNEL_COLUMN = 0, 1, 2


class Params(ctypes.Structure):
    _fields_ = [
        ("nx", ctypes.c_int), ("ny", ctypes.c_int), ("nz", ctypes.c_int), ("q", ctypes.c_int),
        ("bx", ctypes.c_int), ("by", ctypes.c_int), ("bz", ctypes.c_int), ("storage", ctypes.c_int),
        ("wall_x", ctypes.c_int), ("wall_y", ctypes.c_int), ("wall_z", ctypes.c_int), ("kernel", ctypes.c_int),
        ("tau", ctypes.c_double), ("rho0", ctypes.c_double),
        ("fx", ctypes.c_double), ("fy", ctypes.c_double), ("fz", ctypes.c_double),
        ("device", ctypes.c_int), ("seed", ctypes.c_uint),
    ]


class Info(ctypes.Structure):
    _fields_ = [
        ("nx", ctypes.c_int), ("ny", ctypes.c_int), ("nz", ctypes.c_int), ("q", ctypes.c_int),
        ("bx", ctypes.c_int), ("by", ctypes.c_int), ("bz", ctypes.c_int), ("storage", ctypes.c_int),
        ("kernel", ctypes.c_int), ("step", ctypes.c_int64), ("n_bc_classes", ctypes.c_int),
        ("has_solid", ctypes.c_int), ("bytes_per_lu", ctypes.c_double), ("device_bytes", ctypes.c_double),
    ]


class Request(ctypes.Structure):
    _fields_ = [
        ("arrival_s", ctypes.c_double), ("nx", ctypes.c_int), ("ny", ctypes.c_int), ("nz", ctypes.c_int),
        ("nspheres", ctypes.c_int), ("sphere", (ctypes.c_double * 4) * 8), ("gx", ctypes.c_double),
        ("nu", ctypes.c_double), ("steps", ctypes.c_int), ("seed", ctypes.c_uint),
        ("out_moments", ctypes.POINTER(ctypes.c_double)), ("start_s", ctypes.c_double),
        ("finish_s", ctypes.c_double), ("latency_s", ctypes.c_double), ("mass0", ctypes.c_double),
        ("mass1", ctypes.c_double), ("ke0", ctypes.c_double), ("ke1", ctypes.c_double), ("status", ctypes.c_int),
    ]


class ServeParams(ctypes.Structure):
    _fields_ = [
        ("device", ctypes.c_int), ("q", ctypes.c_int), ("storage", ctypes.c_int), ("bx", ctypes.c_int),
        ("by", ctypes.c_int), ("bz", ctypes.c_int), ("hbm_budget_bytes", ctypes.c_double),
        ("quantum_steps", ctypes.c_int), ("graph_steps", ctypes.c_int), ("max_batch", ctypes.c_int),
        ("realtime", ctypes.c_int),
    ]


class ServeStats(ctypes.Structure):
    _fields_ = [
        ("wall_s", ctypes.c_double), ("busy_s", ctypes.c_double), ("lattice_updates", ctypes.c_double),
        ("mlups", ctypes.c_double), ("mlups_busy", ctypes.c_double), ("p50_s", ctypes.c_double),
        ("p99_s", ctypes.c_double), ("max_s", ctypes.c_double), ("peak_bytes", ctypes.c_double),
        ("rounds", ctypes.c_int64), ("failed", ctypes.c_int64),
    ]


class MrlbmError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"mrlbm error {code}: {msg}")
        self.code = code


_dp = ctypes.POINTER(ctypes.c_double)
_u32p = ctypes.POINTER(ctypes.c_uint32)
_vp = ctypes.c_void_p
_lib = None

# name -> argtypes; every entry returns int status unless listed in _RESTYPES
_SIGS = {
    "mrlbm_create": [ctypes.POINTER(Params), ctypes.POINTER(_vp)],
    "mrlbm_destroy": [_vp],
    "mrlbm_set_bc_velocity": [_vp, ctypes.c_int, ctypes.c_double, ctypes.c_double, ctypes.c_double],
    "mrlbm_sample": [_vp, ctypes.c_int, ctypes.POINTER(ctypes.c_int32), _dp],
    "mrlbm_moment_means": [_vp, _dp],
    "mrlbm_mean_flow_accumulate": [_vp],
    "mrlbm_mean_flow_reset": [_vp],
    "mrlbm_mean_flow_read": [_vp, _dp, ctypes.POINTER(ctypes.c_int64)],
    "mrlbm_turbulent_ke": [_vp, _dp],
    "mrlbm_set_node_types": [_vp, _u32p],
    "mrlbm_set_geometry": [_vp, ctypes.POINTER(ctypes.c_uint8), ctypes.c_int, _dp],
    "mrlbm_init_fields": [_vp, _dp, _dp, _dp, _dp, ctypes.c_int],
    "mrlbm_init_uniform": [_vp, ctypes.c_double, ctypes.c_double, ctypes.c_double, ctypes.c_double, ctypes.c_int],
    "mrlbm_init_taylor_green": [_vp, ctypes.c_double, ctypes.c_int],
    "mrlbm_step": [_vp, ctypes.c_int],
    "mrlbm_time_steps": [_vp, ctypes.c_int, _dp],
    "mrlbm_synchronize": [_vp],
    "mrlbm_read_moments": [_vp, _dp],
    "mrlbm_write_moments": [_vp, _dp],
    "mrlbm_reduce": [_vp, _dp, _dp],
    "mrlbm_reduce_checked": [_vp, _dp, _dp, ctypes.POINTER(ctypes.c_int64)],
    "mrlbm_reduce_checked_noalloc": [_vp, _vp],
    "mrlbm_velocity_residual": [_vp, _dp],
    "mrlbm_checkpoint_save": [_vp, ctypes.c_char_p],
    "mrlbm_checkpoint_load": [_vp, ctypes.c_char_p],
    "mrlbm_save_vtk": [_vp, ctypes.c_char_p],
    "mrlbm_checkpoint_info": [ctypes.c_char_p, ctypes.POINTER(Params), ctypes.POINTER(ctypes.c_int64)],
    "mrlbm_save_bin": [_vp, ctypes.c_char_p, ctypes.c_char_p],
    "mrlbm_save_info": [_vp, ctypes.c_char_p, ctypes.c_char_p, ctypes.c_char_p, ctypes.c_double, ctypes.c_int64,
                        ctypes.c_double],
    "mrlbm_get_info": [_vp, ctypes.POINTER(Info)],
    "mrlbm_last_error": [],
    "mrlbm_bc_class_info": [ctypes.c_int, ctypes.c_uint32, ctypes.c_double, _dp, _u32p, _dp],
    "mrlbm_copy_bench": [ctypes.c_int, ctypes.c_size_t, ctypes.c_int, _dp],
    "mrlbm_batch_create": [ctypes.c_int] * 6 + [ctypes.POINTER(_vp)],
    "mrlbm_batch_add": [_vp, _vp],
    "mrlbm_batch_remove": [_vp, _vp],
    "mrlbm_batch_set_grid": [_vp, ctypes.c_int],
    "mrlbm_batch_step": [_vp, ctypes.c_int, ctypes.c_int],
    "mrlbm_batch_time_steps": [_vp, ctypes.c_int, ctypes.c_int, _dp],
    "mrlbm_batch_synchronize": [_vp],
    "mrlbm_batch_size": [_vp, ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_int64)],
    "mrlbm_batch_destroy": [_vp],
    "mrlbm_batch_autotune": [ctypes.c_int] * 4 + [ctypes.POINTER(ctypes.c_int), ctypes.c_int,
                             ctypes.POINTER(ctypes.c_int), ctypes.c_int, ctypes.POINTER(ctypes.c_int), _dp],
    "mrlbm_batch_autotune_walls": [ctypes.c_int] * 4 + [ctypes.POINTER(ctypes.c_int), ctypes.c_int,
                                   ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_int), ctypes.c_int,
                                   ctypes.POINTER(ctypes.c_int), _dp],
    "mrlbm_classify": [ctypes.c_int] * 6 + [ctypes.POINTER(ctypes.c_uint8), _u32p],
    "mrlbm_set_moving_wall": [ctypes.c_int] * 6 + [_u32p],
    "mrlbm_voxelize_spheres": [ctypes.c_int] * 4 + [_dp, ctypes.POINTER(ctypes.c_uint8)],
    "mrlbm_read_voxel_csv": [ctypes.c_char_p] + [ctypes.c_int] * 3 + [ctypes.POINTER(ctypes.c_uint8)],
    "mrlbm_generate_duct_requests": [ctypes.c_uint64, ctypes.c_int, ctypes.c_double, ctypes.POINTER(Request)],
    "mrlbm_serve_simulate": [ctypes.POINTER(ServeParams), ctypes.POINTER(Request), ctypes.c_int, ctypes.c_double,
                             ctypes.c_double, ctypes.POINTER(ServeStats)],
    "mrlbm_serve_run": [ctypes.POINTER(ServeParams), ctypes.POINTER(Request), ctypes.c_int,
                        ctypes.POINTER(ServeStats)],
}
_RESTYPES = {"mrlbm_last_error": ctypes.c_char_p}


def exported_symbols():
    return list(_SIGS)


def load():
    """Load libmrlbm.so; raises if it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} not built: run `make -C mr-lbm_amd` or __graft_entry__.build()")
        L = ctypes.CDLL(LIB_PATH)
        for name, args in _SIGS.items():
            fn = getattr(L, name)
            fn.argtypes = args
            fn.restype = _RESTYPES.get(name, ctypes.c_int)
        _lib = L
    return _lib


def check(rc):
    if rc != 0:
        raise MrlbmError(rc, load().mrlbm_last_error().decode())
    return rc
