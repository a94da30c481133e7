"""Serving: the continuously batched request loop of the library
(mrlbm_serve_run) and its seeded duct request stream."""
import ctypes

import numpy as np

from . import _lib


def duct_requests(n, rate=20.0, seed=5):
    """The seeded Poisson stream of duct solves (mrlbm_generate_duct_requests)."""
    arr = (_lib.Request * n)()
    _lib.check(_lib.load().mrlbm_generate_duct_requests(seed, n, rate, arr))
    return arr


def run(requests, storage=0, q=19, tile=(8, 8, 8), budget_gb=200.0, quantum=100, graph_steps=50, max_batch=0,
        realtime=True, device=0):
    """Serve `requests` (a ctypes array of Request); returns the stats dict.
    Per-request results are written into the array."""
    p = _lib.ServeParams(device, q, storage, tile[0], tile[1], tile[2], budget_gb * 1e9, quantum, graph_steps,
                         max_batch, 1 if realtime else 0)
    st = _lib.ServeStats()
    _lib.check(_lib.load().mrlbm_serve_run(ctypes.byref(p), requests, len(requests), ctypes.byref(st)))
    return {k: getattr(st, k) for k, _ in _lib.ServeStats._fields_}


def simulate(requests, rate_mlups, admit_s=0.0, storage=0, q=19, tile=(8, 8, 8), budget_gb=200.0, quantum=100,
             graph_steps=50, max_batch=0, realtime=True):
    """The same loop on a simulated device (no GPU): virtual clock, rounds
    taking nodes * steps / rate, admissions admit_s each. Returns the stats
    dict; start/finish/latency are written into the array."""
    p = _lib.ServeParams(0, q, storage, tile[0], tile[1], tile[2], budget_gb * 1e9, quantum, graph_steps,
                         max_batch, 1 if realtime else 0)
    st = _lib.ServeStats()
    _lib.check(_lib.load().mrlbm_serve_simulate(ctypes.byref(p), requests, len(requests), float(rate_mlups),
                                                float(admit_s), ctypes.byref(st)))
    return {k: getattr(st, k) for k, _ in _lib.ServeStats._fields_}


def attach_output(req):
    """Allocate a moments buffer for one request; returns the numpy array."""
    n = req.nx * req.ny * req.nz
    buf = np.zeros(10 * n)
    req.out_moments = buf.ctypes.data_as(ctypes.POINTER(ctypes.c_double))
    return buf
