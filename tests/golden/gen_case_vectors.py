"""Write fixtures of reference case directories taken through the case
adapter (include/mrlbm_case.h).

Runs only in a container that holds the reference tree: oracle/ref/Makefile
compiles oracle/ref/case_harness.cpp against each case's own constants.inc,
model.inc, bc_initialization.inc, bc_definition.inc and
flow_initialization.inc (lidDrivenCavity_3D, taylorGreen and squaredDuct
with the size constant N set to 16, parallelPlates_D3Q19/D3Q27 at their own
8 x 128 x 8) into oracle/_ref/libcase_<case>.so. This script records what
the adapter makes of each case - lattice parameters, node-type map,
wall-velocity table, initial rho and u - as tests/golden/case_<case>.npz
(data only). The GPU tests feed them through the C ABI.

    make -C oracle/ref && python tests/golden/gen_case_vectors.py
"""
import ctypes
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
REF = os.path.join(ROOT, "oracle", "_ref")
sys.path.insert(0, os.path.join(ROOT, "mr-lbm_amd"))
from mrlbm import _lib  # noqa: E402  (the Params layout of include/mrlbm.h)

CASES = ["lidDrivenCavity_3D", "taylorGreen", "squaredDuct", "parallelPlates_D3Q19", "parallelPlates_D3Q27"]


def record(name):
    lib = ctypes.CDLL(os.path.join(REF, f"libcase_{name}.so"))
    p = _lib.Params()
    assert lib.case_params(ctypes.byref(p)) == 0
    n = p.nx * p.ny * p.nz
    types = np.zeros(n, np.uint32)
    vel = np.zeros(12)
    rc = lib.case_node_types(types.ctypes.data_as(ctypes.POINTER(ctypes.c_uint32)),
                             vel.ctypes.data_as(ctypes.POINTER(ctypes.c_double)))
    assert rc == 0, rc
    f = [np.zeros(n) for _ in range(4)]
    lib.case_flow_fields(*[a.ctypes.data_as(ctypes.POINTER(ctypes.c_double)) for a in f])
    path = os.path.join(HERE, f"case_{name}.npz")
    np.savez_compressed(path, shape=np.array([p.nx, p.ny, p.nz], np.int32), q=np.int32(p.q),
                        walls=np.array([p.wall_x, p.wall_y, p.wall_z], np.int32), tau=np.float64(p.tau),
                        rho0=np.float64(p.rho0), force=np.array([p.fx, p.fy, p.fz]), types=types,
                        bc_velocity=vel.reshape(4, 3), rho=f[0], ux=f[1], uy=f[2], uz=f[3])
    kinds = {hex(int(t)): int(c) for t, c in zip(*np.unique(types, return_counts=True))}
    print(f"{path}: {p.nx}x{p.ny}x{p.nz} D3Q{p.q} tau={p.tau:.6g} walls={p.wall_x}{p.wall_y}{p.wall_z} "
          f"F=({p.fx},{p.fy},{p.fz}) types={kinds} vel0={vel[:3]}")


def main():
    if not os.path.isdir(REF):
        sys.exit("oracle/_ref missing: run make -C oracle/ref first (needs the reference tree)")
    for c in CASES:
        record(c)


if __name__ == "__main__":
    main()
