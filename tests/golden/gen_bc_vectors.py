"""Write golden vectors of the reference's boundary-condition fragments.

Runs only in a container that holds the reference tree: it loads the
fragment libraries that oracle/ref/Makefile compiles into oracle/_ref/ from
the reference's own cases/*/bc_definition.inc files, evaluates each node
type on seeded random populations and writes tests/golden/bc_<case>.npz
(inputs and outputs only). The tests read the .npz files; they never need
the reference or oracle/_ref.

    make -C oracle/ref && python tests/golden/gen_bc_vectors.py
"""
import ctypes
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
REF = os.path.join(HERE, "..", "..", "oracle", "_ref")

CX = [0, 1, -1, 0, 0, 0, 0, 1, -1, 1, -1, 0, 0, 1, -1, 1, -1, 0, 0, 1, -1, 1, -1, 1, -1, -1, 1]
CY = [0, 0, 0, 1, -1, 0, 0, 1, -1, 0, 0, 1, -1, -1, 1, 0, 0, 1, -1, 1, -1, 1, -1, -1, 1, 1, -1]
CZ = [0, 0, 0, 0, 0, 1, -1, 0, 0, 1, -1, 1, -1, 0, 0, -1, 1, -1, 1, 1, -1, -1, 1, 1, -1, 1, -1]

# node-type words named in the wall fragments (include/mrlbm_nodetype.h)
NAMED = {
    "NORTH": 0xCC, "SOUTH": 0x33, "WEST": 0x55, "EAST": 0xAA, "FRONT": 0xF0, "BACK": 0x0F,
    "NORTH_WEST": 0xDD, "NORTH_EAST": 0xEE, "NORTH_FRONT": 0xFC, "NORTH_BACK": 0xCF,
    "SOUTH_WEST": 0x77, "SOUTH_EAST": 0xBB, "SOUTH_FRONT": 0xF3, "SOUTH_BACK": 0x3F,
    "WEST_FRONT": 0xF5, "WEST_BACK": 0x5F, "EAST_FRONT": 0xFA, "EAST_BACK": 0xAF,
    "NORTH_WEST_FRONT": 0xFD, "NORTH_WEST_BACK": 0xDF, "NORTH_EAST_FRONT": 0xFE, "NORTH_EAST_BACK": 0xEF,
    "SOUTH_WEST_FRONT": 0xF7, "SOUTH_WEST_BACK": 0x7F, "SOUTH_EAST_FRONT": 0xFB, "SOUTH_EAST_BACK": 0xBF,
}


def weights(q):
    w = np.zeros(q)
    for i in range(q):
        k = 0 if i == 0 else (1 if i < 7 else (2 if i < 19 else 3))
        w[i] = [1 / 3, 1 / 18, 1 / 36][k] if q == 19 else [8 / 27, 2 / 27, 1 / 54, 1 / 216][k]
    return w


def sample_pops(rng, q):
    """Populations of a plausible streamed state: equilibrium of a random
    (rho, u) plus a random non-equilibrium part."""
    w = weights(q)
    rho = rng.uniform(0.95, 1.05)
    u = rng.uniform(-0.08, 0.08, 3)
    c = np.stack([CX[:q], CY[:q], CZ[:q]], 1).astype(float)
    cu = c @ u
    feq = w * rho * (1 + 3 * cu + 4.5 * cu * cu - 1.5 * u @ u)
    return feq * (1 + rng.uniform(-0.02, 0.02, q))


def run_case(name, types, nsamp, seed, lid=False):
    lib = ctypes.CDLL(os.path.join(REF, f"libbcref_{name}.so"))
    lib.bcref_eval.argtypes = [ctypes.c_uint, ctypes.c_double, ctypes.c_double,
                               ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_double)]
    q = lib.bcref_q()
    rng = np.random.default_rng(seed)
    rows_t, rows_om, rows_u, rows_p, rows_o = [], [], [], [], []
    for t in types:
        for _ in range(nsamp):
            pop = sample_pops(rng, q)
            om = rng.uniform(0.55, 1.95)
            um = rng.uniform(0.01, 0.1) if lid else 0.0
            out = np.zeros(10)
            lib.bcref_eval(t, om, um, pop.ctypes.data_as(ctypes.POINTER(ctypes.c_double)),
                           out.ctypes.data_as(ctypes.POINTER(ctypes.c_double)))
            if np.isnan(out).any():
                continue  # node type not handled by this fragment
            rows_t.append(t)
            rows_om.append(om)
            rows_u.append(um)
            rows_p.append(pop)
            rows_o.append(out)
    path = os.path.join(HERE, f"bc_{name}.npz")
    np.savez_compressed(path, q=np.int32(q), types=np.array(rows_t, np.uint32), omega=np.array(rows_om),
                        umax=np.array(rows_u), pops=np.array(rows_p), out=np.array(rows_o))
    print(f"{path}: {len(rows_t)} vectors, {len(set(rows_t))} node types")


def main():
    if not os.path.isdir(REF):
        sys.exit("oracle/_ref missing: run make -C oracle/ref first (needs the reference tree)")
    walls = sorted(NAMED.values())
    run_case("lidDrivenCavity_3D", walls, 4, 1, lid=True)
    run_case("parallelPlates_D3Q19", [0x33, 0xCC], 8, 2)
    run_case("parallelPlates_D3Q27", [0x33, 0xCC], 8, 3)
    run_case("voxel", list(range(1, 255)), 2, 4)


if __name__ == "__main__":
    main()
