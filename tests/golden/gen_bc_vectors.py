"""Write golden vectors of the reference's boundary-condition fragments.

Runs only in a container that holds the reference tree: it loads the
fragment libraries that oracle/ref/Makefile compiles into oracle/_ref/ from
the reference's own cases/*/bc_definition.inc files, evaluates each node
type on seeded random populations and writes tests/golden/bc_<case>.npz
(inputs and outputs only), and records the reference's node-type constants
(src/nodeTypeMap.h, compiled into libnodetype_ref.so) as
tests/golden/nodetype_ref.npz. The tests read the .npz files; they never need
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

# wall node types named in the fragments: faces, edges, corners
WALL_NAMES = ["NORTH", "SOUTH", "WEST", "EAST", "FRONT", "BACK",
              "NORTH_WEST", "NORTH_EAST", "NORTH_FRONT", "NORTH_BACK", "SOUTH_WEST", "SOUTH_EAST", "SOUTH_FRONT",
              "SOUTH_BACK", "WEST_FRONT", "WEST_BACK", "EAST_FRONT", "EAST_BACK",
              "NORTH_WEST_FRONT", "NORTH_WEST_BACK", "NORTH_EAST_FRONT", "NORTH_EAST_BACK",
              "SOUTH_WEST_FRONT", "SOUTH_WEST_BACK", "SOUTH_EAST_FRONT", "SOUTH_EAST_BACK"]


def reference_node_types():
    """name -> value of the reference's node-type constants (src/nodeTypeMap.h,
    compiled by oracle/ref/Makefile into libnodetype_ref.so)."""
    lib = ctypes.CDLL(os.path.join(REF, "libnodetype_ref.so"))
    lib.nodetype_ref_name.restype = ctypes.c_char_p
    lib.nodetype_ref_value.restype = ctypes.c_uint32
    return {lib.nodetype_ref_name(i).decode(): int(lib.nodetype_ref_value(i)) for i in range(lib.nodetype_ref_count())}


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
    nt = reference_node_types()
    names = sorted(nt)
    path = os.path.join(HERE, "nodetype_ref.npz")
    np.savez_compressed(path, names=np.array(names), values=np.array([nt[k] for k in names], np.uint32))
    print(f"{path}: {len(names)} node-type constants")
    walls = sorted(nt[k] for k in WALL_NAMES)
    run_case("lidDrivenCavity_3D", walls, 4, 1, lid=True)
    run_case("parallelPlates_D3Q19", [0x33, 0xCC], 8, 2)
    run_case("parallelPlates_D3Q27", [0x33, 0xCC], 8, 3)
    run_case("voxel", list(range(1, 255)), 2, 4)


if __name__ == "__main__":
    main()
