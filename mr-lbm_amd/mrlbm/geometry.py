"""Node-type maps in the reference encoding (include/mrlbm_nodetype.h).

Host-side construction of what a case's bc_initialization.inc and the voxel
import produce (src/lbmInitialization.cu:958-975 hostInitialization_nodeType,
:978-1039 read_xyz_file, :1058-1227 define_voxel_bc / bc_id):
  - domain walls set the octants outside the domain (WEST = 0x55, ...);
  - a solid node is SOLID_NODE (0xFF);
  - a fluid node next to solid nodes gets the octants containing the
    direction of each solid neighbour (bc_id's rule);
  - a moving wall adds BC_VELOCITY_WALL and a velocity-table index.
Walls and solid neighbours combine by OR, so a boundary node never treats a
population streamed from a solid node or from outside the domain as known.
Arrays are (nx, ny, nz), uint32.
"""
import itertools

import numpy as np

BULK = 0x00
SOLID_NODE = 0xFF
BC_VELOCITY_WALL = 0x100
VEL_SHIFT = 11


def octants_of(d):
    """Bit mask of the octants whose closed corner region contains direction d."""
    m = 0
    for k in range(8):
        corner = ((1 if k & 1 else -1), (1 if k & 2 else -1), (1 if k & 4 else -1))
        if all(d[a] == 0 or d[a] == corner[a] for a in range(3)):
            m |= 1 << k
    return m


def wall_types(shape, walls):
    """Octant bits of the domain walls on the wall axes."""
    nx, ny, nz = shape
    t = np.zeros(shape, np.uint32)
    n = (nx, ny, nz)
    for a in range(3):
        if not walls[a]:
            continue
        lo = [slice(None)] * 3
        hi = [slice(None)] * 3
        lo[a] = 0
        hi[a] = n[a] - 1
        dlo = [0, 0, 0]
        dhi = [0, 0, 0]
        dlo[a] = -1
        dhi[a] = 1
        t[tuple(lo)] |= octants_of(dlo)
        t[tuple(hi)] |= octants_of(dhi)
    return t


def classify(shape, walls=(0, 0, 0), solid=None, moving=None):
    """Node-type map.

    solid: bool array (nx, ny, nz) of solid voxels, or None.
    moving: list of (axis, side, vel_index): every non-solid node of the
      domain face `side` (-1 low, +1 high) of `axis` becomes a fixed-velocity
      wall node using UX_BC/UY_BC/UZ_BC[vel_index] (e.g. a cavity lid).
    """
    t = wall_types(shape, walls)
    if solid is not None:
        solid = np.asarray(solid, bool)
        acc = np.zeros(shape, np.uint32)
        for d in itertools.product((-1, 0, 1), repeat=3):
            if d == (0, 0, 0):
                continue
            # neighbour at x + d is solid -> octants containing d
            nb = solid
            valid = np.ones(shape, bool)
            for a in range(3):
                if d[a] == 0:
                    continue
                nb = np.roll(nb, -d[a], axis=a)
                if walls[a]:
                    # no wrap across a wall axis
                    idx = [slice(None)] * 3
                    idx[a] = -1 if d[a] > 0 else 0
                    valid[tuple(idx)] = False
            acc |= np.where(nb & valid, np.uint32(octants_of(d)), np.uint32(0))
        t |= acc
        t[solid] = SOLID_NODE
    if moving:
        n = shape
        for axis, side, k in moving:
            sl = [slice(None)] * 3
            sl[axis] = 0 if side < 0 else n[axis] - 1
            face = t[tuple(sl)]
            fluid = (face & 0xFF) != SOLID_NODE
            face[fluid] |= np.uint32(BC_VELOCITY_WALL | (k << VEL_SHIFT))
            t[tuple(sl)] = face
    return t


def cavity(n, lid_axis=2):
    """Lid-driven cavity n^3: walls on all axes, lid on the high face of
    lid_axis using velocity index 0 (edges and corners of the lid plane
    belong to the lid)."""
    return classify((n, n, n), (1, 1, 1), None, [(lid_axis, +1, 0)])


def spheres(shape, centres, radii):
    """Boolean voxel mask of spheres (node centres inside radius)."""
    x, y, z = np.meshgrid(*[np.arange(s) for s in shape], indexing="ij")
    m = np.zeros(shape, bool)
    for (cx, cy, cz), r in zip(centres, radii):
        m |= (x - cx) ** 2 + (y - cy) ** 2 + (z - cz) ** 2 <= r * r
    return m
