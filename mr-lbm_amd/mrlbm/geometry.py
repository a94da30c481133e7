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


def classify(shape, walls=(0, 0, 0), solid=None, moving=None):
    """Node-type map (native, mrlbm_classify / mrlbm_set_moving_wall).

    solid: bool array (nx, ny, nz) of solid voxels, or None.
    moving: list of (axis, side, vel_index): every non-solid node of the
      domain face `side` (-1 low, +1 high) of `axis` becomes a fixed-velocity
      wall node using UX_BC/UY_BC/UZ_BC[vel_index] (e.g. a cavity lid).
    """
    from . import _lib
    import ctypes
    L = _lib.load()
    nx, ny, nz = shape
    t = np.zeros(nx * ny * nz, np.uint32)
    sp = None
    if solid is not None:
        s = np.ascontiguousarray(np.asarray(solid, bool).reshape(-1, order="F"), dtype=np.uint8)
        sp = s.ctypes.data_as(ctypes.POINTER(ctypes.c_uint8))
    _lib.check(L.mrlbm_classify(nx, ny, nz, walls[0], walls[1], walls[2], sp, t.ctypes.data_as(_lib._u32p)))
    for axis, side, k in moving or []:
        _lib.check(L.mrlbm_set_moving_wall(nx, ny, nz, axis, side, k, t.ctypes.data_as(_lib._u32p)))
    return t.reshape(shape, order="F")


def read_voxel_csv(path, shape):
    """Solid mask from the reference's voxel CSV ("x,y,z" per line)."""
    from . import _lib
    import ctypes
    s = np.zeros(int(np.prod(shape)), np.uint8)
    n = _lib.load().mrlbm_read_voxel_csv(str(path).encode(), *shape, s.ctypes.data_as(ctypes.POINTER(ctypes.c_uint8)))
    if n < 0:
        _lib.check(n)
    return s.reshape(shape, order="F").astype(bool)


def cavity(n, lid_axis=2):
    """Lid-driven cavity n^3: walls on all axes, lid on the high face of
    lid_axis using velocity index 0 (edges and corners of the lid plane
    belong to the lid)."""
    return classify((n, n, n), (1, 1, 1), None, [(lid_axis, +1, 0)])


def spheres(shape, centres, radii):
    """Boolean voxel mask of spheres (node centres inside radius; native)."""
    from . import _lib
    import ctypes
    sph = np.ascontiguousarray([[c[0], c[1], c[2], r] for c, r in zip(centres, radii)], dtype=np.float64).reshape(-1)
    s = np.zeros(int(np.prod(shape)), np.uint8)
    _lib.check(_lib.load().mrlbm_voxelize_spheres(*shape, len(radii), sph.ctypes.data_as(_lib._dp),
                                                  s.ctypes.data_as(ctypes.POINTER(ctypes.c_uint8))))
    return s.reshape(shape, order="F").astype(bool)
