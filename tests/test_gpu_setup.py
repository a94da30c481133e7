"""Device-side set-up: geometry classified on the GPU (mrlbm_set_geometry)
and the uniform initial state filled on the GPU (mrlbm_init_uniform) give
the same node classes, boundary rules and stepped fields, bit for bit, as
the host classification (mrlbm_voxelize_spheres + mrlbm_classify +
mrlbm_set_node_types) and the host-built initial state (mrlbm_init_fields)."""
import numpy as np
import pytest

import mrlbm
from mrlbm import geometry

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)


CASES = [
    # shape, q, walls, spheres (cx, cy, cz, r), random-voxel fraction
    ((64, 32, 48), 19, (0, 1, 1), [(3.0, 16.2, 20.0, 6.5), (40.5, 2.0, 30.0, 4.0), (62.0, 31.0, 47.0, 5.5),
                                   (20.0, 16.0, 24.0, 0.4)], 0.0),
    ((64, 32, 48), 27, (0, 1, 1), [(3.0, 16.2, 20.0, 6.5), (40.5, 2.0, 30.0, 4.0)], 0.0),
    ((32, 32, 32), 19, (0, 0, 0), [], 0.04),
    ((32, 40, 24), 19, (1, 0, 1), [(0.0, 0.0, 12.0, 7.0), (31.0, 39.0, 12.0, 3.0)], 0.02),
    ((48, 48, 48), 19, (1, 1, 1), [(24.0, 24.0, 24.0, 10.0)], 0.0),
]


def _lattice(shape, q, walls, storage=mrlbm.STORE_FP32):
    return mrlbm.Lattice(shape, q=q, tau=0.58, walls=walls, force=(2e-6, 0, 0), tile=(8, 8, 8),
                         kernel=mrlbm.KERNEL_TILE, storage=storage, seed=11)


@pytest.mark.parametrize("case", range(len(CASES)))
def test_device_geometry_equals_host_classification(case):
    shape, q, walls, sph, frac = CASES[case]
    say("case", case, shape, q, walls, len(sph), frac)
    rng = np.random.default_rng(100 + case)
    mask = rng.random(shape) < frac if frac > 0 else None
    centres = [s[:3] for s in sph]
    radii = [s[3] for s in sph]
    solid = np.zeros(shape, bool) if mask is None else mask.copy()
    if sph:
        solid |= geometry.spheres(shape, centres, radii)
    with _lattice(shape, q, walls) as host, _lattice(shape, q, walls) as dev:
        host.set_node_types(geometry.classify(shape, walls, solid if solid.any() else None))
        dev.set_geometry(mask, centres, radii)
        ih, idv = host.info(), dev.info()
        assert (ih["n_bc_classes"], ih["has_solid"]) == (idv["n_bc_classes"], idv["has_solid"])
        for lat in (host, dev):
            lat.init_uniform()
            lat.step(40)
        a, b = host.moments(), dev.moments()
        assert np.isfinite(a).all()
        assert np.array_equal(a, b)


@pytest.mark.parametrize("storage", [mrlbm.STORE_FP32, mrlbm.STORE_FP16, mrlbm.STORE_BF16])
@pytest.mark.parametrize("mode", [mrlbm.INIT_EQUILIBRIUM, mrlbm.INIT_REFERENCE])
def test_device_uniform_state_equals_host_fields(storage, mode):
    shape = (32, 16, 24)
    n = int(np.prod(shape))
    rho, u = 1.013, (0.021, -0.0137, 0.0305)
    say("uniform", storage, mode)
    with _lattice(shape, 19, (0, 1, 0), storage) as a, _lattice(shape, 19, (0, 1, 0), storage) as b:
        for lat in (a, b):
            lat.set_node_types(geometry.classify(shape, (0, 1, 0)))
        a.init_uniform(rho, u, mode)
        b.init_fields(np.full(n, rho), np.full(n, u[0]), np.full(n, u[1]), np.full(n, u[2]), mode)
        assert np.array_equal(a.moments(), b.moments())
        a.step(20)
        b.step(20)
        assert np.array_equal(a.moments(), b.moments())


def test_geometry_rejects_bad_spheres():
    with _lattice((16, 16, 16), 19, (0, 0, 0)) as lat:
        with pytest.raises(mrlbm.MrlbmError):
            lat.set_geometry(None, [(1.0, 2.0, float("nan"))], [2.0])
        with pytest.raises(mrlbm.MrlbmError):
            lat.set_geometry(None, [(1.0, 2.0, 3.0)], [-1.0])


@pytest.mark.parametrize("q", [19, 27])
def test_geometry_after_node_types_and_back(q):
    """A class table from mrlbm_set_node_types holds only that map's classes;
    mrlbm_set_geometry afterwards needs the full octant table, and a node-type
    map after that must replace the device classification again."""
    shape = (32, 24, 16)
    walls = (0, 1, 0)
    centres, radii = [(12.0, 11.5, 8.0)], [4.5]
    say("reuse", q)
    with _lattice(shape, q, walls) as a, _lattice(shape, q, walls) as b, _lattice(shape, q, walls) as c:
        a.set_node_types(geometry.classify(shape, walls))  # two wall classes
        a.set_geometry(None, centres, radii)
        b.set_geometry(None, centres, radii)
        assert a.info()["n_bc_classes"] == b.info()["n_bc_classes"]
        for lat in (a, b):
            lat.init_uniform()
            lat.step(30)
        assert np.isfinite(a.moments()).all()
        assert np.array_equal(a.moments(), b.moments())
        a.set_node_types(geometry.classify(shape, walls))
        c.set_node_types(geometry.classify(shape, walls))
        for lat in (a, c):
            lat.init_uniform()
            lat.step(30)
        assert np.array_equal(a.moments(), c.moments())
