"""Pins of the oracle's boundary rule beyond the golden vectors (CPU only):
the node-type constants against the reference's own header, equivariance of
the per-node boundary solve under the 48 symmetries of the cube for every
octant pattern, and plane Couette flow (a moving wall over a resting one),
whose steady profile is exactly linear."""
import itertools
import os
import re

import numpy as np
import pytest

import _oracle as O

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CX = np.array([0, 1, -1, 0, 0, 0, 0, 1, -1, 1, -1, 0, 0, 1, -1, 1, -1, 0, 0, 1, -1, 1, -1, 1, -1, -1, 1])
CY = np.array([0, 0, 0, 1, -1, 0, 0, 1, -1, 0, 0, 1, -1, -1, 1, 0, 0, 1, -1, 1, -1, 1, -1, -1, 1, 1, -1])
CZ = np.array([0, 0, 0, 0, 0, 1, -1, 0, 0, 1, -1, 1, -1, 0, 0, -1, 1, -1, 1, 1, -1, -1, 1, 1, -1, 1, -1])
BC_VELOCITY_WALL = 0x100


def header_constants(path):
    """name -> value of the '#define NAME (0x...u)' lines of a C header."""
    out = {}
    for line in open(path):
        m = re.match(r"#define\s+([A-Z_0-9]+)\s+\((0x[0-9A-Fa-f]+)u\)", line)
        if m:
            out[m.group(1)] = int(m.group(2), 16)
    return out


def test_nodetype_header_matches_reference():
    """include/mrlbm_nodetype.h holds the reference's node-type values
    (src/nodeTypeMap.h, recorded from a build of that header by
    tests/golden/gen_bc_vectors.py): case files name these macros."""
    d = np.load(os.path.join(HERE, "golden", "nodetype_ref.npz"))
    ref = dict(zip([str(n) for n in d["names"]], [int(v) for v in d["values"]]))
    ours = header_constants(os.path.join(ROOT, "include", "mrlbm_nodetype.h"))
    assert len(ref) == 38
    for name, value in ref.items():
        assert ours.get(name) == value, name


def cube_symmetries():
    """The 48 orthogonal matrices of the cube's symmetry group (axis
    permutations times sign flips)."""
    out = []
    for perm in itertools.permutations(range(3)):
        for signs in itertools.product((1, -1), repeat=3):
            g = np.zeros((3, 3), int)
            for a in range(3):
                g[a, perm[a]] = signs[a]
            out.append(g)
    return out


def octant_vector(k):
    return np.array([2 * (k & 1) - 1, 2 * ((k >> 1) & 1) - 1, 2 * ((k >> 2) & 1) - 1])


def octant_index(v):
    return int((v[0] > 0) | ((v[1] > 0) << 1) | ((v[2] > 0) << 2))


def transform_type(g, t):
    low = 0
    for k in range(8):
        if (t >> k) & 1:
            low |= 1 << octant_index(g @ octant_vector(k))
    return (t & ~0xFF) | low


def population_map(g, q):
    """j[i]: index of the population whose velocity is g c_i."""
    c = np.stack([CX[:q], CY[:q], CZ[:q]], 1)
    lookup = {tuple(v): i for i, v in enumerate(c)}
    return np.array([lookup[tuple(g @ c[i])] for i in range(q)])


def transform_moments(g, out):
    """rho, u (vector) and the second moments (symmetric tensor) under g."""
    u = g @ out[1:4]
    M = np.array([[out[4], out[5], out[6]], [out[5], out[7], out[8]], [out[6], out[8], out[9]]])
    Mg = g @ M @ g.T
    return np.array([out[0], *u, Mg[0, 0], Mg[0, 1], Mg[0, 2], Mg[1, 1], Mg[1, 2], Mg[2, 2]])


def with_unknowns(q, t, pop):
    """Unknown populations as NaN: reading one poisons the result."""
    p = pop.copy()
    mask = O.known_mask(q, t)
    for i in range(q):
        if not (mask >> i) & 1:
            p[i] = np.nan
    return p


@pytest.mark.parametrize("q", [19, 27])
def test_bc_solve_cube_equivariant(q):
    """For every octant pattern 1..254 and every cube symmetry g: solving the
    boundary system for the transformed node (octants, known populations,
    wall velocity) gives the transformed solution. An index, sign or axis
    mistake in the known-set rule or in the moment system breaks it for some
    pattern and symmetry. Moving walls (BC_VELOCITY_WALL) are included."""
    rng = np.random.default_rng(11 + q)
    syms = cube_symmetries()
    assert len(syms) == 48 and all(abs(round(np.linalg.det(g))) == 1 for g in syms)
    maps = [population_map(g, q) for g in syms]
    for t0 in range(1, 255):
        moving = t0 % 3 == 0
        t = t0 | (BC_VELOCITY_WALL if moving else 0)
        u = rng.uniform(-0.05, 0.05, 3) if moving else np.zeros(3)
        om = rng.uniform(0.6, 1.9)
        pop = rng.uniform(0.01, 0.1, q)
        base = O.bc_node(q, t, om, u, with_unknowns(q, t, pop))
        if base is None:
            continue  # singular system (a node too enclosed to solve): same for every image
        assert np.isfinite(base).all(), hex(t)
        for g, j in zip(syms, maps):
            tg = transform_type(g, t)
            pg = np.empty(q)
            pg[j] = pop
            got = O.bc_node(q, tg, om, g @ u, with_unknowns(q, tg, pg))
            assert got is not None, (hex(t), hex(tg))
            np.testing.assert_allclose(got, transform_moments(g, base), rtol=1e-10, atol=1e-13,
                                       err_msg=f"type {t:#x} -> {tg:#x}")


def test_transform_helpers_self_consistent():
    """The test's own symmetry helpers: the identity leaves everything alone,
    the y-mirror swaps SOUTH and NORTH and flips u_y and m_xy."""
    e = np.eye(3, dtype=int)
    assert transform_type(e, 0x33) == 0x33
    my = np.diag([1, -1, 1])
    assert transform_type(my, 0x33) == 0xCC  # SOUTH <-> NORTH
    assert transform_type(my, 0x55) == 0x55  # WEST stays
    out = np.arange(10, dtype=float)
    r = transform_moments(my, out)
    np.testing.assert_array_equal(r, [0, 1, -2, 3, 4, -5, 6, 7, -8, 9])
    j = population_map(my, 19)
    assert j[3] == 4 and j[7] == 13  # +y -> -y, (+x,+y) -> (+x,-y)


@pytest.mark.parametrize("q", [19, 27])
def test_plane_couette_linear_profile(q):
    """A moving wall (BC_VELOCITY_WALL, u = (U, 0, 0)) at y = ny-1 over a
    no-slip wall at y = 0, periodic in x and z: the steady state is exactly
    u_x = U y / (ny - 1), u_y = u_z = 0, uniform density (wet nodes on both
    walls). Started from rest, the transient decays as exp(-nu pi^2 t / H^2)
    (H = ny - 1): 2500 steps at nu = 0.1, H = 8 leave ~1e-17 of it."""
    nx, ny, nz, U, tau = 4, 9, 4, 0.05, 0.8
    types = np.zeros((nx, ny, nz), np.uint32)
    types[:, 0, :] = 0x33                      # SOUTH: no-slip
    types[:, ny - 1, :] = 0xCC | BC_VELOCITY_WALL  # NORTH, wall-velocity index 0
    o = O.Oracle(nx, ny, nz, q=q, tau=tau, walls=(0, 1, 0))
    o.set_bc_velocity(0, (U, 0.0, 0.0))
    o.set_node_types(types)
    o.init_uniform()
    o.step(2500)
    assert o.count_nan() == 0
    m = o.moments()
    f = lambda k: m[k].reshape((nx, ny, nz), order="F")
    ux = f(1) / 3
    y = np.arange(ny)[None, :, None]
    np.testing.assert_allclose(ux, np.broadcast_to(U * y / (ny - 1), ux.shape), rtol=0, atol=1e-13)
    np.testing.assert_allclose(f(2), 0.0, atol=1e-14)
    np.testing.assert_allclose(f(3), 0.0, atol=1e-14)
    np.testing.assert_allclose(f(0), 0.0, atol=1e-12)  # rho - rho0 (FP64 round-off of 2500 steps)


def load_case(name):
    d = np.load(os.path.join(HERE, "golden", f"case_{name}.npz"))
    return {k: d[k] for k in d.files}


def test_reference_cavity_typing_reads_outside_the_domain():
    """The reference's lidDrivenCavity_3D types its lid edges and corners
    NORTH (cases/lidDrivenCavity_3D/bc_initialization.inc:17-30), so those
    nodes' boundary rule uses populations streamed from outside the domain:
    the oracle's default rule (NaN from outside a wall axis) flags it, and
    with the reference's ghost semantics for 8^3 tiles the run is finite."""
    c = load_case("lidDrivenCavity_3D")
    shape = tuple(int(v) for v in c["shape"])
    for ghost, expect_nan in ((None, True), ((8, 8, 8), False)):
        o = O.Oracle(*shape, q=int(c["q"]), tau=float(c["tau"]), walls=tuple(int(v) for v in c["walls"]))
        o.set_bc_velocity(0, c["bc_velocity"][0])
        o.set_node_types(c["types"])
        o.set_ghost_tiles(ghost)
        o.init_fields(c["rho"], c["ux"], c["uy"], c["uz"], mode=1)
        o.step(20)
        assert (o.count_nan() > 0) == expect_nan, ghost


def test_ghost_semantics_change_nothing_without_outside_reads():
    """A cavity whose wall nodes carry the octants of every wall they touch
    (the lid owning its edges) never reads from outside the domain: the
    reference ghost semantics give the same bits as the default rule."""
    from mrlbm import geometry
    n = 16
    types = geometry.cavity(n, lid_axis=2)
    res = []
    for ghost in (None, (8, 8, 8)):
        o = O.Oracle(n, n, n, q=19, tau=0.6, walls=(1, 1, 1))
        o.set_bc_velocity(0, (0.05, 0, 0))
        o.set_node_types(types)
        o.set_ghost_tiles(ghost)
        o.init_uniform()
        o.step(30)
        assert o.count_nan() == 0
        res.append(o.moments())
    assert np.array_equal(res[0], res[1])


def test_stale_ghost_value_of_a_lid_edge_by_hand():
    """One step of the reference cavity from rest: the lid-edge node
    (0, NY-1, 4) (typed NORTH) pulls population 1 (c = +x) from x = -1.
    In the reference that value sits in the X_1 ghost plane of the last tile
    along x, a wall face never rewritten after initialisation: the rest
    population w_1 RHO_0. Replacing it by that value by hand and solving the
    node's boundary system gives the oracle's stored moments."""
    c = load_case("lidDrivenCavity_3D")
    shape = tuple(int(v) for v in c["shape"])
    nx, ny, nz = shape
    tau, u = float(c["tau"]), c["bc_velocity"][0]
    o = O.Oracle(*shape, q=19, tau=tau, walls=(1, 1, 1))
    o.set_bc_velocity(0, u)
    o.set_node_types(c["types"])
    o.set_ghost_tiles((8, 8, 8))
    o.init_fields(c["rho"], c["ux"], c["uy"], c["uz"], mode=1)
    m0 = o.moments().copy()
    o.step(1)
    got = o.moments()[:, 0 + nx * ((ny - 1) + ny * 4)]
    # by hand: every neighbour of the node holds the rest state, so every
    # pulled population is w_i RHO_0 (inside or stale), and the node is NORTH
    x, y, z = 0, ny - 1, 4
    j = x + nx * (y + ny * z)
    assert np.all(m0[:, j] == m0[:, 0])  # uniform rest start
    w = np.array([1 / 3] + [1 / 18] * 6 + [1 / 36] * 12)
    t = int(c["types"][j])
    assert t == 0xCC | BC_VELOCITY_WALL
    out = O.bc_node(19, t, 1.0 / tau, u, with_unknowns(19, t, w.copy()))
    s = np.array([out[0], 3 * out[1], 3 * out[2], 3 * out[3], 4.5 * out[4], 9 * out[5], 9 * out[6], 4.5 * out[7],
                  9 * out[8], 4.5 * out[9]])
    r = O.collide(s, 1.0 / tau)
    r[0] -= 1.0
    np.testing.assert_allclose(got, r, rtol=1e-12, atol=1e-15)
