"""The FP64 oracle, pinned against the reference's own numbers and against
what the mathematics fixes (CPU only)."""
import glob
import os

import numpy as np
import pytest

import _oracle as O
from fields import drho_rel, rel_l2

HERE = os.path.dirname(os.path.abspath(__file__))
CX = np.array([0, 1, -1, 0, 0, 0, 0, 1, -1, 1, -1, 0, 0, 1, -1, 1, -1, 0, 0, 1, -1, 1, -1, 1, -1, -1, 1])
CY = np.array([0, 0, 0, 1, -1, 0, 0, 1, -1, 0, 0, 1, -1, -1, 1, 0, 0, 1, -1, 1, -1, 1, -1, -1, 1, 1, -1])
CZ = np.array([0, 0, 0, 0, 0, 1, -1, 0, 0, 1, -1, 1, -1, 0, 0, -1, 1, -1, 1, 1, -1, -1, 1, 1, -1, 1, -1])


def weights(q):
    k = np.array([0 if i == 0 else (1 if i < 7 else (2 if i < 19 else 3)) for i in range(q)])
    return np.array([1 / 3, 1 / 18, 1 / 36])[k] if q == 19 else np.array([8 / 27, 2 / 27, 1 / 54, 1 / 216])[k]


GOLDEN = sorted(glob.glob(os.path.join(HERE, "golden", "bc_*.npz")))


@pytest.mark.parametrize("path", GOLDEN, ids=[os.path.basename(p) for p in GOLDEN])
def test_bc_matches_reference_fragments(path):
    """Boundary formulas of the reference's case files (lid-driven cavity
    with its moving lid, parallel plates D3Q19/D3Q27, the 254-pattern voxel
    table), evaluated from the reference sources, vs the oracle's per-node
    moment-matching solve. Unknown populations are NaN for the oracle, so
    reading one fails the comparison."""
    d = np.load(path)
    q = int(d["q"])
    assert len(d["types"]) > 0
    for t, om, pop, out in zip(d["types"], d["omega"], d["pops"], d["out"]):
        mask = O.known_mask(q, t)
        p = pop.copy()
        for i in range(q):
            if not (mask >> i) & 1:
                p[i] = np.nan
        r = O.bc_node(q, t, om, out[1:4], p)
        assert r is not None
        np.testing.assert_allclose(r, out, rtol=1e-11, atol=1e-13)


def test_golden_covers_all_voxel_patterns():
    d = np.load(os.path.join(HERE, "golden", "bc_voxel.npz"))
    assert sorted(set(int(t) for t in d["types"])) == list(range(1, 255))


def test_known_mask_faces():
    # SOUTH (y = 0): the populations arriving from -y (c_y = +1) are unknown
    m = O.known_mask(19, 0x33)
    assert [i for i in range(19) if not (m >> i) & 1] == [i for i in range(19) if CY[i] == 1]
    # a fully surrounded node knows only the rest population
    assert O.known_mask(19, 0xFE) & 1


@pytest.mark.parametrize("q", [19, 27])
def test_reconstruction_moment_round_trip(q):
    """Moments of R(rho, U, M) give back (rho, u, m) exactly (second-order
    isotropy of D3Q19 and D3Q27)."""
    rng = np.random.default_rng(0)
    w = weights(q)
    for _ in range(20):
        m = np.concatenate([[rng.uniform(0.9, 1.1)], rng.uniform(-0.2, 0.2, 9)])
        f = O.reconstruct(q, m)
        c = np.stack([CX[:q], CY[:q], CZ[:q]])
        rho = f.sum()
        u = c @ f / rho
        S = lambda a, b: (f * c[a] * c[b]).sum() / rho - (1 / 3 if a == b else 0)
        got = np.array([rho, 3 * u[0], 3 * u[1], 3 * u[2], 4.5 * S(0, 0), 9 * S(0, 1), 9 * S(0, 2), 4.5 * S(1, 1),
                        9 * S(1, 2), 4.5 * S(2, 2)])
        np.testing.assert_allclose(got, m, atol=1e-13)
    # rest state -> f_i = w_i
    np.testing.assert_allclose(O.reconstruct(q, np.array([1.0] + [0] * 9)), w, atol=1e-15)


def test_collision_closed_forms():
    rng = np.random.default_rng(1)
    m = np.concatenate([[1.03], rng.uniform(-0.1, 0.1, 9)])
    # omega = 1, F = 0: second moments relax to equilibrium U U / 2 (diag), U U (off-diag)
    r = O.collide(m, 1.0)
    U = m[1:4]
    np.testing.assert_allclose(r[[4, 7, 9]], U ** 2 / 2, atol=1e-15)
    np.testing.assert_allclose(r[[5, 6, 8]], [U[0] * U[1], U[0] * U[2], U[1] * U[2]], atol=1e-15)
    # omega -> 0 is the identity
    np.testing.assert_allclose(O.collide(m, 1e-300), m, atol=1e-15)
    # the force shifts U by 3F/rho and enters the moments linearly + quadratically via U
    F = np.array([1e-3, -2e-3, 5e-4])
    r = O.collide(m, 0.7, F)
    np.testing.assert_allclose(r[1:4], m[1:4] + 3 * F / m[0], atol=1e-15)


def test_periodic_conservation():
    """Periodic, unforced: mass and momentum conserved to round-off."""
    o = O.Oracle(16, 16, 16, 19, tau=0.6)
    o.init_taylor_green(0.05)
    m0 = o.moments()
    rho0 = 1 + m0[0]
    mom0 = [(rho0 * m0[k] / 3).sum() for k in (1, 2, 3)]
    o.step(50)
    m1 = o.moments()
    rho1 = 1 + m1[0]
    assert abs(rho1.sum() - rho0.sum()) < 1e-10
    for k in (1, 2, 3):
        # the stored velocity is post-collision (includes F/rho); with F = 0 momentum is conserved
        assert abs((rho1 * m1[k] / 3).sum() - mom0[k - 1]) < 1e-11


def test_forced_momentum_growth():
    """Periodic with body force F: total momentum grows by N F per step."""
    F = (1e-5, 0.0, 0.0)
    o = O.Oracle(8, 8, 8, 19, tau=0.7, force=F)
    o.init_uniform()
    steps = 10
    o.step(steps)
    m = o.moments()
    # stored U = 3 (u_pre + F/rho) at every step; momentum after k steps is k N F
    # plus the half-step shift carried in the stored velocity
    px = ((1 + m[0]) * m[1] / 3).sum()
    n = 8 ** 3
    np.testing.assert_allclose(px, n * F[0] * steps, rtol=1e-9)


def test_single_node_translation():
    """A perturbation at one node moves by c_i per step: with omega = 1 the
    first step sends f_i^neq along c_i. Check the density perturbation's
    spread after one step lies in the node's 1-neighbourhood."""
    o = O.Oracle(12, 12, 12, 19, tau=1.0)
    rho = np.ones((12, 12, 12))
    rho[6, 6, 6] = 1.01
    z = np.zeros_like(rho)
    o.init_fields(rho, z, z, z)
    o.step(1)
    d = (o.moments()[0]).reshape(12, 12, 12, order="F")
    nz = np.argwhere(np.abs(d) > 1e-15)
    assert np.all(np.abs(nz - 6).max(axis=1) <= 1)
    assert abs(d.sum() - 0.01) < 1e-12


def test_taylor_green_early_decay():
    """TGV kinetic energy decays as exp(-2 nu 3 k^2 t) (linear regime); the
    lattice error is second order: it drops ~4x when the lattice doubles
    under diffusive scaling (steps ~ n^2, u0 ~ 1/n)."""
    errs = []
    for n in (16, 32):
        nu, u0, steps = 0.05, 0.02 * 32 / n, int(100 * (n / 32) ** 2)
        o = O.Oracle(n, n, n, 19, tau=0.5 + 3 * nu)
        o.init_taylor_green(u0)
        _, ke0 = o.reduce()
        assert abs(ke0 - u0 * u0 / 8) < 1e-12 * ke0 + 1e-18
        o.step(steps)
        _, ke = o.reduce()
        k = 2 * np.pi / n
        expect = np.exp(-2 * nu * 3 * k * k * steps)
        errs.append(abs(ke / ke0 - expect) / expect)
    assert errs[1] < 0.03
    assert errs[0] / errs[1] > 3.0


def test_plane_poiseuille():
    """Force-driven flow between walls at y = 0 and y = NY-1 (wet nodes):
    the steady streamed velocity is exactly the parabola
    u(y) = F y (H - y) / (2 nu), H = NY-1; the stored (post-collision)
    velocity carries the +F/rho shift of collision.inc:2-4."""
    ny, nu, F = 17, 1.0 / 6.0, 1e-6
    o = O.Oracle(4, ny, 4, 19, tau=0.5 + 3 * nu, force=(F, 0, 0), walls=(0, 1, 0))
    types = np.zeros((4, ny, 4), np.uint32)
    types[:, 0, :] = 0x33
    types[:, ny - 1, :] = 0xCC
    o.set_node_types(types)
    o.init_uniform()
    o.step(8000)  # slowest mode exp(-nu pi^2 t / H^2) ~ 1e-22
    m = o.moments()
    ux = (m[1] / 3).reshape(4, ny, 4, order="F")[1, :, 1]
    rho = (1 + m[0]).reshape(4, ny, 4, order="F")[1, :, 1]
    y = np.arange(ny)
    ana = F * y * (ny - 1 - y) / (2 * nu)
    assert np.max(np.abs(ux - F / rho - ana)) < 1e-10 * ana.max()
    assert o.count_nan() == 0


def test_cavity_axis_permutation():
    """The z-lid cavity equals the y-lid cavity with y and z swapped."""
    import sys
    sys.path.insert(0, os.path.join(os.path.dirname(HERE), "mr-lbm_amd"))
    from mrlbm import geometry
    n = 12
    tz = geometry.cavity(n, lid_axis=2)
    ty = geometry.cavity(n, lid_axis=1)
    res = []
    for types, u in ((tz, (0.05, 0, 0)), (ty, (0.05, 0, 0))):
        o = O.Oracle(n, n, n, 19, tau=0.6, walls=(1, 1, 1))
        o.set_bc_velocity(0, u)
        o.set_node_types(types)
        o.init_uniform()
        o.step(30)
        assert o.count_nan() == 0
        res.append(o.moments().reshape(10, n, n, n, order="F"))
    a, b = res
    # swap y<->z: fields transpose axes 1,2; components uy<->uz, myy<->mzz, mxy<->mxz
    perm = [0, 1, 3, 2, 4, 6, 5, 9, 8, 7]
    bt = b[perm].transpose(0, 1, 3, 2)
    np.testing.assert_allclose(a, bt, atol=1e-13)


def test_reference_init_second_moments():
    """Reference init (src/lbmInitialization.cu:102-130): m_ab = (rho0/rho)(cs2 d_ab + u_a u_b) - cs2 d_ab."""
    o = O.Oracle(2, 2, 2, 19)
    rho = np.full((2, 2, 2), 1.02)
    ux = np.full((2, 2, 2), 0.03)
    uy = np.full((2, 2, 2), -0.01)
    uz = np.zeros((2, 2, 2))
    o.init_fields(rho, ux, uy, uz, mode=1)
    m = o.moments()[:, 0]
    c = 1 / 3
    np.testing.assert_allclose(m[4] / 4.5, (c + 0.03 ** 2) / 1.02 - c, atol=1e-15)
    np.testing.assert_allclose(m[5] / 9, 0.03 * -0.01 / 1.02, atol=1e-15)
    # D3Q27 with the reference init: at rest m_xx = -1/27 (subset quirk)
    o27 = O.Oracle(2, 2, 2, 27)
    o27.init_uniform(mode=1)
    np.testing.assert_allclose(o27.moments()[4, 0] / 4.5, -1 / 27, atol=1e-15)


def test_rel_l2_helper_detects_difference():
    a = np.zeros((10, 8))
    a[1] = 0.1
    b = a.copy()
    b[1, 0] += 1e-3
    assert rel_l2(a, b)["u"] > 1e-3


def test_rel_l2_density_scale():
    """The density-deviation error is relative to the deviation where it
    exceeds the velocity, and relative to the velocity where the deviation
    is smaller (round-off in a sum of velocity-sized populations) or zero."""
    rng = np.random.default_rng(1)
    b = np.zeros((10, 64))
    b[1:4] = 3 * rng.normal(0, 0.01, (3, 64))  # stored 3u, |u| ~ 0.01
    big = b.copy()
    big[0] = rng.normal(0, 0.1, 64)  # |drho| >> |u|
    a = big.copy()
    a[0] *= 1 + 2e-5
    assert rel_l2(a, big)["drho"] == pytest.approx(2e-5, rel=1e-6)
    assert drho_rel(a, big) == pytest.approx(2e-5, rel=1e-6)
    flat = b.copy()  # drho = 0 (uniform density): the velocity sets the scale
    a = flat.copy()
    a[0] = 3e-5 * np.linalg.norm(b[1:4]) / 3 / np.sqrt(64)
    assert rel_l2(a, flat)["drho"] == pytest.approx(3e-5, rel=1e-6)


def test_duct_series_limits():
    """The rectangular-duct series (mrlbm.analytic) against its own limits:
    a very wide duct is plane Poiseuille in its middle; t -> 0 is rest; the
    steady profile is reached for t >> 1/(nu lambda_11)."""
    from mrlbm import analytic
    u = analytic.duct_start_up(33, 2001, 1e-6, 0.05)
    pl = analytic.channel_steady(33, 1e-6, 0.05)
    assert np.abs(u[:, 1000] - pl).max() < 1e-4 * pl.max()
    assert np.abs(analytic.duct_start_up(16, 24, 1e-6, 0.05, t=0.0)).max() == 0.0
    steady = analytic.duct_start_up(16, 24, 1e-6, 0.05)
    late = analytic.duct_start_up(16, 24, 1e-6, 0.05, t=1e6)
    assert np.abs(late - steady).max() < 1e-12 * steady.max()


@pytest.mark.parametrize("q", [19, 27])
def test_duct_flow_converges_to_series(q):
    """Force-driven duct (x periodic, walls on y and z) run to steady state
    against the series: the error must fall at least 3x as the cross-section
    doubles (second order gives 4x) and be below 0.3 % at 32 x 48."""
    from mrlbm import analytic, geometry
    nu = 0.05
    errs = []
    for k in (1, 2):
        shape = (4, 16 * k, 24 * k)
        g = 1e-5 / (k * k)
        steps = int(25 / (nu * np.pi ** 2 * (1 / (shape[1] - 1) ** 2 + 1 / (shape[2] - 1) ** 2)))
        o = O.Oracle(*shape, q=q, tau=0.5 + 3 * nu, force=(g, 0, 0), walls=(0, 1, 1))
        o.set_node_types(geometry.classify(shape, (0, 1, 1)))
        o.init_uniform()
        o.step(steps)
        assert o.count_nan() == 0
        ux = (o.moments()[1] / 3).reshape(shape, order="F").mean(axis=0) - g  # stored u carries +F/rho
        ser = analytic.duct_start_up(shape[1], shape[2], g, nu)
        errs.append(np.linalg.norm(ux - ser) / np.linalg.norm(ser))
    print("q", q, "steady duct errors", errs, flush=True)
    assert errs[1] < errs[0] / 3 and errs[1] < 3e-3
