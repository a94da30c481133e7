"""Closed-form flows used to check the solver.

duct_start_up: body-force-driven flow in a rectangular duct started from
rest (walls at y = 0, W and z = 0, H; force g along x; kinematic
viscosity nu), the eigenfunction series of du/dt = nu (u_yy + u_zz) + g:

    u(y, z, t) = sum over odd m, n of
        16 g / (pi^2 m n nu lambda_mn) (1 - exp(-nu lambda_mn t))
        sin(m pi y / W) sin(n pi z / H),
    lambda_mn = pi^2 (m^2 / W^2 + n^2 / H^2).

t = None gives the steady (fully developed) profile. With wet-node walls on
lattice nodes 0 and N-1 the walls are at distance W = NY - 1, H = NZ - 1.
"""
import numpy as np


def duct_start_up(ny, nz, g, nu, t=None, terms=399):
    """Velocity u_x on the (ny, nz) node grid, series truncated at odd
    m, n <= terms (the neglected tail is below 1e-6 of the centre value
    for terms >= 199)."""
    W, H = ny - 1, nz - 1
    m = np.arange(1, terms + 1, 2, dtype=np.float64)
    n = m.copy()
    lam = np.pi ** 2 * (m[:, None] ** 2 / W ** 2 + n[None, :] ** 2 / H ** 2)
    A = 16.0 * g / (np.pi ** 2 * m[:, None] * n[None, :] * nu * lam)
    if t is not None:
        A = A * (1.0 - np.exp(-nu * lam * t))
    sy = np.sin(np.pi * np.outer(np.arange(ny), m) / W)
    sz = np.sin(np.pi * np.outer(np.arange(nz), n) / H)
    return sy @ A @ sz.T


def channel_steady(ny, g, nu):
    """Plane Poiseuille profile between wet-node walls at y = 0 and ny - 1."""
    y = np.arange(ny, dtype=np.float64)
    return g * y * (ny - 1 - y) / (2.0 * nu)
