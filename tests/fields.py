"""Shared helpers: field comparison and standard cases (test side)."""
import numpy as np


def rel_l2(a, b):
    """Per-field-group relative L2 errors of two (10, N) moment arrays in the
    stored convention: density (rho0 + m0), velocity (3 components) and
    second moments (6 components)."""
    out = {}
    rho_a, rho_b = 1.0 + a[0], 1.0 + b[0]
    out["rho"] = np.linalg.norm(rho_a - rho_b) / np.linalg.norm(rho_b)
    du, ub = a[1:4] - b[1:4], b[1:4]
    out["u"] = np.linalg.norm(du) / max(np.linalg.norm(ub), 1e-300)
    dm, mb = a[4:10] - b[4:10], b[4:10]
    out["m"] = np.linalg.norm(dm) / max(np.linalg.norm(mb), 1e-300)
    return out


def tau_from_re(u, n, re):
    return 0.5 + 3.0 * u * n / re
