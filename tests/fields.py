"""Shared helpers: field comparison and standard cases (test side)."""
import numpy as np


def rel_l2(a, b, rho0=1.0):
    """Per-field-group relative L2 errors of two (10, N) moment arrays in the
    stored convention:
      drho  the stored density field rho - rho0 (plane 0) itself, so that an
            error in the density deviation is measured against the deviation
            and not against rho ~ 1;
      rho   the full density rho0 + plane 0;
      u     the velocity (planes 1-3);
      m     the second moments (planes 4-9)."""
    out = {}
    d0 = a[0] - b[0]
    out["drho"] = np.linalg.norm(d0) / max(np.linalg.norm(b[0]), 1e-300)
    out["rho"] = np.linalg.norm(d0) / np.linalg.norm(rho0 + b[0])
    du, ub = a[1:4] - b[1:4], b[1:4]
    out["u"] = np.linalg.norm(du) / max(np.linalg.norm(ub), 1e-300)
    dm, mb = a[4:10] - b[4:10], b[4:10]
    out["m"] = np.linalg.norm(dm) / max(np.linalg.norm(mb), 1e-300)
    return out


def tau_from_re(u, n, re):
    return 0.5 + 3.0 * u * n / re
