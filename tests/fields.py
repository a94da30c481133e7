"""Shared helpers: field comparison and standard cases (test side)."""
import numpy as np


def rel_l2(a, b, rho0=1.0):
    """Per-field-group relative L2 errors of two (10, N) moment arrays in the
    stored convention (a: under test, b: reference):
      drho  the stored density deviation rho - rho0 (plane 0), on its own
            scale or the velocity's, whichever is larger:
            ||d drho|| / max(||drho_b||, ||u_b||). The deviation is summed
            from populations of the size of the velocity (g_i ~ w_i 3 c_i.u),
            so FP32 round-off in it scales with |u|, not with |drho|; where
            |drho| ~ Ma^2 falls below |u| ~ Ma (slow flows, and flows with
            a uniform density such as Poiseuille, where drho_b is ~0), the
            velocity sets the scale. Where |drho| >= |u| this is the plain
            relative error of the deviation;
      rho   the full density rho0 + plane 0;
      u     the velocity (planes 1-3);
      m     the second moments (planes 4-9).
    drho_rel(a, b) gives the plain relative error of the deviation (reported
    beside the gated numbers)."""
    out = {}
    d0 = a[0] - b[0]
    u_scale = np.linalg.norm(b[1:4]) / 3.0  # the stored velocity is 3u
    out["drho"] = np.linalg.norm(d0) / max(np.linalg.norm(b[0]), u_scale, 1e-300)
    out["rho"] = np.linalg.norm(d0) / np.linalg.norm(rho0 + b[0])
    du, ub = a[1:4] - b[1:4], b[1:4]
    out["u"] = np.linalg.norm(du) / max(np.linalg.norm(ub), 1e-300)
    dm, mb = a[4:10] - b[4:10], b[4:10]
    out["m"] = np.linalg.norm(dm) / max(np.linalg.norm(mb), 1e-300)
    return out


def drho_rel(a, b):
    """Plain relative L2 error of the stored density deviation (plane 0)."""
    return float(np.linalg.norm(a[0] - b[0]) / max(np.linalg.norm(b[0]), 1e-300))


def tau_from_re(u, n, re):
    return 0.5 + 3.0 * u * n / re
