"""TTI stiffness from (Vp, Vs, rho, Thomsen epsilon, delta, tilt).

Restates compute_stiffness (src/functions.py:72-109) vectorised over a grid:
VTI stiffness from Thomsen's relations (c11 = c33 (1 + 2 eps),
c13 = sqrt((c33 - c55)^2 + 2 delta c33 (c33 - c55)) - c55), then the 2-D Bond
rotation Cr = M(theta) C M(theta)^T.  Computed in GPa and returned in Pa, as
the reference does (SI = 1e9).
"""
from __future__ import annotations

import numpy as np

SI = 1e9


def compute_stiffness(vp, vs, rho, eps, delta, theta_deg):
    """Returns [C11, C13, C15, C33, C35, C55] (float64, Pa) with the inputs' shape."""
    vp, vs, rho, eps, delta, th = np.broadcast_arrays(*[np.asarray(a, np.float64) for a in
                                                          (vp, vs, rho, eps, delta, theta_deg)])
    c33 = rho * vp ** 2 / SI
    c55 = rho * vs ** 2 / SI
    c11 = c33 * (1.0 + 2.0 * eps)
    c13 = np.sqrt((c33 - c55) ** 2 + 2.0 * delta * c33 * (c33 - c55)) - c55
    zero = np.zeros_like(c33)
    C = np.stack([np.stack([c11, c13, zero], -1),
                  np.stack([c13, c33, zero], -1),
                  np.stack([zero, zero, c55], -1)], -2)
    t = np.radians(th)
    c, s = np.cos(t), np.sin(t)
    s2, c2 = np.sin(2.0 * t), np.cos(2.0 * t)
    M = np.stack([np.stack([c ** 2, s ** 2, s2], -1),
                  np.stack([s ** 2, c ** 2, -s2], -1),
                  np.stack([-0.5 * s2, 0.5 * s2, c2], -1)], -2)
    Cr = np.einsum("...ij,...jk,...lk->...il", M, C, M) * SI
    return [Cr[..., 0, 0], Cr[..., 0, 1], Cr[..., 0, 2], Cr[..., 1, 1], Cr[..., 1, 2], Cr[..., 2, 2]]
