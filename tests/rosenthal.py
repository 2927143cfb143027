"""Rosenthal's quasi-steady moving point source over a half-space: the
analytic melt pool that pins the melt-pool measure (tests/test_oracle_pins.py)
and that the single-track run's pool is compared with
(tests/test_gpu_full_runs.py)."""
import numpy as np


ROSENTHAL = {"k": 0.02, "rho_c": 4e-3, "v": 1000.0, "eta_P": 70.0, "T0": 300.0,
             "Tl": 1609.0}  # W/mm K, J/mm^3 K, mm/s, W, K: the single-track deck's values


def rosenthal_pool():
    """Melt pool of Rosenthal's quasi-steady point source moving along +x over
    a half-space, T = T0 + eta P / (2 pi k R) exp(-v (R + xi) / (2 alpha)),
    xi = x - x_source, R = |(xi, y, z)|: the isotherm T = Tl is axisymmetric
    about the track, so depth = half the width = the largest radius r(xi) of
    the isotherm, and the length is the isotherm's span on the axis.  Returns
    (T(xi, r), depth, width, length), the sizes from 1-D root finding."""
    from scipy.optimize import brentq, minimize_scalar
    c = ROSENTHAL
    a = c["k"] / c["rho_c"]

    def T(xi, r):
        R = np.sqrt(xi * xi + r * r)
        return c["T0"] + c["eta_P"] / (2 * np.pi * c["k"] * R) * np.exp(-c["v"] * (R + xi) / (2 * a))

    Tl = c["Tl"]
    rear = -brentq(lambda s: T(-s, 0.0) - Tl, 1e-6, 10.0)
    front = brentq(lambda s: T(s, 0.0) - Tl, 1e-6, 10.0)

    def radius(xi):
        return brentq(lambda r: T(xi, r) - Tl, 1e-12, 10.0) if T(xi, 1e-12) > Tl else 0.0

    best = minimize_scalar(lambda xi: -radius(xi), bounds=(rear, front), method="bounded",
                           options={"xatol": 1e-10})
    r = -best.fun
    return T, r, 2 * r, front - rear
