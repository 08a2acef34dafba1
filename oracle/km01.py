"""Oracle: Kormann & Meixner (2001) analytic footprint (test infrastructure only).

Restates src/bldfm/ffm_kormann_meixner.py:16-474 in numpy (the reference
module is read, never imported): the stability functions phi_m, phi_c,
psi_m (Eqs. 33-35), the power-law exponents m, n (Eq. 36), the roughness
estimate of Eq. (31) with its wind-direction median smoothing, and the
footprint phi(x, y) of Eqs. (9) and (21) on a regular grid, optionally
rotated into a wind-direction frame.  Pinned by the known answers of
tests/test_ffm_kormann_meixner.py (restated in tests/test_oracle_km01.py)
and by the model's own invariants (non-negativity, mass <= 1, mass
conservation on a large domain).
"""

import warnings

import numpy as np
from scipy import special

KARMAN = 0.4


def phi_m(zm, L):
    zm, L = np.asarray(zm, dtype=float), np.asarray(L, dtype=float)
    return np.where(L < 0, np.abs(1 - 16 * zm / L) ** -0.25, 1 + 5 * zm / L)


def phi_c(zm, L):
    zm, L = np.asarray(zm, dtype=float), np.asarray(L, dtype=float)
    return np.where(L < 0, np.abs(1 - 16 * zm / L) ** -0.5, 1 + 5 * zm / L)


def psi_m(zm, L):
    zm, L = np.asarray(zm, dtype=float), np.asarray(L, dtype=float)
    x = np.abs(1 - 16 * zm / L) ** 0.25
    unstable = -2 * np.log(0.5 * (1 + x)) - np.log(0.5 * (1 + x * x)) + 2 * np.arctan(x) - 0.5 * np.pi
    return np.where(L < 0, unstable, 5 * zm / L)


def m_param(zm, ws, ustar, L):
    return ustar * phi_m(zm, L) / (KARMAN * np.asarray(ws, dtype=float))


def n_param(zm, L):
    zm, L = np.asarray(zm, dtype=float), np.asarray(L, dtype=float)
    return np.where(L < 0, (1 - 24 * zm / L) / (1 - 16 * zm / L), 1 / (1 + 5 * zm / L))


def estimate_z0(zm, ws, wd, ustar, L, half_wd_win=22):
    """Eq. (31) solved for z0, median-smoothed over +-half_wd_win degrees of
    wind direction (src/bldfm/ffm_kormann_meixner.py:16-107)."""
    n = len(zm)
    if not (n == len(ws) == len(wd) == len(ustar) == len(L)):
        raise RuntimeError("Input parameters must be of the same length!")
    z0 = zm * np.exp(psi_m(zm, L) - KARMAN * ws / ustar)
    z0[z0 > 1000] = np.nan
    if half_wd_win < 1:
        return z0
    out = np.full_like(z0, np.nan)
    for kk in range(360):
        w = wd.copy()
        if kk < 90:
            w[wd > 270] = wd[wd > 270] - 360
        elif kk > 270:
            w[wd < 90] = wd[wd < 90] + 360
        sel = (wd >= kk) & (wd < kk + 1)
        win = (w >= kk - half_wd_win) & (w < kk + 1 + half_wd_win)
        out[sel] = np.nanmedian(z0[win]) if sel.any() else np.nan
    return out


def footprint_params(zm, z0, ws, ustar, L, sigma_v):
    """Scalars of the footprint formula, or None when U < 0."""
    pm, pc, ps = phi_m(zm, L), phi_c(zm, L), psi_m(zm, L)
    m = ustar * pm / (KARMAN * ws)
    n = float(n_param(zm, L))
    kappa = KARMAN * zm * ustar / (pc * zm**n)
    U = ustar * (np.log(zm / z0) + ps) / (KARMAN * zm**m)
    if U < 0:
        return None
    r = 2 + m - n
    mu = (1 + m) / r
    Xi = U * zm**r / (r**2 * kappa)
    A = U / (special.gamma(1 / r) * sigma_v) * (kappa * r**2 / U) ** (m / r)
    return dict(m=m, n=n, r=r, mu=mu, Xi=Xi, A=A, gmm=special.gamma(mu), mr=m / r,
                num=(1 / np.sqrt(2 * np.pi)) * Xi**mu, U=U)


def estimate_footprint(zm, z0, ws, ustar, L, sigma_v, grid_domain, grid_res, mxy, wd=None):
    """(grid_x, grid_y, footprint) of one measurement (src/bldfm/ffm_kormann_meixner.py:110-330)."""
    xmin, xmax, ymin, ymax = grid_domain
    gx, gy = np.meshgrid(np.arange(xmin + 0.5 * grid_res, xmax, grid_res),
                         np.arange(ymax - 0.5 * grid_res, ymin, -grid_res))
    f = np.zeros_like(gx)
    p = footprint_params(zm, z0, ws, ustar, L, sigma_v)
    if p is None:
        warnings.warn("U in Eq. (11) of Kormann & Meixner, 2001 is estimated negative; returning an empty footprint")
        return gx, gy, f
    x, y = gx - mxy[0], gy - mxy[1]
    if wd is not None:
        rho = np.sqrt(x * x + y * y)
        th = np.arctan2(y, x) + np.deg2rad(wd) - 0.5 * np.pi
        x, y = rho * np.cos(th), rho * np.sin(th)
    up = x > 0
    xu, yu = x[up], y[up]
    f[up] = (grid_res**2 * p["num"] * p["A"] * xu ** (p["mr"] - 2 - p["mu"])
             * np.exp(-p["Xi"] / xu - 0.5 * (p["gmm"] * yu * p["A"] * xu ** (p["mr"] - 1)) ** 2))
    return gx, gy, f
