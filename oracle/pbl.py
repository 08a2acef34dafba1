"""Oracle: wind components, synthetic sources and PBL profiles (numpy, FP64).

Test infrastructure only (see ``oracle/__init__.py``).
"""

import numpy as np

KARMAN = 0.4  # von Karman constant, src/bldfm/pbl_model.py:60

# planted mistakes of vertical_profiles, each of which a pinning test must catch
VALID_MUTATIONS = (None, "mostm_swap", "oaahoc_cm", "oaahoc_k", "stretch_ignored")


def compute_wind_fields(u_rot, wind_dir):
    """Met direction (deg, blowing FROM, clockwise from north) -> (u, v).

    Follows src/bldfm/utils.py:23-27.
    """
    theta = np.deg2rad(wind_dir)
    return -u_rot * np.sin(theta), -u_rot * np.cos(theta)


def ideal_source(nxy, domain, src_loc=None, shape="diamond"):
    """Synthetic surface-flux field; follows src/bldfm/utils.py:30-77.

    Note the source grid is ``linspace(..., endpoint=True)``
    (src/bldfm/utils.py:55-56), unlike the solver's output grid.
    """
    nx, ny = nxy
    xmx, ymx = domain
    dx = xmx / nx
    if src_loc is None:
        src_loc = (xmx / 2, ymx / 2)
    xs, ys = src_loc
    X, Y = np.meshgrid(np.linspace(0.0, xmx, nx), np.linspace(0.0, ymx, ny))
    q0 = np.zeros([ny, nx])
    if shape == "diamond":
        q0 = np.where(np.abs(X - xs) + np.abs(Y - ys) < xmx / 12, 1.0, 0.0)
    if shape == "circle":
        q0 = np.where(np.sqrt((X - xs) ** 2 + (Y - ys) ** 2) < xmx / 12, 1.0, 0.0)
    if shape == "point":
        sig = 4.0 * dx
        rsq = (X - xs) ** 2 + (Y - ys) ** 2
        q0 = np.exp(-rsq / 2.0 / sig**2) / sig / np.sqrt(2.0 * np.pi)
    return q0


def psi(x):
    """MOST momentum stability correction; src/bldfm/pbl_model.py:207-230.

    Stable (x > 0): 5x.  Otherwise Businger-Dyer with
    xi = Re((1 - 16x)^(1/4)) evaluated as a complex power.
    """
    x = np.asarray(x, dtype=np.float64)
    stable = x > 0.0
    xi = np.where(stable, np.nan, np.power(1.0 - 16.0 * x, 0.25, dtype=complex).real)
    unstable = (
        -2.0 * np.log(0.5 * (1.0 + xi))
        - np.log(0.5 * (1.0 + xi**2))
        + 2.0 * np.arctan(xi)
        - 0.5 * np.pi
    )
    return np.where(stable, 5.0 * x, unstable)


def phi(x):
    """MOST diffusivity stability function; src/bldfm/pbl_model.py:233-250."""
    x = np.asarray(x, dtype=np.float64)
    return np.where(
        x > 0.0, 1.0 + 5.0 * x, np.power(1.0 - 16.0 * x, -0.5, dtype=complex).real
    )


def stretched_grid(n, zm, z0, domain_height=None, stretch=None, _mutation=None):
    """Log-stretched vertical grid, src/bldfm/pbl_model.py:105-129.

    zeta is uniform with step zm/n from 0 past zeta(zmx); the mapping
    z = -h ln(-(zeta - aa)/bb) sends zeta=0 to z0 and zeta=zm to zm.  The
    length is numpy ``arange``'s ceil((stop - start)/step) and the last point
    may overshoot ``domain_height``.
    """
    h = 2.0 * zm if stretch is None or _mutation == "stretch_ignored" else stretch
    zmx = 2.0 * zm if domain_height is None else domain_height
    bb = zm / (np.exp(-z0 / h) - np.exp(-zm / h))
    aa = bb * np.exp(-z0 / h)
    zetamx = aa - bb * np.exp(-zmx / h)
    dzeta = zm / n
    zeta = np.arange(0.0, np.squeeze(zetamx).item() + dzeta, dzeta)
    return -h * np.log(-(zeta - aa) / bb)


def vertical_profiles(
    n,
    meas_height,
    wind,
    ustar=None,
    z0=None,
    mol=1e9,
    prsc=1.0,
    closure="MOST",
    domain_height=None,
    stretch=None,
    z0_min=0.001,
    z0_max=2.0,
    tke=None,
    _mutation=None,
):
    """Grid and profiles (u, v, Kx, Ky, Kz); src/bldfm/pbl_model.py:8-204.

    Exactly one of ``ustar``/``z0`` for the MOST-type closures: the other is
    derived from |U(zm)| by the log law (src/bldfm/pbl_model.py:65-77).
    ``z0_min``/``z0_max`` are accepted and unused, as in the reference.
    ``_mutation`` plants a known mistake for the tests that pin this oracle
    (one of ``VALID_MUTATIONS``); None is the restatement itself.
    """
    assert _mutation in VALID_MUTATIONS, _mutation
    zm = meas_height
    um, vm = wind
    absum = np.sqrt(um**2 + vm**2)
    if closure in ("CONSTANT", "MOST", "MOSTM"):
        if z0 is None:
            z0 = zm * np.exp(-KARMAN * absum / ustar + psi(zm / mol))
        elif ustar is None:
            ustar = absum * KARMAN / (np.log(zm / z0) + psi(zm / mol))
        else:
            raise ValueError("Either z0 or ustar must be provided.")
    elif closure == "OAAHOC":
        cl, cm = 0.845, 0.0856
        if tke is None:
            tke = 1.0
        tke = np.array(tke)[..., np.newaxis]
        z0 = zm * np.exp(-cm * cl * absum * np.sqrt(tke) / ustar**2)
    else:
        raise ValueError(
            f"Invalid closure type: {closure}. "
            "Supported closures are 'MOST', 'CONSTANT', and 'OAAHOC'."
        )

    z = stretched_grid(n, zm, z0, domain_height, stretch, _mutation)
    nlev = z.shape[-1]
    if nlev == 0:
        # the reference's stats logging evaluates max(u) on the empty profiles
        # (src/bldfm/pbl_model.py:184-194)
        raise ValueError("max() arg is an empty sequence")

    if closure == "CONSTANT":
        K = KARMAN * ustar * zm / prsc * np.ones(nlev)
        return z, (um * np.ones(nlev), vm * np.ones(nlev), K, K, K)
    if closure in ("MOST", "MOSTM"):
        absu = ustar / KARMAN * (np.log(z / z0) + psi(z / mol))
        u = um / absum * absu
        v = vm / absum * absu
        K = KARMAN * ustar * z / phi(z / mol) / prsc
        if closure == "MOST":
            return z, (u, v, K, K, K)
        if _mutation == "mostm_swap":  # planted: along- and cross-wind swapped
            return z, (u, v, K * u**2 / (u**2 + v**2), K * v**2 / (u**2 + v**2), K)
        return z, (u, v, K * v**2 / (u**2 + v**2), K * u**2 / (u**2 + v**2), K)
    # OAAHOC (Schumann-Lilly one-and-a-half order closure)
    cl, cm, ch = 0.845, 0.0856, 0.204
    if _mutation == "oaahoc_cm":  # planted: the wind law uses c_h for c_m
        cm = ch
    absu = ustar**2 / cm / cl / np.sqrt(tke) * np.log(z / z0)
    K = ch * cl * z * np.sqrt(tke)
    if _mutation == "oaahoc_k":  # planted: diffusivity without sqrt(tke)
        K = ch * cl * z * tke
    return z, (um / absum * absu, vm / absum * absu, K, K, K)
