"""Scalar prologue of ``vertical_profiles``, evaluated with the host's numpy.

src/bldfm/pbl_model.py:58-120 derives the roughness length z0 (or the
friction velocity u*) and the stretched-grid constants aa, bb and zeta_max
with a handful of scalar numpy exp/log calls.  numpy's exp on AVX-512 hosts is
its own table method and is not correctly rounded (about 4.5% of arguments
come out one ulp low), and a one-ulp change of z0 moves a 512^2 unstable
footprint by about 1e-10.  The drop-in therefore evaluates these few scalars
with the reference's own numpy expressions, in the reference's order and
with the caller's argument types, and hands them to the library
(``bldfm_met.has_prologue``); the per-level profile arrays are evaluated on
the GPU.
"""

import numpy as np

KARMAN = 0.4  # src/bldfm/pbl_model.py:60
OAAHOC_CL, OAAHOC_CM = 0.845, 0.0856  # src/bldfm/pbl_model.py:84-85


def psi(x):
    """Businger-Dyer psi(z/L) as numpy evaluates src/bldfm/pbl_model.py:222-230."""
    with np.errstate(invalid="ignore"):
        xi = np.where(x > 0.0, np.nan, np.power(1.0 - 16.0 * x, 0.25, dtype=complex).real)
        return np.where(
            x > 0.0,
            5.0 * x,
            -2.0 * np.log(0.5 * (1.0 + xi)) - np.log(0.5 * (1.0 + xi**2)) + 2.0 * np.arctan(xi) - 0.5 * np.pi,
        )


def prologue(meas_height, wind, ustar, z0, mol, closure, domain_height, stretch, tke):
    """(z0, ustar, aa, bb, zeta_max) of src/bldfm/pbl_model.py:58-120.

    Raises what the reference raises there: ValueError when both or neither
    of ustar/z0 are usable for a MOST-type closure, ValueError from
    ``.item()`` for an OAAHOC ``tke`` array of more than one value.  The
    closure name is validated by the caller.
    """
    zm, (um, vm) = meas_height, wind
    kap = KARMAN
    absum = np.sqrt(um**2 + vm**2)
    if closure in ("CONSTANT", "MOST", "MOSTM"):
        if z0 is None:
            z0 = zm * np.exp(-kap * absum / ustar + psi(zm / mol))
        elif ustar is None:
            ustar = absum * kap / (np.log(zm / z0) + psi(zm / mol))
        else:
            raise ValueError("Either z0 or ustar must be provided.")
    else:  # OAAHOC
        tke = np.array(1.0 if tke is None else tke)[..., np.newaxis]
        z0 = zm * np.exp(-OAAHOC_CM * OAAHOC_CL * absum * np.sqrt(tke) / ustar**2)
    h = 2.0 * meas_height if stretch is None else stretch
    zmx = 2.0 * meas_height if domain_height is None else domain_height
    with np.errstate(divide="ignore", invalid="ignore", over="ignore"):
        bb = zm / (np.exp(-z0 / h) - np.exp(-zm / h))
        aa = bb * np.exp(-z0 / h)
        zetamx = aa - bb * np.exp(-zmx / h)
    zetamx = np.squeeze(zetamx).item()
    return (float(np.squeeze(z0)), float(np.squeeze(ustar)), float(np.squeeze(aa)), float(np.squeeze(bb)),
            float(zetamx))


def prologue_many(meas_height, um, vm, ustar, z0, mol, closure, domain_height, stretch):
    """``prologue`` for many MOST-type instances at once (arrays of equal
    length; ``ustar`` or ``z0`` None for all of them, ``domain_height`` and
    ``stretch`` arrays or None).  numpy evaluates each element as it
    evaluates a scalar, so the values are those ``prologue`` gives one
    instance at a time.  Returns arrays (z0, ustar, aa, bb, zeta_max)."""
    zm = np.asarray(meas_height, dtype=np.float64)
    um, vm = np.asarray(um, dtype=np.float64), np.asarray(vm, dtype=np.float64)
    mol = np.asarray(mol, dtype=np.float64)
    kap = KARMAN
    absum = np.sqrt(um**2 + vm**2)
    if closure not in ("CONSTANT", "MOST", "MOSTM") or (ustar is None) == (z0 is None):
        raise ValueError("prologue_many serves MOST-type closures with exactly one of ustar/z0")
    with np.errstate(divide="ignore", invalid="ignore", over="ignore"):
        if z0 is None:
            ustar = np.asarray(ustar, dtype=np.float64)
            z0 = zm * np.exp(-kap * absum / ustar + psi(zm / mol))
        else:
            z0 = np.asarray(z0, dtype=np.float64)
            ustar = absum * kap / (np.log(zm / z0) + psi(zm / mol))
        h = 2.0 * zm if stretch is None else np.asarray(stretch, dtype=np.float64)
        zmx = 2.0 * zm if domain_height is None else np.asarray(domain_height, dtype=np.float64)
        bb = zm / (np.exp(-z0 / h) - np.exp(-zm / h))
        aa = bb * np.exp(-z0 / h)
        zetamx = aa - bb * np.exp(-zmx / h)
    return z0, ustar, aa, bb, zetamx
