"""The reference's own physical-property tests, run against the GPU drop-in.

Restates the checks of the reference's tests/test_integration.py:278-500
(footprint finite values, positivity with CONSTANT and MOST closures,
footprint mass fraction on a larger window, peak upwind of the receptor for
southward and westward wind, and the shifted dispersion path) with the same
inputs and bars, calling bldfm_amd's steady_state_transport_solver and
vertical_profiles exactly where the reference calls its own.  The solver's
default precision ("single") is kept, as in those tests.
"""

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _quick_footprint_solve(closure="MOST", wind=(0.0, -5.0), ustar=0.4, z0=None, mol=1e9, nxy=(64, 256),
                           domain=(50.0, 200.0), modes=(64, 128), meas_pt=(25.0, 10.0), meas_height=10.0, nz=16,
                           halo=None):
    """Footprint solve returning (grid, conc, flx, dx, dy) on the reference
    helper's inputs (tests/test_integration.py:194-237 of the reference)."""
    from bldfm_amd.pbl_model import vertical_profiles
    from bldfm_amd.solver import steady_state_transport_solver

    nx, ny = nxy
    xmx, ymx = domain
    kw = dict(n=nz, meas_height=meas_height, wind=wind, closure=closure, mol=mol)
    if z0 is not None:
        kw["z0"] = z0
    else:
        kw["ustar"] = ustar
    z, profs = vertical_profiles(**kw)
    grid, conc, flx = steady_state_transport_solver(np.zeros((ny, nx)), z, profs, domain, nz, modes=modes,
                                                    meas_pt=meas_pt, footprint=True, halo=halo)
    return grid, conc, flx, xmx / nx, ymx / ny


def test_footprint_finite_values():
    _, conc, flx, _, _ = _quick_footprint_solve()
    assert np.isfinite(flx).all() and np.isfinite(conc).all()


def test_footprint_positivity_constant():
    _, _, flx, _, _ = _quick_footprint_solve(closure="CONSTANT")
    print(f"  CONSTANT min {flx.min():.3e}")
    assert np.all(flx >= -1e-15)


def test_footprint_positivity_most():
    _, _, flx, _, _ = _quick_footprint_solve(closure="MOST", z0=0.1, mol=1e9)
    print(f"  MOST min {flx.min():.3e}")
    assert np.all(flx >= -1e-4)


@pytest.mark.parametrize("closure,z0", [("CONSTANT", None), ("MOST", 0.5)])
def test_footprint_mass_fraction(closure, z0):
    _, _, flx, dx, dy = _quick_footprint_solve(closure=closure, z0=z0, mol=1e9, nxy=(64, 512), domain=(50.0, 400.0),
                                               modes=(64, 256), halo=400.0)
    integral = float(np.sum(flx) * dx * dy)
    print(f"  {closure} integral {integral:.4f}")
    assert 0.25 < integral <= 1.05


def test_footprint_peak_upwind_southward():
    meas_pt = (25.0, 10.0)
    (X, Y, _), _, flx, _, _ = _quick_footprint_solve(closure="MOST", wind=(0.0, -5.0), z0=0.1, mol=1e9,
                                                     meas_pt=meas_pt)
    peak = np.unravel_index(np.argmax(flx), flx.shape)
    assert float(Y[peak]) > meas_pt[1]


def test_footprint_peak_upwind_westward():
    meas_pt = (25.0, 25.0)
    (X, Y, _), _, flx, _, _ = _quick_footprint_solve(closure="MOST", wind=(-5.0, 0.0), z0=0.1, mol=1e9,
                                                     nxy=(256, 64), domain=(200.0, 50.0), modes=(128, 64),
                                                     meas_pt=meas_pt)
    peak = np.unravel_index(np.argmax(flx), flx.shape)
    assert float(X[peak]) > meas_pt[0]


def test_solver_non_footprint_shift():
    """Dispersion mode with a non-zero measurement point (the shift path),
    on the reference's _quick_solve inputs (tests/test_integration.py:165-191)."""
    from bldfm_amd.pbl_model import vertical_profiles
    from bldfm_amd.solver import steady_state_transport_solver
    from bldfm_amd.utils import ideal_source

    nxy, nz, domain = (128, 64), 16, (500.0, 250.0)
    srf_flx = ideal_source(nxy, domain, (250.0, 125.0), shape="point")
    z, profs = vertical_profiles(nz, 10.0, (5.0, 0.0), 0.4, closure="MOST")
    _, conc, flx = steady_state_transport_solver(srf_flx, z, profs, domain, nz - 1, modes=(128, 64), footprint=False,
                                                 precision="double", meas_pt=(25.0, 12.5))
    assert conc.shape == flx.shape and np.isfinite(conc).all()
