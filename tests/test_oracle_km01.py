"""CPU checks of the Kormann & Meixner oracle (oracle/km01.py) with the
known answers of the reference's tests/test_ffm_kormann_meixner.py, and of
the drop-in's host-side helpers against it."""

import warnings

import numpy as np
import pytest

from oracle import km01

K = 0.4


@pytest.mark.parametrize("zm, L, pm, pc", [
    (10.0, 1e6, 1.00005, 1.00005),
    (10.0, -100.0, 2.6 ** -0.25, 2.6 ** -0.5),
    (10.0, 100.0, 1.5, 1.5),
])
def test_stability_functions(zm, L, pm, pc):
    assert km01.phi_m(zm, L) == pytest.approx(pm, rel=1e-12)
    assert km01.phi_c(zm, L) == pytest.approx(pc, rel=1e-12)


def test_psi_m():
    assert km01.psi_m(10.0, 1e6) == pytest.approx(5e-5, rel=1e-12)
    assert km01.psi_m(10.0, 100.0) == pytest.approx(0.5, rel=1e-12)
    x = 2.6 ** 0.25
    exp = -2 * np.log(0.5 * (1 + x)) - np.log(0.5 * (1 + x * x)) + 2 * np.arctan(x) - np.pi / 2
    assert km01.psi_m(10.0, -100.0) == pytest.approx(exp, rel=1e-12)


def test_power_law_exponents():
    assert km01.m_param(10.0, 4.0, 0.4, 1e6) == pytest.approx(0.4 * 1.00005 / (K * 4.0), rel=1e-12)
    assert km01.m_param(10.0, 4.0, 0.4, -100.0) == pytest.approx(0.4 * 2.6 ** -0.25 / (K * 4.0), rel=1e-12)
    assert km01.n_param(10.0, 1e6) == pytest.approx(1 / 1.00005, rel=1e-12)
    assert km01.n_param(10.0, -100.0) == pytest.approx(3.4 / 2.6, rel=1e-12)
    assert km01.n_param(10.0, 100.0) == pytest.approx(1 / 1.5, rel=1e-12)


COMMON = dict(zm=10.0, z0=0.1, ws=4.0, ustar=0.4, L=-100.0, sigma_v=0.3, grid_res=5.0, mxy=[0, 0])


def test_footprint_grid_and_sign():
    gx, gy, f = km01.estimate_footprint(**COMMON, grid_domain=[-200, 200, -100, 100], wd=270.0)
    assert gx.shape == gy.shape == f.shape == (40, 80)
    assert gx[0, 0] == -197.5 and gy[0, 0] == 97.5
    assert np.all(f >= 0) and f.max() > 0


def test_footprint_mass_conservation():
    _, _, f = km01.estimate_footprint(zm=10.0, z0=0.1, ws=4.0, ustar=0.4, L=1e6, sigma_v=0.3,
                                      grid_domain=[-500, 3000, -1000, 1000], grid_res=5.0, mxy=[0, 0], wd=270.0)
    assert 0.5 < f.sum() <= 1.05


def test_footprint_closed_form_on_the_axis():
    """Along the wind axis (y = 0) the footprint is scale * x^e1 e^{-Xi/x};
    its maximum sits at x* = Xi / (2 + mu - mr) (d/dx of the log)."""
    p = km01.footprint_params(10.0, 0.1, 4.0, 0.4, 1e6, 0.3)
    xs = np.linspace(1.0, 400.0, 40000)
    prof = xs ** (p["mr"] - 2 - p["mu"]) * np.exp(-p["Xi"] / xs)
    assert xs[np.argmax(prof)] == pytest.approx(p["Xi"] / (2 + p["mu"] - p["mr"]), rel=1e-3)


def test_footprint_rotation_and_negative_u():
    c = dict(zm=10.0, z0=0.1, ws=4.0, ustar=0.4, L=1e6, sigma_v=0.3, grid_domain=[-200, 200, -200, 200],
             grid_res=5.0, mxy=[0, 0])
    a = km01.estimate_footprint(**c, wd=0.0)[2]
    b = km01.estimate_footprint(**c, wd=90.0)[2]
    assert np.unravel_index(a.argmax(), a.shape) != np.unravel_index(b.argmax(), b.shape)
    # wind from the north (0 deg): the footprint lies north of the receptor (upwind)
    gy = km01.estimate_footprint(**c, wd=0.0)[1]
    assert gy[np.unravel_index(a.argmax(), a.shape)] > 0
    with warnings.catch_warnings(record=True) as w:
        warnings.simplefilter("always")
        f = km01.estimate_footprint(zm=10.0, z0=100.0, ws=4.0, ustar=0.4, L=-100.0, sigma_v=0.3,
                                    grid_domain=[-200, 200, -100, 100], grid_res=10.0, mxy=[0, 0], wd=270.0)[2]
    assert np.all(f == 0.0) and len(w) == 1


def test_estimate_z0():
    with pytest.raises(RuntimeError, match="same length"):
        km01.estimate_z0(np.array([10.0, 10.0]), np.array([4.0]), np.array([270.0, 270.0]),
                         np.array([0.4, 0.4]), np.array([-100.0, -100.0]))
    z = km01.estimate_z0(np.array([10.0]), np.array([4.0]), np.array([270.0]), np.array([0.4]),
                         np.array([-100.0]), half_wd_win=0)
    assert z[0] == pytest.approx(10.0 * np.exp(km01.psi_m(10.0, -100.0) - K * 4.0 / 0.4), rel=1e-12)
    rng = np.random.default_rng(0)
    n = 400
    zm, ws, ustar = np.full(n, 10.0), rng.uniform(2, 8, n), rng.uniform(0.2, 0.6, n)
    wd, L = rng.uniform(0, 360, n), rng.uniform(-500, -50, n)
    sm = km01.estimate_z0(zm, ws, wd, ustar, L)
    raw = km01.estimate_z0(zm, ws, wd, ustar, L, half_wd_win=0)
    # each smoothed value is the median of raw values within +-22 degrees
    i = 7
    d = (wd - np.floor(wd[i]) + 180) % 360 - 180
    win = (d >= -22) & (d < 23)
    assert sm[i] == pytest.approx(np.nanmedian(raw[win]), rel=1e-12)


def test_dropin_host_helpers_match_oracle():
    from bldfm_amd import ffm_kormann_meixner as dk

    zm, L = np.array([10.0, 10.0, 5.0]), np.array([-100.0, 100.0, 1e6])
    assert np.allclose(dk._phiM(zm, L), km01.phi_m(zm, L), rtol=1e-15)
    assert np.allclose(dk._phiC(zm, L), km01.phi_c(zm, L), rtol=1e-15)
    assert np.allclose(dk._psiM(zm, L), km01.psi_m(zm, L), rtol=1e-15)
    assert np.allclose(dk._nParam(zm, L), km01.n_param(zm, L), rtol=1e-15)
    assert np.allclose(dk._mParam(zm, [4.0] * 3, [0.4] * 3, L), km01.m_param(zm, 4.0, 0.4, L), rtol=1e-15)
    rng = np.random.default_rng(1)
    n = 200
    args = (np.full(n, 12.0), rng.uniform(2, 8, n), rng.uniform(0, 360, n), rng.uniform(0.2, 0.6, n),
            rng.uniform(-500, -50, n))
    assert np.allclose(dk.estimateZ0(*args), km01.estimate_z0(*args), rtol=1e-15, equal_nan=True)
