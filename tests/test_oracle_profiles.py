"""CPU pins of the oracle's vertical_profiles branches beyond MOST/CONSTANT.

The MOSTM and OAAHOC closures and the ``stretch``/``domain_height`` grid
arguments (src/bldfm/pbl_model.py:81-129,151-176) are checked against what
the mathematics fixes, independently of the oracle's own formulas:

* MOSTM: Kx + Ky = K; wind along x gives Kx = 0, Ky = K, wind along y gives
  Ky = 0, Kx = K (no diffusion along the mean wind);
* OAAHOC: the grid starts at the roughness length the closure defines,
  z0 = zm exp(-c_m c_l |U| sqrt(tke) / u*^2), the wind speed at zm comes back
  as the input |U|, |u| is proportional to ln(z/z0), and K/z is the constant
  c_h c_l sqrt(tke) with Kx = Ky = Kz;
* any stretch h: z[0] = z0, z[n] = zm, the levels are equally spaced in
  e^{-z/h}, and the last level is the first at or above domain_height.

Each planted mistake in the oracle (``_mutation``) must fail one of these
checks.  The reference's own tests/test_pbl_model.py checks are restated at
the end with its inputs.
"""

import numpy as np
import pytest

from oracle import pbl as opbl

CL, CM, CH = 0.845, 0.0856, 0.204  # Schumann-Lilly constants, src/bldfm/pbl_model.py:84-86


def check_mostm(vp):
    for wind, mol in (((3.0, 1.0), -100.0), ((-2.0, 4.5), 60.0), ((0.3, -5.0), 1e9)):
        z, (u, v, Kx, Ky, Kz) = vp(12, 8.0, wind, ustar=0.35, mol=mol, closure="MOSTM")
        np.testing.assert_allclose(Kx + Ky, Kz, rtol=4e-16 * 4)
        assert np.all(Kx > 0) and np.all(Ky > 0)
    # wind along x: no along-wind (x) diffusion; along y: none in y
    z, (u, v, Kx, Ky, Kz) = vp(12, 8.0, (4.0, 0.0), ustar=0.35, mol=-40.0, closure="MOSTM")
    assert np.all(Kx == 0.0) and np.all(v == 0.0)
    np.testing.assert_allclose(Ky, Kz, rtol=1e-15)
    z, (u, v, Kx, Ky, Kz) = vp(12, 8.0, (0.0, -4.0), ustar=0.35, mol=-40.0, closure="MOSTM")
    assert np.all(Ky == 0.0) and np.all(u == 0.0)
    np.testing.assert_allclose(Kx, Kz, rtol=1e-15)
    # MOSTM's K is MOST's K
    z, (u, v, Kx, Ky, Kz) = vp(12, 8.0, (3.0, 1.0), ustar=0.35, mol=-40.0, closure="MOSTM")
    zm_, (_, _, _, _, Km) = vp(12, 8.0, (3.0, 1.0), ustar=0.35, mol=-40.0, closure="MOST")
    assert np.array_equal(Kz, Km) and np.array_equal(z, zm_)


def check_oaahoc(vp):
    for wind, ustar, tke, n, zm in (((3.0, 1.0), 0.4, 2.0, 10, 5.0), ((-1.5, 6.0), 0.6, 0.7, 16, 12.0)):
        z, (u, v, Kx, Ky, Kz) = vp(n, zm, wind, ustar=ustar, closure="OAAHOC", tke=tke)
        speed = np.hypot(*wind)
        z0 = zm * np.exp(-CM * CL * speed * np.sqrt(tke) / ustar**2)
        np.testing.assert_allclose(z[0], z0, rtol=1e-13)
        np.testing.assert_allclose(z[n], zm, rtol=1e-13)
        np.testing.assert_allclose(np.hypot(u[n], v[n]), speed, rtol=1e-12)  # |U(zm)| = |U|
        ratio = np.hypot(u[1:], v[1:]) / np.log(z[1:] / z0)
        np.testing.assert_allclose(ratio, ratio[0], rtol=1e-11)  # |u| = a ln(z/z0)
        np.testing.assert_allclose(Kz / z, CH * CL * np.sqrt(tke), rtol=1e-14)
        assert Kx is Ky is Kz
        # direction of the wind is the input's at every level
        np.testing.assert_allclose(u * wind[1], v * wind[0], rtol=1e-12, atol=1e-14)


def check_stretch(vp):
    for stretch, dh in ((5.0, None), (15.0, 50.0), (40.0, 33.0)):
        n, zm, z0 = 10, 5.0, 0.07
        z, _ = vp(n, zm, (3.0, 1.0), z0=z0, mol=-80.0, stretch=stretch, domain_height=dh)
        np.testing.assert_allclose(z[0], z0, rtol=1e-12)
        np.testing.assert_allclose(z[n], zm, rtol=1e-12)
        e = np.exp(-z / stretch)  # zeta = aa - bb e^{-z/h} is uniform
        np.testing.assert_allclose(np.diff(e), np.diff(e)[0], rtol=1e-9)
        top = 2 * zm if dh is None else dh
        assert z[-2] < top * (1 - 1e-12) and z[-1] >= top * (1 - 1e-12)


CHECKS = {"mostm_swap": check_mostm, "oaahoc_cm": check_oaahoc, "oaahoc_k": check_oaahoc,
          "stretch_ignored": check_stretch}


@pytest.mark.parametrize("check", [check_mostm, check_oaahoc, check_stretch])
def test_oracle_branch_pins(check):
    check(opbl.vertical_profiles)


@pytest.mark.parametrize("mutation", [m for m in opbl.VALID_MUTATIONS if m])
def test_planted_profile_mistakes_are_caught(mutation):
    def mutated(*a, **kw):
        return opbl.vertical_profiles(*a, _mutation=mutation, **kw)

    with pytest.raises(AssertionError):
        CHECKS[mutation](mutated)


# ---- the reference's tests/test_pbl_model.py, restated on the oracle -----


@pytest.mark.parametrize("n,zm,wind,ustar,mol", [(10, 5.0, (3.0, 1.0), 0.4, -100.0), (20, 10.0, (2.0, 2.0), 0.3, -50.0)])
def test_reference_most_closure(n, zm, wind, ustar, mol):
    """tests/test_pbl_model.py:56-96."""
    z, (u, v, Kx, Ky, Kz) = opbl.vertical_profiles(n=n, meas_height=zm, wind=wind, ustar=ustar, closure="MOST", mol=mol)
    assert len(u) == len(v) == len(Kx) == len(Ky) == len(Kz) == len(z)
    assert not np.isclose(u[0], u[-1]) and not np.isclose(v[0], v[-1]) and not np.isclose(Kz[0], Kz[-1])
    assert np.all(z > 0) and np.all(Kx > 0) and np.all(Ky > 0) and np.all(Kz > 0)


def test_reference_oaahoc_closure():
    """tests/test_pbl_model.py:99-123."""
    z, (u, v, Kx, Ky, Kz) = opbl.vertical_profiles(n=10, meas_height=5.0, wind=(3.0, 1.0), ustar=0.4,
                                                   closure="OAAHOC", tke=2.0)
    assert len(u) == len(z) and len(Kx) == len(z)
    assert not np.isclose(u[0], u[-1]) and not np.isclose(Kz[0], Kz[-1])
    assert np.all(z > 0) and np.all(Kz > 0)
    np.testing.assert_array_equal(Kx, Ky)
    np.testing.assert_array_equal(Kx, Kz)


def test_reference_oaahoc_default_tke():
    """tests/test_pbl_model.py:126-141: tke defaults to 1."""
    z, (u, v, Kx, Ky, Kz) = opbl.vertical_profiles(n=10, meas_height=5.0, wind=(3.0, 1.0), ustar=0.4,
                                                   closure="OAAHOC", tke=None)
    assert len(u) == len(z) and np.all(z > 0) and np.all(Kz > 0)
    z1, p1 = opbl.vertical_profiles(n=10, meas_height=5.0, wind=(3.0, 1.0), ustar=0.4, closure="OAAHOC", tke=1.0)
    assert np.array_equal(z, z1) and np.array_equal(u, p1[0])


def test_reference_oaahoc_tke_arrays():
    """src/bldfm/pbl_model.py:96,126: tke[..., newaxis] -- a one-element array
    broadcasts the outputs to tke.shape + (nlev,); several values fail in
    np.squeeze(...).item()."""
    z, (u, v, Kx, Ky, Kz) = opbl.vertical_profiles(10, 5.0, (3.0, 1.0), 0.4, closure="OAAHOC", tke=[2.0])
    zs, ps = opbl.vertical_profiles(10, 5.0, (3.0, 1.0), 0.4, closure="OAAHOC", tke=2.0)
    assert z.shape == (1, len(zs)) and u.shape == (1, len(zs))
    assert np.array_equal(z[0], zs) and np.array_equal(u[0], ps[0])
    with pytest.raises(ValueError, match="size 1"):
        opbl.vertical_profiles(10, 5.0, (3.0, 1.0), 0.4, closure="OAAHOC", tke=[2.0, 3.0])


def test_reference_mostm_closure():
    """tests/test_pbl_model.py:144-165."""
    z, (u, v, Kx, Ky, Kz) = opbl.vertical_profiles(n=10, meas_height=5.0, wind=(3.0, 1.0), ustar=0.4,
                                                   closure="MOSTM", mol=-100.0)
    assert len(u) == len(z) and np.all(z > 0)
    assert np.all(Kx > 0) and np.all(Ky > 0) and np.all(Kz > 0)
    assert not np.allclose(Kx, Ky)
    assert not np.isclose(u[0], u[-1])


def test_reference_z0_only_path():
    """tests/test_pbl_model.py:168-181: u* from z0 by the log law."""
    z, (u, v, Kx, Ky, Kz) = opbl.vertical_profiles(n=10, meas_height=5.0, wind=(3.0, 1.0), ustar=None, z0=0.1,
                                                   closure="MOST", mol=-100.0)
    assert len(u) == len(z) and np.all(z > 0) and np.all(Kz > 0)
    np.testing.assert_allclose(z[0], 0.1, rtol=1e-12)
    np.testing.assert_allclose(np.hypot(u[10], v[10]), np.hypot(3.0, 1.0), rtol=1e-12)


def test_reference_errors():
    """tests/test_pbl_model.py:184-206."""
    with pytest.raises(ValueError, match="Either z0 or ustar"):
        opbl.vertical_profiles(n=10, meas_height=5.0, wind=(3.0, 1.0), ustar=0.4, z0=0.1, closure="MOST")
    with pytest.raises(ValueError, match="Invalid closure type"):
        opbl.vertical_profiles(n=10, meas_height=5.0, wind=(3.0, 1.0), ustar=0.4, closure="INVALID")


def test_reference_stretch_and_domain_height():
    """tests/test_pbl_model.py:209-232."""
    z_default, _ = opbl.vertical_profiles(n=10, meas_height=5.0, wind=(3.0, 1.0), ustar=0.4, closure="MOST")
    z_custom, (u, v, Kx, Ky, Kz) = opbl.vertical_profiles(n=10, meas_height=5.0, wind=(3.0, 1.0), ustar=0.4,
                                                          closure="MOST", stretch=15.0, domain_height=50.0)
    assert z_custom[-1] > z_default[-1]
    assert len(u) == len(z_custom) and np.all(Kz > 0)
