"""CPU tests that pin the oracle before it is trusted as the checker.

Each test pins a part of oracle/ against something independent of its own
formulas: the reference's regression fixtures, closed forms, brute force,
or invariants the mathematics fixes.  Planted mistakes (oracle mutations)
must fail the fixture comparison.
"""

import numpy as np
import pytest

import scenarios
from conftest import TOWER_X, load_golden, rel_l2
from oracle import pbl as opbl
from oracle import solver as osolver

FIXTURES = ["source_area", "plume_3d", "single_footprint", "timeseries_step2"]


def run_oracle(name, **kw):
    return scenarios.run(name, opbl.vertical_profiles, osolver.steady_state_transport_solver, **kw)


@pytest.mark.parametrize("name", FIXTURES)
def test_reference_fixtures(name):
    """tests/references/*.npz of the reference, at its tolerance and far below it."""
    grid, conc, flx = run_oracle(name)
    ref = load_golden(name)
    np.testing.assert_allclose(conc, ref["conc"], atol=1e-6, rtol=1e-5)
    np.testing.assert_allclose(flx, ref["flx"], atol=1e-6, rtol=1e-5)
    # far below the reference's tolerance: the fixtures were rounded through
    # complex64, so a last-digit difference in an input (e.g. the recovered
    # tower x) can flip a float32 rounding of a small mode
    assert np.abs(conc - ref["conc"]).max() <= 1e-10 * np.abs(ref["conc"]).max()
    assert np.abs(flx - ref["flx"]).max() <= 1e-10 * np.abs(ref["flx"]).max()
    for k, g in zip("XYZ", grid):
        np.testing.assert_array_equal(g, ref[k])


def test_tower_x_is_pinned_by_the_fixture():
    """The fixtures' tower x (from a geodesy helper the reference imports) is
    recovered as the minimiser of the fixture misfit."""
    ref = load_golden("single_footprint")
    errs = []
    for dx in (-0.05, 0.0, 0.05):
        _, _, flx = run_oracle("single_footprint", meas_pt=(TOWER_X + dx, 0.0))
        errs.append(np.abs(flx - ref["flx"]).max())
    assert errs[1] < 1e-12 and errs[0] > 1e-7 and errs[2] > 1e-7


@pytest.mark.parametrize("mutation", [m for m in osolver.VALID_MUTATIONS if m])
def test_planted_mistakes_fail_the_fixtures(mutation):
    """Mistakes in the vertical step, the radiation condition and the phase
    shift change source_area beyond the reference's tolerance."""
    q0, z, profs, domain, levels, kw = scenarios.inputs("source_area", opbl.vertical_profiles)
    _, mc, mf = osolver.steady_state_transport_solver(q0, z, profs, domain, levels, _mutation=mutation, **kw)
    ref = load_golden("source_area")
    assert not np.allclose(mf, ref["flx"], atol=1e-6, rtol=1e-5)


def test_index_map_equals_fftshift_slice_sequence():
    """f_j mod ne equals the reference's fftshift / slice / ifftshift
    truncation (src/bldfm/solver.py:139-145) and its pad inverse (:265-278)."""
    from numpy.fft import fftshift, ifftshift

    rng = np.random.default_rng(0)
    for ne in range(2, 41):
        for nl in range(2, ne + 1, 2):
            if (ne - nl) % 2:
                continue
            a = rng.standard_normal(ne)
            dl = (ne - nl) // 2
            t = ifftshift(fftshift(a)[dl : ne - dl])
            assert np.array_equal(t, a[osolver.padded_index(nl, ne)])
            back = ifftshift(np.pad(fftshift(t), (dl, dl)))
            full = np.zeros(ne)
            full[osolver.padded_index(nl, ne)] = t
            assert np.array_equal(back, full)


@pytest.mark.parametrize("footprint", [False, True])
def test_transform_conventions_by_brute_force(footprint):
    """Direct DFT of the truncated spectra over the padded grid reproduces
    conc/flx: pins the mirror (footprint: e^{-ikx}, unnormalised), the
    dispersion inverse (e^{+ikx}), the index maps and the crop."""
    nx, ny, domain = 10, 6, (20.0, 12.0)
    z, profs = opbl.vertical_profiles(8, 5.0, (2.0, 1.0), 0.3, mol=-30.0)
    q0 = np.random.default_rng(1).random((ny, nx))
    kw = dict(modes=(6, 4), meas_pt=(7.0, 5.0), footprint=footprint, halo=4.0, precision="double")
    (_, conc, flx), (tp, tq) = osolver.steady_state_transport_solver(q0, z, profs, domain, 8, _return_spectra=True, **kw)
    dx, dy = domain[0] / nx, domain[1] / ny
    px, py = int(4.0 / dx), int(4.0 / dy)
    nxe, nye = nx + 2 * px, ny + 2 * py
    fx = np.fft.fftfreq(6, 1 / 6).astype(int)
    fy = np.fft.fftfreq(4, 1 / 4).astype(int)
    sgn = -1.0 if footprint else 1.0
    out_c = np.zeros((ny, nx))
    out_f = np.zeros((ny, nx))
    for iy in range(ny):
        for ix in range(nx):
            Y, X = py + iy, px + ix
            ph = np.exp(sgn * 2j * np.pi * (np.outer(fy, np.ones(6)) * Y / nye + np.outer(np.ones(4), fx) * X / nxe))
            out_c[iy, ix] = (tp[0] * ph).sum().real
            out_f[iy, ix] = (tq[0] * ph).sum().real
    assert rel_l2(conc, out_c) < 1e-13 and rel_l2(flx, out_f) < 1e-13


def test_footprint_and_dispersion_conservation():
    """Without halo the crop is the whole periodic grid: the footprint sums to
    1 at every level and the dispersion flux to sum(q0); the mean
    concentration is the (0,0) trapezoid p000 - q00 int dz/Kz."""
    nx, ny, domain = 16, 12, (160.0, 120.0)
    z, profs = opbl.vertical_profiles(10, 8.0, (3.0, -1.0), 0.3, mol=80.0)
    levels = [0, 4, 10, len(z) - 1]
    _, _, flx = osolver.steady_state_transport_solver(np.zeros((ny, nx)), z, profs, domain, levels,
                                                      modes=(16, 12), footprint=True, halo=0.0, precision="double")
    np.testing.assert_allclose(flx.sum(axis=(1, 2)), 1.0, rtol=1e-12)
    q0 = np.random.default_rng(2).random((ny, nx))
    _, conc, flx = osolver.steady_state_transport_solver(q0, z, profs, domain, levels, modes=(16, 12),
                                                         halo=0.0, precision="double", srf_bg_conc=0.7)
    np.testing.assert_allclose(flx.sum(axis=(1, 2)), q0.sum(), rtol=1e-12)
    Kz = profs[4]
    q00 = q0.mean()
    for s, lv in enumerate(levels):
        integral = np.sum(np.diff(z[: lv + 1]) * 0.5 * (1.0 / Kz[:lv] + 1.0 / Kz[1 : lv + 1]))
        assert abs(conc[s].mean() - (0.7 - q00 * integral)) < 1e-12 * max(1.0, abs(conc[s].mean()))


def test_analytic_closed_form_mode_by_mode():
    """analytic=True equals q^0 e^{-lambda h} with lambda^2 = k^2 + i(u kx +
    v ky)/K, computed here independently in extended precision."""
    nx, ny, domain = 16, 8, (100.0, 50.0)
    q0 = opbl.ideal_source((nx, ny), domain, (5.0, 5.0), shape="point")
    z, profs = opbl.vertical_profiles(16, 5.0, (4.0, 1.0), 0.2, closure="CONSTANT")
    (_, _, _), (tp, tq) = osolver.steady_state_transport_solver(q0, z, profs, domain, 9, modes=(16, 8), halo=50.0,
                                                                analytic=True, precision="double", _return_spectra=True)
    u, v, K = profs[0][0], profs[1][0], profs[4][0]
    dx, dy = domain[0] / nx, domain[1] / ny
    px, py = int(50.0 / dx), int(50.0 / dy)
    nxe, nye = nx + 2 * px, ny + 2 * py
    qpad = np.zeros((nye, nxe))
    qpad[py : py + ny, px : px + nx] = q0
    Q = np.fft.fft2(qpad) / (nxe * nye)
    h = np.longdouble(z[9]) - np.longdouble(z[0])
    for jy, fy in enumerate(np.fft.fftfreq(8, 1 / 8).astype(int)):
        for jx, fx in enumerate(np.fft.fftfreq(16, 1 / 16).astype(int)):
            if fx == 0 and fy == 0:
                continue
            kx = 2 * np.pi * fx / (dx * nxe)
            ky = 2 * np.pi * fy / (dy * nye)
            lam = np.sqrt(np.clongdouble(kx * kx + ky * ky + 1j * (u * kx + v * ky) / K))
            expect = Q[fy % nye, fx % nxe] * np.exp(-lam * h)
            assert abs(tq[0, jy, jx] - complex(expect)) <= 1e-12 * abs(Q).max()


def test_numerical_converges_to_analytic():
    """Monotone convergence of the numerical solve to the analytic branch
    (tests/test_integration.py:87-162 of the reference)."""
    nxy, domain = (128, 64), (100.0, 50.0)
    q0 = opbl.ideal_source(nxy, domain, (5.0, 5.0), shape="point")
    zr, pr = opbl.vertical_profiles(32, 5.0, (4.0, 1.0), 0.2, closure="CONSTANT")
    _, _, fa = osolver.steady_state_transport_solver(q0, zr, pr, domain, 32, modes=(256, 256), halo=500.0, analytic=True)
    errs = []
    for modes, nz in (((64, 64), 8), ((128, 128), 16), ((256, 256), 32)):
        z, p = opbl.vertical_profiles(nz, 5.0, (4.0, 1.0), 0.2, closure="CONSTANT")
        _, _, f = osolver.steady_state_transport_solver(q0, z, p, domain, nz, modes=modes, halo=500.0)
        errs.append(np.mean((f - fa) ** 2) / np.mean(fa**2))
    assert errs[0] > errs[1] > errs[2]


def test_psi_known_answers():
    """Businger-Dyer psi (tests/test_ffm_kormann_meixner.py:60-89 values)."""
    np.testing.assert_allclose(opbl.psi(10.0 / 1e6), 5 * 10.0 / 1e6, rtol=1e-12)
    np.testing.assert_allclose(opbl.psi(10.0 / 100.0), 0.5, rtol=1e-12)
    x = 10.0 / -100.0
    xi = (1 - 16 * x) ** 0.25
    expect = -2 * np.log(0.5 * (1 + xi)) - np.log(0.5 * (1 + xi**2)) + 2 * np.arctan(xi) - np.pi * 0.5
    np.testing.assert_allclose(opbl.psi(x), expect, rtol=1e-10)
    # psi(0) = 0 and phi(0) = 1; phi^-1 = xi for the unstable branch (phi_m = xi^-1)
    assert abs(opbl.psi(0.0)) < 1e-15 and opbl.phi(0.0) == 1.0
    np.testing.assert_allclose(opbl.phi(-0.3), (1 + 16 * 0.3) ** -0.5, rtol=1e-14)


@pytest.mark.parametrize("n,zm,dh,expect", [(16, 10.0, None, 27), (19, 10.0, None, 32), (25, 15.0, 100.0, 63), (26, 15.0, 100.0, 65)])
def test_vertical_grid_invariants(n, zm, dh, expect):
    """z[0] = z0, z[n] = zm, level counts of the stretched grid, and the MOST
    log law reproducing |U(zm)| from the derived z0."""
    wind = (3.0, -4.0)
    z, (u, v, Kx, Ky, Kz) = opbl.vertical_profiles(n, zm, wind, ustar=0.3, mol=-40.0, domain_height=dh)
    z0 = zm * np.exp(-0.4 * 5.0 / 0.3 + float(opbl.psi(zm / -40.0)))
    assert len(z) == expect
    np.testing.assert_allclose(z[0], z0, rtol=1e-12)
    np.testing.assert_allclose(z[n], zm, rtol=1e-13)
    np.testing.assert_allclose(np.hypot(u[n], v[n]), 5.0, rtol=1e-12)
    assert Kx is Ky is Kz and np.all(np.diff(z) > 0)


def test_constant_closure_values():
    """CONSTANT: K = 0.4 u* zm / prsc, u and v constant (tests/test_pbl_model.py:24-53)."""
    z, (u, v, Kx, Ky, Kz) = opbl.vertical_profiles(20, 10.0, (2.0, 2.0), 0.3, prsc=0.8, closure="CONSTANT")
    np.testing.assert_allclose(Kz, 0.4 * 0.3 * 10.0 / 0.8, rtol=1e-15)
    np.testing.assert_allclose(u, 2.0, rtol=0)
    np.testing.assert_allclose(v, 2.0, rtol=0)


def test_wind_fields_cardinal_directions():
    for wd, (eu, ev) in {0.0: (0.0, -5.0), 90.0: (-5.0, 0.0), 180.0: (0.0, 5.0), 270.0: (5.0, 0.0)}.items():
        u, v = opbl.compute_wind_fields(5.0, wd)
        assert abs(u - eu) < 1e-14 and abs(v - ev) < 1e-14


def test_hermitian_symmetry_of_the_spectrum():
    """For real sources the paired modes satisfy X(-k) = conj X(k) (to the
    forward FFT's roundoff); the GPU path computes one of each pair."""
    q0, z, profs, domain, levels, kw = scenarios.inputs("c1_dispersion", opbl.vertical_profiles)
    kw = dict(kw, modes=(32, 32))
    (_, _, _), (tp, tq) = osolver.steady_state_transport_solver(q0, z, profs, domain, levels, _return_spectra=True, **kw)
    for X in (tp[0], tq[0]):
        f = np.fft.fftfreq(32, 1 / 32).astype(int)  # -16 is the unpaired Nyquist frequency
        for jy in range(32):
            for jx in range(32):
                if f[jy] == -16 or f[jx] == -16:
                    continue
                a = X[jy, jx]
                b = X[(-f[jy]) % 32, (-f[jx]) % 32]
                assert abs(a - np.conj(b)) <= 1e-12 * abs(X).max()


@pytest.mark.parametrize("nx,ny,modes,halo", [(33, 20, (16, 12), 9.0), (20, 33, (12, 16), 9.0),
                                              (33, 21, (16, 12), 9.0), (33, 20, (16, 12), 0.0)])
def test_odd_padded_length_footprint_follows_the_reference_sequence(nx, ny, modes, halo):
    """With an odd ne - nl the reference still runs a footprint: its
    fftshift / pad by (ne - nl)//2 / ifftshift makes the inverse transform
    ne - 1 points long and the crop [p, ne - p) keeps min(n, ne - 1 - p) of
    them (src/bldfm/solver.py:265-290).  The oracle's index maps equal that
    literal sequence on its own spectra, and a dispersion solve raises the
    reference's IndexError (the truncated source spectrum has nl + 1 modes)."""
    from numpy.fft import fftshift, ifftshift

    z, profs = opbl.vertical_profiles(10, 5.0, (3.0, 1.0), 0.3, mol=-30.0)
    domain = (float(nx), float(ny))
    kw = dict(modes=modes, meas_pt=(10.0, 7.0), footprint=True, precision="double", halo=halo)
    (_, conc, flx), (tp, tq) = osolver.steady_state_transport_solver(np.zeros((ny, nx)), z, profs, domain, 6,
                                                                     _return_spectra=True, **kw)
    px, py = int(halo / (domain[0] / nx)), int(halo / (domain[1] / ny))
    nxe, nye = nx + 2 * px, ny + 2 * py
    dlx, dly = (nxe - modes[0]) // 2, (nye - modes[1]) // 2
    for spec, out in ((tp, conc), (tq, flx)):
        full = ifftshift(np.pad(fftshift(spec, axes=(1, 2)), ((0, 0), (dly, dly), (dlx, dlx))), axes=(1, 2))
        ref = np.squeeze(np.fft.fft2(full, norm="backward").real[:, py : nye - py, px : nxe - px])
        assert ref.shape == out.shape and np.array_equal(ref, out)
    with pytest.raises(IndexError, match="boolean index did not match"):
        osolver.steady_state_transport_solver(np.ones((ny, nx)), z, profs, domain, 6, modes=modes, halo=halo)


@pytest.mark.parametrize("nxe,nye", [(960, 384), (1536, 1536), (384, 384)])
def test_footprint_delta_spectrum_follows_numpy_complex_division(nxe, nye):
    """The reference forms the footprint's delta spectrum as
    np.ones(..., complex128) / nxe / nye (src/bldfm/solver.py:134); numpy's
    complex division by a real-valued divisor is Smith's method, a product
    with the rounded reciprocal, so the value is (1/nxe)(1/nye) -- one ulp
    away from (1/nxe)/nye at 960 x 384, the fixtures' padded grid.  The
    library forms q^0 the same way."""
    ref = (np.ones((2, 2), dtype=np.complex128) / nxe / nye)[0, 0]
    assert ref.imag == 0.0 and ref.real == (1.0 / nxe) * (1.0 / nye)
    if (nxe, nye) == (960, 384):
        assert ref.real != 1.0 / nxe / nye
