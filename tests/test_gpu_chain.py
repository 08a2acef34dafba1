"""GPU: the drop-in chain end to end against the oracle chain.

Every step runs through the drop-ins -- device compute_wind_fields, device
vertical_profiles (scalar prologue in the host's numpy, per-level profiles on
the GPU with correctly rounded / glibc-compatible functions), device solver --
and is compared with the oracle chain (numpy) from the same met scalars:
wind speed and direction, u*, L, zm, domain height.  Bar: rel-L2 <= 1e-10 of
conc and flx.  The stable C2 footprint is where the method amplifies a one-ulp
change of a profile value into ~1e-8, so this bar holds only because the
device profiles equal the oracle's bit for bit there.
"""

import numpy as np
import pytest

import scenarios
from conftest import rel_l2
from oracle import pbl as opbl
from oracle import solver as osolver

pytestmark = pytest.mark.gpu

CHAIN = {
    # name: (wind speed, direction, profile kwargs without wind)
    "c1_footprint": (5.0, 270.0, dict(n=19, meas_height=10.0, ustar=0.4, mol=-50.0)),
    "c1_dispersion": (5.0, 270.0, dict(n=19, meas_height=10.0, ustar=0.4, mol=-50.0)),
    "c2_unstable_footprint": (scenarios.c2_speed(-20.0), 270.0,
                              dict(n=26, meas_height=15.0, ustar=0.3, mol=-20.0, domain_height=100.0)),
    "c2_stable_footprint": (scenarios.c2_speed(100.0), 270.0,
                            dict(n=26, meas_height=15.0, ustar=0.3, mol=100.0, domain_height=100.0)),
    "c2_unstable_dispersion": (scenarios.c2_speed(-20.0), 270.0,
                               dict(n=26, meas_height=15.0, ustar=0.3, mol=-20.0, domain_height=100.0)),
    "c2_stable_dispersion": (scenarios.c2_speed(100.0), 270.0,
                             dict(n=26, meas_height=15.0, ustar=0.3, mol=100.0, domain_height=100.0)),
}


@pytest.fixture(scope="module")
def dropins():
    import torch

    assert torch.cuda.is_available(), "GPU tests need a HIP device"
    from bldfm_amd import _lib, pbl_model, solver, utils

    _lib.load()
    return utils, pbl_model, solver


@pytest.mark.parametrize("name", list(CHAIN))
def test_chain_end_to_end(dropins, name):
    """C1 and C2 footprints and fields; for the 3-D fields every level from
    the surface to the measurement height (levels 0..n)."""
    utils, pbl_model, solver = dropins
    spd, wd, pk = CHAIN[name]
    print(f"[chain] {name}", flush=True)
    gw = utils.compute_wind_fields(spd, wd)
    ow = opbl.compute_wind_fields(spd, wd)
    gz, gp = pbl_model.vertical_profiles(wind=gw, **pk)
    oz, op = opbl.vertical_profiles(wind=ow, **pk)
    q0, _, _, domain, levels, kw = scenarios.inputs(name, opbl.vertical_profiles)
    if "dispersion" in name and "c2" in name:
        levels = list(range(pk["n"] + 1))
    _, gc, gf = solver.steady_state_transport_solver(q0, gz, gp, domain, levels, **kw)
    _, oc, of = osolver.steady_state_transport_solver(q0, oz, op, domain, levels, **kw)
    same = [bool(np.array_equal(a, b)) for a, b in zip((gz,) + gp, (oz,) + op)]
    if gc.ndim == 3:
        ec = max(rel_l2(gc[i], oc[i]) for i in range(gc.shape[0]))
        ef = max(rel_l2(gf[i], of[i]) for i in range(gf.shape[0]))
    else:
        ec, ef = rel_l2(gc, oc), rel_l2(gf, of)
    print(f"  wind equal {gw == tuple(float(a) for a in ow)}, profiles z,u,v,Kx,Ky,Kz equal {same}; "
          f"rel-L2 conc {ec:.2e} flx {ef:.2e}")
    assert ec <= 1e-10 and ef <= 1e-10
    if "c2" in name:
        assert all(same)


def test_device_wind_every_tenth_degree(dropins):
    """Device compute_wind_fields over 0..360 deg in 0.1 deg steps: equal to
    numpy's except where glibc's sin/cos (which numpy calls) are not
    correctly rounded; never more than 3 ulps of U sin / U cos apart."""
    utils, _, _ = dropins
    d = np.round(np.arange(3601) * 0.1, 1)
    for spd in (1.0, 5.0, 7.3):
        u, v = utils.compute_wind_fields(np.full_like(d, spd), d)
        ou, ov = opbl.compute_wind_fields(spd, d)
        frac = (np.mean(u != ou), np.mean(v != ov))
        ulp = max(np.max(np.abs(u - ou) / np.spacing(np.maximum(np.abs(u), np.abs(ou)))),
                  np.max(np.abs(v - ov) / np.spacing(np.maximum(np.abs(v), np.abs(ov)))))
        print(f"  U={spd}: differing u {frac[0]:.3%} v {frac[1]:.3%}, max {ulp:.0f} ulp")
        assert frac[0] < 5e-3 and frac[1] < 5e-3 and ulp <= 3


def test_device_profiles_met_sweep(dropins):
    """Device vertical_profiles over 200 met draws (u* U(0.1,1), 1/L
    U(-0.1,0.05), zm U(2,50), |U| U(1,10), direction U(0,360)) against the
    oracle: values equal except where numpy's own log near 1 rounds
    differently from the correctly rounded value (< 1% of values, each a
    few ulps at most)."""
    _, pbl_model, _ = dropins
    r = np.random.default_rng(9)
    cases = values = bad = exact_cases = 0
    for _ in range(260):
        ust, invl = r.uniform(0.1, 1.0), r.uniform(-0.1, 0.05)
        mol = 1.0 / invl if invl != 0 else 1e9
        zm, spd, wd = r.uniform(2, 50), r.uniform(1, 10), r.uniform(0, 360)
        w = opbl.compute_wind_fields(spd, wd)
        with np.errstate(all="ignore"):
            try:
                oz, op = opbl.vertical_profiles(26, zm, w, ustar=ust, mol=mol)
            except ValueError:
                continue
        if len(oz) < 27 or not np.all(np.isfinite(oz)) or oz[0] >= zm:
            continue
        z, p = pbl_model.vertical_profiles(26, zm, w, ustar=ust, mol=mol)
        cases += 1
        nb = 0
        for a, b in ((z, oz), (p[0], op[0]), (p[1], op[1]), (p[4], op[4])):
            values += a.size
            nb += int(np.sum(a != b))
            assert np.all(np.isclose(a, b, rtol=1e-12, atol=1e-15))
        bad += nb
        exact_cases += nb == 0
    print(f"  {cases} met cases: {exact_cases} bit for bit; {bad} of {values} values differ ({bad / values:.2%})")
    assert cases >= 200 and bad / values < 0.01


@pytest.mark.parametrize(
    "kw",
    [
        dict(n=16, meas_height=10.0, wind=(0.0, -6.0), ustar=0.5),
        dict(n=10, meas_height=5.0, wind=(3.0, 1.0), ustar=0.4, mol=-100.0),
        dict(n=20, meas_height=10.0, wind=(2.0, 2.0), ustar=0.3, mol=-50.0),
        dict(n=26, meas_height=15.0, wind=(-3.5, 0.1), ustar=0.3, mol=100.0, domain_height=100.0),
        dict(n=10, meas_height=5.0, wind=(3.0, 1.0), z0=0.1, mol=-100.0),
        dict(n=10, meas_height=5.0, wind=(3.0, 1.0), ustar=0.4, closure="CONSTANT", prsc=0.8),
        dict(n=10, meas_height=5.0, wind=(3.0, 1.0), ustar=0.4, closure="MOSTM", mol=-100.0),
        dict(n=10, meas_height=5.0, wind=(3.0, 1.0), ustar=0.4, closure="OAAHOC"),
        dict(n=10, meas_height=5.0, wind=(3.0, 1.0), ustar=0.4, stretch=15.0, domain_height=50.0),
    ],
)
def test_device_profiles_bit_for_bit(dropins, kw):
    """The reference's tests/test_pbl_model.py inputs: device profiles equal
    the oracle's bit for bit."""
    _, pbl_model, _ = dropins
    z, p = pbl_model.vertical_profiles(**kw)
    oz, op = opbl.vertical_profiles(**kw)
    assert np.array_equal(z, oz)
    for a, b in zip(p, op):
        assert np.array_equal(a, b)


def test_device_oaahoc_tke_arrays(dropins):
    """A one-element tke array gives the reference's (1, nlev) outputs;
    several values raise the reference's ValueError."""
    _, pbl_model, _ = dropins
    z, p = pbl_model.vertical_profiles(10, 5.0, (3.0, 1.0), 0.4, closure="OAAHOC", tke=[2.0])
    oz, op = opbl.vertical_profiles(10, 5.0, (3.0, 1.0), 0.4, closure="OAAHOC", tke=[2.0])
    assert z.shape == oz.shape == (1, oz.shape[-1]) and p[0].shape == op[0].shape
    np.testing.assert_allclose(z, oz, rtol=1e-15)
    for a, b in zip(p, op):
        np.testing.assert_allclose(a, b, rtol=1e-14)
    with pytest.raises(ValueError, match="size 1"):
        pbl_model.vertical_profiles(10, 5.0, (3.0, 1.0), 0.4, closure="OAAHOC", tke=[2.0, 3.0])


@pytest.mark.parametrize("kw", [dict(mol=np.inf), dict(mol=-np.inf), dict(mol=0.0), dict(mol=1e-300),
                                dict(wind=(0.0, 0.0)), dict(ustar=1e-12), dict(mol=-1e-3), dict(mol=1e300)],
                         ids=str)
def test_device_profiles_on_degenerate_meteorology(dropins, kw):
    """Neutral limits (L = +-inf), L = 0 (the reference's ZeroDivisionError),
    L -> 0+, zero wind (NaN wind components) and u* -> 0 (infinite speeds):
    the same exception, or the same values with NaN and inf in the same
    places, as the oracle.  One deliberate difference: where the reference's
    grid itself is not finite (L = 1e-300 makes the roughness length
    infinite, so z[0] = inf), the drop-in raises ValueError instead of
    returning NaN/inf profiles."""
    _, pbl_model, _ = dropins
    args = dict(n=8, meas_height=10.0, wind=(3.0, 1.0), ustar=0.4, mol=-50.0)
    args.update(kw)
    with np.errstate(all="ignore"):
        try:
            oz, op = opbl.vertical_profiles(**args)
        except Exception as e:  # noqa: BLE001 -- the type is compared
            with pytest.raises(type(e)):
                pbl_model.vertical_profiles(**args)
            return
    if not np.all(np.isfinite(oz)):
        with pytest.raises(ValueError, match="non-finite"):
            pbl_model.vertical_profiles(**args)
        return
    z, p = pbl_model.vertical_profiles(**args)
    for a, b in zip((z,) + tuple(p), (oz,) + tuple(op)):
        assert a.shape == b.shape
        assert np.array_equal(np.isnan(a), np.isnan(b)) and np.array_equal(np.isinf(a), np.isinf(b))
        fin = np.isfinite(b)
        assert np.array_equal(np.sign(a[np.isinf(b)]), np.sign(b[np.isinf(b)]))
        np.testing.assert_allclose(a[fin], b[fin], rtol=1e-12, atol=0)


def test_device_psi_phi_on_special_arguments(dropins):
    """psi and phi (src/bldfm/pbl_model.py:207-250) at 0, -0, +-inf, NaN,
    subnormal and huge arguments and across the unstable/stable switch: the
    same values (bit for bit where finite), NaN and inf in the same places."""
    _, pbl_model, _ = dropins
    x = np.array([0.0, -0.0, np.inf, -np.inf, np.nan, 5e-324, -5e-324, 1e-300, -1e-300, 1e300, -1e300,
                  1e-8, -1e-8, 0.5, -0.5, 3.0, -3.0, 1e5, -1e5])
    with np.errstate(all="ignore"):
        for ours, ref in ((pbl_model.psi, opbl.psi), (pbl_model.phi, opbl.phi)):
            a, b = np.asarray(ours(x), dtype=np.float64), np.asarray(ref(x), dtype=np.float64)
            assert np.array_equal(np.isnan(a), np.isnan(b)), (ours.__name__, a, b)
            fin = ~np.isnan(b)
            assert np.array_equal(a[fin], b[fin]), (ours.__name__, x[fin][a[fin] != b[fin]])
