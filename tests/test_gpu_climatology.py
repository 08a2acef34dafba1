"""Footprint climatology at serving scale: 4 towers x 8760 hourly footprints
at 512^2 in batches of 128, served by the trained surrogate with its 1 % FP64
spot check and fallback, against the all-FP64 climatology of the same hours.

Bars: per-tower climatology rel-L2 <= 1e-3 and integral error <= 1e-3 (the
serving tolerance: a mean of footprints each within 1e-3 cannot be further
off); every hour either served or solved; positive mass peaking near the
tower.  Two hours per tower of the FP64 leg against the oracle chain from the
same meteorology: rel-L2 <= 1e-10.  (The FP64 footprints themselves have
small negative spectral lobes, a property of the reference's truncated-mode
method, so no sign bar is set.)"""

import math

import numpy as np
import pytest

from conftest import rel_l2

pytestmark = pytest.mark.gpu

HOURS = 8760


@pytest.fixture(scope="module")
def full_run():
    from bldfm_amd import climatology as clim

    return clim.run(clim.DEFAULT_TOWERS, hours=HOURS, batch=128, seed=2026, log=lambda m: print(m, flush=True))


def test_four_tower_climatology_within_tolerance(full_run):
    report, means = full_run
    print(f"  {report['surrogate_footprints_per_s']:.0f} footprints/s with the surrogate, "
          f"{report['fp64_footprints_per_s']:.0f} all-FP64 (end to end)")
    assert len(report["towers"]) == 4
    for name, r in report["towers"].items():
        print(name, r, flush=True)
        assert r["climatology_rel_l2"] <= 1e-3
        assert r["climatology_integral_rel_err"] <= 1e-3
        assert r["accepted"] + r["fallback"] == HOURS and r["accepted"] >= HOURS // 2
        cs, cf = means[name]
        assert cs.shape == (512, 512) and cf.sum() > 0
        # the mean over wind directions peaks near the tower (grid centre, 3.9 m cells)
        iy, ix = np.unravel_index(np.argmax(cf), cf.shape)
        assert np.hypot(iy - 256, ix - 256) * 2000.0 / 512 < 300.0


@pytest.mark.parametrize("hours", [(0, 4380)])
def test_fp64_leg_matches_the_oracle(hours):
    """Two hours per tower of the climatology's FP64 solves (same meteorology
    series, same plan inputs) against the oracle's vertical_profiles and
    solver: rel-L2 <= 1e-10."""
    from bldfm_amd import climatology as clim
    from bldfm_amd.batch import FootprintBatch
    from bldfm_amd.surrogate import Hull
    from oracle import pbl as opbl
    from oracle import solver as osolver

    hull = Hull()
    for k, tower in enumerate(clim.DEFAULT_TOWERS):
        met = clim.hourly_met(tower, HOURS, 2026 + k)
        specs = [clim.specs_of(tower, met, h, h + 1, hull)[0] for h in hours]
        fb = FootprintBatch(specs, precision="double")
        fb.run()
        for j, (h, s) in enumerate(zip(hours, specs)):
            z, profs = opbl.vertical_profiles(s.n, s.meas_height, s.wind, ustar=s.ustar, mol=s.mol,
                                              closure=s.closure, domain_height=s.domain_height)
            _, oc, of = osolver.steady_state_transport_solver(
                np.zeros((s.ny, s.nx)), z, profs, s.domain, s.n, modes=s.modes, meas_pt=s.meas_pt, footprint=True,
                precision="double")
            gc, gf = (t[0].cpu().numpy() for t in fb.output(j))
            ec, ef = rel_l2(gc, oc), rel_l2(gf, of)
            print(f"  {tower.name} hour {h}: u*={s.ustar:.3f} 1/L={1 / s.mol:+.4f} dir={math.degrees(math.atan2(-s.wind[0], -s.wind[1])) % 360:.0f}"
                  f" -> conc {ec:.2e} flx {ef:.2e}", flush=True)
            assert ec <= 1e-10 and ef <= 1e-10
