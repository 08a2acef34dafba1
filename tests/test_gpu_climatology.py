"""Footprint climatology at reduced scale: two towers x 256 hourly
footprints at 512^2, served by the trained surrogate with its 1 % FP64
spot check and fallback, against the all-FP64 climatology of the same
hours.  Bar: per-tower climatology rel-L2 <= 1e-3 (the serving tolerance;
a mean of footprints each within 1e-3 cannot be further off), most hours
served by the network, positive mass peaking near the tower.  (The FP64
footprints themselves have small negative spectral lobes, a property of the
reference's truncated-mode method, so no sign bar is set.)"""

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def test_two_tower_climatology_within_tolerance():
    from bldfm_amd import climatology as clim

    report, means = clim.run(clim.DEFAULT_TOWERS[:2], hours=256, batch=128, seed=7,
                             log=lambda m: print(m, flush=True))
    for name, r in report["towers"].items():
        print(name, r, flush=True)
        assert r["climatology_rel_l2"] <= 1e-3
        assert r["climatology_integral_rel_err"] <= 1e-3
        assert r["accepted"] + r["fallback"] == 256 and r["accepted"] >= 128
        cs, cf = means[name]
        assert cs.shape == (512, 512) and cf.sum() > 0
        # the mean over wind directions peaks near the tower (grid centre, 3.9 m cells)
        iy, ix = np.unravel_index(np.argmax(cf), cf.shape)
        assert np.hypot(iy - 256, ix - 256) * 2000.0 / 512 < 300.0
