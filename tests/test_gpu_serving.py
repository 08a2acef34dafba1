"""Serving with one plan: FootprintBatch.update (bldfm_plan_update_met)
uploads new meteorology into an existing plan.  The updated plan -- run
eagerly or by replaying a graph captured before the update -- equals a
fresh plan built for the new meteorology bit for bit; an update that would
change an instance's vertical grid is refused and leaves the plan as it
was; the surrogate's hull flags follow the new meteorology."""

import dataclasses

import pytest

from test_gpu_batch import c3_like_specs

pytestmark = pytest.mark.gpu


def _run(fb, graph=False):
    import torch

    fb.replay() if graph else fb.run()
    torch.cuda.synchronize()
    return fb.conc.clone(), fb.flx.clone()


def _new_met(specs, seed):
    """Same geometry and vertical grid size, other wind, u* and stability."""
    import numpy as np

    rng = np.random.default_rng(seed)
    out = []
    for s in specs:
        f = rng.uniform(0.8, 1.25)
        wd = rng.uniform(0, 2 * np.pi)
        sp = float(np.hypot(*s.wind)) * f
        out.append(dataclasses.replace(s, wind=(-sp * np.sin(wd), -sp * np.cos(wd)), ustar=s.ustar * f,
                                       mol=s.mol * rng.uniform(0.9, 1.1)))
    return out


@pytest.mark.parametrize("graph", [False, True])
def test_updated_plan_equals_fresh_plan(graph):
    import torch

    from bldfm_amd.batch import FootprintBatch

    specs = c3_like_specs(8, seed=20261210)
    specs2 = _new_met(specs, 3)
    fresh = FootprintBatch(specs2, precision="double")
    ref = _run(fresh)
    fb = FootprintBatch(specs, precision="double")
    if graph:
        fb.capture()
    before = _run(fb, graph)
    fb.update(specs2)
    after = _run(fb, graph)
    assert torch.equal(after[0], ref[0]) and torch.equal(after[1], ref[1])
    assert not torch.equal(after[1], before[1])


def test_update_that_changes_the_vertical_grid_is_refused():
    import torch

    from bldfm_amd.batch import FootprintBatch

    specs = c3_like_specs(3, seed=20261211)
    fb = FootprintBatch(specs, precision="double")
    before = _run(fb)
    bad = list(specs)
    bad[1] = dataclasses.replace(specs[1], domain_height=5.0 * specs[1].meas_height)  # more points above zm
    with pytest.raises(ValueError):
        fb.update(bad)
    with pytest.raises(ValueError):
        fb.update([dataclasses.replace(specs[0], nx=specs[0].nx // 2)] + specs[1:])
    again = _run(fb)
    assert torch.equal(again[0], before[0]) and torch.equal(again[1], before[1])


def test_update_reevaluates_the_surrogate_hull():
    import math

    import numpy as np

    from bldfm_amd import surrogate as sur
    from bldfm_amd.batch import FootprintBatch, FootprintSpec
    from bldfm_amd.pbl_model import psi

    net = sur.SurrogateNet.load(sur.DEFAULT_CHECKPOINT)
    h = net.hull

    def spec(zm, ustar=0.35, invl=-0.02, z0=0.1):
        sp = ustar / 0.4 * (math.log(zm / z0) + float(np.asarray(psi(np.array([zm * invl])))[0]))
        return FootprintSpec(nx=128, ny=128, domain=(500.0, 500.0), modes=(128, 128), meas_pt=(250.0, 250.0),
                             n=h.n, meas_height=zm, wind=(0.0, -sp), ustar=ustar, mol=1.0 / invl,
                             domain_height=h.top_ratio * zm)

    fb = FootprintBatch([spec(15.0), spec(20.0)], precision="double", surrogate=net, tol=1e-3)
    fb.run()
    err, mode = fb.surrogate_stats()
    assert all(math.isfinite(e) for e in err)  # both inside the hull: spot-checked
    fb.update([spec(15.0), spec(20.0, ustar=1.5)])  # u* above the hull
    fb.run()
    err, mode = fb.surrogate_stats()
    assert math.isfinite(err[0]) and math.isnan(err[1]) and mode[1] == 1
