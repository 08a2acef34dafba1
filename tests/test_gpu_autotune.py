"""Launch shapes (bldfm_tuning) and the autotuner.

Every candidate shape -- modes per workgroup of the vertical solve and the
three x-pass input layouts -- gives bit-identical results on a mixed batch
(several grids, vertical resolutions and meteorologies, one and several
output levels): the shapes only regroup identical per-lane arithmetic.  The
autotuner returns one of the candidates it timed, within its 1% switching
margin of the fastest, and serves
later calls from its memory and JSON caches; shapes outside the candidates
are refused.
"""

import json

import numpy as np
import pytest

from test_gpu_batch import c3_like_specs

pytestmark = pytest.mark.gpu

CANDIDATES = [dict(shoot_block=b, xpass_layout=lay) for b in (64, 128, 256) for lay in (0, 1, 2)]


@pytest.fixture(scope="module")
def specs():
    s = c3_like_specs(9, seed=20261207)
    s[0].levels = [0, s[0].n, s[0].n + 3]  # several output slots in one instance
    return s


def _outputs(specs, tuning):
    import torch

    from bldfm_amd.batch import FootprintBatch

    fb = FootprintBatch(specs, precision="double", tuning=tuning)
    fb.run()
    torch.cuda.synchronize()
    return fb, fb.conc.clone(), fb.flx.clone()


@pytest.fixture(scope="module")
def reference(specs):
    return _outputs(specs, None)


@pytest.mark.parametrize("tuning", CANDIDATES, ids=lambda t: f"b{t['shoot_block']}_x{t['xpass_layout']}")
def test_every_shape_is_bitwise_identical(specs, reference, tuning):
    import torch

    fb, conc, flx = _outputs(specs, tuning)
    assert fb.tuning == tuning
    assert torch.isfinite(conc).all() and torch.isfinite(flx).all()
    assert torch.equal(conc, reference[1]) and torch.equal(flx, reference[2])


def test_autotune_picks_the_fastest_and_caches(specs, reference, tmp_path):
    from bldfm_amd import autotune

    fb = reference[0]
    autotune.clear_memo()
    path = str(tmp_path / "tune.json")
    r1 = autotune.tune(fb.instances, fb.levels_i32, fb.device, cache_path=path)
    assert not r1["cached"]
    assert r1["shoot_block"] in (64, 128, 256) and r1["xpass_layout"] in (0, 1, 2)
    assert len(r1["ms"]) == 5  # coordinate search: default, two more blocks, two more layouts
    chosen = f"shoot_block={r1['shoot_block']},xpass_layout={r1['xpass_layout']}"
    # the choice is within the 1% switching margin of the fastest candidate
    assert r1["ms"][chosen] == r1["best_ms"] <= min(r1["ms"].values()) / autotune.MARGIN
    r2 = autotune.tune(fb.instances, fb.levels_i32, fb.device, cache_path=path)
    assert r2["cached"] and r2["shoot_block"] == r1["shoot_block"] and r2["xpass_layout"] == r1["xpass_layout"]
    autotune.clear_memo()
    r3 = autotune.tune(fb.instances, fb.levels_i32, fb.device, cache_path=path)
    assert r3["cached"] and r3["best_ms"] == r1["best_ms"]
    assert autotune.shape_key(fb.instances, fb.device) in json.load(open(path))


def test_auto_batch_matches_default(specs, reference):
    import torch

    fb, conc, flx = _outputs(specs, "auto")
    assert fb.tune_report is not None
    assert torch.equal(conc, reference[1]) and torch.equal(flx, reference[2])


@pytest.mark.parametrize("bad", [dict(shoot_block=96), dict(xpass_layout=3), "fastest"])
def test_invalid_shapes_are_refused(specs, bad):
    from bldfm_amd import _lib
    from bldfm_amd.batch import FootprintBatch

    with pytest.raises((ValueError, _lib.BLDFMError)):
        FootprintBatch(specs[:2], tuning=bad)
