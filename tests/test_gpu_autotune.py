"""Launch shapes (bldfm_tuning) and the autotuner.

Every candidate shape -- modes per workgroup of the vertical solve, the
three x-pass input layouts and the FFT workgroup size -- gives bit-identical
results on a mixed batch
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

CANDIDATES = [dict(shoot_block=b, xpass_layout=lay, fft_threads=0) for b in (64, 128, 256) for lay in (0, 1, 2)]
# FFT workgroup sizes above the default (valid for any batch whose transforms fit the 256-thread kernels)
CANDIDATES += [dict(shoot_block=256, xpass_layout=0, fft_threads=192), dict(shoot_block=256, xpass_layout=0, fft_threads=256),
               dict(shoot_block=128, xpass_layout=2, fft_threads=256)]


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


@pytest.mark.parametrize("tuning", CANDIDATES,
                         ids=lambda t: f"b{t['shoot_block']}_x{t['xpass_layout']}_f{t['fft_threads']}")
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
    fts = autotune.fft_thread_candidates(fb.instances)
    assert fts[0] == 0 and len(fts) >= 2 and r1["fft_threads"] in fts
    # coordinate search: default, two more blocks, two more layouts, the other FFT workgroup size
    assert len(r1["ms"]) == 5 + len(fts) - 1
    chosen = (f"shoot_block={r1['shoot_block']},xpass_layout={r1['xpass_layout']},"
              f"fft_threads={r1['fft_threads']}")
    # the choice is within the 1% switching margin of the fastest candidate
    assert r1["ms"][chosen] == r1["best_ms"] <= min(r1["ms"].values()) / autotune.MARGIN
    r2 = autotune.tune(fb.instances, fb.levels_i32, fb.device, cache_path=path)
    assert r2["cached"] and all(r2[k] == r1[k] for k in ("shoot_block", "xpass_layout", "fft_threads"))
    autotune.clear_memo()
    r3 = autotune.tune(fb.instances, fb.levels_i32, fb.device, cache_path=path)
    assert r3["cached"] and r3["best_ms"] == r1["best_ms"]
    assert autotune.shape_key(fb.instances, fb.device) in json.load(open(path))


def test_auto_batch_matches_default(specs, reference):
    import torch

    fb, conc, flx = _outputs(specs, "auto")
    assert fb.tune_report is not None
    assert torch.equal(conc, reference[1]) and torch.equal(flx, reference[2])


@pytest.mark.parametrize("bad", [dict(shoot_block=96), dict(xpass_layout=3), dict(fft_threads=100),
                                 dict(fft_threads=1088), "fastest"])
def test_invalid_shapes_are_refused(specs, bad):
    from bldfm_amd import _lib
    from bldfm_amd.batch import FootprintBatch

    with pytest.raises((ValueError, _lib.BLDFMError)):
        FootprintBatch(specs[:2], tuning=bad)


@pytest.mark.parametrize("precision", ["double", "single"])
def test_layouts_on_dispersion_solves(precision):
    """The x-pass layouts (including the LDS-tiled transpose) and block sizes
    on dispersion instances (e^{+i} transforms) with several output levels,
    in FP64 and with float32 outputs: bit-identical to the default shape."""
    import torch

    from bldfm_amd import _lib, solver
    from bldfm_amd.pbl_model import vertical_profiles
    from bldfm_amd.utils import ideal_source

    dev = _lib.device()
    nx, ny, dom, levels = 96, 80, (480.0, 400.0), [2, 7, 11]
    z, profs = vertical_profiles(12, 8.0, (3.0, -2.0), ustar=0.35, mol=-80.0)
    q0 = ideal_source((nx, ny), dom, shape="diamond")
    g = solver._geometry((ny, nx), np.asarray(z), dom, levels, (nx, ny), (0.0, 0.0), False, None, precision)
    slot, idx = solver._level_maps(g["levels"], g["nz"])
    ints = np.concatenate([slot, idx]).astype(np.int32)
    inst = solver._instance(g, len(levels), False, False, precision, 0.0)
    prof = torch.as_tensor(np.concatenate([np.asarray(z)] + [np.asarray(p) for p in profs]), device=dev)
    src = torch.as_tensor(np.ascontiguousarray(q0, dtype=np.float64), device=dev).reshape(-1)
    odt = torch.float32 if g["out_f32"] else torch.float64
    outs = []
    for tuning in [None] + CANDIDATES:
        plan = _lib.Plan([inst], ints, dev, tuning=tuning)
        conc = torch.empty(plan.out_elems, dtype=odt, device=dev)
        flx = torch.empty_like(conc)
        ws = torch.empty(max(plan.workspace_bytes, 1), dtype=torch.uint8, device=dev)
        plan.execute(prof, src, conc, flx, ws)
        torch.cuda.synchronize()
        outs.append((conc, flx))
    assert torch.isfinite(outs[0][0]).all() and float(outs[0][0].abs().max()) > 0
    for (c, f), t in zip(outs[1:], CANDIDATES):
        assert torch.equal(c, outs[0][0]) and torch.equal(f, outs[0][1]), t


def test_fft_workgroup_sizes_of_the_large_transform_kernels():
    """Transforms above 2048 points run in the 1024-thread kernel family;
    every allowed workgroup size (the default ceil(n/10) rounded up to a wave,
    the autotuner's candidates and the 1024-thread bound) gives the same
    footprints bit for bit."""
    import torch

    from bldfm_amd import autotune
    from bldfm_amd.batch import FootprintBatch, FootprintSpec

    specs = [FootprintSpec(nx=1000, ny=1000, domain=(2000.0, 2000.0), modes=(256, 256), meas_pt=(1000.0, 1000.0),
                           n=12, meas_height=10.0, wind=(3.0, 1.0), ustar=0.3, mol=-40.0)]
    ref = FootprintBatch(specs, precision="double")
    n = max(max(int(i.nx) + 2 * int(i.px), int(i.ny) + 2 * int(i.py)) for i in ref.instances)
    assert n > 2048
    cands = autotune.fft_thread_candidates(ref.instances)
    assert len(cands) >= 2
    ref.run()
    torch.cuda.synchronize()
    c0, f0 = ref.conc.clone(), ref.flx.clone()
    assert torch.isfinite(c0).all() and float(f0.abs().max()) > 0
    for ft in cands[1:] + [1024]:
        fb = FootprintBatch(specs, precision="double", tuning=dict(shoot_block=256, xpass_layout=0, fft_threads=ft))
        fb.run()
        torch.cuda.synchronize()
        assert torch.equal(fb.conc, c0) and torch.equal(fb.flx, f0), ft
