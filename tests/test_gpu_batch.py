"""Batched footprint solves: mixed grids / level counts / meteorology in one
plan, against single solves and the oracle; HIP-graph replays against the
eager result.

Bars: batch vs single-solve drop-in on the same device profiles, bitwise;
batch vs oracle rel-L2 <= 1e-10 (FP64); device profiles vs oracle profiles
rtol 1e-12; every graph replay equal to the eager execution.
"""

import numpy as np
import pytest

from conftest import rel_l2
from oracle import pbl as opbl
from oracle import solver as osolver

pytestmark = pytest.mark.gpu


def c3_like_specs(count, seed=20261205, grids=(32, 64, 128), ns=(8, 16, 32)):
    """Seeded instances drawn like the batched-solver workload (grid from a
    set, n from a set, u*, 1/L, wind speed, direction, zm uniform); the
    default sets are sizes the oracle finishes in seconds."""
    from bldfm_amd.batch import FootprintSpec

    rng = np.random.default_rng(seed)
    specs = []
    while len(specs) < count:
        g = int(rng.choice(grids))
        n = int(rng.choice(ns))
        ustar = rng.uniform(0.1, 1.0)
        inv_l = rng.uniform(-0.1, 0.05)
        mol = 1.0 / inv_l if inv_l != 0 else 1e9
        speed = rng.uniform(1.0, 10.0)
        wd = np.deg2rad(rng.uniform(0, 360))
        zm = rng.uniform(2.0, 50.0)
        dom = float(rng.choice([500.0, 2000.0]))
        # the reference derives z0 from (u*, |U|); skip draws whose z0 is not
        # below zm/10 (their grid does not reach level n; the reference fails)
        z0 = zm * np.exp(-0.4 * speed / ustar + float(opbl.psi(zm / mol)))
        if not z0 < 0.1 * zm:
            continue
        specs.append(FootprintSpec(
            nx=g, ny=g, domain=(dom, dom), modes=(g, g), meas_pt=(dom / 2, dom / 2), n=n,
            meas_height=zm, wind=(-speed * np.sin(wd), -speed * np.cos(wd)), ustar=ustar, mol=mol))
    return specs


@pytest.fixture(scope="module")
def batch():
    import torch

    assert torch.cuda.is_available()
    from bldfm_amd.batch import FootprintBatch

    specs = c3_like_specs(24)
    fb = FootprintBatch(specs, precision="double")
    fb.run()
    torch.cuda.synchronize()
    return fb


def test_batch_vs_single_and_oracle(batch):
    from bldfm_amd import solver

    for i, s in enumerate(batch.specs):
        print(f"[batch] instance {i} grid {s.nx} n {s.n} nz {batch.nz(i)}", flush=True)
        z, profs = batch.profiles(i)
        oz, oprofs = opbl.vertical_profiles(s.n, s.meas_height, s.wind, ustar=s.ustar, mol=s.mol)
        np.testing.assert_allclose(z, oz, rtol=1e-13)
        for a, b in zip(profs, oprofs):
            np.testing.assert_allclose(a, b, rtol=1e-12, atol=1e-15)
        bc, bf = (t[0].cpu().numpy() for t in batch.output(i))
        sc, sf, _ = solver.solve_device(np.zeros((s.ny, s.nx)), z, profs, s.domain, s.n, modes=s.modes,
                                        meas_pt=s.meas_pt, footprint=True, precision="double")
        assert np.array_equal(bc, sc[0].cpu().numpy()) and np.array_equal(bf, sf[0].cpu().numpy())
        _, oc, of = osolver.steady_state_transport_solver(np.zeros((s.ny, s.nx)), z, profs, s.domain, s.n,
                                                          modes=s.modes, meas_pt=s.meas_pt, footprint=True,
                                                          precision="double")
        assert rel_l2(bc, oc) <= 1e-10 and rel_l2(bf, of) <= 1e-10, i


def test_graph_replay_matches_eager(batch):
    import torch

    ref_c = batch.conc.clone()
    ref_f = batch.flx.clone()
    batch.capture()
    for r in range(3):
        batch.conc.zero_()
        batch.flx.zero_()
        torch.cuda.synchronize()
        batch.replay()
        torch.cuda.synchronize()
        assert torch.equal(batch.conc, ref_c) and torch.equal(batch.flx, ref_f), r
        batch.run()  # eager work between replays
        torch.cuda.synchronize()
        assert torch.equal(batch.conc, ref_c)


def test_batch_level_beyond_grid_raises():
    """A level beyond the instance's grid raises IndexError like z[levels]."""
    from bldfm_amd.batch import FootprintBatch, FootprintSpec

    spec = FootprintSpec(nx=32, ny=32, domain=(500.0, 500.0), modes=(32, 32), meas_pt=(250.0, 250.0), n=8,
                         meas_height=10.0, wind=(0.0, -5.0), ustar=0.4, levels=[400])
    with pytest.raises(IndexError):
        FootprintBatch([spec])


def test_c3_full_size_graph_batch_equals_single_solves():
    """The batched-solver workload at full size: 256 footprints drawn from
    grids {128^2, 256^2, 512^2} and n {16, 32, 64} with seeded meteorology,
    one plan, captured in a HIP graph and replayed; every instance equals its
    own single-instance solve bit for bit, and the two smallest agree with
    the oracle to 1e-10."""
    import torch

    from bldfm_amd.batch import FootprintBatch

    specs = c3_like_specs(256, seed=20261209, grids=(128, 256, 512), ns=(16, 32, 64))
    print(f"[c3 full] grids {sorted(set(s.nx for s in specs))}, n {sorted(set(s.n for s in specs))}", flush=True)
    fb = FootprintBatch(specs, precision="double")
    fb.capture()
    fb.replay()
    fb.replay()
    torch.cuda.synchronize()
    for i, spec in enumerate(specs):
        one = FootprintBatch([spec], precision="double")
        one.run()
        torch.cuda.synchronize()
        for a, b in zip(fb.output(i), one.output(0)):
            assert torch.equal(a, b), f"instance {i}"
        del one
    small = [i for i, s in enumerate(specs) if s.nx == 128][:2]
    for i in small:
        s = specs[i]
        z, profs = fb.profiles(i)  # device profiles (pinned against the oracle's in the test above)
        _, oc, of = osolver.steady_state_transport_solver(
            np.zeros((s.ny, s.nx)), z, profs, s.domain, s.n, modes=s.modes, meas_pt=s.meas_pt, footprint=True,
            precision="double")
        gc, gf = (t.cpu().numpy() for t in fb.output(i))
        assert rel_l2(gf.squeeze(), of) <= 1e-10 and rel_l2(gc.squeeze(), oc) <= 1e-10
