"""High-level runs on batched plans (bldfm_amd.interface, the reference's
src/bldfm/interface.py): result format, z0 precedence, and every batched
entry point against the single-solve path on the same inputs (bitwise: the
same kernels run on device-evaluated profiles either way); the disk cache
returns what was stored.
"""

import numpy as np
import pytest

from oracle import pbl as opbl
from oracle import solver as osolver
from conftest import rel_l2

pytestmark = pytest.mark.gpu


def _config(footprint=True, z0=None, n=3, towers=2, precision="double"):
    from bldfm_amd.config_parser import parse_config_dict

    met = {"ustar": [0.3, 0.45, 0.6][:n], "mol": [-50.0, 200.0, -1000.0][:n],
           "wind_speed": [4.0, 6.0, 3.0][:n], "wind_dir": [270.0, 200.0, 45.0][:n]}
    if z0 is not None:
        met["z0"] = z0
    return parse_config_dict({
        "domain": {"nx": 64, "ny": 48, "xmax": 400.0, "ymax": 300.0, "nz": 12, "modes": [64, 48],
                   "ref_lat": 50.95, "ref_lon": 11.586},
        "towers": [{"name": f"T{k}", "lat": 50.95 + 0.0004 * k, "lon": 11.5885 + 0.0003 * k, "z_m": 6.0 + 2 * k}
                   for k in range(towers)],
        "met": met,
        "solver": {"closure": "MOST", "footprint": footprint, "precision": precision},
    })


@pytest.mark.parametrize("footprint", [True, False])
def test_batched_runs_equal_single_runs(footprint):
    from bldfm_amd import interface as itf

    cfg = _config(footprint=footprint)
    multi = itf.run_bldfm_multitower(cfg)
    for strategy in ("towers", "time", "both"):
        par = itf.run_bldfm_parallel(cfg, parallel_over=strategy)
        for t in cfg.towers:
            for a, b in zip(par[t.name], multi[t.name]):
                assert np.array_equal(a["flx"], b["flx"]) and np.array_equal(a["conc"], b["conc"])
    for t in cfg.towers:
        series = itf.run_bldfm_timeseries(cfg, t)
        assert len(series) == cfg.met.n_timesteps
        for i, r in enumerate(series):
            s = itf.run_bldfm_single(cfg, t, met_index=i)
            assert set(r) == {"grid", "conc", "flx", "tower_name", "tower_xy", "timestamp", "params"}
            assert r["tower_name"] == t.name and r["timestamp"] == i and r["params"] == cfg.met.get_step(i)
            assert r["flx"].shape == (cfg.domain.ny, cfg.domain.nx) and r["flx"].dtype == s["flx"].dtype
            for k in range(3):
                assert np.array_equal(r["grid"][k], s["grid"][k])
            assert np.array_equal(r["flx"], s["flx"]) and np.array_equal(r["conc"], s["conc"])
            assert np.array_equal(multi[t.name][i]["flx"], r["flx"])
    with pytest.raises(ValueError, match="Unknown parallel_over"):
        itf.run_bldfm_parallel(cfg, parallel_over="ranks")


def test_single_run_matches_oracle_and_z0_precedence():
    """run_bldfm_single against the oracle workflow; with z0 configured the
    profiles come from z0 (u* derived), as in src/bldfm/interface.py:76-95."""
    from bldfm_amd import interface as itf
    from bldfm_amd.utils import ideal_source

    cfg = _config(z0=0.2, n=1, towers=1)
    t = cfg.towers[0]
    r = itf.run_bldfm_single(cfg, t)
    step = cfg.met.get_step(0)
    wd = np.deg2rad(step["wind_dir"])
    wind = (-step["wind_speed"] * np.sin(wd), -step["wind_speed"] * np.cos(wd))
    z, prof = opbl.vertical_profiles(12, t.z_m, wind, z0=0.2, mol=step["mol"])
    _, oc, of = osolver.steady_state_transport_solver(
        ideal_source((64, 48), (400.0, 300.0)), z, prof, (400.0, 300.0), 12, modes=(64, 48),
        meas_pt=(t.x, t.y), footprint=True, precision="double")
    assert rel_l2(r["flx"], of) < 1e-10 and rel_l2(r["conc"], oc) < 1e-10


def test_cache_hit_returns_stored_result(tmp_path):
    from bldfm_amd import interface as itf
    from bldfm_amd.cache import GreensFunctionCache

    cfg = _config(n=1, towers=1)
    cache = GreensFunctionCache(tmp_path / "cache")
    r1 = itf.run_bldfm_single(cfg, cfg.towers[0], cache=cache)
    assert len(list((tmp_path / "cache").glob("*.npz"))) == 1
    r2 = itf.run_bldfm_single(cfg, cfg.towers[0], cache=cache)
    assert np.array_equal(r1["flx"], r2["flx"]) and np.array_equal(r1["conc"], r2["conc"])


def test_cli_run_writes_results(tmp_path):
    from bldfm_amd import interface as itf
    from bldfm_amd.cli import main

    p = tmp_path / "c.yaml"
    p.write_text("domain: {nx: 64, ny: 48, xmax: 400.0, ymax: 300.0, nz: 12, modes: [64, 48], "
                 "ref_lat: 50.95, ref_lon: 11.586}\n"
                 "towers: [{name: A, lat: 50.9502, lon: 11.5885, z_m: 6.0}]\n"
                 "met: {ustar: [0.3, 0.45], mol: [-50.0, 200.0], wind_speed: [4.0, 6.0], wind_dir: [270.0, 200.0]}\n"
                 "solver: {footprint: true, precision: double}\n")
    out = tmp_path / "r.npz"
    assert main(["run", str(p), "--out", str(out)]) == 0
    from bldfm_amd.config_parser import load_config

    cfg = load_config(p)
    ref = itf.run_bldfm_timeseries(cfg, cfg.towers[0])
    with np.load(out) as d:
        for i in range(2):
            assert np.array_equal(d[f"A/{i}/flx"], ref[i]["flx"])
            assert np.array_equal(d[f"A/{i}/X"], ref[i]["grid"][0])


# ------------------------------------------------------------------ the reference's timeseries checks
# Its tests/test_interface.py:151-250 on its conftest configuration
# (tests/conftest.py:77-110): one tower, three met steps, default precision.

def _timeseries_config():
    from bldfm_amd.config_parser import parse_config_dict

    return parse_config_dict({
        "domain": {"nx": 64, "ny": 128, "xmax": 100.0, "ymax": 700.0, "nz": 16, "modes": [64, 128],
                   "ref_lat": 50.95, "ref_lon": 11.586},
        "towers": [{"name": "tower_A", "lat": 50.95, "lon": 11.5867, "z_m": 10.0}],
        "met": {"ustar": [0.5, 0.4, 0.6], "mol": [-100.0, -200.0, -50.0], "wind_speed": [6.0, 5.0, 7.0],
                "wind_dir": [0.0, 15.0, 345.0]},
        "solver": {"closure": "MOST", "footprint": True},
    })


@pytest.fixture(scope="module")
def timeseries():
    from bldfm_amd import interface as itf

    cfg = _timeseries_config()
    return cfg, itf.run_bldfm_timeseries(cfg, cfg.towers[0])


def test_timeseries_met_params_vary(timeseries):
    cfg, results = timeseries
    for r in results:
        for key in ("ustar", "mol", "wind_speed", "wind_dir"):
            assert key in r["params"]
    for key in ("ustar", "mol", "wind_speed", "wind_dir"):
        assert len({r["params"][key] for r in results}) > 1


def test_timeseries_footprints_evolve_and_match_the_oracle(timeseries):
    """The reference's shape/positivity/evolution checks, and each step
    against the oracle chain from the same met values (the reference's
    default precision "single": rel-L2 <= 1e-6)."""
    cfg, results = timeseries
    t = cfg.towers[0]
    for i, r in enumerate(results):
        assert r["flx"].shape == (cfg.domain.ny, cfg.domain.nx)
        assert isinstance(r["grid"], tuple) and len(r["grid"]) == 3
        assert r["flx"].max() > 0
        step = cfg.met.get_step(i)
        wd = np.deg2rad(step["wind_dir"])
        wind = (-step["wind_speed"] * np.sin(wd), -step["wind_speed"] * np.cos(wd))
        z, prof = opbl.vertical_profiles(16, t.z_m, wind, ustar=step["ustar"], mol=step["mol"])
        _, oc, of = osolver.steady_state_transport_solver(
            np.zeros((128, 64)), z, prof, (100.0, 700.0), 16, modes=(64, 128), meas_pt=(t.x, t.y), footprint=True)
        assert r["flx"].dtype == of.dtype and rel_l2(r["flx"], of) <= 1e-6 and rel_l2(r["conc"], oc) <= 1e-6
    assert any(not np.allclose(results[0]["flx"], results[i]["flx"]) for i in range(1, len(results)))


def _percentile_contour(flx, grid, pct):
    """Level and enclosed area of the smallest region holding pct of the
    positive flux (the quantity abl-tk's extract_percentile_contour returns)."""
    X, Y = grid[0], grid[1]
    cell = abs(X[0, 1] - X[0, 0]) * abs(Y[1, 0] - Y[0, 0])
    v = np.sort(flx[flx > 0].ravel())[::-1]
    k = int(np.searchsorted(np.cumsum(v), pct * v.sum()))
    return v[min(k, v.size - 1)], (k + 1) * cell


def test_timeseries_aggregated_footprint(timeseries):
    _, results = timeseries
    grid = results[0]["grid"]
    mean_flx = np.mean([r["flx"] for r in results], axis=0)
    assert mean_flx.shape == results[0]["flx"].shape and np.isfinite(mean_flx).all()
    level_50, area_50 = _percentile_contour(mean_flx, grid, 0.5)
    level_70, area_70 = _percentile_contour(mean_flx, grid, 0.7)
    assert level_50 > 0 and area_50 > 0 and level_70 > 0 and area_70 > 0
    assert area_70 >= area_50 and level_50 >= level_70
