"""CPU tests of the configuration schema (bldfm_amd.config_parser, the
reference's src/bldfm/config_parser.py) and of the Green's-function cache
key and file format (bldfm_amd.cache, src/bldfm/cache.py)."""

import hashlib

import numpy as np
import pytest

from conftest import TOWER_X
from bldfm_amd.cache import GreensFunctionCache
from bldfm_amd.config_parser import (BLDFMConfig, MetConfig, TowerConfig, latlon_to_xy, load_config,
                                     parse_config_dict)


@pytest.mark.parametrize("lat, lon, ref_lat, ref_lon, ex, ey, tol", [
    (50.0, 11.0, 50.0, 11.0, 0.0, 0.0, 0.01),
    (51.0, 11.0, 50.0, 11.0, 0.0, 111_194.9, 500),
    (50.0, 12.0, 50.0, 11.0, 71_696.0, 0.0, 500),
])
def test_latlon_to_xy(lat, lon, ref_lat, ref_lon, ex, ey, tol):
    x, y = latlon_to_xy(lat, lon, ref_lat, ref_lon)
    assert abs(x - ex) < tol and abs(y - ey) < tol


def test_tower_xy_reproduces_fixture_tower():
    """The tower position the reference's regression fixtures use."""
    t = TowerConfig(name="t", lat=50.95, lon=11.5867, z_m=10.0)
    t.compute_local_xy(50.95, 11.586)
    assert t.x == pytest.approx(TOWER_X, rel=1e-12) and t.y == 0.0


def test_met_config():
    m = MetConfig(ustar=[0.3, 0.4], mol=[-100.0, -200.0], wind_speed=5.0, wind_dir=[270.0, 180.0])
    m.validate()
    assert m.n_timesteps == 2
    s = m.get_step(1)
    assert (s["ustar"], s["mol"], s["wind_speed"], s["wind_dir"], s["timestamp"]) == (0.4, -200.0, 5.0, 180.0, 1)
    with pytest.raises(ValueError, match="same length"):
        MetConfig(ustar=[0.3, 0.4], mol=[-100.0, -200.0, 500.0]).validate()
    with pytest.raises(ValueError, match="at least one of"):
        MetConfig().validate()
    with pytest.raises(ValueError, match="timestamps length"):
        MetConfig(ustar=[0.3, 0.4], timestamps=["a"]).validate()
    assert "z0" in MetConfig(z0=0.1).get_step(0)


def test_parse_config(tmp_path):
    raw = {"domain": {"nx": 256, "ny": 128, "xmax": 1000.0, "ymax": 500.0, "nz": 32, "ref_lat": 50.95,
                      "ref_lon": 11.586},
           "towers": [{"name": "A", "lat": 50.9505, "lon": 11.5865, "z_m": 10.0}],
           "met": {"ustar": 0.4},
           "solver": {"src_loc": [10, 20], "footprint": True}}
    c = parse_config_dict(raw)
    assert isinstance(c, BLDFMConfig)
    assert c.domain.modes == (512, 512) and c.solver.closure == "MOST" and c.output.format == "netcdf"
    assert c.solver.src_loc == (10, 20) and c.solver.footprint and c.parallel.max_workers == 1
    assert 0 < c.towers[0].x < 100 and 0 < c.towers[0].y < 100
    for key in ("domain", "towers", "met"):
        bad = dict(raw)
        del bad[key]
        with pytest.raises(ValueError, match=key):
            parse_config_dict(bad)
    p = tmp_path / "c.yaml"
    p.write_text("domain: {nx: 64, ny: 32, xmax: 200.0, ymax: 100.0, nz: 8, modes: [64, 32]}\n"
                 "towers: [{name: T, lat: 0.0, lon: 0.0, z_m: 5.0}]\n"
                 "met: {ustar: [0.3, 0.4], mol: [-50.0, 100.0], wind_speed: [4.0, 5.0], wind_dir: [270.0, 90.0]}\n")
    c = load_config(p)
    assert c.domain.modes == (64, 32) and c.met.n_timesteps == 2
    with pytest.raises(FileNotFoundError):
        load_config(tmp_path / "missing.yaml")


def _inputs():
    z = np.linspace(0.1, 10.0, 8)
    profiles = tuple(np.linspace(1.0, 2.0, 8) * (k + 1) for k in range(5))
    return z, profiles, (200.0, 100.0), (64, 32), (50.0, 25.0), None, "double"


def test_cache_key_is_the_reference_recipe():
    """SHA-256 over z, the five profiles, domain, modes, meas_pt (numpy
    buffers), str(halo), precision: cache files are shared with the reference."""
    z, profiles, domain, modes, meas_pt, halo, precision = _inputs()
    h = hashlib.sha256()
    for part in [z, *profiles, np.asarray(domain), np.asarray(modes), np.asarray(meas_pt)]:
        h.update(np.asarray(part).tobytes())
    h.update(b"None")
    h.update(b"double")
    assert GreensFunctionCache._compute_key(*_inputs()) == h.hexdigest()


def test_cache_roundtrip(tmp_path):
    c = GreensFunctionCache(tmp_path / "c")
    args = _inputs()
    assert c.get(*args) is None
    grid = tuple(np.random.default_rng(k).normal(size=(4, 4)) for k in range(3))
    conc, flx = np.ones((4, 4)), np.full((4, 4), 2.0)
    c.put(*args, grid, conc, flx)
    g2, c2, f2 = c.get(*args)
    assert all(np.array_equal(a, b) for a, b in zip(g2, grid))
    assert np.array_equal(c2, conc) and np.array_equal(f2, flx)
    other = list(args)
    other[4] = (10.0, 10.0)
    assert c.get(*other) is None
    assert c.clear() == 1 and c.get(*args) is None


def test_cli_dry_run(tmp_path, caplog):
    import logging

    from bldfm_amd.cli import main

    p = tmp_path / "c.yaml"
    p.write_text("domain: {nx: 64, ny: 32, xmax: 200.0, ymax: 100.0, nz: 8, ref_lat: 50.95, ref_lon: 11.586}\n"
                 "towers: [{name: T, lat: 50.9503, lon: 11.5865, z_m: 5.0}]\n"
                 "met: {ustar: 0.4, mol: -50.0}\n")
    with caplog.at_level(logging.INFO, logger="bldfm_amd.cli"):
        assert main(["run", str(p), "--dry-run"]) == 0
    text = caplog.text
    assert "Dry run complete" in text and "Tower 'T'" in text and "Timesteps: 1" in text


def test_autotune_shape_key():
    """The autotuner's cache key ignores instance order and tracks geometry."""
    import torch

    from bldfm_amd import _lib, autotune

    def inst(nx, nz, footprint=1):
        i = _lib.Instance()
        i.nx, i.ny, i.nlx, i.nly, i.nz, i.nslot, i.footprint = nx, nx, nx, nx, nz, 1, footprint
        return i

    cpu = torch.device("cpu")
    a = [inst(64, 20), inst(128, 30)]
    assert autotune.shape_key(a, cpu) == autotune.shape_key(list(reversed(a)), cpu)
    assert autotune.shape_key(a, cpu) != autotune.shape_key([inst(64, 20), inst(128, 31)], cpu)
    assert autotune.shape_key(a, cpu) != autotune.shape_key([inst(64, 20), inst(128, 30, 0)], cpu)


def test_fft_thread_candidates():
    """FFT workgroup sizes the autotuner times: the default (0) and larger
    multiples of 64 up to the kernel family's bound (256 threads up to 2048
    points, else 1024), from the longest padded transform of the batch."""
    from bldfm_amd import _lib, autotune

    def inst(n, pad):
        i = _lib.Instance()
        i.nx, i.ny, i.px, i.py = n, n, pad, pad
        return i

    assert autotune.fft_thread_candidates([inst(512, 512)]) == [0, 256]  # 1536 points: default 192
    assert autotune.fft_thread_candidates([inst(128, 128)]) == [0, 128]  # 384 points: default 64
    assert autotune.fft_thread_candidates([inst(256, 256), inst(128, 128)]) == [0, 192, 256]  # 768: 128
    assert autotune.fft_thread_candidates([inst(2048, 0)]) == [0]  # 2048 points: 256 already
    assert autotune.fft_thread_candidates([inst(1000, 1000)]) == [0, 384, 640]  # 3000 points: 320


def test_autotune_cache_entries_get_new_coordinates(tmp_path):
    """A cached choice written before a launch-shape coordinate existed comes
    back with that coordinate at its default (no device work on a hit)."""
    import json

    import torch

    from bldfm_amd import _lib, autotune

    i = _lib.Instance()
    i.nx, i.ny, i.nlx, i.nly, i.nz, i.nslot, i.footprint = 64, 64, 64, 64, 20, 1, 1
    cpu = torch.device("cpu")
    key = autotune.shape_key([i], cpu)
    path = tmp_path / "tune.json"
    path.write_text(json.dumps({key: {"shoot_block": 128, "xpass_layout": 1, "best_ms": 1.0, "ms": {}}}))
    autotune.clear_memo()
    r = autotune.tune([i], None, cpu, cache_path=str(path))
    assert r["cached"] and r["shoot_block"] == 128 and r["xpass_layout"] == 1 and r["fft_threads"] == 0
    r2 = autotune.tune([i], None, cpu, cache_path=str(path))  # now from memory
    assert r2["cached"] and r2["fft_threads"] == 0
    autotune.clear_memo()
