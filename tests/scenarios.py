"""Solve scenarios shared by the oracle (CPU) and GPU parity tests.

Each scenario is (profile kwargs, solver kwargs, source spec).  The
reference fixtures' inputs follow tests/conftest.py:196-262 and
src/bldfm/interface.py:74-128 of the reference; the benchmark configs follow
the workload definitions (C1: 128^2 modes, 32 levels; C2: 512^2 modes, 65
levels to ~100 m).
"""

import numpy as np

from conftest import TOWER_X


def source(spec, nx, ny, domain):
    kind = spec[0]
    if kind == "zeros":
        return np.zeros([ny, nx])
    if kind == "ideal":
        from oracle.pbl import ideal_source

        return ideal_source((nx, ny), domain, *spec[1:])
    if kind == "gaussians":  # sum of seeded Gaussians
        _, seed, count, sig_lo, sig_hi, a_lo, a_hi = spec
        rng = np.random.default_rng(seed)
        x = np.linspace(0, domain[0], nx, endpoint=False)
        y = np.linspace(0, domain[1], ny, endpoint=False)
        X, Y = np.meshgrid(x, y)
        q = np.zeros([ny, nx])
        for _ in range(count):
            cx, cy = rng.uniform(0, domain[0]), rng.uniform(0, domain[1])
            s = rng.uniform(sig_lo, sig_hi)
            a = rng.uniform(a_lo, a_hi)
            q += a * np.exp(-((X - cx) ** 2 + (Y - cy) ** 2) / (2 * s * s))
        return q
    if kind == "gaussian":
        _, cx, cy, s, a = spec
        x = np.linspace(0, domain[0], nx, endpoint=False)
        y = np.linspace(0, domain[1], ny, endpoint=False)
        X, Y = np.meshgrid(x, y)
        return a * np.exp(-((X - cx) ** 2 + (Y - cy) ** 2) / (2 * s * s))
    if kind == "random":
        return np.random.default_rng(spec[1]).random((ny, nx))
    raise ValueError(kind)


def _psi_bd(x):
    """Businger-Dyer psi (input generation only)."""
    if x > 0:
        return 5.0 * x
    xi = (1.0 - 16.0 * x) ** 0.25
    return -2 * np.log(0.5 * (1 + xi)) - np.log(0.5 * (1 + xi * xi)) + 2 * np.arctan(xi) - 0.5 * np.pi


def c2_speed(mol, ustar=0.3, z0=0.05, zm=15.0):
    """|U(zm)| for which the MOST log law gives both u* and z0 (C2 inputs)."""
    return ustar / 0.4 * (np.log(zm / z0) + _psi_bd(zm / mol))


def met_wind(speed, direction):
    th = np.deg2rad(direction)
    return (-speed * np.sin(th), -speed * np.cos(th))


SCENARIOS = {
    # reference regression fixtures (precision "single", as generated)
    "source_area": (
        dict(n=16, meas_height=10.0, wind=(0.0, -6.0), ustar=0.5),
        dict(nxy=(64, 128), domain=(100.0, 700.0), levels=16, modes=(64, 128), meas_pt=(50.0, 0.0), footprint=True),
        ("zeros",),
    ),
    "plume_3d": (
        dict(n=16, meas_height=10.0, wind=(6.0, 0.0), ustar=0.4),
        dict(nxy=(64, 32), domain=(800.0, 100.0), levels=np.arange(0, 17, 2), modes=(64, 32), meas_pt=(400.0, 50.0)),
        ("ideal",),
    ),
    "single_footprint": (
        dict(n=16, meas_height=10.0, wind=met_wind(6.0, 0.0), ustar=0.5, mol=-100.0, closure="MOST"),
        dict(nxy=(64, 128), domain=(100.0, 700.0), levels=16, modes=[64, 128], meas_pt=(TOWER_X, 0.0), footprint=True),
        ("ideal",),
    ),
    "timeseries_step2": (
        dict(n=16, meas_height=10.0, wind=met_wind(7.0, 345.0), ustar=0.6, mol=-50.0, closure="MOST"),
        dict(nxy=(64, 128), domain=(100.0, 700.0), levels=16, modes=[64, 128], meas_pt=(TOWER_X, 0.0), footprint=True),
        ("ideal",),
    ),
    # C1: 128^2 modes, 32 levels (n=19), Gaussian 300 m upwind of (500, 500)
    "c1_dispersion": (
        dict(n=19, meas_height=10.0, wind=met_wind(5.0, 270.0), ustar=0.4, mol=-50.0),
        dict(nxy=(128, 128), domain=(1000.0, 1000.0), levels=19, modes=(128, 128), precision="double"),
        ("gaussian", 200.0, 500.0, 50.0, 1.0),
    ),
    "c1_footprint": (
        dict(n=19, meas_height=10.0, wind=met_wind(5.0, 270.0), ustar=0.4, mol=-50.0),
        dict(nxy=(128, 128), domain=(1000.0, 1000.0), levels=19, modes=(128, 128), meas_pt=(500.0, 500.0), footprint=True, precision="double"),
        ("zeros",),
    ),
    # C2 stability cases at production size (footprint level zm = 15 m)
    "c2_unstable_footprint": (
        dict(n=26, meas_height=15.0, wind=met_wind(c2_speed(-20.0), 270.0), ustar=0.3, mol=-20.0, domain_height=100.0),
        dict(nxy=(512, 512), domain=(2000.0, 2000.0), levels=26, modes=(512, 512), meas_pt=(1000.0, 1000.0), footprint=True, precision="double"),
        ("zeros",),
    ),
    "c2_stable_footprint": (
        dict(n=26, meas_height=15.0, wind=met_wind(c2_speed(100.0), 270.0), ustar=0.3, mol=100.0, domain_height=100.0),
        dict(nxy=(512, 512), domain=(2000.0, 2000.0), levels=26, modes=(512, 512), meas_pt=(1000.0, 1000.0), footprint=True, precision="double"),
        ("zeros",),
    ),
    # C2 3-D dispersion: 16 seeded Gaussians (sigma U(20,150) m, amplitude U(0.5,2)), every level
    "c2_unstable_dispersion": (
        dict(n=26, meas_height=15.0, wind=met_wind(c2_speed(-20.0), 270.0), ustar=0.3, mol=-20.0, domain_height=100.0),
        dict(nxy=(512, 512), domain=(2000.0, 2000.0), levels=list(range(65)), modes=(512, 512), precision="double"),
        ("gaussians", 20261208, 16, 20.0, 150.0, 0.5, 2.0),
    ),
    "c2_stable_dispersion": (
        dict(n=26, meas_height=15.0, wind=met_wind(c2_speed(100.0), 270.0), ustar=0.3, mol=100.0, domain_height=100.0),
        dict(nxy=(512, 512), domain=(2000.0, 2000.0), levels=list(range(65)), modes=(512, 512), precision="double"),
        ("gaussians", 20261208, 16, 20.0, 150.0, 0.5, 2.0),
    ),
}


def inputs(name, vertical_profiles, **overrides):
    """(srf_flx, z, profiles, domain, levels, solver kwargs) of a scenario,
    with profiles from the given ``vertical_profiles``."""
    pk, sk, src = SCENARIOS[name]
    sk = dict(sk)
    sk.update(overrides)
    nx, ny = sk.pop("nxy")
    z, profs = vertical_profiles(**pk)
    q0 = source(src, nx, ny, sk["domain"])
    domain, levels = sk.pop("domain"), sk.pop("levels")
    return q0, z, profs, domain, levels, sk


def run(name, vertical_profiles, solver, **overrides):
    """Run scenario ``name``: profiles from ``vertical_profiles``, solve with
    ``solver`` (a steady_state_transport_solver)."""
    q0, z, profs, domain, levels, sk = inputs(name, vertical_profiles, **overrides)
    return solver(q0, z, profs, domain, levels, **sk)
