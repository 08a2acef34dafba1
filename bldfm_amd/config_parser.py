"""YAML configuration schema of BLDFM runs (drop-in for src/bldfm/config_parser.py).

The dataclasses, field names, defaults, YAML keys and validation messages
follow the reference (src/bldfm/config_parser.py:23-318) so that existing
configuration files and code that builds configurations load unchanged.
``bldfm_amd.interface`` turns a configuration into batched GPU plans.

Tower coordinates: the reference converts lat/lon to local metres with the
private abl-tk package (``latlon_to_xy``).  Here the same equirectangular
projection about the reference origin is written out,
x = R cos(ref_lat) (lon - ref_lon), y = R (lat - ref_lat) with R = 6371 km and
angles in radians; the x convention reproduces the tower position the
reference's regression fixtures were made with.
"""

import math
from dataclasses import dataclass, field
from pathlib import Path
from typing import List, Optional, Tuple, Union

import yaml

EARTH_RADIUS = 6371000.0


def latlon_to_xy(lat, lon, ref_lat, ref_lon):
    """Local (x, y) in metres of (lat, lon) about (ref_lat, ref_lon)."""
    x = math.radians(lon - ref_lon) * EARTH_RADIUS * math.cos(math.radians(ref_lat))
    y = math.radians(lat - ref_lat) * EARTH_RADIUS
    return x, y


@dataclass
class TowerConfig:
    """A measurement tower (src/bldfm/config_parser.py:23-39)."""

    name: str
    lat: float
    lon: float
    z_m: float
    x: float = 0.0
    y: float = 0.0

    def compute_local_xy(self, ref_lat: float, ref_lon: float):
        self.x, self.y = latlon_to_xy(self.lat, self.lon, ref_lat, ref_lon)


@dataclass
class DomainConfig:
    """Computational domain (src/bldfm/config_parser.py:42-56)."""

    nx: int
    ny: int
    xmax: float
    ymax: float
    nz: int
    modes: Tuple[int, int] = (512, 512)
    halo: Optional[float] = None
    ref_lat: Optional[float] = None
    ref_lon: Optional[float] = None
    output_levels: Optional[List[int]] = None
    full_output: bool = False


@dataclass
class MetConfig:
    """Meteorological forcing, scalars for one time or lists for a series
    (src/bldfm/config_parser.py:59-128)."""

    ustar: Optional[Union[float, List[float]]] = None
    mol: Union[float, List[float]] = 1e9
    wind_speed: Union[float, List[float]] = 5.0
    wind_dir: Union[float, List[float]] = 270.0
    z0: Optional[float] = None
    timestamps: Optional[List[str]] = None

    @property
    def n_timesteps(self) -> int:
        if isinstance(self.ustar, list):
            return len(self.ustar)
        if isinstance(self.wind_speed, list):
            return len(self.wind_speed)
        return 1

    def get_step(self, i: int) -> dict:
        """Scalar met values of step i: ustar, mol, wind_speed, wind_dir,
        z0 (when configured) and timestamp (the index when none are given)."""

        def pick(val):
            if val is None:
                return None
            return val[i] if isinstance(val, list) else val

        out = {k: pick(getattr(self, k)) for k in ("ustar", "mol", "wind_speed", "wind_dir")}
        if self.z0 is not None:
            out["z0"] = self.z0
        out["timestamp"] = self.timestamps[i] if self.timestamps is not None else i
        return out

    def validate(self):
        if self.ustar is None and self.z0 is None:
            raise ValueError("MetConfig requires at least one of 'ustar' or 'z0'")
        lengths = {k: len(getattr(self, k)) for k in ("ustar", "mol", "wind_speed", "wind_dir")
                   if isinstance(getattr(self, k), list)}
        if not lengths:
            return
        if len(set(lengths.values())) > 1:
            raise ValueError(f"Met timeseries arrays must all have the same length. Got: {lengths}")
        n = next(iter(lengths.values()))
        if self.timestamps is not None and len(self.timestamps) != n:
            raise ValueError(
                f"timestamps length ({len(self.timestamps)}) does not match met array length ({n})"
            )


@dataclass
class SolverConfig:
    closure: str = "MOST"
    precision: str = "single"
    footprint: bool = False
    surface_flux_shape: str = "diamond"
    analytic: bool = False
    src_loc: Optional[Tuple[float, float]] = None


@dataclass
class OutputConfig:
    format: str = "netcdf"
    directory: str = "./output"


@dataclass
class ParallelConfig:
    """``num_threads`` and ``max_workers`` are accepted and ignored (one GPU
    runs a whole batch); ``use_cache`` enables the disk cache."""

    num_threads: int = 1
    max_workers: int = 1
    use_cache: bool = False


@dataclass
class BLDFMConfig:
    domain: DomainConfig
    towers: List[TowerConfig]
    met: MetConfig
    solver: SolverConfig = field(default_factory=SolverConfig)
    output: OutputConfig = field(default_factory=OutputConfig)
    parallel: ParallelConfig = field(default_factory=ParallelConfig)

    def __post_init__(self):
        if self.domain.ref_lat is not None and self.domain.ref_lon is not None:
            for t in self.towers:
                t.compute_local_xy(self.domain.ref_lat, self.domain.ref_lon)
        self.met.validate()


def _section(raw, cls, keys):
    if raw is None:
        return cls()
    return cls(**{k: (tuple(raw[k]) if k == "src_loc" and raw.get(k) is not None else raw.get(k, d))
                  for k, d in keys.items()})


def parse_config_dict(raw: dict) -> BLDFMConfig:
    """Configuration from a dictionary (e.g. parsed YAML)."""
    for key in ("domain", "towers", "met"):
        if key not in raw:
            raise ValueError(f"Config must include '{key}' section")
    d = raw["domain"]
    domain = DomainConfig(
        nx=d["nx"], ny=d["ny"], xmax=float(d["xmax"]), ymax=float(d["ymax"]), nz=d["nz"],
        modes=tuple(d.get("modes", [512, 512])), halo=d.get("halo"), ref_lat=d.get("ref_lat"),
        ref_lon=d.get("ref_lon"), output_levels=d.get("output_levels"), full_output=d.get("full_output", False),
    )
    towers = [TowerConfig(name=t["name"], lat=t["lat"], lon=t["lon"], z_m=t["z_m"]) for t in raw["towers"]]
    m = raw["met"]
    met = MetConfig(ustar=m.get("ustar"), mol=m.get("mol", 1e9), wind_speed=m.get("wind_speed", 5.0),
                    wind_dir=m.get("wind_dir", 270.0), z0=m.get("z0"), timestamps=m.get("timestamps"))
    solver = _section(raw.get("solver"), SolverConfig,
                      dict(closure="MOST", precision="single", footprint=False, surface_flux_shape="diamond",
                           analytic=False, src_loc=None))
    output = _section(raw.get("output"), OutputConfig, dict(format="netcdf", directory="./output"))
    parallel = _section(raw.get("parallel"), ParallelConfig, dict(num_threads=1, max_workers=1, use_cache=False))
    return BLDFMConfig(domain=domain, towers=towers, met=met, solver=solver, output=output, parallel=parallel)


def load_config(path: Union[str, Path]) -> BLDFMConfig:
    """Configuration from a YAML file (safe loader)."""
    path = Path(path)
    if not path.exists():
        raise FileNotFoundError(f"Config file not found: {path}")
    with open(path) as f:
        return parse_config_dict(yaml.safe_load(f))
