"""High-level runs (drop-in for src/bldfm/interface.py) on batched GPU plans.

``run_bldfm_single`` repeats the reference's three-step workflow
(src/bldfm/interface.py:31-138) through the GPU drop-ins.  The multi-run
entry points keep their signatures and result format (a dict per run with
grid, conc, flx, tower_name, tower_xy, timestamp, params; a tower-name ->
list mapping for several towers) but replace the reference's loop of
single solves and its ``ProcessPoolExecutor`` fan-out
(src/bldfm/interface.py:141-326) by batched plans: every (tower, timestep)
pair of a chunk is one instance of one plan, with its profiles evaluated
from the met scalars on the device, and one launch sequence solves them
all.  Results are identical to the single-solve path (same kernels).
"""

import ctypes
import logging
import warnings

import numpy as np
import torch

from . import _lib
from .config_parser import BLDFMConfig, TowerConfig
from .pbl_model import vertical_profiles
from .solver import _geometry, _instance, _level_maps, steady_state_transport_solver
from .utils import compute_wind_fields, ideal_source

logger = logging.getLogger("bldfm_amd.interface")

CHUNK = 128  # (tower, timestep) pairs per plan


def _make_cache(config):
    if config.parallel.use_cache and config.solver.footprint:
        from .cache import GreensFunctionCache

        return GreensFunctionCache()
    return None


def _levels(dom):
    if dom.output_levels:
        return dom.output_levels
    if dom.full_output:
        return list(range(dom.nz + 1))
    return dom.nz


def _profile_kwargs(met_step):
    """z0 takes precedence over ustar (src/bldfm/interface.py:76-95)."""
    if met_step.get("z0") is not None:
        return {"z0": met_step["z0"]}
    return {"ustar": met_step["ustar"]}


def _surface_flux(config, surface_flux):
    if surface_flux is not None:
        return surface_flux
    dom, sol = config.domain, config.solver
    return ideal_source((dom.nx, dom.ny), (dom.xmax, dom.ymax), src_loc=sol.src_loc, shape=sol.surface_flux_shape)


def _result(grid, conc, flx, tower, met_step):
    return {"grid": grid, "conc": conc, "flx": flx, "tower_name": tower.name, "tower_xy": (tower.x, tower.y),
            "timestamp": met_step["timestamp"], "params": met_step}


def run_bldfm_single(config: BLDFMConfig, tower: TowerConfig, met_index: int = 0,
                     surface_flux: np.ndarray = None, cache=None) -> dict:
    """One solve for one tower at one timestep (src/bldfm/interface.py:31-138)."""
    dom, sol = config.domain, config.solver
    met_step = config.met.get_step(met_index)
    u, v = compute_wind_fields(met_step["wind_speed"], met_step["wind_dir"])
    z, profiles = vertical_profiles(n=dom.nz, meas_height=tower.z_m, wind=(u, v), mol=met_step["mol"],
                                    closure=sol.closure, **_profile_kwargs(met_step))
    grid, conc, flx = steady_state_transport_solver(
        srf_flx=_surface_flux(config, surface_flux), z=z, profiles=profiles, domain=(dom.xmax, dom.ymax),
        levels=_levels(dom), modes=dom.modes, meas_pt=(tower.x, tower.y), footprint=sol.footprint,
        analytic=sol.analytic, halo=dom.halo, precision=sol.precision, cache=cache)
    return _result(grid, conc, flx, tower, met_step)


class _PairPlan:
    """One plan over a chunk of (tower, met step) pairs sharing a configuration."""

    def __init__(self, config, pairs, srf_flx):
        lib = _lib.load()
        dom, sol = config.domain, config.solver
        self.dev = _lib.device()
        self.config, self.pairs, self.srf_flx = config, pairs, srf_flx
        steps = [config.met.get_step(i) for _, i in pairs]
        u, v = compute_wind_fields(np.array([s["wind_speed"] for s in steps], dtype=np.float64),
                                   np.array([s["wind_dir"] for s in steps], dtype=np.float64))
        u, v = np.atleast_1d(u), np.atleast_1d(v)
        insts, ints, self.geo = [], [], []
        for k, ((tower, _), step) in enumerate(zip(pairs, steps)):
            pk = _profile_kwargs(step)
            met = _lib.make_met(dom.nz, tower.z_m, (u[k], v[k]), pk.get("ustar"), pk.get("z0"), step["mol"], 1.0,
                                sol.closure, None, None, None)
            if pk.get("z0") is not None and pk.get("ustar") is not None:
                raise ValueError("Either z0 or ustar must be provided.")
            nlev = _lib._I32()
            _lib.check(lib.bldfm_profile_query(ctypes.byref(met), ctypes.byref(nlev), None, None), "profile_query")
            nz = nlev.value
            levels = _levels(dom)
            lv = np.array([levels]) if np.ndim(levels) == 0 else np.asarray(levels)
            for li in lv.ravel():  # the reference indexes z[levels]
                if not -nz <= int(li) < nz:
                    raise IndexError(f"index {li} is out of bounds for axis 0 with size {nz}")
            g = _geometry(srf_flx.shape, np.zeros(nz), (dom.xmax, dom.ymax), levels, dom.modes, (tower.x, tower.y),
                          sol.footprint, dom.halo, sol.precision)
            slot, idx = _level_maps(g["levels"], nz)
            inst = _instance(g, len(g["levels"]), sol.footprint, False, sol.precision, 0.0)
            inst.prof_from_met = 1
            inst.met = met
            insts.append(inst)
            ints.append(np.concatenate([slot, idx]))
            self.geo.append(g)
        f32 = {g["out_f32"] for g in self.geo}
        if len(f32) != 1:
            raise ValueError("pairs of one plan must share the output precision")
        self.out_dtype = torch.float32 if f32.pop() else torch.float64
        self.insts = insts
        self.plan = _lib.Plan(insts, np.concatenate(ints).astype(np.int32), self.dev)

    def run(self):
        p = self.plan
        dev = self.dev
        prof = torch.empty(max(p.prof_doubles, 1), dtype=torch.float64, device=dev)
        src = None
        if p.src_doubles > 0:
            one = torch.as_tensor(np.ascontiguousarray(self.srf_flx, dtype=np.float64), device=dev).reshape(-1)
            src = one.repeat(p.src_doubles // one.numel())
        conc = torch.empty(max(p.out_elems, 1), dtype=self.out_dtype, device=dev)
        flx = torch.empty_like(conc)
        ws = torch.empty(max(p.workspace_bytes, 1), dtype=torch.uint8, device=dev)
        p.execute(prof, src, conc, flx, ws)
        prof_h, conc_h, flx_h = prof.cpu().numpy(), conc.cpu().numpy(), flx.cpu().numpy()
        dom = self.config.domain
        out = []
        for k, ((tower, i), g) in enumerate(zip(self.pairs, self.geo)):
            op, _, oo, nz = p.offsets[k]
            z = prof_h[op : op + nz]
            n = len(g["levels"]) * g["nyo"] * g["nxo"]
            shape = (len(g["levels"]), g["nyo"], g["nxo"])
            c = conc_h[oo : oo + n].reshape(shape)
            f = flx_h[oo : oo + n].reshape(shape)
            x = np.linspace(0, dom.xmax, g["nx"], endpoint=False)
            y = np.linspace(0, dom.ymax, g["ny"], endpoint=False)
            Zg, Yg, Xg = np.meshgrid(z[g["levels"]], y, x, indexing="ij")
            grid = (np.squeeze(Xg), np.squeeze(Yg), np.squeeze(Zg))
            out.append(_result(grid, np.squeeze(c), np.squeeze(f), tower, self.config.met.get_step(i)))
        return out


def _run_pairs(config, pairs, surface_flux=None, chunk=CHUNK):
    """Results of (tower, met index) pairs in order, chunk pairs per plan."""
    if config.solver.analytic:  # closed form: one cheap launch per pair
        return [run_bldfm_single(config, t, i, surface_flux) for t, i in pairs]
    srf = _surface_flux(config, surface_flux)
    out = []
    for a in range(0, len(pairs), chunk):
        out.extend(_PairPlan(config, pairs[a : a + chunk], srf).run())
    return out


def run_bldfm_timeseries(config: BLDFMConfig, tower: TowerConfig, surface_flux: np.ndarray = None) -> list:
    """Every met timestep for one tower (src/bldfm/interface.py:141-174), batched."""
    n = config.met.n_timesteps
    cache = _make_cache(config)
    if cache is None:
        return _run_pairs(config, [(tower, i) for i in range(n)], surface_flux)
    # with the disk cache: solve the misses as one batch, then store them
    out = [None] * n
    miss = []
    for i in range(n):
        r = _cached(config, tower, i, cache)
        if r is None:
            miss.append(i)
        else:
            out[i] = r
    for i, r in zip(miss, _run_pairs(config, [(tower, i) for i in miss], surface_flux)):
        _store(config, tower, i, cache, r)
        out[i] = r
    return out


def _cache_key_args(config, tower, i):
    dom, sol = config.domain, config.solver
    step = config.met.get_step(i)
    u, v = compute_wind_fields(step["wind_speed"], step["wind_dir"])
    z, profiles = vertical_profiles(n=dom.nz, meas_height=tower.z_m, wind=(u, v), mol=step["mol"],
                                    closure=sol.closure, **_profile_kwargs(step))
    return (z, profiles, (dom.xmax, dom.ymax), dom.modes, (tower.x, tower.y), dom.halo, sol.precision), step


def _cached(config, tower, i, cache):
    args, step = _cache_key_args(config, tower, i)
    hit = cache.get(*args)
    return None if hit is None else _result(hit[0], hit[1], hit[2], tower, step)


def _store(config, tower, i, cache, r):
    args, _ = _cache_key_args(config, tower, i)
    cache.put(*args, r["grid"], r["conc"], r["flx"])


def run_bldfm_multitower(config: BLDFMConfig, surface_flux: np.ndarray = None) -> dict:
    """All towers and timesteps (src/bldfm/interface.py:177-206), batched over both."""
    if config.parallel.use_cache and config.solver.footprint:
        return {t.name: run_bldfm_timeseries(config, t, surface_flux) for t in config.towers}
    n = config.met.n_timesteps
    pairs = [(t, i) for t in config.towers for i in range(n)]
    flat = _run_pairs(config, pairs, surface_flux)
    return {t.name: flat[k * n : (k + 1) * n] for k, t in enumerate(config.towers)}


def run_bldfm_parallel(config: BLDFMConfig, max_workers: int = None, parallel_over: str = "towers",
                       surface_flux: np.ndarray = None) -> dict:
    """The reference's process-parallel entry point (src/bldfm/interface.py:241-326).

    Every strategy runs the same batched plans on the GPU; ``max_workers``
    is accepted and ignored.  As in the reference, a given ``surface_flux``
    is not used here (its workers build their own ideal source)."""
    if parallel_over not in ("towers", "time", "both"):
        raise ValueError(f"Unknown parallel_over={parallel_over!r}. Choose 'towers', 'time', or 'both'.")
    if surface_flux is not None:
        warnings.warn("surface_flux is not passed to worker processes in parallel mode; each worker generates its "
                      "own ideal source. The provided array will be ignored.")
    n = config.met.n_timesteps
    pairs = [(t, i) for t in config.towers for i in range(n)]
    flat = _run_pairs(config, pairs, None)
    return {t.name: flat[k * n : (k + 1) * n] for k, t in enumerate(config.towers)}
