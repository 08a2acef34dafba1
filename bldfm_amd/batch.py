"""Batched footprint solves: many towers x timesteps in one plan.

The reference parallelises footprints over processes
(``run_bldfm_parallel``, src/bldfm/interface.py:241-326) and memoises them
on disk (src/bldfm/cache.py).  Here a batch of independent footprint solves
with their own grids, level counts and meteorology is one plan: the level
table kernel evaluates each instance's MOST profiles from its met scalars
on the device, one shooting launch covers every instance's modes, and the
two inverse-transform passes cover every instance's rows and columns.  The
whole batch can be captured once into a HIP graph and replayed.

Each instance is the solve

    z, profs = vertical_profiles(n, zm, wind, ustar|z0, mol, prsc, closure,
                                 domain_height, stretch, tke)
    steady_state_transport_solver(zeros((ny, nx)), z, profs, domain, levels,
                                  modes, meas_pt, footprint=True, halo=halo,
                                  precision=precision)

and gives the same values (the batch uses the same device profile and
solver code as those drop-ins).
"""

import ctypes
from dataclasses import dataclass, field
from typing import Optional, Sequence, Tuple

import numpy as np
import torch

from . import _lib


@dataclass
class FootprintSpec:
    """One footprint solve of a batch (arguments as in the reference)."""

    nx: int
    ny: int
    domain: Tuple[float, float]
    modes: Tuple[int, int]
    meas_pt: Tuple[float, float]
    n: int
    meas_height: float
    wind: Tuple[float, float]
    ustar: Optional[float] = None
    z0: Optional[float] = None
    mol: float = 1e9
    prsc: float = 1.0
    closure: str = "MOST"
    domain_height: Optional[float] = None
    stretch: Optional[float] = None
    tke: Optional[float] = None
    halo: Optional[float] = None
    levels: Optional[Sequence[int]] = None  # default: the measurement level n
    extra: dict = field(default_factory=dict)


def _instance(spec: FootprintSpec, precision: str, fast: bool, surrogate: bool = False, met=None):
    xmx, ymx = spec.domain
    nlx, nly = spec.modes
    if (nlx % 2 > 0) or (nly % 2 > 0):
        raise ValueError("modes must consist of even numbers.")
    if precision not in ("single", "double"):
        raise ValueError("precision must be single (default) or double.")
    nx, ny = spec.nx, spec.ny
    dx, dy = xmx / nx, ymx / ny
    halo = max(xmx, ymx) if spec.halo is None else spec.halo
    px, py = int(halo / dx), int(halo / dy)
    nxe, nye = nx + 2 * px, ny + 2 * py
    if nlx > nxe or nly > nye:
        nlx, nly = nxe, nye
    if spec.closure in ("CONSTANT", "MOST", "MOSTM") and spec.ustar is not None and spec.z0 is not None:
        raise ValueError("Either z0 or ustar must be provided.")
    if met is None:
        met = _lib.make_met(spec.n, spec.meas_height, spec.wind, spec.ustar, spec.z0, spec.mol, spec.prsc,
                            spec.closure, spec.domain_height, spec.stretch, spec.tke)
    lib = _lib.load()
    nlev = ctypes.c_int32()
    _lib.check(lib.bldfm_profile_query(ctypes.byref(met), ctypes.byref(nlev), None, None), "profile_query")
    inst = _lib.Instance()
    inst.nx, inst.ny, inst.nlx, inst.nly, inst.px, inst.py = nx, ny, nlx, nly, px, py
    inst.nz = nlev.value
    levels = [spec.n] if spec.levels is None else list(spec.levels)
    for lv in levels:  # the reference indexes z[levels] (src/bldfm/solver.py:296)
        if not -inst.nz <= int(lv) < inst.nz:
            raise IndexError(
                f"index {lv} is out of bounds for axis 0 with size {inst.nz} "
                f"(vertical grid of this instance has {inst.nz} points)"
            )
    inst.nslot = len(levels)
    inst.footprint = 1
    inst.analytic = 0
    inst.single = precision == "single"
    inst.out_f32 = 0
    inst.has_shift = 1
    inst.prof_from_met = 1
    inst.fast = bool(fast)
    inst.surrogate = bool(surrogate)
    inst.dx, inst.dy = dx, dy
    inst.shift_x, inst.shift_y = spec.meas_pt[0] + halo, spec.meas_pt[1] + halo
    inst.srf_bg_conc = 0.0
    inst.met = met
    slot = np.full(inst.nz, -1, dtype=np.int32)
    k = 0
    for i in range(inst.nz):
        if i in levels:
            slot[i] = k
            k += 1
    idx = np.mod(np.asarray(levels, dtype=np.int64), inst.nz).astype(np.int32)
    from .solver import out_extent

    dims = (out_extent(ny, py, nly, nye), out_extent(nx, px, nlx, nxe))
    return inst, np.concatenate([slot, idx]), dims


class FootprintBatch:
    """A batch of footprint solves sharing one plan and one set of buffers.

    ``tuning``: launch shape of the plan -- None for the defaults, a dict of
    ``bldfm_tuning`` fields, or "auto" to time the candidates on the device
    first (``bldfm_amd.autotune``; the choice is cached per batch shape).

    With ``surrogate`` (a trained ``bldfm_amd.surrogate.SurrogateNet``), each
    execution serves the instances inside the network's training hull from
    the network, re-solves one mode in 100 of each in FP64, and solves an
    instance in FP64 when that spot check's relative L2 error exceeds
    ``tol``.  Served instances produce the flux footprint (conc is 0)."""

    def __init__(self, specs: Sequence[FootprintSpec], precision: str = "double", fast: bool = False,
                 surrogate=None, tol: float = 1e-3, tuning=None):
        if len(specs) == 0:
            raise ValueError("empty batch")
        self.specs = list(specs)
        self.device = _lib.device()
        for s in self.specs:
            if s.closure in ("CONSTANT", "MOST", "MOSTM") and s.ustar is not None and s.z0 is not None:
                raise ValueError("Either z0 or ustar must be provided.")
        mets = _lib.make_mets(self.specs)
        built = [_instance(s, precision, fast, surrogate is not None, m) for s, m in zip(self.specs, mets)]
        self.instances = [b[0] for b in built]
        self.out_dims = [b[2] for b in built]  # (rows, columns) of each instance's outputs
        ints = np.concatenate([b[1] for b in built]).astype(np.int32)
        self.levels_i32 = ints
        self.tune_report = None
        if isinstance(tuning, str):
            if tuning != "auto":
                raise ValueError("tuning must be None, 'auto' or a dict of bldfm_tuning fields")
            from . import autotune

            self.tune_report = autotune.tune(self.instances, ints, self.device)
            tuning = {k: self.tune_report[k] for k in ("shoot_block", "xpass_layout", "fft_threads")}
        self.plan = _lib.Plan(self.instances, ints, self.device, tuning=tuning)
        self.tuning = self.plan.tuning
        dev = self.device
        self.prof = torch.empty(max(self.plan.prof_doubles, 1), dtype=torch.float64, device=dev)
        self.conc = torch.empty(max(self.plan.out_elems, 1), dtype=torch.float64, device=dev)
        self.flx = torch.empty_like(self.conc)
        self.workspace = torch.empty(max(self.plan.workspace_bytes, 1), dtype=torch.uint8, device=dev)
        self._graph = None
        self.surrogate = surrogate
        if surrogate is not None:
            self.plan.set_surrogate(surrogate.net, surrogate.hull.c_struct(), tol)

    _GEOMETRY = ("nx", "ny", "domain", "modes", "meas_pt", "halo", "levels")

    def update(self, specs: Sequence[FootprintSpec]):
        """New meteorology (wind, u*, z0, L, measurement height, ...) for the
        batch's instances, same geometry (grid, domain, modes, tower
        position, halo, requested levels) and the same number of vertical
        points each: the plan, its buffers and a captured graph are reused.
        Raises ValueError when a spec changes the geometry or the vertical
        grid size."""
        specs = list(specs)
        if len(specs) != len(self.specs):
            raise ValueError("update needs one spec per instance of the batch")
        for i, (new, old) in enumerate(zip(specs, self.specs)):
            for f in self._GEOMETRY:
                if getattr(new, f) != getattr(old, f):
                    raise ValueError(f"instance {i}: update may not change {f}")
            if old.levels is None and new.n != old.n:
                raise ValueError(f"instance {i}: update may not change n (the output level)")
            if new.closure in ("CONSTANT", "MOST", "MOSTM") and new.ustar is not None and new.z0 is not None:
                raise ValueError("Either z0 or ustar must be provided.")
        mets = _lib.make_mets(specs)
        self.plan.update_met(mets)
        for inst, m in zip(self.instances, mets):
            inst.met = m
        self.specs = specs

    def surrogate_stats(self):
        """After a run: per instance the spot-check relative L2 error (NaN
        when outside the hull) and the path taken (0 surrogate, 1 solver)."""
        return self.plan.surrogate_stats(self.workspace, len(self.specs))

    def __len__(self):
        return len(self.specs)

    def run(self):
        """Execute the whole batch on the current stream (asynchronous)."""
        self.plan.execute(self.prof, None, self.conc, self.flx, self.workspace)

    def capture(self):
        """Capture one execution into a HIP graph; later ``replay()`` calls
        launch the graph."""
        self.run()  # plans and LDS attributes are set before capture
        torch.cuda.synchronize(self.device)
        g = torch.cuda.CUDAGraph()
        with torch.cuda.graph(g):
            self.run()
        self._graph = g
        return g

    def replay(self):
        if self._graph is None:
            self.capture()
        self._graph.replay()

    def nz(self, i):
        return self.plan.offsets[i][3]

    def output(self, i):
        """(conc, flx) device views of instance i, shape (nslot, ny, nx)."""
        _, _, off, _ = self.plan.offsets[i]
        s = self.specs[i]
        ny, nx = self.out_dims[i]
        n = self.instances[i].nslot * ny * nx
        shape = (self.instances[i].nslot, ny, nx)
        return self.conc[off : off + n].view(shape), self.flx[off : off + n].view(shape)

    def profiles(self, i):
        """(z, (u, v, Kx, Ky, Kz)) of instance i as evaluated on the device."""
        off, _, _, nz = self.plan.offsets[i]
        p = self.prof[off : off + 6 * nz].view(6, nz).cpu().numpy()
        return p[0].copy(), tuple(p[k].copy() for k in range(1, 6))

    def lanes_and_steps(self):
        """(shooting lanes, vertical steps) per instance: the work units of the
        shooting kernel (each lane marches two initial-value problems)."""
        out = []
        for inst in self.instances:
            px = (inst.nlx - 1) // 2
            py = (inst.nly - 1) // 2
            nyq = (inst.nlx if inst.nly % 2 == 0 else 0) + ((2 * py + 1) if inst.nlx % 2 == 0 else 0)
            lanes = py * (2 * px + 1) + px + 1 + nyq
            out.append((lanes, inst.nz - 1))
        return out
