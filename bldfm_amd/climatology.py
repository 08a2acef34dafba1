"""Footprint climatologies: hourly footprints of several towers summed on the device.

Serving-scale use of the batched solver and the learned surrogate.  Per
tower, a seeded hourly meteorology series drives footprints at the
measurement height zm on a 512^2 grid; they are solved in batches of 128
(by the surrogate with its 1 % FP64 spot check and FP64 fallback, or by the
FP64 solver) and summed into the tower's climatology, the time mean of its
footprints, without leaving HBM.  The reference assembles such climatologies
on the host from run_bldfm_timeseries results (src/bldfm/interface.py:141-174);
here only the per-tower means come back.

Meteorology per tower (inside the surrogate's training hull): u* an AR(1)
series about 0.35 m/s (lag-1 correlation 0.9, clipped to [0.1, 1] m/s); 1/L
a diurnal cycle between -0.08 m^-1 (early afternoon) and 0.03 m^-1 (night)
plus AR(1) noise, clipped to that range; wind direction a random walk with a
15 degree hourly step; a fixed roughness length and measurement height per
tower, the wind speed at zm from the log law so that the profiles derive
that roughness (src/bldfm/pbl_model.py:65-77).

    python -m bldfm_amd.climatology [--towers 4] [--hours 8760] [--batch 128]
                                    [--tol 1e-3] [--seed 2026] [--out clim.npz]

prints one JSON line: per tower the climatology rel-L2 of the surrogate run
against the all-FP64 run, the accepted / fallback counts, and both runs'
footprints/s.
"""

import argparse
import json
import math
import sys
import time
from dataclasses import dataclass

import numpy as np
import torch

from . import _lib

KARMAN = 0.4


@dataclass
class Tower:
    name: str
    zm: float
    z0: float


DEFAULT_TOWERS = (Tower("T1", 10.0, 0.05), Tower("T2", 15.0, 0.1), Tower("T3", 22.0, 0.3), Tower("T4", 30.0, 0.5))


def hourly_met(tower, hours, seed):
    """Seeded hourly (u*, 1/L, wind speed, direction) of one tower (see the
    module docstring)."""
    from .pbl_model import psi

    rng = np.random.default_rng(seed)
    us = np.empty(hours)
    x = 0.0
    for h in range(hours):
        x = 0.9 * x + math.sqrt(1 - 0.81) * 0.12 * rng.standard_normal()
        us[h] = min(max(0.35 + x, 0.1), 1.0)
    hod = np.arange(hours) % 24
    cycle = 0.5 * (1 + np.cos(2 * np.pi * (hod - 14) / 24))  # 1 at 14 h, 0 at 2 h
    invl = 0.03 - 0.11 * cycle
    e = 0.0
    for h in range(hours):
        e = 0.8 * e + 0.6 * 0.01 * rng.standard_normal()
        invl[h] = min(max(invl[h] + e, -0.08), 0.03)
    wd = np.mod(rng.uniform(0, 360) + np.cumsum(rng.normal(0.0, 15.0, hours)), 360.0)
    speed = us / KARMAN * (np.log(tower.zm / tower.z0) + np.asarray(psi(tower.zm * invl)))
    return {"ustar": us, "invl": invl, "speed": speed, "dir": wd}


def specs_of(tower, met, lo, hi, hull):
    from .batch import FootprintSpec

    out = []
    for h in range(lo, hi):
        d = math.radians(met["dir"][h])
        sp = float(met["speed"][h])
        out.append(FootprintSpec(
            nx=512, ny=512, domain=(2000.0, 2000.0), modes=(512, 512), meas_pt=(1000.0, 1000.0), n=hull.n,
            meas_height=tower.zm, wind=(-sp * math.sin(d), -sp * math.cos(d)), ustar=float(met["ustar"][h]),
            mol=1.0 / float(met["invl"][h]), closure=hull.closure, domain_height=hull.top_ratio * tower.zm))
    return out


def climatology(tower, met, hours, batch=128, surrogate=None, tol=1e-3, hull=None):
    """Time mean of the tower's footprints (512 x 512 float64 on the device)
    and the surrogate's accepted / fallback counts.  One plan serves every
    batch of the same size: later batches only upload their meteorology
    (FootprintBatch.update)."""
    from .batch import FootprintBatch
    from .surrogate import Hull

    hull = hull or (surrogate.hull if surrogate is not None else Hull())
    acc = None
    n_acc = n_fb = 0
    plans = {}  # batch size -> FootprintBatch
    for lo in range(0, hours, batch):
        hi = min(lo + batch, hours)
        specs = specs_of(tower, met, lo, hi, hull)
        fb = plans.get(hi - lo)
        if fb is None:
            fb = plans[hi - lo] = FootprintBatch(specs, precision="double", surrogate=surrogate, tol=tol)
        else:
            fb.update(specs)
        fb.run()
        flx = torch.stack([fb.output(i)[1][0] for i in range(hi - lo)]).sum(0)
        acc = flx if acc is None else acc + flx
        if surrogate is not None:
            _, mode = fb.surrogate_stats()
            n_acc += sum(1 for m in mode if m == 0)
            n_fb += sum(1 for m in mode if m != 0)
    return acc / hours, n_acc, n_fb


def run(towers=DEFAULT_TOWERS, hours=8760, batch=128, tol=1e-3, seed=2026, checkpoint=None, log=None):
    """Climatologies of every tower by the surrogate (with its gate) and by
    the FP64 solver; returns the report and the (surrogate, FP64) means."""
    from . import surrogate as sur

    net = sur.SurrogateNet.load(checkpoint or sur.DEFAULT_CHECKPOINT)
    report = {"hours": hours, "batch": batch, "tol": tol, "towers": {}}
    means = {}
    t_sur = t_fp = 0.0
    for k, tower in enumerate(towers):
        met = hourly_met(tower, hours, seed + k)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        cs, n_acc, n_fb = climatology(tower, met, hours, batch, surrogate=net, tol=tol)
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        cf, _, _ = climatology(tower, met, hours, batch)
        torch.cuda.synchronize()
        t2 = time.perf_counter()
        t_sur += t1 - t0
        t_fp += t2 - t1
        rel = float(torch.linalg.vector_norm(cs - cf) / torch.linalg.vector_norm(cf))
        integ = abs(float(cs.sum() - cf.sum())) / abs(float(cf.sum()))
        report["towers"][tower.name] = {"zm": tower.zm, "z0": tower.z0, "climatology_rel_l2": rel,
                                        "climatology_integral_rel_err": integ, "accepted": n_acc,
                                        "fallback": n_fb}
        means[tower.name] = (cs.cpu().numpy(), cf.cpu().numpy())
        if log:
            log(f"[climatology] {tower.name}: rel-L2 {rel:.2e}, {n_acc} served / {n_fb} fallback, "
                f"{t1 - t0:.1f} s surrogate, {t2 - t1:.1f} s FP64")
    total = hours * len(towers)
    report["surrogate_footprints_per_s"] = total / t_sur
    report["fp64_footprints_per_s"] = total / t_fp
    report["note"] = ("end-to-end rates include the host-side meteorology of each batch of 128, its upload into the "
                      "reused plan, and the sums on the device")
    return report, means


def main(argv=None):
    ap = argparse.ArgumentParser(prog="bldfm_amd.climatology")
    ap.add_argument("--towers", type=int, default=4)
    ap.add_argument("--hours", type=int, default=8760)
    ap.add_argument("--batch", type=int, default=128)
    ap.add_argument("--tol", type=float, default=1e-3)
    ap.add_argument("--seed", type=int, default=2026)
    ap.add_argument("--checkpoint", default=None)
    ap.add_argument("--out", default=None, help="optional .npz of the per-tower climatologies")
    a = ap.parse_args(argv)
    _lib.load()
    towers = DEFAULT_TOWERS[: a.towers]
    report, means = run(towers, a.hours, a.batch, a.tol, a.seed, a.checkpoint,
                        log=lambda m: print(m, file=sys.stderr, flush=True))
    if a.out:
        np.savez(a.out, **{f"{k}/surrogate": v[0] for k, v in means.items()},
                 **{f"{k}/fp64": v[1] for k, v in means.items()})
    print(json.dumps(report), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
