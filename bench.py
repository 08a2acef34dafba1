"""Benchmark: batched FP64 footprint solves at 512^2 modes x 64 vertical steps.

Workload (one "step" = one pass of the hot path over one batch): B footprint
solves of the production configuration -- domain 2000 x 2000 m, 512 x 512
grid and Fourier modes (default halo, padded grid 1536^2), measurement height
15 m, 26 points below zm and a column to 100 m (65 levels = 64 vertical
steps), u* = 0.3 m/s, z0 = 0.05 m, Obukhov length alternating -20 m / +100 m,
seeded wind directions, tower at the domain centre.  Each solve produces the
footprint and the Green's function at zm (complex FP64 spectra, FP64 output).
Inputs (met scalars) are resident on the GPU before the timed region; each
step evaluates profiles, marches every mode, shoots, and transforms back.

Prints one JSON line (the driver's contract) with the roofline of the
dominant kernel (the shooting kernel, FP64 VALU bound) measured with HIP
events on the launch stream, and a CPU baseline: the numpy oracle timed on
a bounded sample of the same workload (16 footprints) in 16 worker processes,
one footprint per host core at a time, like the reference's process pool.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--batch B]
                    [--surrogate CHECKPOINT] [--tol 1e-3] [--no-surrogate] [--no-tune]
                    [--cpu-sample 16] [--cpu-workers 16] [--host-steps 10] [--interface-steps 128]
    python bench.py --copy-calibration   # 1 GiB device copy, the PMC byte reference

The surrogate leg serves the same batch with the trained network
(``bldfm_amd/data``), re-solving 1 % of every footprint's modes in FP64 as
an error gate and solving a footprint in FP64 when that check exceeds
``--tol``; it reports its footprints/s, the accepted/fallback split, the
footprint and footprint-integral errors against the FP64 batch and the
network kernel's bf16 MFMA throughput.

Two host-inclusive legs report what a caller that needs the footprints in
host memory gets: the same batches with every footprint copied to pinned host
memory (copies overlapped with the next batch), and the reference's
``run_bldfm_timeseries`` interface end to end.  Neither is ``value``.
"""

import argparse
import json
import os
import sys
import time

os.environ.setdefault("OMP_NUM_THREADS", "1")
os.environ.setdefault("OPENBLAS_NUM_THREADS", "1")
os.environ.setdefault("MKL_NUM_THREADS", "1")

import numpy as np  # noqa: E402

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "footprints/s at 512²×64: batched FP64 solver vs surrogate at rel-L2≤1e-3"
PEAK_FP64 = 78.6e12  # MI355X FP64 vector, FMA counted as 2 FLOP
FLOP_PER_MODE_STEP = 80  # algorithmic FP64 FLOP per (mode, vertical step), both IVPs


def psi_bd(x):
    if x > 0:
        return 5.0 * x
    xi = (1.0 - 16.0 * x) ** 0.25
    return -2 * np.log(0.5 * (1 + xi)) - np.log(0.5 * (1 + xi * xi)) + 2 * np.arctan(xi) - 0.5 * np.pi


def workload(batch, seed=20261204):
    """Footprint specs of the 512^2 x 64 production configuration."""
    from bldfm_amd.batch import FootprintSpec

    rng = np.random.default_rng(seed)
    specs = []
    for i in range(batch):
        mol = -20.0 if i % 2 == 0 else 100.0
        speed = 0.3 / 0.4 * (np.log(15.0 / 0.05) + psi_bd(15.0 / mol))  # |U(zm)| giving z0 = 0.05
        wd = np.deg2rad(rng.uniform(0.0, 360.0))
        specs.append(
            FootprintSpec(
                nx=512, ny=512, domain=(2000.0, 2000.0), modes=(512, 512), meas_pt=(1000.0, 1000.0),
                n=26, meas_height=15.0, wind=(-speed * np.sin(wd), -speed * np.cos(wd)), ustar=0.3,
                mol=mol, domain_height=100.0,
            )
        )
    return specs


PMC_SUMMARY = os.path.join(ROOT, "profiles", "footprint_batch_512x64_pmc.txt")


def kernel_source_hash():
    """sha256 of the HIP library's sources (bldfm_amd/csrc and the C header),
    which identifies the kernels a profile was taken on."""
    import hashlib

    h = hashlib.sha256()
    files = sorted(os.listdir(os.path.join(ROOT, "bldfm_amd", "csrc")))
    paths = [os.path.join(ROOT, "bldfm_amd", "csrc", f) for f in files if f.endswith((".hip", ".h"))]
    for path in paths + [os.path.join(ROOT, "include", "bldfm_amd.h")]:
        h.update(os.path.basename(path).encode())
        with open(path, "rb") as f:
            h.update(f.read())
    return h.hexdigest()


def measured_traffic(kernel="k_shoot"):
    """HBM bytes per launch of ``kernel`` from the committed rocprofv3 --pmc
    summary of this workload (FETCH_SIZE + WRITE_SIZE, copy-calibrated), and
    whether that summary was taken on the current kernel sources (its
    ``source_sha256`` line).  A summary of other sources gives no figure."""
    value, sha = None, None
    try:
        for line in open(PMC_SUMMARY):
            if line.startswith("source_sha256:"):
                sha = line.split(":")[1].strip()
            if line.startswith(f"traffic_bytes_per_launch {kernel}:"):
                value = float(line.split(":")[1])
    except OSError:
        return None, "no PMC summary"
    if sha != kernel_source_hash():
        return None, "the committed PMC summary was taken on other kernel sources"
    return value, "rocprofv3 FETCH_SIZE+WRITE_SIZE of these kernel sources, copy-calibrated (profiles/)"


def timed(batch_obj, steps, warmup):
    import torch

    for _ in range(warmup):
        batch_obj.run()
    torch.cuda.synchronize()
    batch_obj.plan.timing_begin(steps)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        batch_obj.run()
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    phases, n = batch_obj.plan.timing_end()
    return t1 - t0, phases, n


MFMA_FLOP_PER_POINT = 2 * (96 * 256 + 4 * 256 * 256)  # bf16 MFMA work of one network evaluation
PEAK_BF16 = 2.5e15  # MI355X dense bf16 MFMA
PEAK_HBM = 8.0e12  # MI355X HBM3E bytes/s


def surrogate_leg(specs, fb, args):
    """The same batch served by the trained surrogate with the 1 % FP64
    spot-check gate and FP64 fallback at rel-L2 > tol; footprint errors
    against the exact batch ``fb`` (same specs, already run)."""
    import torch

    from bldfm_amd import surrogate as sur
    from bldfm_amd.batch import FootprintBatch

    if not os.path.exists(args.surrogate):
        return {"error": f"no trained network at {args.surrogate}"}
    net = sur.SurrogateNet.load(args.surrogate)
    fbs = FootprintBatch(specs, precision="double", fast=False, surrogate=net, tol=args.tol, tuning=fb.tuning)
    elapsed, phases, n = timed(fbs, args.steps, args.warmup)
    fps = len(specs) * args.steps / elapsed
    err, mode = fbs.surrogate_stats()
    rel, integ = [], []
    for i in range(len(specs)):
        a, b = fbs.output(i)[1], fb.output(i)[1]
        rel.append(float(torch.linalg.vector_norm(a - b) / torch.linalg.vector_norm(b)))
        integ.append(abs(float(a.sum() - b.sum())) / abs(float(b.sum())))
    acc = [i for i, m in enumerate(mode) if m == 0]
    points = sum(-(-l // 128) * 128 for l, _ in fbs.lanes_and_steps())
    sur_ms = phases["surrogate"] / max(n, 1)
    achieved = MFMA_FLOP_PER_POINT * points / (sur_ms * 1e-3) if acc else None
    del fbs
    pick = lambda v, idx: [v[i] for i in idx]
    return {
        "value": round(fps, 2),
        "unit": "footprints/s",
        "ratio_vs_fp64_solver": None,  # filled by the caller
        "tol": args.tol,
        "spot_check": "1 mode in 100 re-solved in FP64 per footprint",
        "accepted": len(acc),
        "fallback": len(specs) - len(acc),
        "footprint_rel_l2_accepted": {"max": max(pick(rel, acc)) if acc else None,
                                      "median": float(np.median(pick(rel, acc))) if acc else None},
        "footprint_rel_l2_all": {"max": max(rel), "median": float(np.median(rel))},
        "spot_estimate_median": float(np.nanmedian(err)),
        "footprint_integral_rel_err_accepted": {"max": max(pick(integ, acc)) if acc else None,
                                                "median": float(np.median(pick(integ, acc))) if acc else None},
        "phase_ms_per_step": {k: round(v / max(n, 1), 4) for k, v in phases.items()},
        "roofline": {
            "bound": "mfma",
            "kernel": "k_sur_plan (bf16 MFMA network on every mode) + spot check + gate",
            "achieved": round(achieved / 1e12, 2) if achieved else None,
            "peak": PEAK_BF16 / 1e12,
            "unit": "TFLOP/s",
            "frac": round(achieved / PEAK_BF16, 4) if achieved else None,
            "flop_per_point": MFMA_FLOP_PER_POINT,
            "points_per_launch": points,
        },
        "network": os.path.relpath(args.surrogate, ROOT),
    }


def _oracle_footprint(job):
    """One footprint by the oracle (numpy restatement of the reference
    algorithm): the CPU baseline's unit of work, run in a worker process."""
    ny, nx, domain, n, modes, meas_pt, z, profs = job
    from oracle import solver as osolver

    t0 = time.perf_counter()
    _, oc, of = osolver.steady_state_transport_solver(
        np.zeros((ny, nx)), z, profs, domain, n, modes=modes, meas_pt=meas_pt, footprint=True, precision="double",
    )
    return oc, of, time.perf_counter() - t0


def cpu_baseline(specs, gpu_batch, nsample, workers, pool_ctx=None):
    """Oracle footprints on `workers` host cores (one footprint per worker
    process at a time, like the reference's process-pool fan-out), on the
    first ``nsample`` footprints of the workload: footprints per second of
    wall time.  Also returns the largest rel-L2 between those GPU footprints
    and the oracle's."""
    from oracle import pbl as opbl

    jobs = []
    for i in range(nsample):
        s = specs[i]
        z, profs = gpu_batch.profiles(i)  # same inputs for both solvers
        jobs.append((s.ny, s.nx, s.domain, s.n, s.modes, s.meas_pt, z, profs))
        # profiles: device evaluation vs the oracle's
        oz, _ = opbl.vertical_profiles(s.n, s.meas_height, s.wind, ustar=s.ustar, mol=s.mol,
                                       domain_height=s.domain_height)
        assert len(oz) == len(z)
    t0 = time.perf_counter()
    if workers > 1:
        with pool_ctx.Pool(workers) as pool:
            res = pool.map(_oracle_footprint, jobs, chunksize=1)
    else:
        res = [_oracle_footprint(j) for j in jobs]
    wall = time.perf_counter() - t0
    worst = 0.0
    for i, (oc, of, _) in enumerate(res):
        gc, gf = gpu_batch.output(i)
        gc, gf = gc[0].cpu().numpy(), gf[0].cpu().numpy()
        for a, b in ((gc, oc), (gf, of)):
            worst = max(worst, float(np.linalg.norm(a - b) / np.linalg.norm(b)))
    cpu_s = sum(r[2] for r in res)
    return nsample / wall, worst, cpu_s


def fft_roofline(fb, phases, n):
    """HBM roofline of the two inverse passes: algorithmic bytes (band in,
    cropped columns out; cropped rows in, conc and flx out) over their
    HIP-event time."""
    def band(nl, ne):
        return nl + 1 if (nl % 2 == 0 and nl < ne) else nl

    by = bx = 0
    for inst in fb.instances:
        nxe, nye = inst.nx + 2 * inst.px, inst.ny + 2 * inst.py
        nxb, nyb = band(inst.nlx, nxe), band(inst.nly, nye)
        ob = 4 if inst.out_f32 else 8
        by += inst.nslot * nxb * (nyb + inst.ny) * 16
        bx += inst.nslot * inst.ny * (nxb * 16 + inst.nx * 2 * ob)
    ty = phases["inverse_y"] / max(n, 1) * 1e-3
    tx = phases["inverse_x"] / max(n, 1) * 1e-3
    out = {"bound": "hbm", "peak": PEAK_HBM / 1e9, "unit": "GB/s"}
    for name, b, t in (("k_inv_cols", by, ty), ("k_inv_rows", bx, tx)):
        out[name] = {"bytes": b, "ms": round(t * 1e3, 4), "achieved": round(b / t / 1e9, 1),
                     "frac": round(b / t / PEAK_HBM, 4)}
    return out


def host_inclusive_leg(fb, steps):
    """The FP64 batch with every footprint's conc and flx copied to pinned
    host memory: two output buffers, each batch's device-to-host copy on a
    second stream overlapping the next batch's kernels.  Footprints per
    second of wall time, host copies included."""
    import torch

    comp = torch.cuda.current_stream()
    copy = torch.cuda.Stream()
    outs = [(fb.conc, fb.flx), (torch.empty_like(fb.conc), torch.empty_like(fb.flx))]
    host = [tuple(torch.empty(fb.conc.numel(), dtype=fb.conc.dtype, pin_memory=True) for _ in range(2))
            for _ in range(2)]
    done_comp = [torch.cuda.Event() for _ in range(2)]
    done_copy = [torch.cuda.Event() for _ in range(2)]

    def step(k):
        b = k % 2
        comp.wait_event(done_copy[b])  # the copy that last read these buffers has finished
        fb.plan.execute(fb.prof, None, outs[b][0], outs[b][1], fb.workspace)
        done_comp[b].record(comp)
        copy.wait_event(done_comp[b])
        with torch.cuda.stream(copy):
            host[b][0].copy_(outs[b][0], non_blocking=True)
            host[b][1].copy_(outs[b][1], non_blocking=True)
        done_copy[b].record(copy)

    for b in range(2):
        done_copy[b].record(copy)
    for k in range(2):
        step(k)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for k in range(steps):
        step(k)
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    # device-to-host bandwidth alone, on the same buffers
    t1 = time.perf_counter()
    for k in range(steps):
        host[k % 2][0].copy_(outs[k % 2][0], non_blocking=True)
        host[k % 2][1].copy_(outs[k % 2][1], non_blocking=True)
    torch.cuda.synchronize()
    d2h = 2 * fb.conc.numel() * fb.conc.element_size()
    ok = all(torch.equal(host[b][i], outs[b][i].cpu()) for b in range(2) for i in range(2))
    del outs, host
    return {"value": round(len(fb) * steps / dt, 2), "unit": "footprints/s",
            "bytes_to_host_per_step": d2h, "d2h_GBps": round(d2h * steps / (time.perf_counter() - t1) / 1e9, 2),
            "copies_verified": bool(ok),
            "how": "two output buffers; each batch's conc+flx copied to pinned host memory on a second stream "
                   "while the next batch runs"}


def interface_leg(specs, nsteps):
    """``run_bldfm_timeseries`` (the reference's interface) over ``nsteps``
    hourly steps of the same configuration: plan set-up, solves, copies to
    the host and the reference's result dictionaries (grid arrays per result)
    included.  Footprints per second of wall time."""
    from bldfm_amd.config_parser import BLDFMConfig, DomainConfig, MetConfig, SolverConfig, TowerConfig
    from bldfm_amd.interface import run_bldfm_timeseries

    s = specs[0]
    rng = np.random.default_rng(7)
    cfg = BLDFMConfig(
        domain=DomainConfig(nx=s.nx, ny=s.ny, xmax=s.domain[0], ymax=s.domain[1], nz=s.n, modes=s.modes),
        towers=[TowerConfig(name="t", lat=0.0, lon=0.0, z_m=s.meas_height, x=s.meas_pt[0], y=s.meas_pt[1])],
        met=MetConfig(ustar=[float(v) for v in rng.uniform(0.2, 0.5, nsteps)],
                      mol=[float(v) for v in rng.choice([-20.0, 100.0], nsteps)],
                      wind_speed=[float(v) for v in rng.uniform(2.0, 6.0, nsteps)],
                      wind_dir=[float(v) for v in rng.uniform(0.0, 360.0, nsteps)]),
        solver=SolverConfig(closure="MOST", precision="double", footprint=True),
    )
    run_bldfm_timeseries(cfg, cfg.towers[0])  # plan caches and device warm-up
    t0 = time.perf_counter()
    res = run_bldfm_timeseries(cfg, cfg.towers[0])
    dt = time.perf_counter() - t0
    return {"value": round(nsteps / dt, 2), "unit": "footprints/s", "timesteps": nsteps,
            "output_shape": list(np.shape(res[0]["flx"])),
            "how": "run_bldfm_timeseries on the same grid, modes and tower (the interface's own column: "
                   "nz points to 2 zm, output at zm): met -> device profiles -> batched solve -> host copies "
                   "-> result dicts with their grid arrays"}


def copy_calibration():
    """A plain device copy of 1 GiB (three times): profiled with rocprofv3
    --pmc FETCH_SIZE / WRITE_SIZE, it gives the factors that turn those
    counters into HBM bytes."""
    import torch

    n = 1 << 27
    a = torch.ones(n, dtype=torch.float64, device="cuda")
    b = torch.empty_like(a)
    torch.cuda.synchronize()
    for _ in range(3):
        b.copy_(a)
    torch.cuda.synchronize()
    print(json.dumps({"copy_bytes": a.numel() * 8, "copies": 3}))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=30)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--batch", type=int, default=128)
    ap.add_argument("--cpu-sample", type=int, default=16, help="footprints of the CPU baseline (0: skip it)")
    ap.add_argument("--cpu-workers", type=int, default=16,
                    help="worker processes (host cores) of the CPU baseline; 1 runs it in this process")
    ap.add_argument("--no-fast", action="store_true")
    ap.add_argument("--surrogate", default=os.path.join(ROOT, "bldfm_amd", "data", "surrogate_most_n26.pt"))
    ap.add_argument("--tol", type=float, default=1e-3)
    ap.add_argument("--no-surrogate", action="store_true")
    ap.add_argument("--no-tune", action="store_true", help="default launch shape instead of the autotuned one")
    ap.add_argument("--host-steps", type=int, default=10,
                    help="batches of the leg that copies every footprint to host memory (0: skip it)")
    ap.add_argument("--interface-steps", type=int, default=128,
                    help="timesteps of the run_bldfm_timeseries leg (0: skip it)")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="only run a plain 1 GiB device copy (reference for FETCH_SIZE / WRITE_SIZE)")
    args = ap.parse_args()
    if args.copy_calibration:
        copy_calibration()
        return
    if args.gpus != 1:
        print(f"bench.py: runs on one GPU; --gpus {args.gpus} is not supported "
              "(footprint batches would shard across GPUs with no exchange between them)", file=sys.stderr)
        sys.exit(2)
    args.cpu_sample = max(0, min(args.cpu_sample, args.batch))
    workers = max(1, min(args.cpu_workers, args.cpu_sample)) if args.cpu_sample > 0 else 1
    pool_ctx = None
    if workers > 1:
        # The baseline's worker processes are forked by a fork server started
        # here, before this process touches the GPU: no process that has
        # initialised the GPU forks or execs.
        import multiprocessing as mp
        import multiprocessing.forkserver as mpfs

        pool_ctx = mp.get_context("forkserver")
        mpfs.ensure_running()

    import torch

    from bldfm_amd.batch import FootprintBatch

    torch.cuda.set_device(0)
    specs = workload(args.batch)
    print(f"[bench] batch={args.batch} steps={args.steps} warmup={args.warmup}", file=sys.stderr, flush=True)
    # launch shape chosen by the autotuner (timed before, outside the measured steps)
    fb = FootprintBatch(specs, precision="double", fast=False, tuning=None if args.no_tune else "auto")
    if fb.tune_report is not None:
        print(f"[bench] autotuned launch shape {fb.tuning}: "
              + ", ".join(f"{k} {v:.3f} ms" for k, v in fb.tune_report["ms"].items()), file=sys.stderr, flush=True)
    elapsed, phases, n = timed(fb, args.steps, args.warmup)
    fps = args.batch * args.steps / elapsed
    ms_step = 1e3 * elapsed / args.steps
    units = sum(l * s for l, s in fb.lanes_and_steps())  # (mode, step) units per launch
    shoot_ms = phases["shoot"] / max(n, 1)
    achieved = FLOP_PER_MODE_STEP * units / (shoot_ms * 1e-3)
    print(f"[bench] exact: {fps:.1f} footprints/s, {ms_step:.3f} ms/step, phases(ms/step)="
          + ", ".join(f"{k}={v / max(n, 1):.3f}" for k, v in phases.items()), file=sys.stderr, flush=True)

    extra = {}
    if not args.no_fast:
        fbf = FootprintBatch(specs, precision="double", fast=True, tuning=fb.tuning)
        el_f, ph_f, n_f = timed(fbf, args.steps, args.warmup)
        extra["fast_fma_value"] = round(args.batch * args.steps / el_f, 2)
        extra["fast_fma_shoot_ms"] = round(ph_f["shoot"] / max(n_f, 1), 4)
        d = 0.0
        for i in range(min(4, args.batch)):
            for a, b in zip(fbf.output(i), fb.output(i)):
                a, b = a.cpu().numpy(), b.cpu().numpy()
                d = max(d, float(np.linalg.norm(a - b) / np.linalg.norm(b)))
        extra["fast_fma_rel_l2_vs_exact"] = d
        del fbf

    if not args.no_surrogate:
        sl = surrogate_leg(specs, fb, args)
        if "value" in sl:
            sl["ratio_vs_fp64_solver"] = round(sl["value"] / fps, 4)
            print(f"[bench] surrogate: {sl['value']:.1f} footprints/s ({sl['ratio_vs_fp64_solver']:.3f}x the "
                  f"FP64 solver), {sl['accepted']} accepted / {sl['fallback']} fallback", file=sys.stderr, flush=True)
        extra["surrogate"] = sl

    # the host-inclusive legs are reported beside the line; a failure in one
    # of them is recorded in its entry and does not cost the bench line
    if args.host_steps > 0:
        try:
            extra["host_inclusive"] = host_inclusive_leg(fb, args.host_steps)
            print(f"[bench] host-inclusive: {extra['host_inclusive']['value']:.1f} footprints/s "
                  f"(device-to-host {extra['host_inclusive']['d2h_GBps']:.1f} GB/s)", file=sys.stderr, flush=True)
        except Exception as e:  # noqa: BLE001
            extra["host_inclusive"] = {"error": f"{type(e).__name__}: {e}"}
    if args.interface_steps > 0:
        try:
            extra["interface"] = interface_leg(specs, args.interface_steps)
            print(f"[bench] interface: {extra['interface']['value']:.1f} footprints/s", file=sys.stderr, flush=True)
        except Exception as e:  # noqa: BLE001
            extra["interface"] = {"error": f"{type(e).__name__}: {e}"}

    cpu_fps, worst, cpu_s = (cpu_baseline(specs, fb, args.cpu_sample, workers, pool_ctx) if args.cpu_sample > 0
                             else (None, None, None))
    if cpu_fps is not None:
        print(f"[bench] cpu baseline: {cpu_fps:.3f} footprints/s on {workers} cores "
              f"({args.cpu_sample} footprints, {cpu_s:.1f} s of CPU time)", file=sys.stderr, flush=True)
    traffic, traffic_src = measured_traffic() if args.batch == 128 else (None, "PMC summary is of batch 128")
    out = {
        "metric": METRIC,
        "value": round(fps, 2),
        "unit": "footprints/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_step, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (seeded met per footprint; delta-source footprints)",
        "config": {
            "workload": "batched FP64 footprint solver, 512x512 grid and modes, 2000 m domain, "
                        "padded 1536^2, 65 levels (64 steps) to 101.5 m, zm=15 m, L=-20/+100 m",
            "footprints_per_step": args.batch,
            "rounding": "reference (numba/numpy) rounding reproduced; no fused multiply-add in the march",
        },
        "roofline": {
            "bound": "fp64-valu",
            "kernel": "k_shoot<1,false,true> (vertical march + shooting)",
            "achieved": round(achieved / 1e12, 3),
            "peak": PEAK_FP64 / 1e12,
            "unit": "TFLOP/s",
            "frac": round(achieved / PEAK_FP64, 4),
            "traffic": traffic,
            "traffic_unit": "bytes per launch",
            "traffic_source": traffic_src,
            "kernel_source_sha256": kernel_source_hash(),
            "algorithmic_bytes": 16 * sum(2 * (l - 1) for l, _ in fb.lanes_and_steps()),
            "flop_per_unit": FLOP_PER_MODE_STEP,
            "units_per_launch": units,
            "avg_launch_ms": round(shoot_ms, 4),
            "issue_ceiling_frac_without_fma": 0.5,
        },
        "phase_ms_per_step": {k: round(v / max(n, 1), 4) for k, v in phases.items()},
        "fft_roofline": fft_roofline(fb, phases, n),
        "tuning": {"choice": fb.tuning,
                   "candidates_ms": ({k: round(v, 4) for k, v in fb.tune_report["ms"].items()}
                                     if fb.tune_report is not None else None)},
        "cpu_baseline": {
            "value": round(cpu_fps, 4) if cpu_fps is not None else None,
            "unit": "footprints/s",
            "cores": workers,
            "kind": "port",
            "sample": (f"{args.cpu_sample} footprints of the same workload (numpy oracle, {workers} worker "
                       f"process{'es' if workers > 1 else ''}, one footprint each at a time; "
                       f"{cpu_s:.1f} s of CPU time)" if cpu_fps is not None else None),
        },
        "parity": {"max_rel_l2_vs_oracle": worst, "footprints_checked": args.cpu_sample},
    }
    out.update(extra)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
