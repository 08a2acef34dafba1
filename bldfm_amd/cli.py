"""Command line (counterpart of ``bldfm run``, src/bldfm/cli.py:20-62).

    python -m bldfm_amd.cli run config.yaml [--dry-run] [--out results.npz]

Loads the YAML configuration with the reference's schema, reports domain,
towers and timesteps, and (unless ``--dry-run``) runs every tower and
timestep as batched GPU plans (``bldfm_amd.interface``).  Results go to an
``.npz`` file with, per tower and timestep, ``conc``, ``flx`` and the
``X, Y, Z`` grid under keys ``<tower>/<timestep>/<name>``; the reference's
plots and NetCDF writer belong to its private abl-tk package and are not
reproduced.
"""

import argparse
import logging
import sys

import numpy as np

log = logging.getLogger("bldfm_amd.cli")


def cmd_run(args):
    from .config_parser import load_config

    config = load_config(args.config)
    log.info("Loaded config: %s", args.config)
    log.info("  Domain: %dx%d, nz=%d", config.domain.nx, config.domain.ny, config.domain.nz)
    log.info("  Towers: %d", len(config.towers))
    log.info("  Timesteps: %d", config.met.n_timesteps)
    if args.dry_run:
        log.info("Dry run complete. Config is valid.")
        for t in config.towers:
            log.info("  Tower '%s': lat=%s, lon=%s, z_m=%s, x=%.1f, y=%.1f", t.name, t.lat, t.lon, t.z_m, t.x, t.y)
        return 0
    from .interface import run_bldfm_multitower

    results = run_bldfm_multitower(config)
    arrays = {}
    for name, runs in results.items():
        for r in runs:
            key = f"{name}/{r['timestamp']}"
            arrays[f"{key}/conc"] = r["conc"]
            arrays[f"{key}/flx"] = r["flx"]
            for axis, g in zip("XYZ", r["grid"]):
                arrays[f"{key}/{axis}"] = g
    np.savez(args.out, **arrays)
    log.info("All runs complete: %d results written to %s", sum(len(v) for v in results.values()), args.out)
    return 0


def main(argv=None):
    ap = argparse.ArgumentParser(prog="bldfm_amd", description="BLDFM footprint runs on the GPU")
    sub = ap.add_subparsers(dest="command")
    run = sub.add_parser("run", help="run a YAML configuration")
    run.add_argument("config")
    run.add_argument("--dry-run", action="store_true", help="validate the configuration only")
    run.add_argument("--out", default="bldfm_results.npz", help="output .npz file")
    args = ap.parse_args(argv)
    if args.command != "run":
        ap.print_help()
        return 1
    logging.basicConfig(level=logging.INFO, format="%(message)s")
    return cmd_run(args)


if __name__ == "__main__":
    sys.exit(main())
