"""Green's-function disk cache with the reference's key and file format.

Drop-in for ``GreensFunctionCache`` (src/bldfm/cache.py:19-83).  A
footprint solve depends only on the vertical grid, the profiles and the
geometry, so its ``(grid, conc, flx)`` can be memoised.  Files are
``<key>.npz`` with arrays X, Y, Z, conc, flx, and the key is the SHA-256 of
the same byte sequence the reference hashes (z, the five profiles, domain,
modes, meas_pt as numpy buffers, then ``str(halo)`` and ``precision``), so a
cache directory written by either implementation serves the other.  As in
the reference the key leaves out ``levels``, the surface grid shape and
``analytic``: callers that vary those must not share a directory.

On this path the batched plans (``bldfm_amd.batch``) make the cache
unnecessary for throughput; it is kept for workflows that persist results
across runs.  Files are read with ``allow_pickle=False``.
"""

import hashlib
from pathlib import Path

import numpy as np


class GreensFunctionCache:
    """Disk cache of footprint-mode solver outputs, keyed by their inputs."""

    def __init__(self, cache_dir=".bldfm_cache"):
        self.cache_dir = Path(cache_dir)
        self.cache_dir.mkdir(parents=True, exist_ok=True)

    @staticmethod
    def _compute_key(z, profiles, domain, modes, meas_pt, halo, precision):
        h = hashlib.sha256()
        for part in (z, *profiles, domain, modes, meas_pt):
            h.update(np.asarray(part).tobytes())
        h.update(str(halo).encode())
        h.update(precision.encode())
        return h.hexdigest()

    def _path(self, *key_args):
        return self.cache_dir / f"{self._compute_key(*key_args)}.npz"

    def get(self, z, profiles, domain, modes, meas_pt, halo, precision):
        """``(grid, conc, flx)`` of a stored solve, or None."""
        path = self._path(z, profiles, domain, modes, meas_pt, halo, precision)
        if not path.exists():
            return None
        with np.load(path, allow_pickle=False) as d:
            return (d["X"], d["Y"], d["Z"]), d["conc"], d["flx"]

    def put(self, z, profiles, domain, modes, meas_pt, halo, precision, grid, conc, flx):
        X, Y, Z = grid
        np.savez(self._path(z, profiles, domain, modes, meas_pt, halo, precision), X=X, Y=Y, Z=Z, conc=conc,
                 flx=flx)

    def clear(self):
        """Remove every cached file; returns how many were removed."""
        n = 0
        for f in self.cache_dir.glob("*.npz"):
            f.unlink()
            n += 1
        return n
