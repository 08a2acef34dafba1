"""Drop-in utilities of src/bldfm/utils.py.

``compute_wind_fields`` runs on the GPU (``bldfm_wind_fields``).
``ideal_source`` and ``point_measurement`` build test inputs and evaluate a
dot product on the host exactly as the reference does
(src/bldfm/utils.py:30-92); they are not part of the solve.  ``parallelize``
keeps the reference's decorator name (src/bldfm/utils.py:95-106): there is
no JIT to drive here, so it returns the function unchanged.
"""

import numpy as np
import torch

from . import _lib


def compute_wind_fields(u_rot, wind_dir):
    """(u, v) from wind speed and met direction (degrees, blowing FROM).

    u = -U sin(theta), v = -U cos(theta), src/bldfm/utils.py:23-27.  Scalars
    in give Python floats out; arrays give arrays.
    """
    lib = _lib.load()
    dev = _lib.device()
    spd = np.asarray(u_rot, dtype=np.float64)
    wd = np.asarray(wind_dir, dtype=np.float64)
    spd, wd = np.broadcast_arrays(spd, wd)
    s = torch.as_tensor(np.ascontiguousarray(spd).ravel(), device=dev)
    d = torch.as_tensor(np.ascontiguousarray(wd).ravel(), device=dev)
    u = torch.empty_like(s)
    v = torch.empty_like(s)
    _lib.check(lib.bldfm_wind_fields(_lib.ptr(s), _lib.ptr(d), s.numel(), _lib.ptr(u), _lib.ptr(v), _lib.stream_handle(dev)), "wind_fields")
    u = u.cpu().numpy().reshape(spd.shape)
    v = v.cpu().numpy().reshape(spd.shape)
    if u.ndim == 0:
        return np.float64(u), np.float64(v)
    return u, v


def ideal_source(nxy, domain, src_loc=None, shape="diamond"):
    """Synthetic surface flux: "diamond", "circle" (radius xmax/12) or "point"
    (Gaussian with sigma = 4 dx), on ``linspace(..., endpoint=True)`` like
    src/bldfm/utils.py:30-77."""
    nx, ny = nxy
    xmx, ymx = domain
    dx = xmx / nx
    xs, ys = (xmx / 2, ymx / 2) if src_loc is None else src_loc
    X, Y = np.meshgrid(np.linspace(0.0, xmx, nx), np.linspace(0.0, ymx, ny))
    q0 = np.zeros([ny, nx])
    if shape == "diamond":
        q0 = np.where(np.abs(X - xs) + np.abs(Y - ys) < xmx / 12, 1.0, 0.0)
    if shape == "circle":
        q0 = np.where(np.sqrt((X - xs) ** 2 + (Y - ys) ** 2) < xmx / 12, 1.0, 0.0)
    if shape == "point":
        sig = 4.0 * dx
        q0 = np.exp(-((X - xs) ** 2 + (Y - ys) ** 2) / 2.0 / sig**2) / sig / np.sqrt(2.0 * np.pi)
    return q0


def point_measurement(f, g):
    """Footprint-weighted sum sum(f*g) (src/bldfm/utils.py:80-92)."""
    return np.sum(f * g)


def parallelize(func):
    """Identity decorator kept for API compatibility with the reference."""
    return func
