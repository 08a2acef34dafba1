"""Kormann & Meixner (2001) analytic footprint (drop-in for
src/bldfm/ffm_kormann_meixner.py).

``estimateFootprint`` keeps the reference's signature, grid conventions and
empty-footprint warning; the footprint grid is evaluated on the GPU
(``bldfm_km01_footprint``, one thread per cell).  ``estimate_footprints``
evaluates many footprints (e.g. a timeseries of met conditions) in one
launch.  The scalar helpers ``_phiM``, ``_phiC``, ``_psiM``, ``_mParam``,
``_nParam`` (Eqs. 33-36) and ``estimateZ0`` (Eq. 31 with the reference's
wind-direction median smoothing) are host arithmetic on the inputs, as in
the reference.
"""

import ctypes
import warnings

import numpy as np
import torch

from . import _lib

von_karman = 0.4


def _phiM(zm, mo_len):
    zm, L = np.asarray(zm, dtype=float), np.asarray(mo_len, dtype=float)
    return np.where(L < 0, np.abs(1 - 16 * zm / L) ** -0.25, 1 + 5 * zm / L)


def _phiC(zm, mo_len):
    zm, L = np.asarray(zm, dtype=float), np.asarray(mo_len, dtype=float)
    return np.where(L < 0, np.abs(1 - 16 * zm / L) ** -0.5, 1 + 5 * zm / L)


def _psiM(zm, mo_len):
    zm, L = np.asarray(zm, dtype=float), np.asarray(mo_len, dtype=float)
    x = np.abs(1 - 16 * zm / L) ** 0.25
    unstable = -2 * np.log(0.5 * (1 + x)) - np.log(0.5 * (1 + x * x)) + 2 * np.arctan(x) - np.pi * 0.5
    return np.where(L < 0, unstable, 5 * zm / L)


def _mParam(zm, ws, ustar, mo_len):
    return np.asarray(ustar, dtype=float) * _phiM(zm, mo_len) / (von_karman * np.asarray(ws, dtype=float))


def _nParam(zm, mo_len):
    zm, L = np.asarray(zm, dtype=float), np.asarray(mo_len, dtype=float)
    return np.where(L < 0, (1 - 24 * zm / L) / (1 - 16 * zm / L), 1 / (1 + 5 * zm / L))


def estimateZ0(zm, ws, wd, ustar, mo_len, half_wd_win=22):
    """Roughness length from Eq. (31), median-smoothed in a window of
    +-half_wd_win degrees of wind direction at 1-degree steps
    (src/bldfm/ffm_kormann_meixner.py:16-107)."""
    n = len(zm)
    if n != len(ws) or n != len(wd) or n != len(ustar) or n != len(mo_len):
        raise RuntimeError("Input parameters must be of the same length!")
    zm, ws, wd = (np.asarray(a, dtype=float) for a in (zm, ws, wd))
    z0 = zm * np.exp(_psiM(zm, mo_len) - von_karman * ws / np.asarray(ustar, dtype=float))
    z0[z0 > 1000] = np.nan
    if half_wd_win < 1:
        return z0
    out = np.full_like(z0, np.nan)
    for kk in range(360):
        w = wd.copy()
        if kk < 90:
            w[wd > 270] = wd[wd > 270] - 360
        elif kk > 270:
            w[wd < 90] = wd[wd < 90] + 360
        sel = (wd >= kk) & (wd < kk + 1)
        if sel.any():
            win = (w >= kk - half_wd_win) & (w < kk + 1 + half_wd_win)
            out[sel] = np.nanmedian(z0[win])
    return out


def _record(zm, z0, ws, ustar, mo_len, sigma_v, grid_domain, grid_res, mxy, wd):
    r = _lib.Km01()
    r.zm, r.z0, r.ws, r.ustar, r.mo_len, r.sigma_v = (float(v) for v in (zm, z0, ws, ustar, mo_len, sigma_v))
    r.xmin, r.xmax, r.ymin, r.ymax = (float(v) for v in grid_domain)
    r.res = float(grid_res)
    r.mx, r.my = float(mxy[0]), float(mxy[1])
    r.has_wd = wd is not None
    r.wd = float(wd) if wd is not None else 0.0
    return r


def estimate_footprints(records):
    """Footprints of many measurements in one launch.  ``records`` is a list
    of dicts with estimateFootprint's arguments; returns a list of
    ``(grid_x, grid_y, grid_ffm)``."""
    lib = _lib.load()
    dev = _lib.device()
    n = len(records)
    recs = (_lib.Km01 * n)(*[_record(**{k: r.get(k) for k in (
        "zm", "z0", "ws", "ustar", "mo_len", "sigma_v", "grid_domain", "grid_res", "mxy", "wd")}) for r in records])
    shapes = []
    for i in range(n):
        nx, ny = _lib._I32(), _lib._I32()
        _lib.check(lib.bldfm_km01_grid(ctypes.byref(recs[i]), ctypes.byref(nx), ctypes.byref(ny)), "km01_grid")
        shapes.append((ny.value, nx.value))
    total = sum(a * b for a, b in shapes)
    ffm = torch.empty(max(total, 1), dtype=torch.float64, device=dev)
    status = (ctypes.c_int32 * n)()
    _lib.check(lib.bldfm_km01_footprint(dev.index, recs, n, _lib.ptr(ffm), status, _lib.stream_handle(dev)),
               "km01_footprint")
    host = ffm.cpu().numpy()
    out, off = [], 0
    for i, (ny, nx) in enumerate(shapes):
        r = recs[i]
        gx, gy = np.meshgrid(np.arange(r.xmin + 0.5 * r.res, r.xmax, r.res),
                             np.arange(r.ymax - 0.5 * r.res, r.ymin, -r.res))
        if status[i]:
            warnings.warn(
                "U in Eq. (11) of Kormann & Meixner, 2001 is estimated negative, physically impossible! "
                "Return empty footprint.\n"
                f"zm = {r.zm:.3f}, z0 = {r.z0:.3f}, ws = {r.ws:.3f}, ustar = {r.ustar:.3f}, "
                f"L = {r.mo_len:.3f}, sigma_v = {r.sigma_v:.3f}")
        out.append((gx, gy, host[off : off + ny * nx].reshape(ny, nx).copy()))
        off += ny * nx
    return out


def estimateFootprint(zm, z0, ws, ustar, mo_len, sigma_v, grid_domain, grid_res, mxy, wd=None):
    """Footprint of one measurement (src/bldfm/ffm_kormann_meixner.py:110-330):
    returns ``(grid_x, grid_y, grid_ffm)`` of cell centres and flux portions."""
    return estimate_footprints([dict(zm=zm, z0=z0, ws=ws, ustar=ustar, mo_len=mo_len, sigma_v=sigma_v,
                                     grid_domain=grid_domain, grid_res=grid_res, mxy=mxy, wd=wd)])[0]
