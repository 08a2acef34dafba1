"""Drop-in ``vertical_profiles``, ``psi`` and ``phi`` evaluated on the GPU.

Same signatures, return values and error messages as
src/bldfm/pbl_model.py:8-250.  The scalar prologue (derived z0 or u*,
stretched-grid constants, number of points) and the per-point profile
formulas run in the HIP library (``bldfm_profile_query``,
``bldfm_profiles``, ``bldfm_stability``); the same device functions feed
the batched solver, which evaluates profiles inside its level-table kernel.
"""

import ctypes
import logging

import numpy as np
import torch

from . import _lib

logger = logging.getLogger("bldfm_amd.pbl_model")


def vertical_profiles(
    n,
    meas_height,
    wind,
    ustar=None,
    z0=None,
    mol=1e9,
    prsc=1.0,
    closure="MOST",
    domain_height=None,
    stretch=None,
    z0_min=0.001,
    z0_max=2.0,
    tke=None,
):
    """Vertical grid and profiles (u, v, Kx, Ky, Kz) of the surface layer.

    Mirrors src/bldfm/pbl_model.py:8-204: ``n`` grid points between z0 and
    ``meas_height`` on a log-stretched grid up to ``domain_height`` (default
    2*meas_height); closures "MOST" (Monin-Obukhov), "MOSTM" (MOST without
    along-wind diffusion), "CONSTANT" and "OAAHOC" (Schumann-Lilly with
    ``tke``).  For the MOST-type closures exactly one of ``ustar``/``z0`` is
    given and the other follows from |wind| at meas_height.  ``z0_min`` and
    ``z0_max`` are accepted and unused, as in the reference.  Kx, Ky and Kz
    are one shared array except for "MOSTM".
    """
    if closure in ("CONSTANT", "MOST", "MOSTM"):
        if z0 is not None and ustar is not None:
            raise ValueError("Either z0 or ustar must be provided.")
        if z0 is None and ustar is None:
            raise TypeError("vertical_profiles needs ustar or z0 (both are None)")
    elif closure == "OAAHOC":
        if tke is None:
            logger.warning("No tke provided. Setting TKE to 1.0.")
            tke = 1.0
        if ustar is None:
            raise TypeError("the OAAHOC closure needs ustar")
    # the scalar prologue runs here in numpy (bldfm_amd.prologue); it raises
    # the reference's ValueError for an OAAHOC tke array of several values
    met = _lib.make_met(n, meas_height, wind, ustar, z0, mol, prsc, closure, domain_height, stretch, tke)
    lib = _lib.load()
    dev = _lib.device()
    nlev, z0o, ustaro = ctypes.c_int32(), ctypes.c_double(), ctypes.c_double()
    rc = lib.bldfm_profile_query(ctypes.byref(met), ctypes.byref(nlev), ctypes.byref(z0o), ctypes.byref(ustaro))
    if rc != 0 and "grid is empty" in _lib.last_error():
        # the reference's stats logging calls max() on the empty profiles
        raise ValueError("max() arg is an empty sequence")
    _lib.check(rc)
    out = torch.empty(6 * nlev.value, dtype=torch.float64, device=dev)
    _lib.check(lib.bldfm_profiles(ctypes.byref(met), _lib.ptr(out), _lib.stream_handle(dev)), "profiles")
    z, u, v, kx, ky, kz = out.cpu().numpy().reshape(6, nlev.value)
    z, u, v = z.copy(), u.copy(), v.copy()
    if closure == "MOSTM":
        Kx, Ky, Kz = kx.copy(), ky.copy(), kz.copy()
    else:
        Kx = Ky = Kz = kz.copy()
    if closure == "OAAHOC" and np.ndim(tke) > 0:
        # tke[..., newaxis] broadcasts the grid and profiles to tke.shape + (nlev,)
        shape = np.shape(tke) + (nlev.value,)
        z, u, v = z.reshape(shape), u.reshape(shape), v.reshape(shape)
        Kx = Ky = Kz = Kx.reshape(shape)

    logger.info("Stats from vertical_profiles")
    logger.info("z0    = %.3f m", float(np.ravel(z)[0]))
    logger.info("ustar = %.3f m s-1", ustaro.value)
    return z, (u, v, Kx, Ky, Kz)


def _elementwise(x, which):
    lib = _lib.load()
    dev = _lib.device()
    xa = np.asarray(x, dtype=np.float64)
    xt = torch.as_tensor(np.ascontiguousarray(xa).ravel(), device=dev)
    out = torch.empty_like(xt)
    args = (_lib.ptr(out), None) if which == "psi" else (None, _lib.ptr(out))
    _lib.check(lib.bldfm_stability(_lib.ptr(xt), xt.numel(), args[0], args[1], _lib.stream_handle(dev)), which)
    return out.cpu().numpy().reshape(xa.shape)


def psi(x):
    """MOST momentum stability correction psi(z/L) (src/bldfm/pbl_model.py:207-230)."""
    return _elementwise(x, "psi")


def phi(x):
    """MOST diffusivity stability function phi(z/L) (src/bldfm/pbl_model.py:233-250)."""
    return _elementwise(x, "phi")
