"""Drop-in FFT backend: the interface of src/bldfm/fft_manager.py on the GPU.

``FFTManager.fft2`` / ``ifft2`` and the module-level ``fft2`` / ``ifft2``
keep numpy's semantics (transform over the last two axes, leading axes
batched, ``norm`` in {"backward", "forward", "ortho"}) and run the project's
LDS Stockham kernels through ``bldfm_fft2``.  There is no FFTW wisdom to
load or save: twiddle tables are built once per length in extended
precision and kept by the library.  ``num_threads`` and
``cache_keepalive`` are accepted for compatibility and have no effect.
Inputs may be numpy arrays (numpy out) or torch tensors on the GPU (torch
out, no host copies).
"""

import math

import numpy as np
import torch

from . import _lib


def _scale(norm, n, forward):
    if norm is None or norm == "backward":
        return 1.0 if forward else 1.0 / n
    if norm == "forward":
        return 1.0 / n if forward else 1.0
    if norm == "ortho":
        return 1.0 / math.sqrt(n)
    raise ValueError(f'Invalid norm value {norm}; should be "backward", "ortho" or "forward".')


def _fft2(x, norm, forward):
    lib = _lib.load()
    as_numpy = not torch.is_tensor(x)
    if as_numpy:
        xa = np.asarray(x)
        single = xa.dtype in (np.complex64, np.float32)
        dev = _lib.device()
        t = torch.as_tensor(np.ascontiguousarray(xa, dtype=np.complex128), device=dev)
    else:
        single = x.dtype in (torch.complex64, torch.float32)
        t = x.to(torch.complex128).contiguous()
        dev = t.device
    if t.ndim < 2:
        raise ValueError("fft2 needs at least two dimensions")
    n0, n1 = t.shape[-2], t.shape[-1]
    batch = int(np.prod(t.shape[:-2], dtype=np.int64)) if t.ndim > 2 else 1
    out = torch.empty_like(t)
    if t.numel() > 0:
        ws = torch.empty(lib.bldfm_fft2_workspace_bytes(n0, n1, batch), dtype=torch.uint8, device=dev)
        sign = -1 if forward else 1
        _lib.check(
            lib.bldfm_fft2(dev.index, n0, n1, batch, _lib.ptr(t), _lib.ptr(out), sign,
                           _scale(norm, n0 * n1, forward), _lib.ptr(ws), _lib.stream_handle(dev)),
            "fft2",
        )
    if single:
        out = out.to(torch.complex64)
    return out.cpu().numpy() if as_numpy else out


class FFTManager:
    """GPU counterpart of the reference's FFTManager (src/bldfm/fft_manager.py:12-114)."""

    def __init__(self, wisdom_file="fftw_wisdom.pkl", num_threads=1, cache_keepalive=30):
        self.wisdom_file = wisdom_file
        self.num_threads = num_threads
        self.cache_keepalive = cache_keepalive

    def fft2(self, input_data, norm="backward"):
        """2-D forward transform (e^{-i}), numpy ``norm`` semantics."""
        return _fft2(input_data, norm, True)

    def ifft2(self, input_data, norm="backward"):
        """2-D inverse transform (e^{+i}), numpy ``norm`` semantics."""
        return _fft2(input_data, norm, False)

    def clear_cache(self):
        """Nothing to clear per call: twiddle tables are owned by the library."""


_fft_manager = None


def get_fft_manager(num_threads=1, cache_keepalive=30):
    """Process-wide manager, re-created when ``num_threads`` changes
    (src/bldfm/fft_manager.py:121-139)."""
    global _fft_manager
    if _fft_manager is not None and _fft_manager.num_threads != num_threads:
        _fft_manager = None
    if _fft_manager is None:
        _fft_manager = FFTManager(num_threads=num_threads, cache_keepalive=cache_keepalive)
    return _fft_manager


def reset_fft_manager():
    """Drop the singleton (src/bldfm/fft_manager.py:142-145)."""
    global _fft_manager
    _fft_manager = None


def fft2(input_data, norm="backward"):
    return get_fft_manager().fft2(input_data, norm)


def ifft2(input_data, norm="backward"):
    return get_fft_manager().ifft2(input_data, norm)
