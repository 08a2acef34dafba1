"""GPU: the drop-in 2-D FFT (src/bldfm/fft_manager.py:78-102 semantics) on
seeded random shapes against numpy.

Lengths are drawn from four pools: small (2-64), medium composites and
primes up to 4096 (the generic prime stages), large primes up to the
single-workgroup limit of 10033 points, and lengths above it that run as
four-step transforms (n = n1 n2, both factors within the limit).  Batches of
one or two, every norm, forward and inverse.  Bar: rel-L2 <= 1e-11 against
numpy's pocketfft (both are FP64 transforms; a length-p prime stage sums p
terms, hence the bar above the usual 1e-13).
"""

import numpy as np
import pytest

from conftest import rel_l2

pytestmark = pytest.mark.gpu

PRIMES = [2003, 3001, 4093, 5003, 7919, 9973, 10007]
LONG = [10034, 10240, 12288, 15625, 20000, 30030]


def _shapes(seed, count):
    rng = np.random.default_rng(seed)
    for _ in range(count):
        pool = int(rng.integers(0, 4))
        if pool == 0:
            a, b = int(rng.integers(2, 65)), int(rng.integers(2, 65))
        elif pool == 1:
            a, b = int(rng.integers(65, 4097)), int(rng.integers(2, 257))
        elif pool == 2:
            a, b = int(rng.choice(PRIMES)), int(rng.integers(2, 9))
        else:
            a, b = int(rng.choice(LONG)), int(rng.integers(2, 5))
        if rng.integers(0, 2):
            a, b = b, a
        yield (int(rng.integers(1, 3)), a, b), ["backward", "forward", "ortho"][int(rng.integers(0, 3))], int(rng.integers(0, 2**31))


def test_random_shapes_match_numpy():
    import torch

    assert torch.cuda.is_available(), "GPU tests need a HIP device"
    from bldfm_amd import fft_manager

    worst = 0.0
    for shape, norm, seed in _shapes(20261205, 60):
        r = np.random.default_rng(seed)
        x = r.standard_normal(shape) + 1j * r.standard_normal(shape)
        for ours, ref in ((fft_manager.fft2, np.fft.fft2), (fft_manager.ifft2, np.fft.ifft2)):
            e = rel_l2(ours(x, norm=norm), ref(x, norm=norm))
            worst = max(worst, e)
            assert e <= 1e-11, (shape, norm, ours.__name__, e)
        print(f"  {shape} {norm}: ok", flush=True)
    print(f"  worst rel-L2 {worst:.2e}")


@pytest.mark.parametrize("dtype", [np.complex64, np.float32, np.float64, np.complex128])
def test_output_dtype_follows_numpy(dtype):
    """Output dtype as numpy gives it (single-precision inputs stay single),
    values within the input precision's roundoff."""
    from bldfm_amd import fft_manager

    r = np.random.default_rng(3)
    x = r.standard_normal((2, 48, 40))
    if np.issubdtype(dtype, np.complexfloating):
        x = x + 1j * r.standard_normal((2, 48, 40))
    x = x.astype(dtype)
    tol = 1e-6 if dtype in (np.complex64, np.float32) else 1e-13
    for ours, ref in ((fft_manager.fft2, np.fft.fft2), (fft_manager.ifft2, np.fft.ifft2)):
        a, b = ours(x), ref(x)
        assert a.dtype == b.dtype, (dtype, a.dtype, b.dtype)
        assert rel_l2(a, b) <= tol
