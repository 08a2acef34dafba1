"""Device and host memory of the drop-in, after the reference's
tests/test_memory.py:64-119 (single solve < 500 MB; < 20 % growth over 10
sequential solves), extended to what a GPU library can leak: plans and FFT
twiddle blocks.  100 solves over 40 distinct grid shapes (more than the
drop-in's 32-plan cache holds) and plain 2-D FFTs of 161 distinct lengths
(more than the 16 MiB twiddle cache holds) twice over: the second round may
not grow device memory, and the library's own cache stays under its cap."""

import gc

import numpy as np
import psutil
import pytest

pytestmark = pytest.mark.gpu

MB = 1024**2


@pytest.fixture(scope="module")
def mods():
    import torch

    assert torch.cuda.is_available()
    from bldfm_amd import _lib, fft_manager, pbl_model, solver, utils

    _lib.load()
    return torch, _lib, fft_manager, pbl_model, solver, utils


def _solve(mods, nxy=(128, 64), modes=(128, 64), nz=16, footprint=True):
    """The reference's conftest-scale solve (tests/test_memory.py:34-56)."""
    torch, _lib, _, pbl_model, solver, utils = mods
    domain = (500.0, 250.0)
    srf_flx = utils.ideal_source(nxy, domain, (250.0, 125.0), shape="point")
    z, profs = pbl_model.vertical_profiles(nz, 10.0, (5.0, 0.0), 0.4, closure="MOST")
    return solver.steady_state_transport_solver(srf_flx, z, profs, domain, nz - 1, modes=modes, footprint=footprint)


def _device_used(torch):
    torch.cuda.synchronize()
    free, total = torch.cuda.mem_get_info()
    return total - free


def test_single_solve_memory(mods):
    torch = mods[0]
    _solve(mods)  # load the library and warm the allocator
    gc.collect()
    rss0 = psutil.Process().memory_info().rss
    torch.cuda.reset_peak_memory_stats()
    base = torch.cuda.memory_allocated()
    _solve(mods)
    peak = torch.cuda.max_memory_allocated() - base
    rss = psutil.Process().memory_info().rss - rss0
    print(f"\n  MEMORY device peak {peak / MB:.1f} MB, host RSS delta {rss / MB:.1f} MB")
    assert peak < 500 * MB and rss < 500 * MB


def test_memory_growth_over_ten_solves(mods):
    torch = mods[0]
    _solve(mods)
    gc.collect()
    rss0, dev0 = psutil.Process().memory_info().rss, _device_used(torch)
    for _ in range(10):
        _solve(mods)
    gc.collect()
    rss1, dev1 = psutil.Process().memory_info().rss, _device_used(torch)
    print(f"\n  MEMORY RSS {rss0 / MB:.0f} -> {rss1 / MB:.0f} MB, device {dev0 / MB:.0f} -> {dev1 / MB:.0f} MB")
    assert rss1 < 1.2 * rss0 and dev1 < 1.2 * dev0


def test_caches_stay_bounded(mods):
    torch, _lib, fft_manager, _, solver, _ = mods
    shapes = [((64 + 8 * k, 32 + 4 * (k % 5)), (32 + 2 * (k % 8), 32)) for k in range(40)]
    lengths = list(range(6000, 10001, 25))  # 161 twiddle blocks of 96-160 KB: ~20 MB, above the cap
    used = []
    for rnd in range(2):
        for i in range(100):
            nxy, modes = shapes[i % 40]
            _solve(mods, nxy=nxy, modes=modes, nz=12 + i % 3)
        for n in lengths:
            fft_manager.fft2(np.ones((2, n), dtype=complex))
        gc.collect()
        torch.cuda.empty_cache()
        used.append(_device_used(torch))
        cache, _, _ = _lib.mem_info()
        print(f"\n  round {rnd}: device used {used[-1] / MB:.0f} MB, library cache {cache / MB:.1f} MB, "
              f"{len(solver._plan_cache)} cached plans")
        assert cache <= 16 * MB and len(solver._plan_cache) <= 32
    assert used[1] <= used[0] + 64 * MB
