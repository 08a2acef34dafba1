"""GPU: the no-allocation entry point bldfm_plan_execute_static.

It runs a plan with the library's static per-GPU scratch as workspace.
Checked through the C ABI against bldfm_plan_execute with a caller
workspace: bit for bit, on two calls in a row with different plans (the
second smaller), then the first plan again (the array keeps the last call's
contents, which may not leak into results); after the first call on the
device (whose symbol lookup makes the runtime place the array) device memory
in use does not grow; a plan whose workspace exceeds the array is refused
with the size error."""

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _specs(k, grid, modes, zms):
    from bldfm_amd.batch import FootprintSpec

    out = []
    for i, zm in enumerate(zms[:k]):
        wd = np.deg2rad(20.0 + 65.0 * i)
        sp = 3.0 + 0.7 * i
        out.append(FootprintSpec(nx=grid, ny=grid, domain=(1000.0, 1000.0), modes=(modes, modes),
                                 meas_pt=(500.0, 500.0), n=16 + 2 * i, meas_height=zm,
                                 wind=(-sp * np.sin(wd), -sp * np.cos(wd)), ustar=0.25 + 0.05 * i,
                                 mol=(-40.0, 80.0, 1e9, -15.0)[i % 4]))
    return out


def _used(torch):
    torch.cuda.synchronize()
    free, total = torch.cuda.mem_get_info()
    return total - free


def test_static_execute_matches_workspace_execute():
    import torch

    assert torch.cuda.is_available()
    from bldfm_amd import _lib
    from bldfm_amd.batch import FootprintBatch

    _lib.load()
    big = FootprintBatch(_specs(4, 256, 256, [10.0, 15.0, 8.0, 20.0]))
    small = FootprintBatch(_specs(1, 128, 128, [12.0]))
    assert small.plan.workspace_bytes < big.plan.workspace_bytes <= 64 * 2**20
    refs = {}
    for name, fb in (("big", big), ("small", small)):
        fb.run()
        refs[name] = (fb.conc.clone(), fb.flx.clone())
    # torch loads the code of its own kernels on first use: run the ones the
    # loop below uses before measuring
    for fb in (big, small):
        fb.conc.fill_(np.nan)
        assert not torch.equal(fb.conc, refs["big" if fb is big else "small"][0])
    before = _used(torch)
    big.plan.execute_static(big.prof, None, big.conc, big.flx)  # set-up: symbol address of this device's array
    used0 = _used(torch)
    print(f"  first call on the device: {before / 2**20:.1f} -> {used0 / 2**20:.1f} MiB in use", flush=True)
    for name, fb in (("big", big), ("small", small), ("big", big)):
        fb.conc.fill_(np.nan)
        fb.flx.fill_(np.nan)
        fb.plan.execute_static(fb.prof, None, fb.conc, fb.flx)
        torch.cuda.synchronize()
        print(f"  {name}: static-scratch execution vs workspace execution", flush=True)
        assert torch.equal(fb.conc, refs[name][0]) and torch.equal(fb.flx, refs[name][1])
    used1 = _used(torch)
    print(f"  device memory in use {used0 / 2**20:.1f} -> {used1 / 2**20:.1f} MiB")
    assert used1 <= used0 + 4 * 2**20  # far below the 64 MiB an allocation per call would take


def test_static_execute_refuses_large_plans():
    import torch

    assert torch.cuda.is_available()
    from bldfm_amd import _lib
    from bldfm_amd.batch import FootprintBatch

    fb = FootprintBatch(_specs(4, 512, 512, [10.0, 15.0, 8.0, 20.0]) * 3)  # 12 x 512^2: ~100 MB of workspace
    assert fb.plan.workspace_bytes > 64 * 2**20
    with pytest.raises(_lib.BLDFMError, match="exceeds the static scratch"):
        fb.plan.execute_static(fb.prof, None, fb.conc, fb.flx)
    fb.run()  # the plan still runs with its own workspace
    torch.cuda.synchronize()
    assert torch.isfinite(fb.flx).all()
