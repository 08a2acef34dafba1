"""The CPU-baseline leg of bench.py on tiny footprints.

bench.cpu_baseline times oracle footprints (the numpy restatement of the
reference algorithm) in worker processes forked by a fork server, compares
them with the batch's footprints and reports the largest rel-L2.  Here the
batch is a stand-in that holds the oracle's own footprints, so the pooled
and the in-process runs must both report exactly zero error, and the pooled
workers must return the footprints the parent computes.
"""

import multiprocessing as mp
import os

import numpy as np
import pytest
import torch

import bench
from bldfm_amd.batch import FootprintSpec
from oracle import pbl as opbl
from oracle import solver as osolver


class _OracleBatch:
    """FootprintBatch stand-in: profiles and footprints from the oracle."""

    def __init__(self, specs):
        self.prof, self.out = [], []
        for s in specs:
            z, profs = opbl.vertical_profiles(s.n, s.meas_height, s.wind, ustar=s.ustar, mol=s.mol,
                                              domain_height=s.domain_height)
            _, c, f = osolver.steady_state_transport_solver(
                np.zeros((s.ny, s.nx)), z, profs, s.domain, s.n, modes=s.modes, meas_pt=s.meas_pt,
                footprint=True, precision="double")
            self.prof.append((z, profs))
            self.out.append((torch.from_numpy(c)[None], torch.from_numpy(f)[None]))

    def profiles(self, i):
        return self.prof[i]

    def output(self, i):
        return self.out[i]


def _specs(k):
    return [FootprintSpec(nx=32, ny=32, domain=(400.0, 400.0), modes=(32, 32), meas_pt=(200.0, 200.0), n=8,
                          meas_height=10.0, wind=(3.0, 0.5 * i), ustar=0.3, mol=-50.0 if i % 2 == 0 else 80.0,
                          domain_height=40.0) for i in range(k)]


def test_cpu_baseline_in_process_and_pooled():
    if torch.cuda.is_initialized() or os.path.exists("/dev/kfd"):
        # a fork server must start before the process touches a GPU; bench.py
        # starts it first thing, a test process on a GPU machine cannot
        pytest.skip("pooled baseline is exercised by bench.py itself on GPU machines")
    specs = _specs(3)
    fb = _OracleBatch(specs)
    fps1, worst1, cpu1 = bench.cpu_baseline(specs, fb, 3, 1)
    fps2, worst2, cpu2 = bench.cpu_baseline(specs, fb, 3, 2, mp.get_context("forkserver"))
    assert worst1 == 0.0 and worst2 == 0.0
    assert fps1 > 0 and fps2 > 0 and cpu1 > 0 and cpu2 > 0


def test_oracle_footprint_worker_unit():
    s = _specs(1)[0]
    fb = _OracleBatch([s])
    z, profs = fb.profiles(0)
    oc, of, dt = bench._oracle_footprint((s.ny, s.nx, s.domain, s.n, s.modes, s.meas_pt, z, profs))
    assert dt > 0
    np.testing.assert_array_equal(oc, fb.output(0)[0][0].numpy())
    np.testing.assert_array_equal(of, fb.output(0)[1][0].numpy())


def test_traffic_only_for_the_profiled_kernel_sources(tmp_path, monkeypatch):
    """roofline.traffic comes from the committed PMC summary only when the
    summary's source_sha256 is the hash of the kernel sources bench.py runs."""
    h = bench.kernel_source_hash()
    assert h == bench.kernel_source_hash() and len(h) == 64
    pmc = tmp_path / "pmc.txt"
    monkeypatch.setattr(bench, "PMC_SUMMARY", str(pmc))
    assert bench.measured_traffic()[0] is None  # no summary
    pmc.write_text(f"source_sha256: {h}\ntraffic_bytes_per_launch k_shoot: 5.7e+08\n")
    value, src = bench.measured_traffic()
    assert value == 5.7e8 and "these kernel sources" in src
    pmc.write_text(f"source_sha256: {'0' * 64}\ntraffic_bytes_per_launch k_shoot: 5.7e+08\n")
    value, src = bench.measured_traffic()
    assert value is None and "other kernel sources" in src


def test_kernel_source_hash_follows_the_sources(tmp_path, monkeypatch):
    """Changing any kernel source file changes the hash."""
    import shutil

    root = tmp_path / "r"
    shutil.copytree(os.path.join(bench.ROOT, "bldfm_amd", "csrc"), root / "bldfm_amd" / "csrc")
    (root / "include").mkdir()
    shutil.copy(os.path.join(bench.ROOT, "include", "bldfm_amd.h"), root / "include")
    monkeypatch.setattr(bench, "ROOT", str(root))
    h0 = bench.kernel_source_hash()
    assert h0 == bench.kernel_source_hash()
    with open(root / "bldfm_amd" / "csrc" / "bldfm_march.h", "a") as f:
        f.write("\n")
    assert bench.kernel_source_hash() != h0
