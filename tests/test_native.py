"""CPU tests of the native boundary: the C-ABI library builds, loads and
exports every symbol include/bldfm_amd.h declares; the index maps the kernels
use pass a bounds/coverage check compiled from the same header; host-side
ABI logic (profile grid query) matches the oracle; the drop-ins raise the
reference's errors before touching a device and fail loudly without one.
No GPU compute is called here.
"""

import ctypes
import re
import shutil
import subprocess
import sys
from pathlib import Path

import numpy as np
import pytest

ROOT = Path(__file__).resolve().parent.parent
HEADER = ROOT / "include" / "bldfm_amd.h"


@pytest.fixture(scope="module")
def lib():
    from bldfm_amd import _lib

    if not _lib.LIB_PATH.exists():
        from bldfm_amd import build

        build.build(verbose=False)
    return _lib.load()


def header_functions():
    text = re.sub(r"/\*.*?\*/", "", HEADER.read_text(), flags=re.S)
    return sorted(set(re.findall(r"\b(bldfm_[a-z0-9_]+)\s*\(", text)))


def test_header_symbols_exported_and_bound(lib):
    from bldfm_amd import _lib

    names = header_functions()
    assert len(names) >= 15
    for name in names:
        assert hasattr(lib, name), name
        assert name in _lib.SIGNATURES, f"{name} has no ctypes signature"
    assert lib.bldfm_abi_version() == 2


def test_exports_are_exactly_the_header(lib):
    """No C symbol outside the header is exported with the bldfm_ prefix and C linkage."""
    nm = shutil.which("nm")
    if nm is None:
        pytest.skip("nm not available")
    out = subprocess.run([nm, "-D", "--defined-only", str(ROOT / "bldfm_amd" / "_bldfm_hip.so")],
                         capture_output=True, text=True, check=True).stdout
    exported = {l.split()[-1] for l in out.splitlines() if l.split()[-1].startswith("bldfm_")}
    assert exported == set(header_functions())


def test_index_maps_harness():
    """Compile tests/native/index_maps_check.cpp (the kernels' index maps,
    from bldfm_types.h) with sanitizers and run it."""
    cxx = shutil.which("g++")
    if cxx is None:
        pytest.skip("g++ not available")
    exe = ROOT / "tests" / "native" / "_index_maps_check"
    subprocess.run([cxx, "-O1", "-std=c++17", "-fsanitize=address,undefined", "-fno-sanitize-recover=all",
                    str(ROOT / "tests" / "native" / "index_maps_check.cpp"), "-o", str(exe)], check=True)
    try:
        res = subprocess.run([str(exe)], capture_output=True, text=True, timeout=120)
    finally:
        exe.unlink(missing_ok=True)
    assert res.returncode == 0 and res.stdout.startswith("OK"), res.stdout + res.stderr


@pytest.mark.parametrize(
    "kw",
    [
        dict(n=16, meas_height=10.0, wind=(0.0, -6.0), ustar=0.5),
        dict(n=19, meas_height=10.0, wind=(5.0, 0.0), ustar=0.4, mol=-50.0),
        dict(n=26, meas_height=15.0, wind=(-3.5, 0.2), ustar=0.3, mol=100.0, domain_height=100.0),
        dict(n=25, meas_height=15.0, wind=(-3.5, 0.2), ustar=0.3, mol=-20.0, domain_height=100.0),
        dict(n=10, meas_height=5.0, wind=(3.0, 1.0), z0=0.1, mol=-100.0),
        dict(n=10, meas_height=5.0, wind=(3.0, 1.0), ustar=0.4, closure="OAAHOC", tke=2.0),
        dict(n=10, meas_height=5.0, wind=(3.0, 1.0), ustar=0.4, stretch=15.0, domain_height=50.0),
    ],
)
def test_profile_query_matches_oracle(lib, kw):
    """Host side of bldfm_profile_query: level count and derived z0/u*."""
    from bldfm_amd import _lib
    from oracle import pbl as opbl

    args = dict(ustar=None, z0=None, mol=1e9, prsc=1.0, closure="MOST", domain_height=None, stretch=None, tke=None)
    args.update({k: v for k, v in kw.items() if k not in ("n", "meas_height", "wind")})
    z, _ = opbl.vertical_profiles(**kw)
    # the host's numpy prologue (the drop-in's default) and the library's own
    for prologue in (True, False):
        met = _lib.make_met(kw["n"], kw["meas_height"], kw["wind"], prologue=prologue, **args)
        assert met.has_prologue == prologue
        n, z0, us = ctypes.c_int32(), ctypes.c_double(), ctypes.c_double()
        _lib.check(lib.bldfm_profile_query(ctypes.byref(met), ctypes.byref(n), ctypes.byref(z0), ctypes.byref(us)))
        assert n.value == len(z)
        np.testing.assert_allclose(z0.value, z[0], rtol=1e-12)


def test_profile_query_errors(lib):
    from bldfm_amd import _lib

    met = _lib.make_met(10, 5.0, (3.0, 1.0), 0.4, 0.1, 1e9, 1.0, "MOST", None, None, None)
    rc = lib.bldfm_profile_query(ctypes.byref(met), None, None, None)
    assert rc == -1 and b"Either z0 or ustar" in lib.bldfm_last_error()


def test_solver_errors_before_device():
    """The reference's ValueErrors (src/bldfm/solver.py:90-91, 187-188) are
    raised before any device work."""
    from bldfm_amd.solver import steady_state_transport_solver
    from oracle import pbl as opbl

    z, profs = opbl.vertical_profiles(16, 10.0, (5.0, 0.0), 0.4)
    q0 = np.zeros((64, 128))
    with pytest.raises(ValueError, match="modes must consist of even numbers"):
        steady_state_transport_solver(q0, z, profs, (500.0, 250.0), 15, modes=(63, 64))
    with pytest.raises(ValueError, match="precision must be"):
        steady_state_transport_solver(q0, z, profs, (500.0, 250.0), 15, modes=(64, 64), precision="quad")
    with pytest.raises(IndexError):
        steady_state_transport_solver(q0, z, profs, (500.0, 250.0), len(z) + 3, modes=(64, 64))


def test_pbl_errors_before_device():
    """src/bldfm/pbl_model.py:76-77, 99-103 messages."""
    from bldfm_amd.pbl_model import vertical_profiles

    with pytest.raises(ValueError, match="Either z0 or ustar"):
        vertical_profiles(n=10, meas_height=5.0, wind=(3.0, 1.0), ustar=0.4, z0=0.1, closure="MOST")
    with pytest.raises(ValueError, match="Invalid closure type"):
        vertical_profiles(n=10, meas_height=5.0, wind=(3.0, 1.0), ustar=0.4, closure="INVALID")


def test_no_cpu_fallback_without_gpu():
    """Without a HIP device the drop-ins raise; nothing computes on the CPU."""
    import torch

    if torch.cuda.is_available():
        pytest.skip("a GPU is visible")
    from bldfm_amd.pbl_model import vertical_profiles
    from bldfm_amd.solver import steady_state_transport_solver
    from bldfm_amd.utils import compute_wind_fields
    from oracle import pbl as opbl

    with pytest.raises(RuntimeError, match="HIP GPU"):
        vertical_profiles(16, 10.0, (0.0, -6.0), 0.5)
    with pytest.raises(RuntimeError, match="HIP GPU"):
        compute_wind_fields(5.0, 270.0)
    z, profs = opbl.vertical_profiles(16, 10.0, (0.0, -6.0), 0.5)
    with pytest.raises(RuntimeError, match="HIP GPU"):
        steady_state_transport_solver(np.zeros((128, 64)), z, profs, (100.0, 700.0), 16, modes=(64, 128), footprint=True)


def test_product_does_not_import_the_oracle():
    """bldfm_amd/ never imports, reads or runs oracle/."""
    for path in (ROOT / "bldfm_amd").rglob("*.py"):
        text = path.read_text()
        assert "oracle" not in re.sub(r"#.*", "", text).replace('"""', ""), path


def test_bench_refuses_multi_gpu():
    res = subprocess.run([sys.executable, str(ROOT / "bench.py"), "--gpus", "2"], capture_output=True, text=True,
                         timeout=120)
    assert res.returncode == 2 and "one GPU" in res.stderr
