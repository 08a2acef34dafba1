"""Argument checks at the boundary that return before any device work.

Empty batches, null plans and buffers, empty or negative transform sizes
and mismatched meteorology counts are refused with return code -1 and a
message in bldfm_last_error (the code the Python layer turns into
ValueError); an empty FootprintBatch raises ValueError before any plan is
built.  None of these calls reaches the HIP runtime, so they run without a
GPU.
"""

import ctypes

import pytest

from bldfm_amd import _lib


@pytest.fixture(scope="module")
def lib():
    return _lib.load()


def _refused(lib, rc, text):
    assert rc == -1
    assert text in lib.bldfm_last_error().decode()


def test_empty_or_null_plan_arguments(lib):
    out = ctypes.c_void_p()
    _refused(lib, lib.bldfm_plan_create(0, None, 0, None, ctypes.byref(out)), "bad plan arguments")
    assert not out.value
    inst = (_lib.Instance * 1)()
    lv = (ctypes.c_int32 * 2)()
    _refused(lib, lib.bldfm_plan_create(0, inst, 0, lv, ctypes.byref(out)), "bad plan arguments")
    _refused(lib, lib.bldfm_plan_create(0, inst, -3, lv, ctypes.byref(out)), "bad plan arguments")
    bad = _lib.make_tuning(shoot_block=96)
    _refused(lib, lib.bldfm_plan_create_tuned(0, inst, 1, lv, ctypes.byref(bad), ctypes.byref(out)),
             "shoot_block must be 64, 128 or 256")
    bad = _lib.make_tuning(xpass_layout=5)
    _refused(lib, lib.bldfm_plan_create_tuned(0, inst, 1, lv, ctypes.byref(bad), ctypes.byref(out)),
             "xpass_layout must be 0, 1 or 2")


def test_null_plan_queries_and_execution(lib):
    ws = ctypes.c_size_t()
    a, b, c = ctypes.c_int64(), ctypes.c_int64(), ctypes.c_int64()
    _refused(lib, lib.bldfm_plan_sizes(None, ctypes.byref(ws), ctypes.byref(a), ctypes.byref(b), ctypes.byref(c)),
             "null plan")
    nz = ctypes.c_int32()
    _refused(lib, lib.bldfm_plan_offsets(None, 0, ctypes.byref(a), ctypes.byref(b), ctypes.byref(c),
                                         ctypes.byref(nz)), "bad plan or instance index")
    _refused(lib, lib.bldfm_plan_execute(None, None, None, None, None, None, None), "null plan")
    _refused(lib, lib.bldfm_plan_timing_begin(None, 4), "bad timing arguments")
    _refused(lib, lib.bldfm_plan_update_met(None, None, 0, None), "update_met needs one met per plan instance")


def test_empty_and_negative_sizes(lib):
    _refused(lib, lib.bldfm_fft2(0, 0, 8, 1, None, None, 1, 1.0, None, None), "bad fft2 arguments")
    _refused(lib, lib.bldfm_fft2(0, 8, -1, 1, None, None, 1, 1.0, None, None), "bad fft2 arguments")
    _refused(lib, lib.bldfm_stability(None, -1, None, None, None), "bad stability arguments")
    _refused(lib, lib.bldfm_sur_met(None, 0, None), "bad sur_met arguments")


def test_empty_batch_raises():
    from bldfm_amd.batch import FootprintBatch

    with pytest.raises(ValueError, match="empty batch"):
        FootprintBatch([])


@pytest.mark.parametrize("short", [0, 2, 4])
def test_short_profile_arrays_raise_the_references_error(short):
    """A profile array shorter than z raises the reference's IndexError at
    its first top-level index (src/bldfm/solver.py:164-173: Kz, Kx, Ky, u, v
    in that order), before any device work."""
    import numpy as np

    from bldfm_amd import solver
    from oracle import pbl as opbl
    from oracle import solver as osolver

    z, profs = opbl.vertical_profiles(n=8, meas_height=10.0, wind=(3.0, 1.0), ustar=0.4, mol=-50.0)
    profs = list(profs)
    profs[short] = profs[short][:-2]
    q0 = np.zeros((16, 16))
    kw = dict(modes=(8, 8), meas_pt=(50.0, 50.0), footprint=True, precision="double")
    with pytest.raises(IndexError) as oe:
        osolver.steady_state_transport_solver(q0, z, tuple(profs), (100.0, 100.0), 8, **kw)
    with pytest.raises(IndexError) as ge:
        solver.steady_state_transport_solver(q0, z, tuple(profs), (100.0, 100.0), 8, **kw)
    assert str(ge.value) == str(oe.value)
