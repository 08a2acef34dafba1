"""ctypes binding of the BLDFM HIP library (``include/bldfm_amd.h``).

The library is loaded once per process, after torch (so it binds to the HIP
runtime torch loaded).  There is no CPU fallback: if the library is missing
or no HIP device is visible, every entry point raises.
"""

import ctypes
import threading
from pathlib import Path

import numpy as np
import torch

LIB_PATH = Path(__file__).resolve().parent / "_bldfm_hip.so"

CLOSURES = {"MOST": 0, "CONSTANT": 1, "MOSTM": 2, "OAAHOC": 3}


class Met(ctypes.Structure):
    """Mirror of ``bldfm_met``."""

    _fields_ = [
        ("n", ctypes.c_int32),
        ("closure", ctypes.c_int32),
        ("has_ustar", ctypes.c_int32),
        ("has_z0", ctypes.c_int32),
        ("has_domain_height", ctypes.c_int32),
        ("has_stretch", ctypes.c_int32),
        ("meas_height", ctypes.c_double),
        ("um", ctypes.c_double),
        ("vm", ctypes.c_double),
        ("ustar", ctypes.c_double),
        ("z0", ctypes.c_double),
        ("mol", ctypes.c_double),
        ("prsc", ctypes.c_double),
        ("domain_height", ctypes.c_double),
        ("stretch", ctypes.c_double),
        ("tke", ctypes.c_double),
        ("has_prologue", ctypes.c_int32),
        ("reserved", ctypes.c_int32),
        ("grid_aa", ctypes.c_double),
        ("grid_bb", ctypes.c_double),
        ("grid_zetamax", ctypes.c_double),
    ]


class Instance(ctypes.Structure):
    """Mirror of ``bldfm_instance``."""

    _fields_ = [
        ("nx", ctypes.c_int32),
        ("ny", ctypes.c_int32),
        ("nlx", ctypes.c_int32),
        ("nly", ctypes.c_int32),
        ("px", ctypes.c_int32),
        ("py", ctypes.c_int32),
        ("nz", ctypes.c_int32),
        ("nslot", ctypes.c_int32),
        ("footprint", ctypes.c_int32),
        ("analytic", ctypes.c_int32),
        ("single", ctypes.c_int32),
        ("out_f32", ctypes.c_int32),
        ("has_shift", ctypes.c_int32),
        ("prof_from_met", ctypes.c_int32),
        ("fast", ctypes.c_int32),
        ("surrogate", ctypes.c_int32),
        ("dx", ctypes.c_double),
        ("dy", ctypes.c_double),
        ("shift_x", ctypes.c_double),
        ("shift_y", ctypes.c_double),
        ("srf_bg_conc", ctypes.c_double),
        ("met", Met),
    ]


class SurHull(ctypes.Structure):
    """Mirror of ``bldfm_sur_hull``."""

    _fields_ = [
        ("ustar_lo", ctypes.c_double),
        ("ustar_hi", ctypes.c_double),
        ("invl_lo", ctypes.c_double),
        ("invl_hi", ctypes.c_double),
        ("z0_lo", ctypes.c_double),
        ("z0_hi", ctypes.c_double),
        ("z0_zm_max", ctypes.c_double),
        ("zm_lo", ctypes.c_double),
        ("zm_hi", ctypes.c_double),
        ("kmax", ctypes.c_double),
        ("top_ratio", ctypes.c_double),
        ("n", ctypes.c_int32),
        ("closure", ctypes.c_int32),
    ]


class Km01(ctypes.Structure):
    """Mirror of ``bldfm_km01``."""

    _fields_ = [(f, ctypes.c_double) for f in ("zm", "z0", "ws", "ustar", "mo_len", "sigma_v", "xmin", "xmax",
                                               "ymin", "ymax", "res", "mx", "my", "wd")] + [
        ("has_wd", ctypes.c_int32), ("reserved", ctypes.c_int32)]


_P = ctypes.c_void_p
_I32 = ctypes.c_int32
_I64 = ctypes.c_int64
_D = ctypes.c_double

# name -> (restype, argtypes); every symbol the header declares
SIGNATURES = {
    "bldfm_abi_version": (ctypes.c_int, []),
    "bldfm_last_error": (ctypes.c_char_p, []),
    "bldfm_mem_info": (ctypes.c_int, [ctypes.c_int, ctypes.POINTER(ctypes.c_size_t), ctypes.POINTER(ctypes.c_size_t),
                                      ctypes.POINTER(ctypes.c_size_t)]),
    "bldfm_profile_query": (ctypes.c_int, [ctypes.POINTER(Met), ctypes.POINTER(_I32), ctypes.POINTER(_D), ctypes.POINTER(_D)]),
    "bldfm_profiles": (ctypes.c_int, [ctypes.POINTER(Met), _P, _P]),
    "bldfm_plan_create": (ctypes.c_int, [ctypes.c_int, ctypes.POINTER(Instance), _I32, ctypes.POINTER(_I32), ctypes.POINTER(_P)]),
    "bldfm_plan_update_met": (ctypes.c_int, [_P, ctypes.POINTER(Met), _I32, ctypes.c_void_p]),
    "bldfm_plan_create_tuned": (ctypes.c_int, [ctypes.c_int, ctypes.POINTER(Instance), _I32, ctypes.POINTER(_I32),
                                               ctypes.c_void_p, ctypes.POINTER(_P)]),
    "bldfm_plan_destroy": (ctypes.c_int, [_P]),
    "bldfm_plan_sizes": (ctypes.c_int, [_P, ctypes.POINTER(ctypes.c_size_t), ctypes.POINTER(_I64), ctypes.POINTER(_I64), ctypes.POINTER(_I64)]),
    "bldfm_plan_offsets": (ctypes.c_int, [_P, _I32, ctypes.POINTER(_I64), ctypes.POINTER(_I64), ctypes.POINTER(_I64), ctypes.POINTER(_I32)]),
    "bldfm_plan_execute": (ctypes.c_int, [_P, _P, _P, _P, _P, _P, _P]),
    "bldfm_plan_execute_static": (ctypes.c_int, [_P, _P, _P, _P, _P, _P]),
    "bldfm_plan_timing_begin": (ctypes.c_int, [_P, _I32]),
    "bldfm_plan_timing_end": (ctypes.c_int, [_P, ctypes.POINTER(_D), ctypes.POINTER(_I32)]),
    "bldfm_stability": (ctypes.c_int, [_P, _I64, _P, _P, _P]),
    "bldfm_wind_fields": (ctypes.c_int, [_P, _P, _I64, _P, _P, _P]),
    "bldfm_math_probe": (ctypes.c_int, [_I32, _P, _P, _I64, _P, _P]),
    "bldfm_fft2_workspace_bytes": (ctypes.c_size_t, [_I32, _I32, _I32]),
    "bldfm_fft2": (ctypes.c_int, [ctypes.c_int, _I32, _I32, _I32, _P, _P, _I32, _D, _P, _P]),
    "bldfm_sur_sizes": (ctypes.c_int, [ctypes.POINTER(_I64), ctypes.POINTER(_I64), ctypes.POINTER(_I64), ctypes.POINTER(_I32)]),
    "bldfm_sur_met": (ctypes.c_int, [ctypes.POINTER(Met), _I32, _P]),
    "bldfm_transfer_tables": (ctypes.c_int, [ctypes.c_int, ctypes.POINTER(Met), _I32, _I32, _P, _P, _P]),
    "bldfm_transfer_solve": (ctypes.c_int, [_P, _I32, _P, _P, _I32, _P, _I32, _P, _P, _I32, _P]),
    "bldfm_sur_pack": (ctypes.c_int, [_P, _P, _P]),
    "bldfm_sur_forward": (ctypes.c_int, [_P, _P, _I32, _P, _P, _I64, _P, _P]),
    "bldfm_sur_train_grad": (ctypes.c_int, [_P, _P, _I32, _P, _P, _P, _I64, _P, _P, _P, _P, _P]),
    "bldfm_sur_adam": (ctypes.c_int, [_P, _P, _P, _P, _P, ctypes.c_float, ctypes.c_float, ctypes.c_float, ctypes.c_float, _I32, _P]),
    "bldfm_plan_set_surrogate": (ctypes.c_int, [_P, _P, ctypes.POINTER(SurHull), _D]),
    "bldfm_plan_sur_stats": (ctypes.c_int, [_P, _P, _P, _P, _P]),
    "bldfm_km01_grid": (ctypes.c_int, [ctypes.POINTER(Km01), ctypes.POINTER(_I32), ctypes.POINTER(_I32)]),
    "bldfm_km01_footprint": (ctypes.c_int, [ctypes.c_int, ctypes.POINTER(Km01), _I32, _P, _P, _P]),
}

_lock = threading.Lock()
_lib = None


class BLDFMError(RuntimeError):
    """Error reported by the HIP library."""


def load():
    """Load the HIP library (raises if it is missing: there is no fallback)."""
    global _lib
    with _lock:
        if _lib is not None:
            return _lib
        if not LIB_PATH.exists():
            raise RuntimeError(
                f"BLDFM HIP library {LIB_PATH} is not built; run `python -m bldfm_amd.build`"
            )
        lib = ctypes.CDLL(str(LIB_PATH))
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(lib, name)
            fn.restype = res
            fn.argtypes = args
        _lib = lib
        return lib


def mem_info(device=None):
    """(bytes held by the library's caches, free, total) on the device."""
    dev = device if device is not None else torch.cuda.current_device()
    c, f, t = ctypes.c_size_t(), ctypes.c_size_t(), ctypes.c_size_t()
    check(load().bldfm_mem_info(int(dev), ctypes.byref(c), ctypes.byref(f), ctypes.byref(t)), "mem_info")
    return c.value, f.value, t.value


def last_error():
    """The library's last error message."""
    return load().bldfm_last_error().decode(errors="replace")


def check(rc, what="BLDFM call"):
    """Raise ValueError for argument errors, BLDFMError for the rest."""
    if rc == 0:
        return
    msg = load().bldfm_last_error().decode(errors="replace")
    if rc == -1:
        raise ValueError(msg)
    raise BLDFMError(f"{what} failed ({rc}): {msg}")


def device():
    """The torch device the solver runs on; raises without a HIP GPU."""
    if not torch.cuda.is_available():
        raise RuntimeError("BLDFM needs a HIP GPU (torch.cuda.is_available() is False)")
    return torch.device("cuda", torch.cuda.current_device())


def stream_handle(dev=None):
    return ctypes.c_void_p(torch.cuda.current_stream(dev).cuda_stream)


def ptr(t):
    return ctypes.c_void_p(t.data_ptr()) if t is not None else ctypes.c_void_p(0)


def make_met(n, meas_height, wind, ustar, z0, mol, prsc, closure, domain_height, stretch, tke, prologue=True):
    """Fill a ``Met`` from vertical_profiles arguments (closure validated here).

    With ``prologue`` (the default) the scalar prologue -- the derived z0 or
    u* and the grid constants -- is evaluated here with numpy exactly as the
    reference evaluates it (``bldfm_amd.prologue``), and the library only
    evaluates the per-level profiles; otherwise the library derives them with
    correctly rounded functions.
    """
    if closure not in CLOSURES:
        raise ValueError(
            f"Invalid closure type: {closure}. "
            "Supported closures are 'MOST', 'CONSTANT', and 'OAAHOC'."
        )
    um, vm = wind
    m = Met()
    m.n = int(n)
    m.closure = CLOSURES[closure]
    m.has_ustar = ustar is not None
    m.has_z0 = z0 is not None
    m.has_domain_height = domain_height is not None
    m.has_stretch = stretch is not None
    m.meas_height = float(meas_height)
    m.um = float(um)
    m.vm = float(vm)
    m.ustar = float(ustar) if ustar is not None else 0.0
    m.z0 = float(z0) if z0 is not None else 0.0
    m.mol = float(mol)
    m.prsc = float(prsc)
    m.domain_height = float(domain_height) if domain_height is not None else 0.0
    m.stretch = float(stretch) if stretch is not None else 0.0
    m.tke = float(np.squeeze(tke)) if tke is not None and np.size(tke) == 1 else 1.0
    if prologue and (closure == "OAAHOC" or (ustar is None) != (z0 is None)):
        from .prologue import prologue as _prologue

        pz0, pus, aa, bb, zetamx = _prologue(meas_height, wind, ustar, z0, mol, closure, domain_height, stretch, tke)
        m.has_prologue = 1
        m.has_ustar = m.has_z0 = 1
        m.z0, m.ustar = pz0, pus
        m.grid_aa, m.grid_bb, m.grid_zetamax = aa, bb, zetamx
    return m


def make_mets(specs):
    """``make_met`` for a list of specs (objects with the vertical_profiles
    fields n, meas_height, wind, ustar, z0, mol, prsc, closure,
    domain_height, stretch, tke).  The numpy prologue of MOST-type instances
    that share which of u*/z0 is given and whether domain_height/stretch are
    set is evaluated in one vectorised pass (same values as one at a time)."""
    from .prologue import prologue_many

    mets = [make_met(s.n, s.meas_height, s.wind, s.ustar, s.z0, s.mol, s.prsc, s.closure, s.domain_height,
                     s.stretch, s.tke, prologue=False) for s in specs]
    groups = {}
    for i, s in enumerate(specs):
        if s.closure in ("CONSTANT", "MOST", "MOSTM") and (s.ustar is None) != (s.z0 is None):
            key = (s.closure, s.ustar is None, s.domain_height is None, s.stretch is None)
            groups.setdefault(key, []).append(i)
        else:
            mets[i] = make_met(s.n, s.meas_height, s.wind, s.ustar, s.z0, s.mol, s.prsc, s.closure,
                               s.domain_height, s.stretch, s.tke)
    for (closure, z0_given, no_dh, no_st), idx in groups.items():
        sp = [specs[i] for i in idx]
        col = lambda f: [getattr(s, f) for s in sp]  # noqa: E731
        z0, us, aa, bb, zmx = prologue_many(
            col("meas_height"), [s.wind[0] for s in sp], [s.wind[1] for s in sp],
            None if z0_given else col("ustar"), col("z0") if z0_given else None, col("mol"), closure,
            None if no_dh else col("domain_height"), None if no_st else col("stretch"))
        for j, i in enumerate(idx):
            m = mets[i]
            m.has_prologue = 1
            m.has_ustar = m.has_z0 = 1
            m.z0, m.ustar = float(z0[j]), float(us[j])
            m.grid_aa, m.grid_bb, m.grid_zetamax = float(aa[j]), float(bb[j]), float(zmx[j])
    return mets


class Tuning(ctypes.Structure):
    """``bldfm_tuning``: launch shape of a plan (see bldfm_amd.autotune)."""

    _fields_ = [("shoot_block", _I32), ("xpass_layout", _I32), ("fft_threads", _I32), ("reserved", _I32 * 5)]

    def as_dict(self):
        return {"shoot_block": self.shoot_block or 256, "xpass_layout": self.xpass_layout,
                "fft_threads": self.fft_threads}


def make_tuning(shoot_block=256, xpass_layout=0, fft_threads=0):
    """fft_threads 0: the plan's default workgroup size for its FFT passes."""
    t = Tuning()
    t.shoot_block = int(shoot_block)
    t.xpass_layout = int(xpass_layout)
    t.fft_threads = int(fft_threads)
    return t


class Plan:
    """Owns a ``bldfm_plan``; sizes and per-instance offsets are cached.
    ``tuning`` (a dict with shoot_block / xpass_layout, or a Tuning) sets
    the launch shape; None keeps the defaults."""

    def __init__(self, instances, levels_i32, dev=None, tuning=None):
        lib = load()
        self.device = dev if dev is not None else device()
        n = len(instances)
        arr = (Instance * n)(*instances)
        lv = (ctypes.c_int32 * len(levels_i32))(*[int(v) for v in levels_i32])
        handle = ctypes.c_void_p()
        if isinstance(tuning, dict):
            tuning = make_tuning(**tuning)
        self.tuning = tuning.as_dict() if tuning is not None else make_tuning().as_dict()
        tp = ctypes.byref(tuning) if tuning is not None else None
        check(lib.bldfm_plan_create_tuned(self.device.index, arr, n, lv, tp, ctypes.byref(handle)), "plan_create")
        self.handle = handle
        ws = ctypes.c_size_t()
        prof, src, out = _I64(), _I64(), _I64()
        check(lib.bldfm_plan_sizes(handle, ctypes.byref(ws), ctypes.byref(prof), ctypes.byref(src), ctypes.byref(out)))
        self.workspace_bytes = ws.value
        self.prof_doubles = prof.value
        self.src_doubles = src.value
        self.out_elems = out.value
        self.offsets = []
        for i in range(n):
            op, os_, oo, nz = _I64(), _I64(), _I64(), _I32()
            check(lib.bldfm_plan_offsets(handle, i, ctypes.byref(op), ctypes.byref(os_), ctypes.byref(oo), ctypes.byref(nz)))
            self.offsets.append((op.value, os_.value, oo.value, nz.value))

    def execute(self, prof, src, conc, flx, workspace, stream=None):
        lib = load()
        s = stream if stream is not None else stream_handle(self.device)
        check(
            lib.bldfm_plan_execute(self.handle, ptr(prof), ptr(src), ptr(conc), ptr(flx), ptr(workspace), s),
            "plan_execute",
        )

    def execute_static(self, prof, src, conc, flx, stream=None):
        """``execute`` with the library's static per-GPU scratch as workspace
        (no device allocation; plans up to BLDFM_STATIC_SCRATCH_BYTES of
        workspace; overlapping calls on one GPU race)."""
        s = stream if stream is not None else stream_handle(self.device)
        check(load().bldfm_plan_execute_static(self.handle, ptr(prof), ptr(src), ptr(conc), ptr(flx), s),
              "plan_execute_static")

    PHASES = ("level_table", "forward_fft", "surrogate", "shoot", "nyquist_fix", "inverse_y", "inverse_x")

    def timing_begin(self, max_execs):
        check(load().bldfm_plan_timing_begin(self.handle, int(max_execs)), "timing_begin")

    def timing_end(self):
        """Summed milliseconds per phase over the timed executions, and their count."""
        ms = (ctypes.c_double * len(self.PHASES))()
        n = _I32()
        check(load().bldfm_plan_timing_end(self.handle, ms, ctypes.byref(n)), "timing_end")
        return dict(zip(self.PHASES, list(ms))), n.value

    def update_met(self, mets):
        """Replace the meteorology of the plan's instances (a list of Met,
        one per instance; each keeps its number of vertical points)."""
        n = len(mets)
        arr = (Met * n)(*mets)
        check(load().bldfm_plan_update_met(self.handle, arr, n, stream_handle(self.device)), "plan_update_met")

    def set_surrogate(self, net, hull, tol):
        """Serve surrogate-flagged instances with the network buffer ``net``
        (a device tensor, or None to switch the surrogate off)."""
        check(load().bldfm_plan_set_surrogate(self.handle, ptr(net), ctypes.byref(hull) if hull is not None else None,
                                              float(tol)), "plan_set_surrogate")

    def surrogate_stats(self, workspace, n):
        """Per-instance spot-check error (NaN if not served) and path (0 surrogate, 1 solver)."""
        err = (ctypes.c_double * n)()
        mode = (ctypes.c_int32 * n)()
        check(load().bldfm_plan_sur_stats(self.handle, ptr(workspace), err, mode, stream_handle(self.device)),
              "plan_sur_stats")
        return list(err), list(mode)

    def __del__(self):
        h = getattr(self, "handle", None)
        if h is not None and _lib is not None:
            _lib.bldfm_plan_destroy(h)
            self.handle = None
