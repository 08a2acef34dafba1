"""CPU tests of the device math that feeds the profile and wind formulas.

tests/native/crmath_shim.cpp compiles bldfm_crmath.h and bldfm_math.h (the
code the kernels run) for the host with contraction off; these tests pin it:

* the correctly rounded functions against a 70-digit Decimal evaluation
  (exp and ln from Decimal itself; atan, sin, cos from their series after
  exact argument reduction) -- bit for bit on every sample;
* the glibc-compatible exp and log1p against this host's libm, and the
  complex-power model against numpy's complex power, bit for bit;
* psi, phi, the wind components and whole profiles against the oracle
  (numpy), with the scalar prologue evaluated by ``bldfm_amd.prologue`` as the
  drop-in does.  numpy's own vectorised log (near 1) and exp are not
  correctly rounded on AVX-512 hosts; the residual one-ulp differences they
  cause are measured and bounded here.
"""

import ctypes
import shutil
import subprocess
from decimal import Decimal, getcontext
from pathlib import Path

import numpy as np
import pytest

from oracle import pbl as opbl

ROOT = Path(__file__).resolve().parent.parent
SHIM = ROOT / "tests" / "native" / "crmath_shim.cpp"
CLOSURE = {"MOST": 0, "CONSTANT": 1, "MOSTM": 2, "OAAHOC": 3}


@pytest.fixture(scope="module")
def shim(tmp_path_factory):
    cxx = shutil.which("g++")
    if cxx is None:
        pytest.skip("g++ not available")
    so = tmp_path_factory.mktemp("crm") / "crm.so"
    subprocess.run([cxx, "-O2", "-std=c++17", "-ffp-contract=off", "-fPIC", "-shared", "-Wno-unknown-pragmas",
                    "-D__HIP_PLATFORM_AMD__", "-I/opt/rocm/include", str(SHIM), "-o", str(so)], check=True)
    return ctypes.CDLL(str(so))


def ulps(a, b):
    """|a - b| in units of the last place of the larger magnitude."""
    a, b = np.asarray(a, dtype=np.float64), np.asarray(b, dtype=np.float64)
    return np.abs(a - b) / np.spacing(np.maximum(np.abs(a), np.abs(b)))


def ev(lib, op, x):
    x = np.ascontiguousarray(x, dtype=np.float64)
    out = np.empty_like(x)
    lib.crm_eval(ctypes.c_int(op), x.ctypes.data_as(ctypes.c_void_p), ctypes.c_int64(x.size),
                 out.ctypes.data_as(ctypes.c_void_p))
    return out


def host_wind(lib, spd, wd):
    spd = np.ascontiguousarray(spd, dtype=np.float64)
    wd = np.ascontiguousarray(wd, dtype=np.float64)
    u, v = np.empty_like(wd), np.empty_like(wd)
    lib.crm_wind(spd.ctypes.data_as(ctypes.c_void_p), wd.ctypes.data_as(ctypes.c_void_p), ctypes.c_int64(wd.size),
                 u.ctypes.data_as(ctypes.c_void_p), v.ctypes.data_as(ctypes.c_void_p))
    return u, v


def host_profiles(lib, n, meas_height, wind, ustar=None, z0=None, mol=1e9, prsc=1.0, closure="MOST",
                  domain_height=None, stretch=None, tke=None):
    """vertical_profiles through the host build, prologue from bldfm_amd.prologue."""
    from bldfm_amd.prologue import prologue

    pz0, pus, aa, bb, zmx = prologue(meas_height, wind, ustar, z0, mol, closure, domain_height, stretch, tke)
    cap = 4096
    out = np.zeros(6 * cap)
    D = ctypes.c_double
    nl = lib.crm_profiles_given(
        ctypes.c_int(n), D(meas_height), D(wind[0]), D(wind[1]), D(pus), D(pz0), D(mol), D(prsc),
        ctypes.c_int(CLOSURE[closure]), ctypes.c_int(stretch is not None), D(stretch or 0.0),
        D(1.0 if tke is None else float(np.squeeze(tke))), D(aa), D(bb), D(zmx),
        out.ctypes.data_as(ctypes.c_void_p), ctypes.c_int(cap))
    assert nl > 0, nl
    o = out.reshape(6, cap)[:, :nl].copy()
    return o[0], tuple(o[1:])


# ------------------------------------------------------- Decimal references
getcontext().prec = 70


def _dec_pi():
    def atan_inv(n):
        x = Decimal(1) / n
        x2, s, t, k, sign = x * x, Decimal(0), x, 1, 1
        while True:
            term = t / k
            if term < Decimal(10) ** -68:
                return s
            s += sign * term
            t *= x2
            k += 2
            sign = -sign

    return 4 * (4 * atan_inv(5) - atan_inv(239))  # Machin


PI = _dec_pi()


def dec_atan(x):
    x = Decimal(x)
    neg, x = x < 0, abs(x)
    base, sg = Decimal(0), 1
    if x > 1:
        x, base, sg = 1 / x, PI / 2, -1
    for _ in range(6):  # atan x = 2 atan(x / (1 + sqrt(1 + x^2)))
        x = x / (1 + (1 + x * x).sqrt())
    s, t, k, sign = Decimal(0), x, 1, 1
    while abs(t / k) >= Decimal(10) ** -68:
        s += sign * t / k
        t *= x * x
        k += 2
        sign = -sign
    r = base + sg * 64 * s
    return -r if neg else r


def dec_sincos(x):
    x = Decimal(x)
    x = x - 2 * PI * (x / (2 * PI)).to_integral_value()
    s, c, t, k = Decimal(0), Decimal(0), Decimal(1), 0
    while abs(t) >= Decimal(10) ** -68 or k < 2:
        # t = x^k / k!
        if k % 2 == 0:
            c += (-1) ** (k // 2) * t
        else:
            s += (-1) ** (k // 2) * t
        k += 1
        t = t * x / k
    return s, c


def test_pi_reference_digits():
    assert str(PI).startswith("3.14159265358979323846264338327950288419716939937510")


def test_split_constants_sum_to_their_values(shim):
    """The constants of bldfm_crmath.h: ln 2 in three parts, pi/2 in 33-bit
    parts and as double-doubles, recomputed here in 70 digits."""
    text = (ROOT / "bldfm_amd" / "csrc" / "bldfm_crmath.h").read_text()

    def c(name):
        import re

        return Decimal(float.fromhex(re.search(rf"#define {name} (\S+)", text).group(1)))

    ln2 = Decimal(2).ln()
    assert abs(c("CR_LN2_HI") + c("CR_LN2_MID") + c("CR_LN2_LO") - ln2) < Decimal(2) ** -160
    assert abs(c("CR_PIO2_1") + c("CR_PIO2_2") + c("CR_PIO2_3") - PI / 2) < Decimal(2) ** -120
    assert abs(c("CR_PIO2_HI") + c("CR_PIO2_LO") - PI / 2) < Decimal(2) ** -106
    assert abs(c("CR_PIO4_HI") + c("CR_PIO4_LO") - PI / 4) < Decimal(2) ** -107
    for name in ("CR_PIO2_1", "CR_PIO2_2"):  # 33 significant bits: k * part is exact for |k| < 2^20
        m, _ = np.frexp(float(c(name)))
        assert float(m * 2**33) == int(m * 2**33)


@pytest.mark.parametrize(
    "op,name,ref,gen",
    [
        (0, "exp", lambda x: Decimal(x).exp(),
         lambda r: np.concatenate([r.uniform(-40, 10, 800), r.uniform(-1e-3, 1e-3, 200), r.uniform(-700, 700, 200)])),
        (1, "log", lambda x: Decimal(x).ln(),
         lambda r: np.concatenate([r.uniform(1e-3, 300, 800), 1 + r.uniform(-0.07, 0.07, 300), 10 ** r.uniform(-300, 300, 200)])),
        (2, "log1p", lambda x: (1 + Decimal(x)).ln(),
         lambda r: np.concatenate([r.uniform(0, 3, 800), r.uniform(-0.07, 0.07, 300), 10 ** r.uniform(-20, -1, 200)])),
        (3, "atan", dec_atan,
         lambda r: np.concatenate([r.uniform(0.9, 10, 600), r.uniform(-3, 3, 300), 10 ** r.uniform(-10, 10, 200)])),
        (4, "sin", lambda x: dec_sincos(x)[0], lambda r: np.concatenate([r.uniform(0, 7, 700), r.uniform(-1e3, 1e3, 200)])),
        (5, "cos", lambda x: dec_sincos(x)[1], lambda r: np.concatenate([r.uniform(0, 7, 700), r.uniform(-1e3, 1e3, 200)])),
    ],
)
def test_correct_rounding_against_decimal(shim, op, name, ref, gen):
    xs = gen(np.random.default_rng(op + 100))
    got = ev(shim, op, xs)
    want = np.array([float(ref(float(x))) for x in xs])  # float(Decimal) rounds correctly
    bad = np.flatnonzero(got != want)
    assert bad.size == 0, [(xs[i].hex(), got[i].hex(), want[i].hex()) for i in bad[:5]]


def _has_fma():
    try:
        return " fma " in Path("/proc/cpuinfo").read_text()
    except OSError:
        return False


@pytest.mark.skipif(not _has_fma(), reason="glibc's exp follows its FMA variant on FMA hosts only")
def test_glibc_exp_and_log1p_bit_for_bit(shim):
    libm = ctypes.CDLL("libm.so.6")
    for f in ("exp", "log1p"):
        getattr(libm, f).restype = ctypes.c_double
        getattr(libm, f).argtypes = [ctypes.c_double]
    r = np.random.default_rng(5)
    x = np.concatenate([r.uniform(-5, 5, 40000), r.uniform(-500, 500, 5000), r.uniform(-1e-3, 1e-3, 2000)])
    assert np.array_equal(ev(shim, 6, x), np.array([libm.exp(float(a)) for a in x]))
    x = np.concatenate([r.uniform(0, 3, 40000), r.uniform(-0.29, 0.41, 5000), 10 ** r.uniform(-12, 3, 5000)])
    assert np.array_equal(ev(shim, 7, x), np.array([libm.log1p(float(a)) for a in x]))


def test_complex_power_model_matches_numpy(shim):
    """numpy's complex power of a real base >= 1 (psi, phi) is glibc's
    cexp(p clog(v)); the model reproduces it bit for bit."""
    r = np.random.default_rng(6)
    v = np.concatenate([1 + r.uniform(0, 1, 20000), r.uniform(2, 200, 20000), [1.0, 2.0, 1.0 + 2**-52]])
    for op, p in ((8, 0.25), (9, -0.5)):
        got, want = ev(shim, op, v), np.power(v, p, dtype=complex).real
        # the log of bases >= 2 is correctly rounded here and glibc's is not
        # always; e^{p ln v} scales an ulp of ln v by |p ln v| (up to ~1.3 here)
        assert np.mean(got != want) < 5e-4
        assert np.all(ulps(got, want) <= 4)
    small = 1 + r.uniform(0, 1, 20000)  # bases in [1, 2): log1p path, bit for bit
    assert np.array_equal(ev(shim, 8, small), np.power(small, 0.25, dtype=complex).real)
    assert np.array_equal(ev(shim, 9, small), np.power(small, -0.5, dtype=complex).real)


def test_psi_phi_bit_for_bit_on_the_stability_grid(shim):
    x = np.concatenate([np.linspace(-5, 2, 301), [0.0, 1e-9, -1e-9]])
    assert np.array_equal(ev(shim, 10, x), opbl.psi(x))
    assert np.array_equal(ev(shim, 11, x), opbl.phi(x))
    xx = np.random.default_rng(7).uniform(-6, 0, 20000)
    for op, f in ((10, opbl.psi), (11, opbl.phi)):
        got, want = ev(shim, op, xx), f(xx)
        # numpy's log near 1 and glibc's log above 2; psi's sum cancels, which
        # turns a one-ulp difference of a term into a few ulps of the result
        assert np.mean(got != want) < 2e-3
        assert np.all(ulps(got, want) <= 8)


def test_wind_directions_every_tenth_degree(shim):
    d = np.round(np.arange(3601) * 0.1, 1)
    for spd in (1.0, 5.0, 7.3):
        u, v = host_wind(shim, np.full_like(d, spd), d)
        ou, ov = opbl.compute_wind_fields(spd, d)
        # glibc's sin/cos (numpy calls them) are not correctly rounded either;
        # a one-ulp difference of sin becomes up to 3 ulps of -U sin
        assert np.mean(u != ou) < 5e-3 and np.mean(v != ov) < 5e-3
        assert np.all(ulps(u, ou) <= 3) and np.all(ulps(v, ov) <= 3)
    for wd in (0.0, 90.0, 180.0, 270.0, 345.0):
        u, v = host_wind(shim, np.array([6.0]), np.array([wd]))
        assert (u[0], v[0]) == tuple(float(a) for a in opbl.compute_wind_fields(6.0, wd))


REFERENCE_PROFILE_CASES = [
    # the reference's tests/test_pbl_model.py inputs and the scenario met of the parity tests
    dict(n=16, meas_height=10.0, wind=(0.0, -6.0), ustar=0.5),
    dict(n=10, meas_height=5.0, wind=(3.0, 1.0), ustar=0.4, mol=-100.0),
    dict(n=20, meas_height=10.0, wind=(2.0, 2.0), ustar=0.3, mol=-50.0),
    dict(n=26, meas_height=15.0, wind=(-3.5, 0.1), ustar=0.3, mol=100.0, domain_height=100.0),
    dict(n=10, meas_height=5.0, wind=(3.0, 1.0), z0=0.1, mol=-100.0),
    dict(n=10, meas_height=5.0, wind=(3.0, 1.0), ustar=0.4, closure="CONSTANT", prsc=0.8),
    dict(n=20, meas_height=10.0, wind=(2.0, 2.0), ustar=0.3, closure="CONSTANT", prsc=0.8),
    dict(n=10, meas_height=5.0, wind=(3.0, 1.0), ustar=0.4, closure="MOSTM", mol=-100.0),
    dict(n=10, meas_height=5.0, wind=(3.0, 1.0), ustar=0.4, closure="OAAHOC"),
    dict(n=10, meas_height=5.0, wind=(3.0, 1.0), ustar=0.4, stretch=15.0, domain_height=50.0),
]


@pytest.mark.parametrize("kw", REFERENCE_PROFILE_CASES)
def test_profiles_bit_for_bit(shim, kw):
    z, p = host_profiles(shim, **kw)
    oz, op = opbl.vertical_profiles(**kw)
    assert np.array_equal(z, oz)
    for a, b in zip(p, op):
        assert np.array_equal(a, b)


def test_profiles_oaahoc_tke2_one_level_one_ulp(shim):
    """OAAHOC with tke=2 (tests/test_pbl_model.py:99-123 of the reference):
    one grid level's log argument lies in numpy's near-1 range, where numpy's
    vectorised log is not correctly rounded; everything else is bitwise."""
    kw = dict(n=10, meas_height=5.0, wind=(3.0, 1.0), ustar=0.4, closure="OAAHOC", tke=2.0)
    z, p = host_profiles(shim, **kw)
    oz, op = opbl.vertical_profiles(**kw)
    for a, b in zip((z,) + p, (oz,) + op):
        assert np.sum(a != b) <= 1
        assert np.all(ulps(a, b) <= 2)


def test_profiles_met_sweep(shim):
    """200 met cases drawn like the batched workload: profile values equal the
    oracle's except where numpy's own log near 1 rounds differently."""
    r = np.random.default_rng(9)
    cases = values = bad = 0
    for _ in range(260):
        ust, invl = r.uniform(0.1, 1.0), r.uniform(-0.1, 0.05)
        mol = 1.0 / invl if invl != 0 else 1e9
        zm, spd, wd = r.uniform(2, 50), r.uniform(1, 10), r.uniform(0, 360)
        w = opbl.compute_wind_fields(spd, wd)
        with np.errstate(all="ignore"):
            try:
                oz, op = opbl.vertical_profiles(26, zm, w, ustar=ust, mol=mol)
            except ValueError:  # empty grid (z0 far above zm)
                continue
            if len(oz) < 27 or not np.all(np.isfinite(oz)) or oz[0] >= zm:
                continue  # z0 above zm: not a surface layer
            z, p = host_profiles(shim, 26, zm, w, ustar=ust, mol=mol)
        cases += 1
        for a, b in ((z, oz), (p[0], op[0]), (p[1], op[1]), (p[4], op[4])):
            values += a.size
            bad += int(np.sum(a != b))
            assert np.all(np.isclose(a, b, rtol=1e-12, atol=1e-15))
    print(f"\n  {cases} met cases, {values} profile values, {bad} differ ({bad / values:.2%})")
    assert cases >= 200 and bad / values < 0.01


def test_vectorised_prologue_equals_scalar_prologue():
    """The batched drop-ins evaluate the numpy prologue over whole batches
    (bldfm_amd.prologue.prologue_many); each element equals the one-instance
    evaluation bit for bit, for both the u*-given and the z0-given forms."""
    from bldfm_amd.prologue import prologue, prologue_many

    r = np.random.default_rng(12)
    n = 3000
    zm, us, z0 = r.uniform(2, 50, n), r.uniform(0.1, 1.0, n), 10 ** r.uniform(-2, -0.5, n)
    mol = 1.0 / np.where(np.abs(iv := r.uniform(-0.1, 0.05, n)) < 1e-6, 1e-9, iv)
    sp, wd = r.uniform(1, 10, n), np.deg2rad(r.uniform(0, 360, n))
    um, vm = -sp * np.sin(wd), -sp * np.cos(wd)
    dh = zm * r.uniform(2, 8, n)
    for given in ("ustar", "z0"):
        many = prologue_many(zm, um, vm, us if given == "ustar" else None, z0 if given == "z0" else None, mol,
                             "MOST", dh, None)
        for i in range(0, n, 7):
            one = prologue(float(zm[i]), (float(um[i]), float(vm[i])), float(us[i]) if given == "ustar" else None,
                           float(z0[i]) if given == "z0" else None, float(mol[i]), "MOST", float(dh[i]), None, None)
            assert tuple(float(a[i]) for a in many) == one or any(np.isnan(one)), (given, i)
