// Device/host numerics shared by the BLDFM kernels: FP64 complex helpers and
// the Monin-Obukhov profile formulas.
#pragma once
#include <hip/hip_runtime.h>
#include <math.h>
#include "bldfm_types.h"
#include "bldfm_crmath.h"

// ------------------------------------------------------------------ complex
BLDFM_HD double2 cmk(double re, double im) { double2 r; r.x = re; r.y = im; return r; }
BLDFM_HD double2 cadd(double2 a, double2 b) { return cmk(a.x + b.x, a.y + b.y); }
BLDFM_HD double2 csub(double2 a, double2 b) { return cmk(a.x - b.x, a.y - b.y); }
BLDFM_HD double2 cmul(double2 a, double2 b) { return cmk(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x); }
BLDFM_HD double2 cscale(double s, double2 a) { return cmk(s * a.x, s * a.y); }
BLDFM_HD double2 cconj(double2 a) { return cmk(a.x, -a.y); }
// a * conj(b) when cj, else a * b
BLDFM_HD double2 cmul_cj(double2 a, double2 b, bool cj) { return cmul(a, cj ? cconj(b) : b); }

// Complex division by Smith's method (R. L. Smith, CACM 5(8):435, 1962),
// the scaling numpy's complex divide uses.
BLDFM_HD double2 cdiv(double2 a, double2 b) {
#pragma clang fp contract(off)
  if (fabs(b.x) >= fabs(b.y)) {
    const double rat = b.y / b.x;
    const double scl = 1.0 / (b.x + b.y * rat);
    return cmk((a.x + a.y * rat) * scl, (a.y - a.x * rat) * scl);
  }
  const double rat = b.x / b.y;
  const double scl = 1.0 / (b.y + b.x * rat);
  return cmk((a.x * rat + a.y) * scl, (a.y * rat - a.x) * scl);
}

// Principal square root: t = sqrt((|x| + |z|)/2); for x >= 0 the root is
// (t, y/(2t)), else (|y|/(2t), copysign(t, y)).
BLDFM_HD double2 csqrt_principal(double2 z) {
  const double x = z.x, y = z.y;
  if (x == 0.0 && y == 0.0) return cmk(0.0, y);
  const double t = sqrt(0.5 * (fabs(x) + hypot(x, y)));
  if (x >= 0.0) return cmk(t, y / (2.0 * t));
  return cmk(fabs(y) / (2.0 * t), copysign(t, y));
}

// e^{z} = e^{x} (cos y, sin y)
BLDFM_HD double2 cexp_(double2 z) {
  double s, c;
  sincos(z.y, &s, &c);
  const double e = exp(z.x);
  return cmk(e * c, e * s);
}

// Round a complex value through complex64 (precision="single" spectra).
BLDFM_HD double2 round_c64(double2 a) { return cmk((double)(float)a.x, (double)(float)a.y); }

// ------------------------------------------------ uncontracted (exact) helpers
// Every product and sum below is rounded on its own (no fused multiply-add),
// in the operation order numpy and numba use for complex arithmetic:
// (a+ib)(c+id) = (ac - bd) + i(ad + bc).  Used where the reference's rounding
// must be reproduced bit for bit (the shooting cancellation amplifies any
// ULP-level difference).
BLDFM_HD double xmul(double a, double b) {
#pragma clang fp contract(off)
  return a * b;
}
BLDFM_HD double2 xcmul(double2 a, double2 b) {
#pragma clang fp contract(off)
  return cmk(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
BLDFM_HD double2 xcadd(double2 a, double2 b) {
#pragma clang fp contract(off)
  return cmk(a.x + b.x, a.y + b.y);
}
BLDFM_HD double2 xcsub(double2 a, double2 b) {
#pragma clang fp contract(off)
  return cmk(a.x - b.x, a.y - b.y);
}
// real * complex with the real promoted to (s + 0i): (s*x - 0*y, s*y + 0*x),
// which equals (s*x, s*y) for finite values.
BLDFM_HD double2 xcscale(double s, double2 a) {
#pragma clang fp contract(off)
  return cmk(s * a.x, s * a.y);
}
// a*b + c*d for complex values, each product rounded, then the sum.
BLDFM_HD double2 xcmuladd(double2 a, double2 b, double2 c, double2 d) {
#pragma clang fp contract(off)
  const double2 ab = cmk(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
  const double2 cd = cmk(c.x * d.x - c.y * d.y, c.x * d.y + c.y * d.x);
  return cmk(ab.x + cd.x, ab.y + cd.y);
}

// Complex product rounded as numpy's vectorised (AVX-512/FMA3) complex
// multiply rounds it: re = fma(ar, br, -(ai*bi)), im = fma(ar, bi, ai*br).
// The reference evaluates its shooting coefficient, the combination of the
// two solutions and the phase shift with numpy array arithmetic
// (src/bldfm/solver.py:228-262), so these steps use this rounding.
BLDFM_HD double2 ncmul(double2 a, double2 b) {
#pragma clang fp contract(off)
  return cmk(fma(a.x, b.x, -(a.y * b.y)), fma(a.x, b.y, a.y * b.x));
}

// hypot(x, y) by C. F. Borges' correction of the naive formula ("An Improved
// Algorithm for hypot(a,b)", ACM TOMS 2021, variant without fused
// multiply-add): h = sqrt(ax^2 + ay^2) corrected by -(t1 + t2)/(2h).  This is
// the rounding the x86-64 glibc 2.35 hypot produces.
BLDFM_HD double hypot_borges(double x, double y) {
#pragma clang fp contract(off)
  double ax = fabs(x), ay = fabs(y);
  if (ax < ay) { const double t = ax; ax = ay; ay = t; }
  if (ay == 0.0 || !(ax <= 1.7976931348623157e308)) return ax + ay;
  double scale = 1.0;
  if (ax > 0x1p+500) { ax *= 0x1p-600; ay *= 0x1p-600; scale = 0x1p+600; }
  else if (ay < 0x1p-500) { ax *= 0x1p+600; ay *= 0x1p+600; scale = 0x1p-600; }
  double h = sqrt(ax * ax + ay * ay);
  double t1, t2;
  if (h <= 2.0 * ay) {
    const double delta = h - ay;
    t1 = ax * (2.0 * delta - ax);
    t2 = (delta - 2.0 * (ax - ay)) * delta;
  } else {
    const double delta = h - ax;
    t1 = 2.0 * delta * (ax - 2.0 * ay);
    t2 = (4.0 * delta - ay) * ay + delta * delta;
  }
  h -= (t1 + t2) / (2.0 * h);
  return h * scale;
}

// Principal complex square root with the branch structure of C99 csqrt as
// numpy evaluates it: d = |z|; for x > 0, r = sqrt((d + x)/2), s = (y/r)/2;
// for x < 0, s = sqrt((d - x)/2), r = |(y/s)/2|; result (r, copysign(s, y)).
BLDFM_HD double2 csqrt_exact(double2 z) {
#pragma clang fp contract(off)
  const double x = z.x, y = z.y;
  if (y == 0.0) {
    if (x < 0.0) return cmk(0.0, copysign(sqrt(-x), y));
    return cmk(fabs(sqrt(x)), copysign(0.0, y));
  }
  if (x == 0.0) {
    const double r = sqrt(0.5 * fabs(y));
    return cmk(r, copysign(r, y));
  }
  const double d = hypot_borges(x, y);
  double r, s;
  if (x > 0.0) {
    r = sqrt(0.5 * (d + x));
    s = 0.5 * (y / r);
  } else {
    s = sqrt(0.5 * (d - x));
    r = fabs(0.5 * (y / s));
  }
  return cmk(r, copysign(s, y));
}

// ------------------------------------------------------------------ profiles
// Closure ids (src/bldfm/pbl_model.py:65-182).
enum Closure : int32_t { CL_MOST = 0, CL_CONSTANT = 1, CL_MOSTM = 2, CL_OAAHOC = 3 };

#define BLDFM_KARMAN 0.4
#define BLDFM_PI 3.141592653589793

// The profile formulas below round as the reference's numpy/glibc evaluation
// does: every product and sum on its own, in the reference's order (no
// contraction), with the correctly rounded or glibc-compatible functions of
// bldfm_crmath.h in place of the device libm.  EXACT = false swaps in the
// plain libm functions for callers that round the result to float anyway
// (the surrogate's input features), where the double-double cost buys nothing.
template <bool EXACT> BLDFM_HD double pf_log(double x) { return EXACT ? cr_log(x) : log(x); }
template <bool EXACT> BLDFM_HD double pf_atan(double x) { return EXACT ? cr_atan(x) : atan(x); }
template <bool EXACT> BLDFM_HD double pf_cpow(double v, double p) {
  return EXACT ? gl_cpow_real(v, p) : exp(p * log(v));
}

// MOST momentum stability correction psi(z/L) (src/bldfm/pbl_model.py:207-230):
// 5x for x > 0, else with xi = Re((1 - 16x)^(1/4)) (complex power)
// -2 ln((1 + xi)/2) - ln((1 + xi^2)/2) + 2 atan(xi) - pi/2.
template <bool EXACT = true>
BLDFM_HD double most_psi(double x) {
  CR_EXACT
  if (x > 0.0) return 5.0 * x;
  const double xi = pf_cpow<EXACT>(1.0 - 16.0 * x, 0.25);
  return -2.0 * pf_log<EXACT>(0.5 * (1.0 + xi)) - pf_log<EXACT>(0.5 * (1.0 + xi * xi)) + 2.0 * pf_atan<EXACT>(xi) -
         0.5 * BLDFM_PI;
}

// MOST diffusivity stability function phi(z/L) (src/bldfm/pbl_model.py:233-250).
template <bool EXACT = true>
BLDFM_HD double most_phi(double x) {
  CR_EXACT
  if (x > 0.0) return 1.0 + 5.0 * x;
  return pf_cpow<EXACT>(1.0 - 16.0 * x, -0.5);
}

// Met wind (speed, direction blowing FROM, degrees) -> (u, v)
// (src/bldfm/utils.py:23-27): theta = dir * (pi/180) as numpy's deg2rad
// rounds it, u = -U sin(theta), v = -U cos(theta).
BLDFM_HD void wind_uv(double spd, double dir, double* u, double* v) {
  CR_EXACT
  const double th = dir * (BLDFM_PI / 180.0);
  double s, c;
  cr_sincos(th, &s, &c);
  *u = -spd * s;
  *v = -spd * c;
}

// Scalar prologue of vertical_profiles: derived roughness or friction
// velocity and the stretched-grid constants (src/bldfm/pbl_model.py:58-129).
struct ProfSetup {
  double zm, um, vm, absum, ustar, z0, mol, prsc, tke;
  double h, aa, bb, dzeta;
  int32_t closure, nlev;
};

// Returns 0, or a negative error code: -1 both/neither of ustar,z0 given
// for a MOST-type closure, -2 invalid closure, -3 empty grid.
BLDFM_HD int prof_setup(int n, double zm, double um, double vm, int has_ustar, double ustar,
                        int has_z0, double z0, double mol, double prsc, int closure,
                        int has_domain_height, double domain_height, int has_stretch,
                        double stretch, double tke, ProfSetup* s) {
  CR_EXACT
  const double kap = BLDFM_KARMAN;
  s->zm = zm; s->um = um; s->vm = vm; s->mol = mol; s->prsc = prsc; s->closure = closure;
  s->tke = tke;
  s->absum = sqrt(um * um + vm * vm);
  if (closure == CL_MOST || closure == CL_CONSTANT || closure == CL_MOSTM) {
    if (!has_z0 && has_ustar) {
      z0 = zm * cr_exp(-kap * s->absum / ustar + most_psi(zm / mol));
    } else if (!has_ustar && has_z0) {
      ustar = s->absum * kap / (cr_log(zm / z0) + most_psi(zm / mol));
    } else {
      return -1;
    }
  } else if (closure == CL_OAAHOC) {
    const double cl = 0.845, cm = 0.0856;
    z0 = zm * cr_exp(-cm * cl * s->absum * sqrt(tke) / (ustar * ustar));
  } else {
    return -2;
  }
  s->ustar = ustar;
  s->z0 = z0;
  s->h = has_stretch ? stretch : 2.0 * zm;
  const double zmx = has_domain_height ? domain_height : 2.0 * zm;
  s->bb = zm / (cr_exp(-z0 / s->h) - cr_exp(-zm / s->h));
  s->aa = s->bb * cr_exp(-z0 / s->h);
  const double zetamx = s->aa - s->bb * cr_exp(-zmx / s->h);
  s->dzeta = zm / n;
  // numpy arange length: ceil((stop - start) / step)
  const double len = ceil((zetamx + s->dzeta) / s->dzeta);
  if (!(len >= 1.0)) return -3;
  s->nlev = (int32_t)len;
  return 0;
}

// The same prologue from a caller's values of the derived z0 / u* and the
// grid constants aa, bb, zeta_max (the reference's own numpy evaluation of
// src/bldfm/pbl_model.py:65-120); h, dzeta and the numpy arange length
// ceil((zeta_max + dzeta)/dzeta) are formed here.  Returns 0, -2 for an
// invalid closure, -3 for an empty grid, -4 when the length is not finite.
BLDFM_HD int prof_setup_given(int n, double zm, double um, double vm, double ustar, double z0, double mol,
                              double prsc, int closure, int has_stretch, double stretch, double tke,
                              double aa, double bb, double zetamx, ProfSetup* s) {
  CR_EXACT
  if (closure != CL_MOST && closure != CL_CONSTANT && closure != CL_MOSTM && closure != CL_OAAHOC) return -2;
  s->zm = zm; s->um = um; s->vm = vm; s->mol = mol; s->prsc = prsc; s->closure = closure;
  s->tke = tke;
  s->absum = sqrt(um * um + vm * vm);
  s->ustar = ustar;
  s->z0 = z0;
  s->h = has_stretch ? stretch : 2.0 * zm;
  s->aa = aa;
  s->bb = bb;
  s->dzeta = zm / n;
  const double len = ceil((zetamx + s->dzeta) / s->dzeta);
  if (!(len == len) || len > 1e9) return -4;
  if (!(len >= 1.0)) return -3;
  s->nlev = (int32_t)len;
  return 0;
}

// Height and profile values at grid point i (src/bldfm/pbl_model.py:127-182).
// out = {z, u, v, Kx, Ky, Kz}
template <bool EXACT = true>
BLDFM_HD void prof_point(const ProfSetup& s, int i, double out[6]) {
  CR_EXACT
  const double kap = BLDFM_KARMAN;
  const double zeta = 0.0 + i * s.dzeta;
  const double z = -s.h * pf_log<EXACT>(-(zeta - s.aa) / s.bb);
  double u, v, kx, ky, kz;
  if (s.closure == CL_CONSTANT) {
    const double km = kap * s.ustar * s.zm / s.prsc;
    u = s.um; v = s.vm; kx = ky = kz = km;
  } else if (s.closure == CL_MOST || s.closure == CL_MOSTM) {
    const double absu = s.ustar / kap * (pf_log<EXACT>(z / s.z0) + most_psi<EXACT>(z / s.mol));
    u = s.um / s.absum * absu;
    v = s.vm / s.absum * absu;
    const double K = kap * s.ustar * z / most_phi<EXACT>(z / s.mol) / s.prsc;
    if (s.closure == CL_MOST) {
      kx = ky = kz = K;
    } else {
      kx = K * (v * v) / (u * u + v * v);
      ky = K * (u * u) / (u * u + v * v);
      kz = K;
    }
  } else {  // OAAHOC
    const double cl = 0.845, cm = 0.0856, ch = 0.204;
    const double absu = s.ustar * s.ustar / cm / cl / sqrt(s.tke) * pf_log<EXACT>(z / s.z0);
    u = s.um / s.absum * absu;
    v = s.vm / s.absum * absu;
    kx = ky = kz = ch * cl * z * sqrt(s.tke);
  }
  out[0] = z; out[1] = u; out[2] = v; out[3] = kx; out[4] = ky; out[5] = kz;
}
