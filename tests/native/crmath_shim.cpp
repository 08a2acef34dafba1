// Host build of the correctly rounded and glibc-compatible math
// (bldfm_crmath.h) and of the profile and wind formulas (bldfm_math.h) for
// CPU tests: the same source the kernels compile, evaluated over arrays
// through ctypes.  Built with -ffp-contract=off so that no multiply-add is
// fused on the host either.
#include <stdint.h>

#include "../../bldfm_amd/csrc/bldfm_math.h"

extern "C" {

// op: 0 exp, 1 log, 2 log1p, 3 atan, 4 sin, 5 cos (correctly rounded);
// 6 glibc exp, 7 glibc log1p, 8 glibc cpow(x, 1/4), 9 glibc cpow(x, -1/2);
// 10 psi, 11 phi
void crm_eval(int op, const double* x, int64_t n, double* out) {
  for (int64_t i = 0; i < n; ++i) {
    double s, c;
    switch (op) {
      case 0: out[i] = cr_exp(x[i]); break;
      case 1: out[i] = cr_log(x[i]); break;
      case 2: out[i] = cr_log1p(x[i]); break;
      case 3: out[i] = cr_atan(x[i]); break;
      case 4: cr_sincos(x[i], &s, &c); out[i] = s; break;
      case 5: cr_sincos(x[i], &s, &c); out[i] = c; break;
      case 6: out[i] = gl_exp(x[i]); break;
      case 7: out[i] = gl_log1p(x[i]); break;
      case 8: out[i] = gl_cpow_real(x[i], 0.25); break;
      case 9: out[i] = gl_cpow_real(x[i], -0.5); break;
      case 10: out[i] = most_psi(x[i]); break;
      default: out[i] = most_phi(x[i]); break;
    }
  }
}

void crm_wind(const double* spd, const double* dir, int64_t n, double* u, double* v) {
  for (int64_t i = 0; i < n; ++i) wind_uv(spd[i], dir[i], &u[i], &v[i]);
}

// vertical_profiles on the host: returns the number of levels (or the
// prof_setup error code < 0) and fills out (6 x cap: z, u, v, Kx, Ky, Kz)
// when the levels fit.
int crm_profiles(int n, double zm, double um, double vm, int has_ustar, double ustar, int has_z0,
                 double z0, double mol, double prsc, int closure, int has_dh, double dh, int has_st,
                 double st, double tke, double* out, int cap) {
  ProfSetup s;
  const int rc = prof_setup(n, zm, um, vm, has_ustar, ustar, has_z0, z0, mol, prsc, closure, has_dh, dh,
                            has_st, st, tke, &s);
  if (rc) return rc;
  if (s.nlev <= cap)
    for (int i = 0; i < s.nlev; ++i) {
      double o[6];
      prof_point(s, i, o);
      for (int c = 0; c < 6; ++c) out[(int64_t)c * cap + i] = o[c];
    }
  return s.nlev;
}
}

extern "C" {
// The scalar prologue alone: {z0, ustar, h, aa, bb, dzeta, nlev}.
int crm_setup(int n, double zm, double um, double vm, int has_ustar, double ustar, int has_z0, double z0,
              double mol, double prsc, int closure, int has_dh, double dh, int has_st, double st, double tke,
              double* out7) {
  ProfSetup s;
  const int rc = prof_setup(n, zm, um, vm, has_ustar, ustar, has_z0, z0, mol, prsc, closure, has_dh, dh,
                            has_st, st, tke, &s);
  if (rc) return rc;
  out7[0] = s.z0; out7[1] = s.ustar; out7[2] = s.h; out7[3] = s.aa; out7[4] = s.bb; out7[5] = s.dzeta;
  out7[6] = s.nlev;
  return 0;
}
}

extern "C" {
// vertical_profiles from a caller's prologue (z0, ustar, aa, bb, zeta_max);
// returns the number of levels or the prof_setup_given error code < 0.
int crm_profiles_given(int n, double zm, double um, double vm, double ustar, double z0, double mol, double prsc,
                       int closure, int has_st, double st, double tke, double aa, double bb, double zetamx,
                       double* out, int cap) {
  ProfSetup s;
  const int rc = prof_setup_given(n, zm, um, vm, ustar, z0, mol, prsc, closure, has_st, st, tke, aa, bb, zetamx, &s);
  if (rc) return rc;
  if (s.nlev <= cap)
    for (int i = 0; i < s.nlev; ++i) {
      double o[6];
      prof_point(s, i, o);
      for (int c = 0; c < 6; ++c) out[(int64_t)c * cap + i] = o[c];
    }
  return s.nlev;
}
}
