// Device helpers shared by the solver kernels and the surrogate kernels:
// uniform (scalar-load) views of plan metadata, lane decoding, the vertical
// march step, the radiation condition, the shooting coefficient and the
// packed spectral output of one mode.
#pragma once
#include "bldfm_internal.h"

// Uniform per-tile data read through the constant address space (scalar
// loads into SGPRs): tile records, instance descriptors, FFT plans.
typedef const __attribute__((address_space(4))) InstDesc cdesc_t;
typedef const __attribute__((address_space(4))) FftPlan cplan_t;
typedef const __attribute__((address_space(4))) TileRec ctile_t;

__device__ __forceinline__ TileRec tile_rec(const TileRec* recs, int64_t t) {
  const ctile_t* r = (const ctile_t*)(recs + t);
  TileRec o;
  o.inst = r->inst;
  o.local = r->local;
  return o;
}

// ---------------------------------------------------------------- shooting
struct Lane {
  int kind, fx, fy, nidx;
  double lx, ly;
};

__device__ __forceinline__ bool lane_setup(const cdesc_t& D, int lane, Lane& L) {
  if (lane >= D.lanes) return false;
  L.fx = 0; L.fy = 0; L.nidx = -1;
  L.kind = decode_lane(lane, D.nlx, D.nly, &L.fx, &L.fy, &L.nidx);
  if (L.kind == LANE_IDLE) return false;
  // lx = ((2 pi)/dx)/nxe * (f * 1/(nl*(1/nl)))   (numpy fftfreq, solver.py:148-153)
  L.lx = xmul(D.sx, xmul((double)L.fx, D.vx));
  L.ly = xmul(D.sy, xmul((double)L.fy, D.vy));
  return true;
}

// Packed output of one mode (after complex64 rounding and phase shift).
__device__ __forceinline__ void emit_mode(const cdesc_t& D, const Lane& L, double2* __restrict__ ws,
                                          int s, double2 Xp, double2 Xq, double2 sh) {
  if (D.single) { Xp = round_c64(Xp); Xq = round_c64(Xq); }
  if (D.has_shift) { Xp = ncmul(Xp, sh); Xq = ncmul(Xq, sh); }
  if (L.kind == LANE_HALF) {
    const int cx = compact_of_freq(L.fx, D.nxb), cy = compact_of_freq(L.fy, D.nyb);
    const int mx = compact_mirror(cx, D.nxb), my = compact_mirror(cy, D.nyb);
    double2* Z = ws + D.off_z + (int64_t)s * D.nxb * D.nyb;
    // Z(k) = Xp + i Xq ; Z(-k) = conj(Xp) + i conj(Xq)
    Z[(int64_t)cx * D.nyb + cy] = cmk(Xp.x - Xq.y, Xp.y + Xq.x);
    Z[(int64_t)mx * D.nyb + my] = cmk(Xp.x + Xq.y, Xq.x - Xp.y);
  } else {  // LANE_NYQ: keep X for the Hermitian fix-up
    double2* N = ws + D.off_nyq + ((int64_t)s * (D.nlx + D.nly) + L.nidx) * 2;
    N[0] = Xp;
    N[1] = Xq;
  }
}

__device__ __forceinline__ void emit_origin(const cdesc_t& D, double2* __restrict__ ws, int s,
                                            double p00, double q00) {
  if (D.single) { p00 = (double)(float)p00; q00 = (double)(float)q00; }
  ws[D.off_z + (int64_t)s * D.nxb * D.nyb] = cmk(p00, q00);
}

// e^{i(lx*shx + ly*shy)} (src/bldfm/solver.py:254-262)
__device__ __forceinline__ double2 phase_shift(const cdesc_t& D, const Lane& L) {
  double s = 0.0, c = 1.0;
  if (D.has_shift) {
    const double th = xmul(L.lx, D.shx) + xmul(L.ly, D.shy);
    sincos(th, &s, &c);
  }
  return cmk(c, s);
}

__device__ __forceinline__ double2 ivp2_init(const cdesc_t& D, const Lane& L,
                                             const double2* __restrict__ ws) {
  if (D.footprint) return cmk(D.q00, 0.0);
  const int jx = trunc_of_freq(L.fx, D.nlx), jy = trunc_of_freq(L.fy, D.nly);
  return ws[D.off_q0 + (int64_t)jx * D.nly + jy];
}

// One step of both initial-value problems,
// [p; q] <- [[a, b], [c, a]] [p; q]  (src/bldfm/solver.py:357-368).
// EXACT: every product and sum rounded separately, in the reference's order
//   T = (-(Kx*lx2 + Ky*ly2), -(u*lx + v*ly))
//   a = 1 - (sa*T)*h2,  b = b0 - (sb*T)*h3,  c = T*h - (sc*T^2)*h3
// otherwise: the same algebra with fused multiply-adds.
// The level table and slot map are uniform across a workgroup and read-only
// during the march: reading them through the constant address space lets
// the compiler fetch them with scalar loads into SGPRs.
typedef const __attribute__((address_space(4))) double ctab_t;
typedef const __attribute__((address_space(4))) int32_t cint_t;

// One level-table row in registers (SGPRs: the row is uniform).
struct Row {
  double kx, ky, u, v, sa, h2, b0, sb, h3, h, sc, A, B1, C;
};

__device__ __forceinline__ Row load_row(ctab_t* __restrict__ T) {
  Row r;
  r.kx = T[0]; r.ky = T[1]; r.u = T[2]; r.v = T[3];
  r.sa = T[4]; r.h2 = T[5]; r.b0 = T[6]; r.sb = T[7]; r.h3 = T[8]; r.h = T[9]; r.sc = T[10];
  r.A = T[12]; r.B1 = T[13]; r.C = T[14];
  return r;
}

template <bool EXACT>
__device__ __forceinline__ void march_step(const Row& R, double lx, double ly, double lx2, double ly2,
                                           double2& p1, double2& q1, double2& p2, double2& q2);

template <>
__device__ __forceinline__ void march_step<true>(const Row& R, double lx, double ly, double lx2,
                                                 double ly2, double2& p1, double2& q1, double2& p2,
                                                 double2& q2) {
#pragma clang fp contract(off)
  const double tr = -(R.kx * lx2 + R.ky * ly2);
  const double ti = -(R.u * lx + R.v * ly);
  const double ar = 1.0 - (R.sa * tr) * R.h2, ai = -((R.sa * ti) * R.h2);
  const double br = R.b0 - (R.sb * tr) * R.h3, bi = -((R.sb * ti) * R.h3);
  const double t2r = tr * tr - ti * ti;
  const double tri = tr * ti;
  const double t2i = tri + tri;
  const double cr = tr * R.h - (R.sc * t2r) * R.h3, ci = ti * R.h - (R.sc * t2i) * R.h3;
  const double2 a = cmk(ar, ai), b = cmk(br, bi), c = cmk(cr, ci);
  const double2 np1 = xcmuladd(a, p1, b, q1);
  const double2 nq1 = xcmuladd(c, p1, a, q1);
  const double2 np2 = xcmuladd(a, p2, b, q2);
  const double2 nq2 = xcmuladd(c, p2, a, q2);
  p1 = np1; q1 = nq1; p2 = np2; q2 = nq2;
}

template <>
__device__ __forceinline__ void march_step<false>(const Row& R, double lx, double ly, double lx2,
                                                  double ly2, double2& p1, double2& q1, double2& p2,
                                                  double2& q2) {
  const double tr = -(R.kx * lx2 + R.ky * ly2);
  const double ti = -(R.u * lx + R.v * ly);
  const double2 a = cmk(1.0 - R.A * tr, -(R.A * ti));
  const double2 b = cmk(R.b0 - R.B1 * tr, -(R.B1 * ti));
  const double t2r = tr * tr - ti * ti, t2i = 2.0 * tr * ti;
  const double2 c = cmk(R.h * tr - R.C * t2r, R.h * ti - R.C * t2i);
  const double2 np1 = cadd(cmul(a, p1), cmul(b, q1));
  const double2 nq1 = cadd(cmul(c, p1), cmul(a, q1));
  const double2 np2 = cadd(cmul(a, p2), cmul(b, q2));
  const double2 nq2 = cadd(cmul(c, p2), cmul(a, q2));
  p1 = np1; q1 = nq1; p2 = np2; q2 = nq2;
}

// Radiation condition q = Kz lambda p at the top and the shooting coefficient
// alpha = -(q2 - Kz lambda p2)/(q1 - Kz lambda p1)
// (src/bldfm/solver.py:164-174, 228-230), rounded as numpy rounds it.
__device__ __forceinline__ double2 top_lambda(const double* __restrict__ tt, double lx, double ly,
                                              double lx2, double ly2) {
#pragma clang fp contract(off)
  return csqrt_exact(cmk(tt[0] * lx2 + tt[1] * ly2, tt[2] * lx + tt[3] * ly));
}

__device__ __forceinline__ double2 shoot_alpha(const double* __restrict__ tt, double2 lam, double2 p1,
                                               double2 q1, double2 p2, double2 q2) {
  const double2 kl = xcscale(tt[4], lam);
  const double2 num = xcsub(q2, ncmul(kl, p2));
  const double2 den = xcsub(q1, ncmul(kl, p1));
  return cdiv(cmk(-num.x, -num.y), den);
}
