// HIP kernels of the BLDFM steady-state solver for gfx950 (MI355X).
//
// Pipeline per solve instance (all FP64):
//   k_level_table  per-step constants of the vertical march (profiles from
//                  caller arrays, or evaluated here from met scalars)
//   k_fwd_rows/cols  dispersion mode only: pruned forward 2-D DFT of the
//                  halo-padded surface flux onto the retained modes
//   k_shoot_*      one lane per Fourier mode: two initial-value problems,
//                  radiation condition, shooting, (0,0) trapezoid, phase
//                  shift, Hermitian packing of p^ and q^ into Z = Hp + i Hq
//   k_nyq_fix      Hermitian part on the unpaired Nyquist row/column
//   k_inv_cols/rows  pruned inverse 2-D DFT of Z: y-pass over the band
//                  columns, x-pass over the cropped rows; Re -> conc, Im -> flx
//
// Work of a batch of instances is laid out as tiles with per-kernel prefix
// sums; every kernel is a grid-stride loop over its tiles.
#include "bldfm_internal.h"
#include "bldfm_fft.h"
#include "bldfm_march.h"

#define FFT_SMALL_E 8     // values per thread per stage, 256-thread FFT kernels (n <= 2048)

// Workgroups are dispatched round-robin over the 8 XCDs (each with its own
// L2).  This bijection of [0, n) gives every XCD a contiguous range of tiles,
// so neighbouring tiles that share cache lines run behind the same L2.
__device__ __forceinline__ int64_t xcd_swizzle(int64_t b, int64_t n) {
  const int64_t q = n >> 3, r = n & 7, x = b & 7, i = b >> 3;
  return x * q + (x < r ? x : r) + i;
}

// ---------------------------------------------------------------- level table
// Entry i < nz-1 (step from z_i to z_{i+1}, coefficients of the LOWER level,
// src/bldfm/solver.py:357-364), with kappa = 1/Kz_i, h = z_{i+1} - z_i:
//   [0..3]  Kx, Ky, u, v
//   [4..11] the reference's scalar sub-expressions, rounded as numba rounds
//           them: sa = 0.5*kappa, h2 = h*h, b0 = (-kappa)*h,
//           sb = (1/6)*(kappa*kappa), h3 = h*(h*h), h, sc = (1/6)*kappa,
//           W = 0.5/Kz_i + 0.5/Kz_{i+1}  (trapezoid weight of the (0,0) mode)
//   [12..14] fused-path coefficients A = sa*h2, B1 = sb*h3, C = sc*h3
// so that T = -(Kx lx^2 + Ky ly^2) - i(u lx + v ly), a = d = 1 - sa T h2,
// b = b0 - sb T h3, c = T h - sc T^2 h3.
// Entry nz-1 (top, src/bldfm/solver.py:164-174):
//   {Kx/Kz, Ky/Kz, u/Kz, v/Kz, Kz, 1/Kz} (products with 1/Kz as the reference forms them)
// One level-table row from the profile values (z,u,v,Kx,Ky,Kz) of level i
// (cur) and level i+1 (nxt, unused for the top row).
__device__ __forceinline__ void fill_level_row(const double* cur, const double* nxt, bool top,
                                               double* __restrict__ T) {
#pragma clang fp contract(off)
  const double kz = cur[5];
  const double kinv = 1.0 / kz;
  if (!top) {
    const double h = nxt[0] - cur[0];
    const double h2 = h * h;
    const double h3 = h * h2;
    const double sixth = 1.0 / 6.0;
    const double sa = 0.5 * kinv, sb = sixth * (kinv * kinv), sc = sixth * kinv;
    T[0] = cur[3];
    T[1] = cur[4];
    T[2] = cur[1];
    T[3] = cur[2];
    T[4] = sa;
    T[5] = h2;
    T[6] = -kinv * h;
    T[7] = sb;
    T[8] = h3;
    T[9] = h;
    T[10] = sc;
    T[11] = 0.5 / kz + 0.5 / nxt[5];
    T[12] = sa * h2;
    T[13] = sb * h3;
    T[14] = sc * h3;
    T[15] = 0.0;
  } else {
    T[0] = cur[3] * kinv;
    T[1] = cur[4] * kinv;
    T[2] = cur[1] * kinv;
    T[3] = cur[2] * kinv;
    T[4] = kz;
    T[5] = kinv;
    for (int c = 6; c < BLDFM_LEVEL_STRIDE; ++c) T[c] = 0.0;
  }
}

__global__ __launch_bounds__(256) void k_level_table(PlanDev P, ExecBufs B) {
#pragma clang fp contract(off)
  // levels evaluated from met in this tile: 256 of them and the one above
  __shared__ double lev[6][257];
  for (int64_t t = blockIdx.x; t < P.tot_tab; t += gridDim.x) {
    const TileRec tr = tile_rec(P.t_tab, t);
    const int inst = tr.inst;
    const cdesc_t& D = *(const cdesc_t*)(P.desc + inst);
    const int nz = D.nz;
    const int tid = (int)threadIdx.x, T = (int)blockDim.x;
    const int i = tr.local * T + tid;
    double* prof = B.prof + D.off_prof;
    double cur[6], nxt[6];
    if (P.met_flag[inst]) {  // uniform over the workgroup (one instance per tile)
      // every level once (the correctly rounded profile functions are the
      // costly part): each thread its level, the last thread also the next
      const ProfSetup& S = P.met[inst];
      if (t != blockIdx.x) __syncthreads();  // the previous tile's reads of lev are done
      if (i < nz) {
        prof_point(S, i, cur);
        for (int c = 0; c < 6; ++c) {
          prof[(int64_t)c * nz + i] = cur[c];
          lev[c][tid] = cur[c];
        }
      }
      if (tid == T - 1 && i + 1 < nz) {
        prof_point(S, i + 1, nxt);
        for (int c = 0; c < 6; ++c) lev[c][T] = nxt[c];
      }
      __syncthreads();
      if (i + 1 < nz)
        for (int c = 0; c < 6; ++c) nxt[c] = lev[c][tid + 1];
    } else {
      if (i < nz)
        for (int c = 0; c < 6; ++c) cur[c] = prof[(int64_t)c * nz + i];
      if (i + 1 < nz)
        for (int c = 0; c < 6; ++c) nxt[c] = prof[(int64_t)c * nz + i + 1];
    }
    if (i < nz) fill_level_row(cur, nxt, i == nz - 1, B.tab + D.off_tab + (int64_t)i * BLDFM_LEVEL_STRIDE);
  }
}

// Degenerate mode (0,0): flux conserved, p by the trapezoid rule in z
// (src/bldfm/solver.py:190-191, 239-251): p00 <- p00 - (q00*dz)*W.
__device__ void origin_numeric(const cdesc_t& D, const double* __restrict__ tab,
                               const int32_t* __restrict__ slot, double2* __restrict__ ws) {
#pragma clang fp contract(off)
  const double q00 = D.footprint ? D.q00 : ws[D.off_q0].x;
  double p00 = D.p000;
  for (int i = 0; i < D.nz; ++i) {
    const int sl = slot[i];
    if (sl >= 0) emit_origin(D, ws, sl, p00, q00);
    if (i < D.nz - 1) {
      const double* T = tab + (int64_t)i * BLDFM_LEVEL_STRIDE;
      p00 = p00 - (q00 * T[9]) * T[11];
    }
  }
  for (int s = D.nfilled; s < D.nslot; ++s) emit_origin(D, ws, s, s == 0 ? D.p000 : 0.0, q00);
}

// Numerical branch.  MAXS register snapshots (RECOMP == false), or a second
// march from the bottom after the shooting coefficient is known
// (RECOMP == true) for many output levels.  Both give the same values.
template <int MAXS, bool RECOMP, bool EXACT, int BLOCK>
__global__ __launch_bounds__(BLOCK) void k_shoot(PlanDev P, ExecBufs B, int variant) {
  for (int64_t t = blockIdx.x; t < P.tot_shoot; t += gridDim.x) {
    const TileRec tr = tile_rec(P.t_shoot, t);
    const cdesc_t& D = *(const cdesc_t*)(P.desc + tr.inst);
    if (D.variant != variant || D.exact != (int)EXACT) continue;
    if (B.mode && B.mode[tr.inst] == 0) continue;  // accepted surrogate instance
    Lane L;
    const int lane = tr.local * BLOCK + (int)threadIdx.x;
    if (!lane_setup(D, lane, L)) continue;
    const double* __restrict__ tab = B.tab + D.off_tab;
    ctab_t* __restrict__ ctab = (ctab_t*)tab;
    cint_t* __restrict__ slot = (cint_t*)(P.ints + D.off_slot);
    const int nz = D.nz;
    if (L.kind == LANE_ORIGIN) {
      origin_numeric(D, tab, P.ints + D.off_slot, B.ws);
      continue;
    }
    const double lx = L.lx, ly = L.ly, lx2 = xmul(lx, lx), ly2 = xmul(ly, ly);
    const double2 q0 = ivp2_init(D, L, B.ws);
    double2 p1 = cmk(1.0, 0.0), q1 = cmk(0.0, 0.0), p2 = cmk(0.0, 0.0), q2 = q0;
    double2 sp1[MAXS], sq1[MAXS], sp2[MAXS], sq2[MAXS];
#pragma unroll
    for (int s = 0; s < MAXS; ++s) sp1[s] = sq1[s] = sp2[s] = sq2[s] = cmk(0.0, 0.0);
    // the next row is fetched while the current step computes; unrolled by
    // two so the prefetched row and the (p, q) state alternate between two
    // register sets instead of being copied every step.  The snapshot levels
    // come from the descriptor (uniform compares), so no step waits on a
    // load issued in the same step.
    int snap[MAXS];
#pragma unroll
    for (int s = 0; s < MAXS; ++s) snap[s] = D.snap_level[s];
    Row nxt = load_row(ctab);
#pragma unroll 2
    for (int i = 0; i < nz - 1; ++i) {
      if (!RECOMP) {
#pragma unroll
        for (int s = 0; s < MAXS; ++s)
          if (i == snap[s]) { sp1[s] = p1; sq1[s] = q1; sp2[s] = p2; sq2[s] = q2; }
      }
      const Row cur = nxt;
      if (i + 1 < nz - 1) nxt = load_row(ctab + (int64_t)(i + 1) * BLDFM_LEVEL_STRIDE);
      march_step<EXACT>(cur, lx, ly, lx2, ly2, p1, q1, p2, q2);
    }
    if (!RECOMP) {
#pragma unroll
      for (int s = 0; s < MAXS; ++s)
        if (nz - 1 == snap[s]) { sp1[s] = p1; sq1[s] = q1; sp2[s] = p2; sq2[s] = q2; }
    }
    const double* tt = tab + (int64_t)(nz - 1) * BLDFM_LEVEL_STRIDE;
    const double2 alpha = shoot_alpha(tt, top_lambda(tt, lx, ly, lx2, ly2), p1, q1, p2, q2);
    const double2 sh = phase_shift(D, L);
    if (!RECOMP) {
#pragma unroll
      for (int s = 0; s < MAXS; ++s)
        if (s < D.nfilled)
          emit_mode(D, L, B.ws, s, xcadd(ncmul(alpha, sp1[s]), sp2[s]),
                    xcadd(ncmul(alpha, sq1[s]), sq2[s]), sh);
    } else {
      p1 = cmk(1.0, 0.0); q1 = cmk(0.0, 0.0); p2 = cmk(0.0, 0.0); q2 = q0;
      Row nxt2 = load_row(ctab);
      for (int i = 0; i <= D.max_level; ++i) {
        const int sl = slot[i];
        if (sl >= 0)
          emit_mode(D, L, B.ws, sl, xcadd(ncmul(alpha, p1), p2), xcadd(ncmul(alpha, q1), q2), sh);
        if (i < D.max_level) {
          const Row cur = nxt2;
          if (i + 1 < D.max_level) nxt2 = load_row(ctab + (int64_t)(i + 1) * BLDFM_LEVEL_STRIDE);
          march_step<EXACT>(cur, lx, ly, lx2, ly2, p1, q1, p2, q2);
        }
      }
    }
    for (int s = D.nfilled; s < D.nslot; ++s) emit_mode(D, L, B.ws, s, cmk(0.0, 0.0), cmk(0.0, 0.0), sh);
  }
}

// Analytic branch for constant profiles (src/bldfm/solver.py:193-202):
// q^ = q^0 e^{-lambda h}, p^ = (q^ / Kz) / lambda, p^00 = p000 - (q^00 / Kz) h.
template <int BLOCK>
__global__ __launch_bounds__(BLOCK) void k_shoot_analytic(PlanDev P, ExecBufs B, int variant) {
#pragma clang fp contract(off)
  for (int64_t t = blockIdx.x; t < P.tot_shoot; t += gridDim.x) {
    const TileRec tr = tile_rec(P.t_shoot, t);
    const cdesc_t& D = *(const cdesc_t*)(P.desc + tr.inst);
    if (D.variant != variant) continue;
    if (B.mode && B.mode[tr.inst] == 0) continue;
    Lane L;
    const int lane = tr.local * BLOCK + (int)threadIdx.x;
    if (!lane_setup(D, lane, L)) continue;
    const int nz = D.nz;
    const double* tt = B.tab + D.off_tab + (int64_t)(nz - 1) * BLDFM_LEVEL_STRIDE;
    const double kinv = tt[5];
    const double* z = B.prof + D.off_prof;
    const int32_t* lvidx = P.ints + D.off_slot + nz;
    if (L.kind == LANE_ORIGIN) {
      const double q00 = D.footprint ? D.q00 : B.ws[D.off_q0].x;
      for (int s = 0; s < D.nslot; ++s) {
        const double h = z[lvidx[s]] - z[0];
        emit_origin(D, B.ws, s, D.p000 - (q00 * kinv) * h, q00);
      }
      continue;
    }
    const double lx = L.lx, ly = L.ly;
    const double2 lam = top_lambda(tt, lx, ly, lx * lx, ly * ly);
    const double2 q0 = ivp2_init(D, L, B.ws);
    const double2 sh = phase_shift(D, L);
    for (int s = 0; s < D.nslot; ++s) {
      const double h = z[lvidx[s]] - z[0];
      double2 Xq = ncmul(q0, cexp_(cmk(-lam.x * h, -lam.y * h)));
      if (D.single) Xq = round_c64(Xq);
      const double2 Xp = cdiv(xcscale(kinv, Xq), lam);
      emit_mode(D, L, B.ws, s, Xp, Xq, sh);
    }
  }
}

// Hermitian part on the Nyquist row/column:
// Z(c) = (X(c) + conj X(-c))/2 packed as Hp + i Hq.
__global__ __launch_bounds__(256) void k_nyq_fix(PlanDev P, ExecBufs B) {
  for (int64_t t = blockIdx.x; t < P.tot_nyq; t += gridDim.x) {
    const TileRec tr = tile_rec(P.t_nyq, t);
    const cdesc_t& D = *(const cdesc_t*)(P.desc + tr.inst);
    const int cnt = nyq_fix_count(D.nlx, D.nly, D.nxb, D.nyb);
    const int bps = (cnt + 255) / 256;
    const int local = tr.local;
    const int s = local / bps;
    const int p = (local % bps) * 256 + threadIdx.x;
    if (p >= cnt) continue;
    int cx, cy;
    nyq_fix_decode(p, D.nlx, D.nly, D.nxb, D.nyb, &cx, &cy);
    const int mx = compact_mirror(cx, D.nxb), my = compact_mirror(cy, D.nyb);
    const int n1 = nyq_index_at(cx, cy, D.nlx, D.nly, D.nxb, D.nyb);
    const int n2 = nyq_index_at(mx, my, D.nlx, D.nly, D.nxb, D.nyb);
    const double2* N = B.ws + D.off_nyq + (int64_t)s * (D.nlx + D.nly) * 2;
    double2 A = cmk(0.0, 0.0), Q = cmk(0.0, 0.0);
    if (n1 >= 0) { A = N[2 * n1]; Q = N[2 * n1 + 1]; }
    if (n2 >= 0) { A = cadd(A, cconj(N[2 * n2])); Q = cadd(Q, cconj(N[2 * n2 + 1])); }
    B.ws[D.off_z + ((int64_t)s * D.nxb + cx) * D.nyb + cy] = cmk(0.5 * (A.x - Q.y), 0.5 * (A.y + Q.x));
  }
}

// ---------------------------------------------------------------- point solver
// Level tables of nmet met samples for the point solver, `rows` rows per
// sample (rows >= the sample's level count).
__global__ __launch_bounds__(256) void k_tables_met(const ProfSetup* __restrict__ S, int nmet, int rows,
                                                    double* __restrict__ tab) {
  const int64_t g = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (g >= (int64_t)nmet * rows) return;
  const int m = (int)(g / rows), i = (int)(g % rows);
  const ProfSetup& s = S[m];
  const int nz = s.nlev;
  if (i >= nz) return;
  double cur[6], nxt[6];
  prof_point(s, i, cur);
  if (i + 1 < nz) prof_point(s, i + 1, nxt);
  fill_level_row(cur, nxt, i == nz - 1, tab + g * BLDFM_LEVEL_STRIDE);
}

// Flux transfer G = q^(z_lvl) / q^(0) of single modes: the shooting solve
// of the footprint mode (both IVPs, radiation condition, shooting
// coefficient) for a unit surface flux.  Blocks of 256 points share one met
// sample (its table rows are uniform, read through scalar loads).  Invalid
// block metadata yields NaN instead of an out-of-range read.
template <bool EXACT>
__global__ __launch_bounds__(256) void k_transfer(const double* __restrict__ tab, int rows,
                                                  const int32_t* __restrict__ nzs,
                                                  const int32_t* __restrict__ lvls, int nmet,
                                                  const int32_t* __restrict__ blk_met, int nblk,
                                                  const double2* __restrict__ kxy, double2* __restrict__ g) {
  for (int b = blockIdx.x; b < nblk; b += gridDim.x) {
    const int64_t p = (int64_t)b * 256 + threadIdx.x;
    const int m = blk_met[b];
    const int nz = (m >= 0 && m < nmet) ? nzs[m] : 0;
    const int lvl = (m >= 0 && m < nmet) ? lvls[m] : -1;
    if (nz < 2 || nz > rows || lvl < 0 || lvl >= nz) {
      g[p] = cmk(__builtin_nan(""), __builtin_nan(""));
      continue;
    }
    const double* T = tab + (int64_t)m * rows * BLDFM_LEVEL_STRIDE;
    ctab_t* ct = (ctab_t*)T;
    const double2 k = kxy[p];
    const double lx = k.x, ly = k.y, lx2 = xmul(lx, lx), ly2 = xmul(ly, ly);
    double2 p1 = cmk(1.0, 0.0), q1 = cmk(0.0, 0.0), p2 = cmk(0.0, 0.0), q2 = cmk(1.0, 0.0);
    double2 sq1 = q1, sq2 = q2;
    Row nxt = load_row(ct);
    for (int i = 0; i < nz - 1; ++i) {
      if (i == lvl) { sq1 = q1; sq2 = q2; }
      const Row cur = nxt;
      if (i + 1 < nz - 1) nxt = load_row(ct + (int64_t)(i + 1) * BLDFM_LEVEL_STRIDE);
      march_step<EXACT>(cur, lx, ly, lx2, ly2, p1, q1, p2, q2);
    }
    if (lvl == nz - 1) { sq1 = q1; sq2 = q2; }
    const double* tt = T + (int64_t)(nz - 1) * BLDFM_LEVEL_STRIDE;
    const double2 alpha = shoot_alpha(tt, top_lambda(tt, lx, ly, lx2, ly2), p1, q1, p2, q2);
    g[p] = (lx == 0.0 && ly == 0.0) ? cmk(1.0, 0.0) : xcadd(ncmul(alpha, sq1), sq2);
  }
}

// ---------------------------------------------------------------- surrogate gate
// Per-instance gate state before the surrogate runs: sums cleared, accepted
// (mode 0) if flagged and inside the training hull, else solved (mode 1).
__global__ __launch_bounds__(256) void k_sur_clear(PlanDev P, ExecBufs B) {
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i >= P.ninst) return;
  B.gate[i] = cmk(0.0, 0.0);
  B.mode[i] = P.sur_ok[i] ? 0 : 1;
}

// Spot check: every sur_stride-th lane of an accepted instance is solved in
// FP64 and compared with the surrogate's packed output X_q = -i Z(k).
template <bool EXACT>
__global__ __launch_bounds__(256) void k_sur_spot(PlanDev P, ExecBufs B) {
  __shared__ double sd[4], ss[4];
  for (int64_t t = blockIdx.x; t < P.tot_spot; t += gridDim.x) {
    const TileRec tr = tile_rec(P.t_spot, t);
    const int inst = tr.inst;
    const cdesc_t& D = *(const cdesc_t*)(P.desc + inst);
    if (B.mode[inst] != 0 || D.exact != (int)EXACT) continue;  // uniform over the workgroup
    const int off = (int)((inst * 61LL) % P.sur_stride);
    const int lane = off + (tr.local * 256 + (int)threadIdx.x) * P.sur_stride;
    double d2 = 0.0, s2 = 0.0;
    Lane L;
    if (lane_setup(D, lane, L) && L.kind == LANE_HALF) {
      const double* __restrict__ tab = B.tab + D.off_tab;
      ctab_t* __restrict__ ctab = (ctab_t*)tab;
      const int nz = D.nz, lvl = D.max_level;
      const double lx = L.lx, ly = L.ly, lx2 = xmul(lx, lx), ly2 = xmul(ly, ly);
      double2 p1 = cmk(1.0, 0.0), q1 = cmk(0.0, 0.0), p2 = cmk(0.0, 0.0), q2 = cmk(D.q00, 0.0);
      double2 sq1 = q1, sq2 = q2;
      Row nxt = load_row(ctab);
      for (int i = 0; i < nz - 1; ++i) {
        if (i == lvl) { sq1 = q1; sq2 = q2; }
        const Row cur = nxt;
        if (i + 1 < nz - 1) nxt = load_row(ctab + (int64_t)(i + 1) * BLDFM_LEVEL_STRIDE);
        march_step<EXACT>(cur, lx, ly, lx2, ly2, p1, q1, p2, q2);
      }
      if (lvl == nz - 1) { sq1 = q1; sq2 = q2; }
      const double* tt = tab + (int64_t)(nz - 1) * BLDFM_LEVEL_STRIDE;
      const double2 alpha = shoot_alpha(tt, top_lambda(tt, lx, ly, lx2, ly2), p1, q1, p2, q2);
      double2 xq = xcadd(ncmul(alpha, sq1), sq2);
      if (D.has_shift) xq = ncmul(xq, phase_shift(D, L));
      const int cx = compact_of_freq(L.fx, D.nxb), cy = compact_of_freq(L.fy, D.nyb);
      const double2 z = B.ws[D.off_z + (int64_t)cx * D.nyb + cy];
      const double er = z.y - xq.x, ei = -z.x - xq.y;
      d2 = er * er + ei * ei;
      s2 = xq.x * xq.x + xq.y * xq.y;
    }
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1) {
      d2 += __shfl_xor(d2, o);
      s2 += __shfl_xor(s2, o);
    }
    if ((threadIdx.x & 63) == 0) { sd[threadIdx.x >> 6] = d2; ss[threadIdx.x >> 6] = s2; }
    __syncthreads();
    if (threadIdx.x == 0) {
      atomicAdd(&B.gate[inst].x, sd[0] + sd[1] + sd[2] + sd[3]);
      atomicAdd(&B.gate[inst].y, ss[0] + ss[1] + ss[2] + ss[3]);
    }
    __syncthreads();
  }
}

// Gate decision: an accepted instance whose spot-check relative L2 error
// exceeds sur_tol (or is not finite) falls back to the FP64 solver.
__global__ __launch_bounds__(256) void k_sur_gate(PlanDev P, ExecBufs B) {
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i >= P.ninst || B.mode[i] != 0) return;
  const double2 s = B.gate[i];
  // no checked lane (or a zero reference) is no evidence: the solver serves it
  const double err = s.y > 0.0 ? sqrt(s.x / s.y) : INFINITY;
  if (!(err <= P.sur_tol)) B.mode[i] = 1;
}

// ---------------------------------------------------------------- KM01
// Kormann & Meixner (2001) analytic footprint on a regular grid
// (src/bldfm/ffm_kormann_meixner.py:110-330, Eqs. 9 and 21): one thread per
// cell, blockIdx.y = footprint.  Cell (j, i) sits at x = xmin + res/2 + i res,
// y = ymax - res/2 - j res (numpy arange order), is shifted to the receptor
// and, with a wind direction, rotated into the along/cross-wind frame.
__global__ __launch_bounds__(256) void k_km01(const Km01Dev* __restrict__ P, double* __restrict__ out) {
  const Km01Dev& D = P[blockIdx.y];
  const int64_t c = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (c >= (int64_t)D.nx * D.ny) return;
  const int j = (int)(c / D.nx), i = (int)(c % D.nx);
  double x = (D.xmin + 0.5 * D.res + i * D.res) - D.mx;
  double y = (D.ymax - 0.5 * D.res - j * D.res) - D.my;
  if (D.has_wd) {
    const double rho = sqrt(x * x + y * y);
    const double th = atan2(y, x) + D.wd_rad - 0.5 * BLDFM_PI;
    x = rho * cos(th);
    y = rho * sin(th);
  }
  double f = 0.0;
  if (D.valid && x > 0.0) {
    const double g = D.gA * y * pow(x, D.e2);
    f = D.scale * pow(x, D.e1) * exp(-D.Xi / x - 0.5 * (g * g));
  }
  out[D.off + c] = f;
}

// ---------------------------------------------------------------- FFT passes
extern __shared__ double2 smem[];

// Sources (first stage) and sinks (last stage) of the transforms.  A band of
// nb compact values (truncated order: f >= 0 first, then f < 0) occupies
// padded indices [0, lo) and [n - (nb - lo), n) with lo = (nb + 1)/2; the
// gap between the two halves is zero.
struct Band {
  int lo, hi_start, nb;
  __device__ __forceinline__ Band(int nb_, int n) : lo((nb_ + 1) / 2), hi_start(n - (nb_ - (nb_ + 1) / 2)), nb(nb_) {}
  // compact index of padded index i, or -1 in the gap
  __device__ __forceinline__ int compact(int i) const { return i < lo ? i : (i >= hi_start ? i - hi_start + lo : -1); }
};

// Compact band at element stride `stride` in global memory.
struct BandIn {
  const double2* __restrict__ src;
  int64_t stride;
  Band band;
  __device__ __forceinline__ double2 operator()(int i) const {
    const int c = band.compact(i);
    return c < 0 ? cmk(0.0, 0.0) : src[(int64_t)c * stride];
  }
};

// Band of the spectrum, scaled, to a compact array.
struct BandOut {
  double2* __restrict__ dst;
  double fct;
  Band band;
  __device__ __forceinline__ void operator()(int o, double2 v) const {
    const int c = band.compact(o);
    if (c >= 0) dst[c] = cscale(fct, v);
  }
};

// cnt values (element stride `stride`) placed at [off, off + cnt), zeros
// elsewhere; real input when `re` is set.
template <class T_>
struct PadIn {
  const T_* __restrict__ src;
  int64_t stride;
  int off, cnt;
  __device__ __forceinline__ double2 operator()(int i) const {
    const unsigned t = (unsigned)(i - off);
    if (t >= (unsigned)cnt) return cmk(0.0, 0.0);
    if constexpr (sizeof(T_) == sizeof(double)) return cmk(src[(int64_t)t * stride], 0.0);
    else return src[(int64_t)t * stride];
  }
};

// Outputs [off, off + cnt) to an array with element stride `stride`.
struct CropOut {
  double2* __restrict__ dst;
  int64_t stride;
  int off, cnt;
  __device__ __forceinline__ void operator()(int o, double2 v) const {
    const unsigned t = (unsigned)(o - off);
    if (t < (unsigned)cnt) dst[(int64_t)t * stride] = v;
  }
};

// Outputs [off, off + cnt) split into conc = cm * Re and flx = Im.
template <class T_>
struct SplitOut {
  T_* __restrict__ conc;
  T_* __restrict__ flx;
  double cm;
  int off, cnt;
  __device__ __forceinline__ void operator()(int o, double2 v) const {
    const unsigned t = (unsigned)(o - off);
    if (t < (unsigned)cnt) {
      conc[t] = (T_)(cm * v.x);
      flx[t] = (T_)v.y;
    }
  }
};

// One transform of a tile loop; `reuse`: not the workgroup's first tile.
template <int MAXE, class In, class Out>
__device__ __forceinline__ void fft_dir(bool inv, const cplan_t& F, const double2* __restrict__ tw, int tid, int T,
                                        In in, Out out, bool reuse) {
  const bldfm_fft::Lds buf = bldfm_fft::lds_view(smem, F.n);
  if (inv) bldfm_fft::lds_fft<MAXE, true>(buf, (const FftPlan*)&F, tw + F.tw_off, tid, T, in, out, reuse);
  else bldfm_fft::lds_fft<MAXE, false>(buf, (const FftPlan*)&F, tw + F.tw_off, tid, T, in, out, reuse);
}

// The four passes, each as a factory of (plan, source, sink, direction) for
// one line of one instance: the single-workgroup kernels below and the
// four-step kernels (k_long) of split lengths run the same definitions.
enum PassId { PASS_FWD_ROWS = 0, PASS_FWD_COLS = 1, PASS_INV_COLS = 2, PASS_INV_ROWS = 3 };

// Forward, rows: one surface-flux row y placed at [px, px+nx) of nxe points,
// e^{-i} DFT, scale 1/nxe (numpy norm="forward"), keep the nlx retained
// columns in truncated order.
template <class Fn>
__device__ __forceinline__ void pass_fwd_rows(const PlanDev& P, const ExecBufs& B, const cdesc_t& D, int64_t line,
                                              Fn&& fn) {
  const cplan_t& F = *(const cplan_t*)(P.plans + D.plan_x);
  const int y = (int)line;
  fn(F, PadIn<double>{B.src + D.off_src + (int64_t)y * D.nx, 1, D.px, D.nx},
     BandOut{B.ws + D.off_aw + (int64_t)y * D.nlx, 1.0 / (double)F.n, Band(D.nlx, F.n)}, false);
}

// Forward, columns: retained column jx over y placed at [py, py+ny) of nye
// points, e^{-i} DFT, scale 1/nye, keep the nly retained rows -> q^0 (x-major).
template <class Fn>
__device__ __forceinline__ void pass_fwd_cols(const PlanDev& P, const ExecBufs& B, const cdesc_t& D, int64_t line,
                                              Fn&& fn) {
  const cplan_t& F = *(const cplan_t*)(P.plans + D.plan_y);
  const int jx = (int)line;
  fn(F, PadIn<double2>{B.ws + D.off_aw + jx, D.nlx, D.py, D.ny},
     BandOut{B.ws + D.off_q0 + (int64_t)jx * D.nly, 1.0 / (double)F.n, Band(D.nly, F.n)}, false);
}

// Inverse, y-pass: band column cx of slot s (line s*nxb + cx; compact rows
// scattered to their padded indices), unnormalised DFT (e^{-i} footprint,
// e^{+i} dispersion), keep the cropped rows [py, py+ny).
template <class Fn>
__device__ __forceinline__ void pass_inv_cols(const PlanDev& P, const ExecBufs& B, const cdesc_t& D, int64_t line,
                                              Fn&& fn) {
  const cplan_t& F = *(const cplan_t*)(P.plans + D.plan_y);
  const BandIn in{B.ws + D.off_z + line * D.nyb, 1, Band(D.nyb, F.n)};
  // layout 1: column cx of slot s lands at v[s][y][cx]
  const int64_t s = line / D.nxb, cx = line % D.nxb;
  const CropOut out = P.xlayout == 1 ? CropOut{B.ws + D.off_v + s * D.ny * D.nxb + cx, D.nxb, D.py, D.ny}
                                     : CropOut{B.ws + D.off_v + line * D.ny, 1, D.py, D.ny};
  fn(F, in, out, !D.footprint);
}

// Inverse, x-pass: cropped row y of slot s (line s*ny + y) over the band
// columns, DFT, keep [px, px+nx); conc = Re, flx = Im.  Instances served by
// the surrogate carry no concentration spectrum (conc 0).
template <class Fn>
__device__ __forceinline__ void pass_inv_rows(const PlanDev& P, const ExecBufs& B, const cdesc_t& D, int inst,
                                              int64_t line, Fn&& fn) {
  const int s = (int)(line / D.ny), y = (int)(line % D.ny);
  const cplan_t& F = *(const cplan_t*)(P.plans + D.plan_x);
  const BandIn in = P.xlayout == 0
                        ? BandIn{B.ws + D.off_v + (int64_t)s * D.nxb * D.ny + y, D.ny, Band(D.nxb, F.n)}
                        : BandIn{B.ws + (P.xlayout == 1 ? D.off_v : D.off_vt) + ((int64_t)s * D.ny + y) * D.nxb,
                                 1, Band(D.nxb, F.n)};
  const int64_t o = D.off_out + line * D.nx;
  const double cm = (B.mode && B.mode[inst] == 0) ? 0.0 : 1.0;
  if (D.out_f32)
    fn(F, in, SplitOut<float>{(float*)B.conc + o, (float*)B.flx + o, cm, D.px, D.nx}, !D.footprint);
  else
    fn(F, in, SplitOut<double>{(double*)B.conc + o, (double*)B.flx + o, cm, D.px, D.nx}, !D.footprint);
}

template <class Fn>
__device__ __forceinline__ void run_pass(int pass, const PlanDev& P, const ExecBufs& B, const cdesc_t& D, int inst,
                                         int64_t line, Fn&& fn) {
  switch (pass) {
    case PASS_FWD_ROWS: pass_fwd_rows(P, B, D, line, fn); break;
    case PASS_FWD_COLS: pass_fwd_cols(P, B, D, line, fn); break;
    case PASS_INV_COLS: pass_inv_cols(P, B, D, line, fn); break;
    default: pass_inv_rows(P, B, D, inst, line, fn); break;
  }
}

// Single-workgroup transforms of one pass, one line per tile (XCD-swizzled
// so that each XCD works on a contiguous range of lines).
template <int MAXE, int MAXT, int PASS>
__global__ __launch_bounds__(MAXT) void k_pass(PlanDev P, ExecBufs B) {
  const int T = blockDim.x, tid = threadIdx.x;
  const TileRec* recs = PASS == PASS_FWD_ROWS   ? P.t_fwd_rows
                        : PASS == PASS_FWD_COLS ? P.t_fwd_cols
                        : PASS == PASS_INV_COLS ? P.t_inv_cols
                                                : P.t_inv_rows;
  const int64_t tot = PASS == PASS_FWD_ROWS   ? P.tot_fwd_rows
                      : PASS == PASS_FWD_COLS ? P.tot_fwd_cols
                      : PASS == PASS_INV_COLS ? P.tot_inv_cols
                                              : P.tot_inv_rows;
  for (int64_t b = blockIdx.x; b < tot; b += gridDim.x) {
    const int64_t t = gridDim.x >= tot ? xcd_swizzle(b, tot) : b;
    const TileRec tr = tile_rec(recs, t);
    const cdesc_t& D = *(const cdesc_t*)(P.desc + tr.inst);
    run_pass(PASS, P, B, D, tr.inst, tr.local, [&](const cplan_t& F, auto in, auto out, bool inv) {
      fft_dir<MAXE>(inv, F, P.tw, tid, T, in, out, b != blockIdx.x);
    });
  }
}

// Four-step sources and sinks: step A reads the strided sub-sequence
// in(j1 n2 + j2) and stores the twiddled length-n1 result at scr[j2 n1 + k1];
// step B reads scr[j2 n1 + k1] over j2 and writes X[k1 + n1 k2] through the
// pass's own sink.
template <class In>
struct SubIn {
  In in;
  int n2, j2;
  __device__ __forceinline__ double2 operator()(int j1) const { return in(j1 * n2 + j2); }
};
struct TwiddleOut {
  double2* __restrict__ dst;      // scr + j2 n1
  const double2* __restrict__ w;  // W_n^t, t < n
  int j2;
  bool inv;
  __device__ __forceinline__ void operator()(int k1, double2 v) const {
    double2 t = w[j2 * k1];  // j2 k1 < n: no reduction modulo n
    if (inv) t.y = -t.y;
    dst[k1] = cmul(v, t);
  }
};
struct ScrIn {
  const double2* __restrict__ src;  // scr + k1
  int n1;
  __device__ __forceinline__ double2 operator()(int j2) const { return src[(int64_t)j2 * n1]; }
};
template <class Out>
struct SubOut {
  Out out;
  int n1, k1;
  __device__ __forceinline__ void operator()(int k2, double2 v) const { out(k1 + n1 * k2, v); }
};

// One step of the four-step transforms of the split lines of pass `pass`
// (tile local = line * (n2 or n1) + sub-transform index).
template <int MAXE, int MAXT, bool STEPB>
__global__ __launch_bounds__(MAXT) void k_long(PlanDev P, ExecBufs B, int pass) {
  const int T = blockDim.x, tid = threadIdx.x;
  const int k = 2 * pass + (STEPB ? 1 : 0);
  const TileRec* recs = P.t_long[k];
  const int64_t tot = P.tot_long[k];
  for (int64_t b = blockIdx.x; b < tot; b += gridDim.x) {
    const TileRec tr = tile_rec(recs, b);
    const cdesc_t& D = *(const cdesc_t*)(P.desc + tr.inst);
    const bool x_axis = pass == PASS_FWD_ROWS || pass == PASS_INV_ROWS;
    const cplan_t& FL = *(const cplan_t*)(P.plans + (x_axis ? D.plan_x : D.plan_y));
    const int n = FL.n, n1 = FL.n1, n2 = FL.n2;
    const int64_t per = STEPB ? n1 : n2;
    const int64_t line = (int64_t)tr.local / per;
    const int sub = (int)((int64_t)tr.local % per);
    double2* scr = B.ws + D.off_long + line * n;
    const bool reuse = b != blockIdx.x;
    run_pass(pass, P, B, D, tr.inst, line, [&](const cplan_t&, auto in, auto out, bool inv) {
      if constexpr (!STEPB) {
        const cplan_t& F1 = *(const cplan_t*)(P.plans + FL.plan1);
        fft_dir<MAXE>(inv, F1, P.tw, tid, T, SubIn<decltype(in)>{in, n2, sub},
                      TwiddleOut{scr + (int64_t)sub * n1, P.tw + FL.tw_off, sub, inv}, reuse);
      } else {
        const cplan_t& F2 = *(const cplan_t*)(P.plans + FL.plan2);
        fft_dir<MAXE>(inv, F2, P.tw, tid, T, ScrIn{scr + sub, n1}, SubOut<decltype(out)>{out, n1, sub}, reuse);
      }
    });
  }
}

// x-pass layout 2: v[s][cx][y] -> vt[s][y][cx] in 32 x 32 tiles through LDS.
// Reads run along y and writes along cx, both coalesced; the padding column
// makes the transposed LDS reads (stride 33 x 16 bytes) conflict free.
__global__ __launch_bounds__(256) void k_transpose_v(PlanDev P, ExecBufs B) {
  __shared__ double2 tile[32][33];
  const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;
  for (int64_t b = blockIdx.x; b < P.tot_tr; b += gridDim.x) {
    const TileRec tr = tile_rec(P.t_tr, b);
    const cdesc_t& D = *(const cdesc_t*)(P.desc + tr.inst);
    const int tcx = (D.nxb + 31) >> 5, tcy = (D.ny + 31) >> 5;
    const int s = tr.local / (tcx * tcy), r = tr.local % (tcx * tcy);
    const int cx0 = (r / tcy) * 32, y0 = (r % tcy) * 32;
    const double2* src = B.ws + D.off_v + (int64_t)s * D.nxb * D.ny;
    double2* dst = B.ws + D.off_vt + (int64_t)s * D.ny * D.nxb;
    if (b != blockIdx.x) __syncthreads();  // the previous tile's reads of `tile` are done
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      const int cx = cx0 + ty + 8 * k, y = y0 + tx;
      if (cx < D.nxb && y < D.ny) tile[ty + 8 * k][tx] = src[(int64_t)cx * D.ny + y];
    }
    __syncthreads();
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      const int y = y0 + ty + 8 * k, cx = cx0 + tx;
      if (cx < D.nxb && y < D.ny) dst[(int64_t)y * D.nxb + cx] = tile[tx][ty + 8 * k];
    }
  }
}

// ---------------------------------------------------------------- launcher
static inline unsigned grid_for(int64_t tiles, int cap) {
  return (unsigned)(tiles < cap ? (tiles > 0 ? tiles : 1) : cap);
}

#define LAUNCH_CHECK()                       \
  do {                                       \
    hipError_t e_ = hipGetLastError();       \
    if (e_ != hipSuccess) return -(int)e_ - 1000; \
  } while (0)

int bldfm_launch_all(const PlanDev& P, const ExecBufs& B, int grid_cap, hipStream_t stream,
                     hipEvent_t* ev) {
#define MARK(k)                                            \
  if (ev) {                                                \
    if (hipEventRecord(ev[k], stream) != hipSuccess) return -1999; \
  }
  MARK(0)
  if (P.tot_tab > 0) {
    k_level_table<<<grid_for(P.tot_tab, grid_cap), 256, 0, stream>>>(P, B);
    LAUNCH_CHECK();
  }
  MARK(1)
  const unsigned ft = (unsigned)P.fft_threads;
  const size_t lds = (size_t)P.fft_lds_bytes;
  // single-workgroup lines of a pass, then its split (four-step) lines
#define FFT_LAUNCH(PASS, tiles)                                                                        \
  if ((tiles) > 0) {                                                                                   \
    if (P.fft_small) k_pass<FFT_SMALL_E, 256, PASS><<<grid_for(tiles, grid_cap), ft, lds, stream>>>(P, B); \
    else k_pass<BLDFM_FFT_MAXE, 1024, PASS><<<grid_for(tiles, grid_cap), ft, lds, stream>>>(P, B);        \
    LAUNCH_CHECK();                                                                                    \
  }                                                                                                    \
  for (int st_ = 0; st_ < 2; ++st_) {                                                                  \
    const int64_t tl_ = P.tot_long[2 * (PASS) + st_];                                                  \
    if (tl_ <= 0) continue;                                                                            \
    const unsigned g_ = grid_for(tl_, grid_cap);                                                       \
    if (P.fft_small) {                                                                                 \
      if (st_) k_long<FFT_SMALL_E, 256, true><<<g_, ft, lds, stream>>>(P, B, PASS);                    \
      else k_long<FFT_SMALL_E, 256, false><<<g_, ft, lds, stream>>>(P, B, PASS);                       \
    } else {                                                                                           \
      if (st_) k_long<BLDFM_FFT_MAXE, 1024, true><<<g_, ft, lds, stream>>>(P, B, PASS);                \
      else k_long<BLDFM_FFT_MAXE, 1024, false><<<g_, ft, lds, stream>>>(P, B, PASS);                   \
    }                                                                                                  \
    LAUNCH_CHECK();                                                                                    \
  }
  if (P.any_dispersion) {
    FFT_LAUNCH(PASS_FWD_ROWS, P.tot_fwd_rows)
    FFT_LAUNCH(PASS_FWD_COLS, P.tot_fwd_cols)
  }
  MARK(2)
  const bool sur = P.sur_net != nullptr && B.mode != nullptr && P.tot_sur > 0;
  if (sur) {
    const unsigned gi = (unsigned)((P.ninst + 255) / 256);
    k_sur_clear<<<gi, 256, 0, stream>>>(P, B);
    LAUNCH_CHECK();
    const int rc = bldfm_launch_sur_plan(P, B, grid_cap, stream);
    if (rc) return rc;
    if (P.tot_spot > 0) {
      k_sur_spot<true><<<grid_for(P.tot_spot, grid_cap), 256, 0, stream>>>(P, B);
      LAUNCH_CHECK();
      k_sur_spot<false><<<grid_for(P.tot_spot, grid_cap), 256, 0, stream>>>(P, B);
      LAUNCH_CHECK();
    }
    k_sur_gate<<<gi, 256, 0, stream>>>(P, B);
    LAUNCH_CHECK();
  }
  MARK(3)
  const unsigned gs = grid_for(P.tot_shoot, grid_cap);
#define SHOOT(V, MAXS, RECOMP, EX)                                                           \
  if (P.variant_used[V][EX]) {                                                               \
    if (P.shoot_block == 64) k_shoot<MAXS, RECOMP, EX, 64><<<gs, 64, 0, stream>>>(P, B, V);     \
    else if (P.shoot_block == 128) k_shoot<MAXS, RECOMP, EX, 128><<<gs, 128, 0, stream>>>(P, B, V); \
    else k_shoot<MAXS, RECOMP, EX, 256><<<gs, 256, 0, stream>>>(P, B, V);                     \
    LAUNCH_CHECK();                                                                          \
  }
  SHOOT(0, 1, false, true)
  SHOOT(1, 4, false, true)
  SHOOT(2, 1, true, true)
  SHOOT(0, 1, false, false)
  SHOOT(1, 4, false, false)
  SHOOT(2, 1, true, false)
#undef SHOOT
  if (P.variant_used[3][0] || P.variant_used[3][1]) {
    if (P.shoot_block == 64) k_shoot_analytic<64><<<gs, 64, 0, stream>>>(P, B, 3);
    else if (P.shoot_block == 128) k_shoot_analytic<128><<<gs, 128, 0, stream>>>(P, B, 3);
    else k_shoot_analytic<256><<<gs, 256, 0, stream>>>(P, B, 3);
    LAUNCH_CHECK();
  }
  MARK(4)
  if (P.tot_nyq > 0) {
    k_nyq_fix<<<grid_for(P.tot_nyq, grid_cap), 256, 0, stream>>>(P, B);
    LAUNCH_CHECK();
  }
  MARK(5)
  FFT_LAUNCH(PASS_INV_COLS, P.tot_inv_cols)
  MARK(6)
  if (P.xlayout == 2 && P.tot_tr > 0) {
    k_transpose_v<<<grid_for(P.tot_tr, grid_cap), 256, 0, stream>>>(P, B);
    LAUNCH_CHECK();
  }
  FFT_LAUNCH(PASS_INV_ROWS, P.tot_inv_rows)
  MARK(7)
#undef FFT_LAUNCH
#undef MARK
  return 0;
}

// ---------------------------------------------------------------- profiles
// Array part of vertical_profiles (src/bldfm/pbl_model.py:127-182):
// out = [z | u | v | Kx | Ky | Kz], nlev each.
__global__ __launch_bounds__(256) void k_profiles(ProfSetup S, double* __restrict__ out) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= S.nlev) return;
  double o[6];
  prof_point(S, i, o);
  for (int c = 0; c < 6; ++c) out[(int64_t)c * S.nlev + i] = o[c];
}

int bldfm_launch_profiles(const ProfSetup& S, double* out, hipStream_t stream) {
  k_profiles<<<(S.nlev + 255) / 256, 256, 0, stream>>>(S, out);
  LAUNCH_CHECK();
  return 0;
}

// ---------------------------------------------------------------- plain 2-D FFT
// numpy fft2/ifft2 on `batch` contiguous (n0, n1) complex arrays: x-pass
// over the rows into the workspace, then y-pass over the columns, output
// multiplied by `scale`.
struct StridedIn {
  const double2* __restrict__ src;
  int64_t stride;
  __device__ __forceinline__ double2 operator()(int i) const { return src[(int64_t)i * stride]; }
};
struct StridedOut {
  double2* __restrict__ dst;
  int64_t stride;
  double scale;
  __device__ __forceinline__ void operator()(int o, double2 v) const { dst[(int64_t)o * stride] = cscale(scale, v); }
};

template <int MAXE>
__device__ __forceinline__ void fft_plain(bool inv, const FftPlan* F, int n, const double2* __restrict__ tw, int tid,
                                          int T, StridedIn in, StridedOut out, bool reuse) {
  const bldfm_fft::Lds buf = bldfm_fft::lds_view(smem, n);
  if (inv) bldfm_fft::lds_fft<MAXE, true>(buf, F, tw, tid, T, in, out, reuse);
  else bldfm_fft::lds_fft<MAXE, false>(buf, F, tw, tid, T, in, out, reuse);
}

template <int MAXE, int MAXT>
__global__ __launch_bounds__(MAXT) void k_fft2_rows(const FftPlan* F, const double2* __restrict__ tw,
                                                  const double2* __restrict__ in, double2* __restrict__ mid,
                                                  int32_t n, int64_t nrows, bool inv) {
  const int T = blockDim.x, tid = threadIdx.x;
  for (int64_t r = blockIdx.x; r < nrows; r += gridDim.x)
    fft_plain<MAXE>(inv, F, n, tw, tid, T, StridedIn{in + r * n, 1}, StridedOut{mid + r * n, 1, 1.0},
                    r != blockIdx.x);
}

template <int MAXE, int MAXT>
__global__ __launch_bounds__(MAXT) void k_fft2_cols(const FftPlan* F, const double2* __restrict__ tw,
                                                  const double2* __restrict__ mid, double2* __restrict__ out,
                                                  int32_t n0, int32_t n1, int64_t ncols, double scale, bool inv) {
  const int T = blockDim.x, tid = threadIdx.x;
  for (int64_t c = blockIdx.x; c < ncols; c += gridDim.x) {
    const int64_t b = c / n1, col = c % n1;
    const int64_t o = b * (int64_t)n0 * n1 + col;
    fft_plain<MAXE>(inv, F, n0, tw, tid, T, StridedIn{mid + o, n1}, StridedOut{out + o, n1, scale},
                    c != blockIdx.x);
  }
}

// Split (four-step) lines of the plain 2-D FFT: `rows` selects the rows of
// the (n0, n1) arrays (line r at r*n1, stride 1), else their columns (line c
// at (c / n1)*n0*n1 + c % n1, stride n1); scr holds nlines x n values.
template <int MAXE, int MAXT, bool STEPB>
__global__ __launch_bounds__(MAXT) void k_fft2_long(FftSplit S, const double2* __restrict__ in,
                                                  double2* __restrict__ out, double2* __restrict__ scr, int rows,
                                                  int32_t n0, int32_t n1, int64_t nlines, double scale, bool inv) {
  const int T = blockDim.x, tid = threadIdx.x;
  const int n = S.F->n, s1 = S.F->n1, s2 = S.F->n2;
  const int64_t per = STEPB ? s1 : s2;
  for (int64_t b = blockIdx.x; b < nlines * per; b += gridDim.x) {
    const int64_t line = b / per;
    const int sub = (int)(b % per);
    const int64_t base = rows ? line * n1 : (line / n1) * (int64_t)n0 * n1 + line % n1;
    const int64_t stride = rows ? 1 : n1;
    double2* sc = scr + line * n;
    const bldfm_fft::Lds buf0 = bldfm_fft::lds_view(smem, STEPB ? s2 : s1);
    if constexpr (!STEPB) {
      const SubIn<StridedIn> src{StridedIn{in + base, stride}, s2, sub};
      const TwiddleOut dst{sc + (int64_t)sub * s1, S.tw, sub, inv};
      if (inv) bldfm_fft::lds_fft<MAXE, true>(buf0, S.F1, S.tw1, tid, T, src, dst, b != blockIdx.x);
      else bldfm_fft::lds_fft<MAXE, false>(buf0, S.F1, S.tw1, tid, T, src, dst, b != blockIdx.x);
    } else {
      const ScrIn src{sc + sub, s1};
      const SubOut<StridedOut> dst{StridedOut{out + base, stride, scale}, s1, sub};
      if (inv) bldfm_fft::lds_fft<MAXE, true>(buf0, S.F2, S.tw2, tid, T, src, dst, b != blockIdx.x);
      else bldfm_fft::lds_fft<MAXE, false>(buf0, S.F2, S.tw2, tid, T, src, dst, b != blockIdx.x);
    }
  }
}

int bldfm_launch_fft2(const FftSplit& A0, int32_t n0, const FftSplit& A1, int32_t n1, int32_t batch,
                      const double2* in, double2* mid, double2* scr, double2* out, int32_t sign, double scale,
                      int threads, int lds_bytes, int grid_cap, hipStream_t stream) {
  const bool inv = sign > 0;
  const int64_t nrows = (int64_t)batch * n0, ncols = (int64_t)batch * n1;
  // kernel family by the longest LDS transform (a split axis runs its two
  // factors): the 256-thread kernels hold 8 values per thread, so they take
  // lengths up to BLDFM_FFT_SMALL_N; a longer transform may still run on 256
  // threads (10 values each) in the other family
  auto lds_n = [](const FftSplit& A, int32_t n) { return A.F1 ? (A.n1 > A.n2 ? A.n1 : A.n2) : n; };
  const int nmax = lds_n(A0, n0) > lds_n(A1, n1) ? lds_n(A0, n0) : lds_n(A1, n1);
  const bool small = nmax <= BLDFM_FFT_SMALL_N;
  if ((small && threads > 256) || threads > 1024 ||
      (int64_t)threads * (small ? FFT_SMALL_E : BLDFM_FFT_MAXE) < nmax)
    return -1000 - (int)hipErrorInvalidConfiguration;
  // rows (axis 1, length n1): in -> mid
  if (A1.F1) {
    const unsigned ga = grid_for(nrows * A1.n2, grid_cap), gb = grid_for(nrows * A1.n1, grid_cap);
    if (small) {
      k_fft2_long<FFT_SMALL_E, 256, false><<<ga, threads, lds_bytes, stream>>>(A1, in, nullptr, scr, 1, n0, n1, nrows, 1.0, inv);
      LAUNCH_CHECK();
      k_fft2_long<FFT_SMALL_E, 256, true><<<gb, threads, lds_bytes, stream>>>(A1, nullptr, mid, scr, 1, n0, n1, nrows, 1.0, inv);
    } else {
      k_fft2_long<BLDFM_FFT_MAXE, 1024, false><<<ga, threads, lds_bytes, stream>>>(A1, in, nullptr, scr, 1, n0, n1, nrows, 1.0, inv);
      LAUNCH_CHECK();
      k_fft2_long<BLDFM_FFT_MAXE, 1024, true><<<gb, threads, lds_bytes, stream>>>(A1, nullptr, mid, scr, 1, n0, n1, nrows, 1.0, inv);
    }
  } else if (small) {
    k_fft2_rows<FFT_SMALL_E, 256><<<grid_for(nrows, grid_cap), threads, lds_bytes, stream>>>(A1.F, A1.tw, in, mid, n1, nrows, inv);
  } else {
    k_fft2_rows<BLDFM_FFT_MAXE, 1024><<<grid_for(nrows, grid_cap), threads, lds_bytes, stream>>>(A1.F, A1.tw, in, mid, n1, nrows, inv);
  }
  LAUNCH_CHECK();
  // columns (axis 0, length n0): mid -> out, scaled
  if (A0.F1) {
    const unsigned ga = grid_for(ncols * A0.n2, grid_cap), gb = grid_for(ncols * A0.n1, grid_cap);
    if (small) {
      k_fft2_long<FFT_SMALL_E, 256, false><<<ga, threads, lds_bytes, stream>>>(A0, mid, nullptr, scr, 0, n0, n1, ncols, 1.0, inv);
      LAUNCH_CHECK();
      k_fft2_long<FFT_SMALL_E, 256, true><<<gb, threads, lds_bytes, stream>>>(A0, nullptr, out, scr, 0, n0, n1, ncols, scale, inv);
    } else {
      k_fft2_long<BLDFM_FFT_MAXE, 1024, false><<<ga, threads, lds_bytes, stream>>>(A0, mid, nullptr, scr, 0, n0, n1, ncols, 1.0, inv);
      LAUNCH_CHECK();
      k_fft2_long<BLDFM_FFT_MAXE, 1024, true><<<gb, threads, lds_bytes, stream>>>(A0, nullptr, out, scr, 0, n0, n1, ncols, scale, inv);
    }
  } else if (small) {
    k_fft2_cols<FFT_SMALL_E, 256><<<grid_for(ncols, grid_cap), threads, lds_bytes, stream>>>(A0.F, A0.tw, mid, out, n0, n1, ncols, scale, inv);
  } else {
    k_fft2_cols<BLDFM_FFT_MAXE, 1024><<<grid_for(ncols, grid_cap), threads, lds_bytes, stream>>>(A0.F, A0.tw, mid, out, n0, n1, ncols, scale, inv);
  }
  LAUNCH_CHECK();
  return 0;
}

// Allow dynamic LDS beyond the default per-workgroup cap for the FFT kernels.
void bldfm_set_fft_lds_attr(int bytes) {
  const void* fns[] = {
      (const void*)k_pass<FFT_SMALL_E, 256, PASS_FWD_ROWS>, (const void*)k_pass<FFT_SMALL_E, 256, PASS_FWD_COLS>,
      (const void*)k_pass<FFT_SMALL_E, 256, PASS_INV_COLS>, (const void*)k_pass<FFT_SMALL_E, 256, PASS_INV_ROWS>,
      (const void*)k_long<FFT_SMALL_E, 256, false>, (const void*)k_long<FFT_SMALL_E, 256, true>,
      (const void*)k_fft2_rows<FFT_SMALL_E, 256>, (const void*)k_fft2_cols<FFT_SMALL_E, 256>,
      (const void*)k_fft2_long<FFT_SMALL_E, 256, false>, (const void*)k_fft2_long<FFT_SMALL_E, 256, true>,
      (const void*)k_pass<BLDFM_FFT_MAXE, 1024, PASS_FWD_ROWS>, (const void*)k_pass<BLDFM_FFT_MAXE, 1024, PASS_FWD_COLS>,
      (const void*)k_pass<BLDFM_FFT_MAXE, 1024, PASS_INV_COLS>, (const void*)k_pass<BLDFM_FFT_MAXE, 1024, PASS_INV_ROWS>,
      (const void*)k_long<BLDFM_FFT_MAXE, 1024, false>, (const void*)k_long<BLDFM_FFT_MAXE, 1024, true>,
      (const void*)k_fft2_rows<BLDFM_FFT_MAXE, 1024>, (const void*)k_fft2_cols<BLDFM_FFT_MAXE, 1024>,
      (const void*)k_fft2_long<BLDFM_FFT_MAXE, 1024, false>, (const void*)k_fft2_long<BLDFM_FFT_MAXE, 1024, true>};
  for (const void* f : fns) (void)hipFuncSetAttribute(f, hipFuncAttributeMaxDynamicSharedMemorySize, bytes);
}

// ---------------------------------------------------------------- elementwise
// psi(x), phi(x) of MOST (src/bldfm/pbl_model.py:207-250).
__global__ __launch_bounds__(256) void k_stability(const double* __restrict__ x, int64_t n,
                                                   double* __restrict__ psi, double* __restrict__ phi) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
    const double v = x[i];
    if (psi) psi[i] = most_psi(v);
    if (phi) phi[i] = most_phi(v);
  }
}

// Met wind (speed, direction blowing FROM, degrees) -> (u, v)
// (src/bldfm/utils.py:23-27): u = -U sin(theta), v = -U cos(theta).
__global__ __launch_bounds__(256) void k_wind(const double* __restrict__ spd, const double* __restrict__ dir,
                                              int64_t n, double* __restrict__ u, double* __restrict__ v) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
    wind_uv(spd[i], dir[i], &u[i], &v[i]);
  }
}

int bldfm_launch_stability(const double* x, int64_t n, double* psi, double* phi, hipStream_t stream) {
  if (n <= 0) return 0;
  k_stability<<<grid_for((n + 255) / 256, 4096), 256, 0, stream>>>(x, n, psi, phi);
  LAUNCH_CHECK();
  return 0;
}

int bldfm_launch_wind(const double* spd, const double* dir, int64_t n, double* u, double* v, hipStream_t stream) {
  if (n <= 0) return 0;
  k_wind<<<grid_for((n + 255) / 256, 4096), 256, 0, stream>>>(spd, dir, n, u, v);
  LAUNCH_CHECK();
  return 0;
}

// ---------------------------------------------------------------- math probe
// Evaluates the device math helpers on n inputs for bitwise checks:
// op 0 hypot_borges(a, b); op 1 csqrt_exact(a + ib) -> (out[2i], out[2i+1]);
// op 2 cdiv((a0 + i a1), (b0 + i b1)) with a, b holding complex pairs;
// op 3 ncmul of complex pairs; op 4 xcmul of complex pairs; op 5 sqrt(a).
__global__ __launch_bounds__(256) void k_math_probe(int op, const double* __restrict__ a,
                                                    const double* __restrict__ b, int64_t n,
                                                    double* __restrict__ out) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
    if (op == 0) {
      out[i] = hypot_borges(a[i], b[i]);
    } else if (op == 1) {
      const double2 r = csqrt_exact(cmk(a[i], b[i]));
      out[2 * i] = r.x;
      out[2 * i + 1] = r.y;
    } else if (op == 5) {
      out[i] = sqrt(a[i]);
    } else {
      const double2 x = cmk(a[2 * i], a[2 * i + 1]), y = cmk(b[2 * i], b[2 * i + 1]);
      const double2 r = op == 2 ? cdiv(x, y) : (op == 3 ? ncmul(x, y) : xcmul(x, y));
      out[2 * i] = r.x;
      out[2 * i + 1] = r.y;
    }
  }
}

int bldfm_launch_math_probe(int op, const double* a, const double* b, int64_t n, double* out, hipStream_t stream) {
  if (n <= 0) return 0;
  k_math_probe<<<grid_for((n + 255) / 256, 4096), 256, 0, stream>>>(op, a, b, n, out);
  LAUNCH_CHECK();
  return 0;
}

int bldfm_launch_tables(const ProfSetup* S, int nmet, int rows, double* tab, hipStream_t stream) {
  const int64_t n = (int64_t)nmet * rows;
  k_tables_met<<<(unsigned)((n + 255) / 256), 256, 0, stream>>>(S, nmet, rows, tab);
  LAUNCH_CHECK();
  return 0;
}

int bldfm_launch_transfer(const double* tab, int rows, const int32_t* nzs, const int32_t* lvls, int nmet,
                          const int32_t* blk_met, int nblk, const double2* kxy, double2* g, int exact,
                          hipStream_t stream) {
  const unsigned grid = (unsigned)(nblk < (1 << 20) ? nblk : (1 << 20));
  if (exact)
    k_transfer<true><<<grid, 256, 0, stream>>>(tab, rows, nzs, lvls, nmet, blk_met, nblk, kxy, g);
  else
    k_transfer<false><<<grid, 256, 0, stream>>>(tab, rows, nzs, lvls, nmet, blk_met, nblk, kxy, g);
  LAUNCH_CHECK();
  return 0;
}

int bldfm_launch_km01(const Km01Dev* P, int n, int64_t max_cells, double* out, hipStream_t stream) {
  if (n <= 0 || max_cells <= 0) return 0;
  dim3 grid((unsigned)((max_cells + 255) / 256), (unsigned)n);
  k_km01<<<grid, 256, 0, stream>>>(P, out);
  LAUNCH_CHECK();
  return 0;
}
