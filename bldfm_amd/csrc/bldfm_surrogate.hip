// HIP kernels of the learned spectral surrogate for gfx950 (MI355X):
// bf16 MFMA (v_mfma_f32_32x32x16_bf16) with fp32 accumulation.
//
// One workgroup (four waves) carries SUR_PTS = 128 points through the whole
// network.  The activation tile (128 points x 256 features, bf16, 64 KB)
// stays in LDS between layers; weights are read in MFMA fragment order from
// global memory (L2-resident).  Products are formed transposed,
//   Z^T (256 x 128) = W (256 x K) X^T (K x 128),
// so wave w owns output features [64w, 64w+64) for all 128 points, and each
// lane's accumulators hold four consecutive features of one point, which go
// back to the LDS tile as 8-byte stores.
//
// Training kernels (point lists):
//   k_sur_fwd<1>  features, five hidden layers, output, loss and dLoss/dPhi;
//                 stores x_l and z_l (bf16, feature-major) for the backward
//   k_sur_bwd     dz_l for l = 4..0 (the transposed weights, same tile
//                 scheme), bias and output-layer gradients
//   k_sur_wgrad   dW_l = dz_l x_l^T, LDS-staged tiles over point chunks,
//                 fp32 atomics
//   k_sur_adam    Adam on the fp32 master vector, then the bf16/fp32 repack
// Inference inside a plan: k_sur_plan writes the packed footprint spectrum
// of accepted instances exactly where the shooting kernel would.
#include "bldfm_internal.h"
#include "bldfm_march.h"
#include "bldfm_sur.h"

typedef __attribute__((ext_vector_type(8))) __bf16 bf16x8;
typedef __attribute__((ext_vector_type(4))) __bf16 bf16x4;
typedef __attribute__((ext_vector_type(16))) float f32x16;
typedef const __attribute__((address_space(4))) SurStatic cstat_t;
typedef const __attribute__((address_space(4))) SurMet cmet_t;

#define SUR_TWO_PI 6.283185307179586f  // loss phase wrap

// SiLU and its derivative with the hardware exp2 and reciprocal (about one
// ulp each; the results are rounded to bf16 or feed bf16 products).
__device__ __forceinline__ float sigmoid_f(float z) { return __builtin_amdgcn_rcpf(1.0f + __expf(-z)); }

__device__ __forceinline__ float silu_f(float z) { return z * sigmoid_f(z); }

__device__ __forceinline__ float dsilu_f(float z) {
  const float s = sigmoid_f(z);
  return s + z * s * (1.0f - s);
}

// Offsets (elements) of the stored training activations: x0 is SUR_IN rows,
// x1..x5 SUR_W rows each, every row N points long.
__device__ __forceinline__ int64_t xs_off(int l, int64_t N) {
  return l == 0 ? 0 : (int64_t)SUR_IN * N + (int64_t)(l - 1) * SUR_W * N;
}

__device__ __forceinline__ float wave_sum(float v) {
#pragma unroll
  for (int o = 32; o >= 1; o >>= 1) v += __shfl_xor(v, o);
  return v;
}

// ---------------------------------------------------------------- prior
// WKB estimate of -log G: Phi0 = sum_j h_j sqrt(|k|^2 + i kpar w_j), the
// principal root (Re >= 0) in the form that avoids cancellation.
__device__ __forceinline__ float2 wkb_phi0(cmet_t& M, float lx, float ly) {
  const float k2 = lx * lx + ly * ly;
  const float kpar = lx * M.wx + ly * M.wy;
  const int nw = (int)M.nw;
  float sr = 0.f, si = 0.f;
  for (int j = 0; j < nw; ++j) {
    const float b = kpar * M.wkb[j][1];
    const float re = sqrtf(0.5f * (sqrtf(k2 * k2 + b * b) + k2));
    const float im = re > 0.f ? 0.5f * b / re : 0.f;
    sr += M.wkb[j][0] * re;
    si += M.wkb[j][0] * im;
  }
  return make_float2(sr, si);
}

// ---------------------------------------------------------------- features
// Writes the SUR_IN inputs of one point into its LDS row; part 0 writes the
// first 16 frequencies and the prior Phi0 to p0s[pt], part 1 the other 16
// frequencies, the coordinates, met values, Phi0/8 and zero padding.  With
// xs, the bf16 values are also stored feature-major.
__device__ __forceinline__ void sur_features(cstat_t& S, cmet_t& M, float lx, float ly, int part,
                                             __bf16* act, float2* p0s, int pt, __bf16* __restrict__ xs,
                                             int64_t N, int64_t gp) {
  auto put = [&](int f, float v) {
    const __bf16 b = (__bf16)v;
    act[sur_act_off(pt, f)] = b;
    if (xs) xs[(int64_t)f * N + gp] = b;
  };
  const int j0 = part * (SUR_NRFF / 2);
  if (S.mode == (float)SUR_MODE_SPEC) {  // RFF of (kx, ky), 6 met scalars, no prior
    const float c0 = lx * S.cscale[0], c1 = ly * S.cscale[1];
    for (int j = j0; j < j0 + SUR_NRFF / 2; ++j) {
      float rev = S.rff[j][0] * c0 + S.rff[j][1] * c1;
      rev -= rintf(rev);
      put(j, __builtin_amdgcn_sinf(rev));
      put(SUR_NRFF + j, __builtin_amdgcn_cosf(rev));
    }
    if (part == 1) {
      for (int d = 0; d < SUR_NMET_SPEC; ++d) put(2 * SUR_NRFF + d, (M.met[d] - S.met_mu[d]) / S.met_sd[d]);
      for (int f = 2 * SUR_NRFF + SUR_NMET_SPEC; f < SUR_IN; ++f) put(f, 0.0f);
    } else {
      p0s[pt] = make_float2(0.0f, 0.0f);
    }
    return;
  }
  const float2 phi0 = wkb_phi0(M, lx, ly);
  const float kpar = (lx * M.wx + ly * M.wy) * M.zm;
  const float kperp = fabsf(ly * M.wx - lx * M.wy) * M.zm;
  float c[SUR_NC];
  c[0] = kpar * S.cscale[0];
  c[1] = kperp * S.cscale[1];
  c[2] = copysignf(sqrtf(fabsf(kpar)), kpar) * S.cscale[2];
  c[3] = sqrtf(kperp) * S.cscale[3];
  for (int j = j0; j < j0 + SUR_NRFF / 2; ++j) {
    // phase in revolutions, reduced exactly to [-1/2, 1/2]; the hardware
    // sine and cosine take revolutions
    float rev = S.rff[j][0] * c[0] + S.rff[j][1] * c[1] + S.rff[j][2] * c[2] + S.rff[j][3] * c[3];
    rev -= rintf(rev);
    put(j, __builtin_amdgcn_sinf(rev));
    put(SUR_NRFF + j, __builtin_amdgcn_cosf(rev));
  }
  if (part == 1) {
    for (int d = 0; d < SUR_NC; ++d) put(2 * SUR_NRFF + d, c[d]);
    for (int d = 0; d < SUR_NMET; ++d) put(2 * SUR_NRFF + SUR_NC + d, (M.met[d] - S.met_mu[d]) / S.met_sd[d]);
    put(SUR_F_PHI0, phi0.x * SUR_PHI0_SCALE);
    put(SUR_F_PHI0 + 1, phi0.y * SUR_PHI0_SCALE);
    for (int f = SUR_F_PHI0 + 2; f < SUR_IN; ++f) put(f, 0.0f);
  } else {
    p0s[pt] = phi0;
  }
}

// ---------------------------------------------------------------- MFMA tile
// acc[a][pb] = sum_k W[64 wave + 32 a + row][k] * act[32 pb + col][k] over
// KB16 steps of 16, W in forward fragment order (sur_pack_fwd / sur_pack_bwd).
// Element j of acc[a][pb] in lane (half, r) is output feature
// 64 wave + 32 a + 8 (j/4) + 4 half + j%4 of point 32 pb + r.
// With bias, the accumulators start from it (the epilogue then adds nothing).
template <int KB16, int UNROLL>
__device__ __forceinline__ void mfma_tile(const bf16x8* __restrict__ Wp, const __bf16* act, int wave,
                                          int lane, f32x16 (&acc)[2][4], const float* bias = nullptr) {
  const int r = lane & 31, half = lane >> 5;
#pragma unroll
  for (int a = 0; a < 2; ++a)
#pragma unroll
    for (int jq = 0; jq < 4; ++jq) {
      float4 b4 = make_float4(0.f, 0.f, 0.f, 0.f);
      if (bias) b4 = *(const float4*)&bias[64 * wave + 32 * a + 8 * jq + 4 * half];
#pragma unroll
      for (int b = 0; b < 4; ++b) {
        acc[a][b][4 * jq + 0] = b4.x;
        acc[a][b][4 * jq + 1] = b4.y;
        acc[a][b][4 * jq + 2] = b4.z;
        acc[a][b][4 * jq + 3] = b4.w;
      }
    }
  const bf16x8* A0p = Wp + (int64_t)(wave * 2 + 0) * KB16 * 64 + lane;
  const bf16x8* A1p = Wp + (int64_t)(wave * 2 + 1) * KB16 * 64 + lane;
  // weight fragments come from L2: unrolled, the scheduler issues later
  // steps' loads under earlier steps' MFMAs (UNROLL = KB16 where registers allow)
#pragma unroll UNROLL
  for (int kb = 0; kb < KB16; ++kb) {
    const bf16x8 A0 = A0p[kb * 64];
    const bf16x8 A1 = A1p[kb * 64];
    const int f = kb * 16 + half * 8;
    bf16x8 Bv[4];
#pragma unroll
    for (int pb = 0; pb < 4; ++pb) Bv[pb] = *(const bf16x8*)(act + sur_act_off(pb * 32 + r, f));
#pragma unroll
    for (int pb = 0; pb < 4; ++pb) {
      acc[0][pb] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(A0, Bv[pb], acc[0][pb], 0, 0, 0);
      acc[1][pb] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(A1, Bv[pb], acc[1][pb], 0, 0, 0);
    }
  }
}

__device__ __forceinline__ void lds_put4(__bf16* act, int pt, int o0, const float* v) {
  bf16x4 w;
  w[0] = (__bf16)v[0];
  w[1] = (__bf16)v[1];
  w[2] = (__bf16)v[2];
  w[3] = (__bf16)v[3];
  *(bf16x4*)(act + sur_act_off(pt, o0)) = w;
}

// Copy of the network's biases and output layer into LDS (dyn), read by the
// epilogues instead of L2.
__device__ __forceinline__ void stage_dyn(const uint8_t* __restrict__ net, SurDyn* dsh) {
  const float4* src = (const float4*)(net + SUR_OFF_DYN);
  for (int i = threadIdx.x; i < (int)(sizeof(SurDyn) / 16); i += blockDim.x) ((float4*)dsh)[i] = src[i];
}

// Epilogue of hidden layer l for one wave's accumulators: x = SiLU(z) into the
// LDS tile (LAST: folded into the two fp32 outputs, per-wave partial sums in
// red).  FOLD: z is the accumulator itself (the products started from the
// biases).  TRAIN: also stores z_l and x_{l+1} (bf16) feature-major at point
// offset p0.
template <bool TRAIN, bool FOLD, bool LAST>
__device__ __forceinline__ void layer_epilogue(const f32x16 (&acc)[2][4], const SurDyn* dyn, int l, __bf16* act,
                                               float* red, __bf16* __restrict__ xs, __bf16* __restrict__ zs,
                                               int64_t N, int64_t p0, int wave, int lane) {
  const int r = lane & 31, half = lane >> 5;
  float po[4][2];
#pragma unroll
  for (int pb = 0; pb < 4; ++pb) po[pb][0] = po[pb][1] = 0.0f;
#pragma unroll
  for (int a = 0; a < 2; ++a) {
#pragma unroll
    for (int jq = 0; jq < 4; ++jq) {
      const int o0 = 64 * wave + 32 * a + 8 * jq + 4 * half;
      float bb[4] = {0.f, 0.f, 0.f, 0.f};
      if (!FOLD) {
        const float4 b4 = *(const float4*)&dyn->b[l][o0];
        bb[0] = b4.x, bb[1] = b4.y, bb[2] = b4.z, bb[3] = b4.w;
      }
      float4 w0 = make_float4(0.f, 0.f, 0.f, 0.f), w1 = w0;
      if (LAST) {
        w0 = *(const float4*)&dyn->wo[0][o0];
        w1 = *(const float4*)&dyn->wo[1][o0];
      }
#pragma unroll
      for (int pb = 0; pb < 4; ++pb) {
        const int pt = 32 * pb + r;
        float z[4], x[4];
#pragma unroll
        for (int e = 0; e < 4; ++e) z[e] = FOLD ? acc[a][pb][4 * jq + e] : acc[a][pb][4 * jq + e] + bb[e];
#pragma unroll
        for (int e = 0; e < 4; ++e) x[e] = silu_f(z[e]);
        if (TRAIN) {
          __bf16* zr = zs + ((int64_t)l * SUR_W + o0) * N + p0 + pt;
          __bf16* xr = xs + xs_off(l + 1, N) + (int64_t)o0 * N + p0 + pt;
#pragma unroll
          for (int e = 0; e < 4; ++e) {
            zr[e * N] = (__bf16)z[e];
            xr[e * N] = (__bf16)x[e];
          }
        }
        if (!LAST) {
          lds_put4(act, pt, o0, x);
        } else {
          po[pb][0] += w0.x * x[0] + w0.y * x[1] + w0.z * x[2] + w0.w * x[3];
          po[pb][1] += w1.x * x[0] + w1.y * x[1] + w1.z * x[2] + w1.w * x[3];
        }
      }
    }
  }
  if (LAST) {
#pragma unroll
    for (int pb = 0; pb < 4; ++pb) {
      const float v0 = po[pb][0] + __shfl_xor(po[pb][0], 32);
      const float v1 = po[pb][1] + __shfl_xor(po[pb][1], 32);
      if (half == 0) {
        red[(wave * SUR_PTS + 32 * pb + r) * 2 + 0] = v0;
        red[(wave * SUR_PTS + 32 * pb + r) * 2 + 1] = v1;
      }
    }
  }
}

// Hidden layers and output of the points whose inputs are in `act`, with
// biases and output weights from `dyn` (LDS).  Leaves Phi = Phi0 + network
// output of local point t (t < SUR_PTS) in phis[t].  TRAIN: stores z_l
// and x_{l+1} (bf16) feature-major at point offset p0.  FOLD (serving): the
// products start from the biases, so the epilogue is SiLU and the store alone.
template <bool TRAIN, bool FOLD = false>
__device__ __forceinline__ void sur_hidden(const uint8_t* __restrict__ net, const SurDyn* dyn, __bf16* act,
                                           float* red, const float2* p0s, float2* phis,
                                           __bf16* __restrict__ xs, __bf16* __restrict__ zs, int64_t N,
                                           int64_t p0) {
  const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
  constexpr int U = TRAIN ? 1 : SUR_W / 32;  // the training kernel has no registers to spare
  f32x16 acc[2][4];
  // layer 0, the middle layers and the last layer as straight-line code: with
  // the epilogue kinds in branches of one loop body the compiler hoists their
  // common arithmetic above the branch, and the hoisted values spill
  mfma_tile<SUR_IN / 16, TRAIN ? 1 : SUR_IN / 16>((const bf16x8*)(net + SUR_OFF_W0P), act, wave, lane, acc,
                                                  FOLD ? dyn->b[0] : nullptr);
  __syncthreads();  // every wave has read this layer's input tile
  layer_epilogue<TRAIN, FOLD, false>(acc, dyn, 0, act, red, xs, zs, N, p0, wave, lane);
  __syncthreads();
  for (int l = 1; l < SUR_NL - 1; ++l) {
    mfma_tile<SUR_W / 16, U>((const bf16x8*)(net + SUR_OFF_WP(l)), act, wave, lane, acc, FOLD ? dyn->b[l] : nullptr);
    __syncthreads();
    layer_epilogue<TRAIN, FOLD, false>(acc, dyn, l, act, red, xs, zs, N, p0, wave, lane);
    __syncthreads();
  }
  mfma_tile<SUR_W / 16, U>((const bf16x8*)(net + SUR_OFF_WP(SUR_NL - 1)), act, wave, lane, acc,
                           FOLD ? dyn->b[SUR_NL - 1] : nullptr);
  __syncthreads();
  layer_epilogue<TRAIN, FOLD, true>(acc, dyn, SUR_NL - 1, act, red, xs, zs, N, p0, wave, lane);
  __syncthreads();
  if (tid < SUR_PTS) {
    float s0 = dyn->bo[0] + p0s[tid].x, s1 = dyn->bo[1] + p0s[tid].y;
#pragma unroll
    for (int w = 0; w < 4; ++w) {
      s0 += red[(w * SUR_PTS + tid) * 2 + 0];
      s1 += red[(w * SUR_PTS + tid) * 2 + 1];
    }
    phis[tid] = make_float2(s0, s1);
  }
  __syncthreads();
}

// ---------------------------------------------------------------- training
// Loss of a block of SUR_PTS points of one met sample, normalised by the
// block's solver energy S = sum |G|^2 and by the number of blocks:
//   sum_i |G_i - G^_i|^2 + |G_i|^2 |Phi^_i + log G_i|^2   (imaginary part wrapped to [-pi, pi])
// with G^ = exp(-Phi^): the footprint's L2 error plus a log-space term that
// keeps the phase well conditioned on the energetic modes.  Points whose
// G is roundoff (see below) are left out.
template <bool TRAIN>
__global__ __launch_bounds__(SUR_THREADS, 2) void k_sur_fwd(const uint8_t* __restrict__ net,
                                                         const SurMet* __restrict__ mets, int nmet,
                                                         const int32_t* __restrict__ blk_met,
                                                         const double2* __restrict__ kxy,
                                                         const double2* __restrict__ gt, int64_t N,
                                                         __bf16* __restrict__ xs, __bf16* __restrict__ zs,
                                                         float2* __restrict__ dphi, float2* __restrict__ phi,
                                                         float* __restrict__ loss) {
  __shared__ __attribute__((aligned(16))) __bf16 act[SUR_PTS * SUR_W];
  __shared__ float red[4 * SUR_PTS * 2];
  __shared__ float2 phis[SUR_PTS];
  __shared__ float2 p0s[SUR_PTS];
  __shared__ __attribute__((aligned(16))) SurDyn dsh;
  __shared__ float wsum[8];
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int64_t p0 = (int64_t)blockIdx.x * SUR_PTS;
  int m = blk_met[blockIdx.x];
  if (m < 0 || m >= nmet) m = 0;  // the host validates; never index outside mets
  cstat_t& S = *(cstat_t*)(net + SUR_OFF_STATIC);
  cmet_t& M = *(cmet_t*)(mets + m);
  const int pt = tid & (SUR_PTS - 1), part = tid >> 7;
  const double2 k = kxy[p0 + pt];
  sur_features(S, M, (float)k.x, (float)k.y, part, act, p0s, pt, TRAIN ? xs : nullptr, N, p0 + pt);
  stage_dyn(net, &dsh);
  __syncthreads();
  sur_hidden<TRAIN>(net, &dsh, act, red, p0s, phis, xs, zs, N, p0);
  if (phi && tid < SUR_PTS) phi[p0 + tid] = phis[tid];
  if (!TRAIN) return;

  // Points whose solver transfer is roundoff rather than physics carry no
  // loss: not finite, |G| > 1.5, or |G| above 100 times the WKB prior (plus
  // 1e-8) -- in deep, stable columns the shooting's cancellation leaves
  // noise where G is exponentially small.
  float gr = 0.f, gi = 0.f, s2 = 0.f;
  bool valid = false;
  if (tid < SUR_PTS) {
    const double2 g = gt[p0 + tid];
    const double a2 = g.x * g.x + g.y * g.y;
    const double b = 100.0 * exp(-(double)p0s[tid].x) + 1e-8;
    valid = a2 <= 2.25 && a2 <= b * b;  // false for NaN
    if (valid) {
      gr = (float)g.x;
      gi = (float)g.y;
      s2 = gr * gr + gi * gi;
    }
  }
  float ws = wave_sum(s2);
  if (lane == 0) wsum[wave] = ws;
  __syncthreads();
  const float Sb = wsum[0] + wsum[1] + wsum[2] + wsum[3];
  __syncthreads();
  float li = 0.f;
  if (tid < SUR_PTS && !valid) dphi[p0 + tid] = make_float2(0.f, 0.f);
  if (tid < SUR_PTS && valid && S.mode == (float)SUR_MODE_SPEC) {  // outputs are (Re G, Im G)
    const float2 P = phis[tid];
    const float er = P.x - gr, ei = P.y - gi;
    li = er * er + ei * ei;
    const float sc = 1.0f / ((Sb + 1e-30f) * (float)gridDim.x);
    li *= sc;
    dphi[p0 + tid] = make_float2(2.0f * er * sc, 2.0f * ei * sc);
  } else if (tid < SUR_PTS && valid) {
    const float2 P = phis[tid];
    const float mag = __expf(-fmaxf(P.x, -20.0f));
    float sn, cs;
    sincosf(P.y, &sn, &cs);
    const float pr = mag * cs, pim = -mag * sn;  // G^ = exp(-Phi)
    const float er = pr - gr, ei = pim - gi;
    float dre = -2.0f * (er * pr + ei * pim);
    float dim = 2.0f * (er * pim - ei * pr);
    li = er * er + ei * ei;
    if (s2 > 1e-30f) {
      const float tr = -0.5f * __logf(s2), ti = -atan2f(gi, gr);
      const float dr = P.x - tr;
      float di = P.y - ti;
      di -= SUR_TWO_PI * rintf(di / SUR_TWO_PI);
      li += s2 * (dr * dr + di * di);
      dre += 2.0f * s2 * dr;
      dim += 2.0f * s2 * di;
    }
    const float sc = 1.0f / ((Sb + 1e-30f) * (float)gridDim.x);
    li *= sc;
    dphi[p0 + tid] = make_float2(dre * sc, dim * sc);
  }
  const float lw = wave_sum(li);
  if (lane == 0) wsum[4 + wave] = lw;
  __syncthreads();
  if (tid == 0) atomicAdd(loss, wsum[4] + wsum[5] + wsum[6] + wsum[7]);
}

// Bias gradient of one layer from the dz tile in LDS: thread t sums column
// t over the block's points (one atomic per feature and block).
__device__ __forceinline__ void bias_grad(const __bf16* act, float* __restrict__ gb) {
  const int t = threadIdx.x;
  float s = 0.f;
  for (int pt = 0; pt < SUR_PTS; ++pt) s += (float)act[sur_act_off(pt, t)];
  unsafeAtomicAdd(gb + t, s);
}

// dz_l for l = 4..0 of one block of points, plus the bias and output-layer
// gradients of the block (accumulated with atomics).
__global__ __launch_bounds__(SUR_THREADS, 2) void k_sur_bwd(const uint8_t* __restrict__ net, int64_t N,
                                                         const __bf16* __restrict__ xs,
                                                         const __bf16* __restrict__ zs,
                                                         const float2* __restrict__ dphi,
                                                         __bf16* __restrict__ dzs, float* __restrict__ grad) {
  __shared__ __attribute__((aligned(16))) __bf16 act[SUR_PTS * SUR_W];
  __shared__ float2 dp[SUR_PTS];
  const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63, r = lane & 31, half = lane >> 5;
  const int64_t p0 = (int64_t)blockIdx.x * SUR_PTS;
  const SurDyn* __restrict__ dyn = (const SurDyn*)(net + SUR_OFF_DYN);
  if (tid < SUR_PTS) dp[tid] = dphi[p0 + tid];
  __syncthreads();
  // output layer: dWo[c][f] = sum_p dPhi_c x5[f], dbo = sum_p dPhi
  {
    const bf16x8* x5 = (const bf16x8*)(xs + xs_off(SUR_NL, N) + (int64_t)tid * N + p0);
    float s0 = 0.f, s1 = 0.f;
    for (int q = 0; q < SUR_PTS / 8; ++q) {
      const bf16x8 v = x5[q];
#pragma unroll
      for (int i = 0; i < 8; ++i) {
        const float2 d = dp[q * 8 + i];
        s0 += d.x * (float)v[i];
        s1 += d.y * (float)v[i];
      }
    }
    unsafeAtomicAdd(grad + SUR_P_WO + tid, s0);
    unsafeAtomicAdd(grad + SUR_P_WO + SUR_W + tid, s1);
    if (tid < 2) {
      float sb = 0.f;
      for (int q = 0; q < SUR_PTS; ++q) sb += tid == 0 ? dp[q].x : dp[q].y;
      unsafeAtomicAdd(grad + SUR_P_BO + tid, sb);
    }
  }
  // dz_4 = (Wo^T dPhi) * silu'(z_4)
#pragma unroll
  for (int a = 0; a < 2; ++a) {
#pragma unroll
    for (int jq = 0; jq < 4; ++jq) {
      const int o0 = 64 * wave + 32 * a + 8 * jq + 4 * half;
      const float4 w0 = *(const float4*)&dyn->wo[0][o0];
      const float4 w1 = *(const float4*)&dyn->wo[1][o0];
      const float wa[4] = {w0.x, w0.y, w0.z, w0.w}, wb[4] = {w1.x, w1.y, w1.z, w1.w};
#pragma unroll
      for (int pb = 0; pb < 4; ++pb) {
        const int pt = 32 * pb + r;
        const float2 d = dp[pt];
        float v[4];
#pragma unroll
        for (int e = 0; e < 4; ++e) {
          const int64_t at = ((int64_t)(SUR_NL - 1) * SUR_W + o0 + e) * N + p0 + pt;
          const float dz = (d.x * wa[e] + d.y * wb[e]) * dsilu_f((float)zs[at]);
          const __bf16 b = (__bf16)dz;
          dzs[at] = b;
          v[e] = (float)b;
        }
        lds_put4(act, pt, o0, v);
      }
    }
  }
  __syncthreads();
  bias_grad(act, grad + SUR_P_B(SUR_NL - 1));
  f32x16 acc[2][4];
  for (int l = SUR_NL - 1; l >= 1; --l) {
    mfma_tile<SUR_W / 16, SUR_W / 16>((const bf16x8*)(net + SUR_OFF_WT(l)), act, wave, lane, acc);
    __syncthreads();
#pragma unroll
    for (int a = 0; a < 2; ++a) {
#pragma unroll
      for (int jq = 0; jq < 4; ++jq) {
        const int o0 = 64 * wave + 32 * a + 8 * jq + 4 * half;
#pragma unroll
        for (int pb = 0; pb < 4; ++pb) {
          const int pt = 32 * pb + r;
          float v[4];
#pragma unroll
          for (int e = 0; e < 4; ++e) {
            const int64_t at = ((int64_t)(l - 1) * SUR_W + o0 + e) * N + p0 + pt;
            const float dz = acc[a][pb][4 * jq + e] * dsilu_f((float)zs[at]);
            const __bf16 b = (__bf16)dz;
            dzs[at] = b;
            v[e] = (float)b;
          }
          lds_put4(act, pt, o0, v);
        }
      }
    }
    __syncthreads();
    bias_grad(act, grad + SUR_P_B(l - 1));
  }
}

// dW_l[o][k] += sum_p dz_l[o][p] x_l[k][p] over one chunk of points.
// Workgroup tile 128 (o) x 128 (k): both operands are staged through LDS
// 64 points at a time (16-byte chunks XOR-swizzled by row), and wave
// (wo, wk) accumulates its 64 x 64 part from there.
#define WG_PSTEP 64
__global__ __launch_bounds__(SUR_THREADS, 2) void k_sur_wgrad(int64_t N, int chunk, const __bf16* __restrict__ xs,
                                                              const __bf16* __restrict__ dzs,
                                                              float* __restrict__ grad) {
  __shared__ __attribute__((aligned(16))) __bf16 sa[128 * WG_PSTEP];
  __shared__ __attribute__((aligned(16))) __bf16 sb[128 * WG_PSTEP];
  const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63, r = lane & 31, half = lane >> 5;
  const int nch = (int)(N / chunk);
  int b = blockIdx.x, l, ot, kt, c;
  if (b < 2 * nch) {
    l = 0; ot = b / nch; kt = 0; c = b % nch;
  } else {
    b -= 2 * nch;
    l = 1 + b / (4 * nch);
    b %= 4 * nch;
    ot = b / (2 * nch);
    b %= 2 * nch;
    kt = b / nch;
    c = b % nch;
  }
  const int kin = l == 0 ? SUR_IN : SUR_W;
  const __bf16* __restrict__ X = xs + xs_off(l, N);
  const __bf16* __restrict__ dZ = dzs + (int64_t)l * SUR_W * N;
  const int wo = wave >> 1, wk = wave & 1;
  bool kv[2];
#pragma unroll
  for (int q = 0; q < 2; ++q) kv[q] = kt * 128 + wk * 64 + q * 32 < kin;
  f32x16 acc[2][2];
#pragma unroll
  for (int a = 0; a < 2; ++a)
#pragma unroll
    for (int q = 0; q < 2; ++q)
#pragma unroll
      for (int j = 0; j < 16; ++j) acc[a][q][j] = 0.0f;
  bf16x8 zero;
#pragma unroll
  for (int i = 0; i < 8; ++i) zero[i] = (__bf16)0.0f;
  const int64_t pbeg = (int64_t)c * chunk, pend = pbeg + chunk;
  for (int64_t p = pbeg; p < pend; p += WG_PSTEP) {
    for (int q = tid; q < 128 * (WG_PSTEP / 8); q += SUR_THREADS) {
      const int row = q / (WG_PSTEP / 8), ch = q % (WG_PSTEP / 8);
      const int so = row * WG_PSTEP + ((ch ^ (row & 7)) << 3);
      *(bf16x8*)(sa + so) = *(const bf16x8*)(dZ + (int64_t)(ot * 128 + row) * N + p + ch * 8);
      const int krow = kt * 128 + row;
      *(bf16x8*)(sb + so) = krow < kin ? *(const bf16x8*)(X + (int64_t)krow * N + p + ch * 8) : zero;
    }
    __syncthreads();
#pragma unroll
    for (int ks = 0; ks < WG_PSTEP / 16; ++ks) {
      const int ch = ks * 2 + half;
      const int sw = (ch ^ (r & 7)) << 3;
      const bf16x8 A0 = *(const bf16x8*)(sa + (wo * 64 + r) * WG_PSTEP + sw);
      const bf16x8 A1 = *(const bf16x8*)(sa + (wo * 64 + 32 + r) * WG_PSTEP + sw);
      if (kv[0]) {
        const bf16x8 B0 = *(const bf16x8*)(sb + (wk * 64 + r) * WG_PSTEP + sw);
        acc[0][0] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(A0, B0, acc[0][0], 0, 0, 0);
        acc[1][0] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(A1, B0, acc[1][0], 0, 0, 0);
      }
      if (kv[1]) {
        const bf16x8 B1 = *(const bf16x8*)(sb + (wk * 64 + 32 + r) * WG_PSTEP + sw);
        acc[0][1] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(A0, B1, acc[0][1], 0, 0, 0);
        acc[1][1] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(A1, B1, acc[1][1], 0, 0, 0);
      }
    }
    __syncthreads();
  }
  float* __restrict__ gW = grad + SUR_P_W(l);
#pragma unroll
  for (int a = 0; a < 2; ++a) {
#pragma unroll
    for (int q = 0; q < 2; ++q) {
      if (!kv[q]) continue;
      const int k = kt * 128 + wk * 64 + q * 32 + r;
#pragma unroll
      for (int j = 0; j < 16; ++j) {
        const int o = ot * 128 + wo * 64 + a * 32 + 8 * (j / 4) + 4 * half + (j % 4);
        unsafeAtomicAdd(gW + (int64_t)o * kin + k, acc[a][q][j]);
      }
    }
  }
}

// Adam step on the fp32 master parameters (update != 0), then the repack of
// every parameter into the network buffer: bf16 forward and transposed
// fragments of W_l, fp32 biases and output layer.  Clears the gradient.
__global__ __launch_bounds__(256) void k_sur_adam(float* __restrict__ p, float* __restrict__ g,
                                                  float* __restrict__ m, float* __restrict__ v,
                                                  uint8_t* __restrict__ net, int update, float lr, float b1,
                                                  float b2, float eps, float c1, float c2) {
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i >= SUR_NPARAM) return;
  float w = p[i];
  if (update) {
    const float gi = g[i];
    const float mi = b1 * m[i] + (1.0f - b1) * gi;
    const float vi = b2 * v[i] + (1.0f - b2) * gi * gi;
    m[i] = mi;
    v[i] = vi;
    w -= lr * (mi * c1) / (sqrtf(vi * c2) + eps);
    p[i] = w;
    g[i] = 0.0f;
  }
  SurDyn* dyn = (SurDyn*)(net + SUR_OFF_DYN);
  if (i < SUR_P_B0) {
    ((__bf16*)(net + SUR_OFF_W0P))[sur_pack_fwd(i / SUR_IN, i % SUR_IN, SUR_IN)] = (__bf16)w;
  } else if (i < SUR_P_L1) {
    dyn->b[0][i - SUR_P_B0] = w;
  } else if (i < SUR_P_WO) {
    const int j = i - SUR_P_L1, l = 1 + j / SUR_P_LSTRIDE, q = j % SUR_P_LSTRIDE;
    if (q < SUR_W * SUR_W) {
      const int o = q / SUR_W, k = q % SUR_W;
      ((__bf16*)(net + SUR_OFF_WP(l)))[sur_pack_fwd(o, k, SUR_W)] = (__bf16)w;
      ((__bf16*)(net + SUR_OFF_WT(l)))[sur_pack_bwd(o, k)] = (__bf16)w;
    } else {
      dyn->b[l][q - SUR_W * SUR_W] = w;
    }
  } else if (i < SUR_P_BO) {
    dyn->wo[(i - SUR_P_WO) / SUR_W][(i - SUR_P_WO) % SUR_W] = w;
  } else {
    dyn->bo[i - SUR_P_BO] = w;
  }
}

// ---------------------------------------------------------------- inference in a plan
// Footprint spectrum of accepted surrogate instances: per lane
// G = exp(-Phi) with Re Phi >= 0 (|G| <= 1; in mode SUR_MODE_SPEC the two outputs
// are Re G, Im G), X_q = q00 G, X_p = 0, packed and
// phase-shifted like the shooting kernel's output; the (0,0) mode is exact (G = 1).
__global__ __launch_bounds__(SUR_THREADS, 2) void k_sur_plan(PlanDev P, ExecBufs B) {
  __shared__ __attribute__((aligned(16))) __bf16 act[SUR_PTS * SUR_W];
  __shared__ float red[4 * SUR_PTS * 2];
  __shared__ float2 phis[SUR_PTS];
  __shared__ float2 p0s[SUR_PTS];
  __shared__ __attribute__((aligned(16))) SurDyn dsh;
  const int tid = threadIdx.x;
  const TileRec tr = tile_rec(P.t_sur, blockIdx.x);
  const int inst = tr.inst;
  if (B.mode[inst] != 0) return;  // uniform over the workgroup
  cstat_t& S = *(cstat_t*)(P.sur_net + SUR_OFF_STATIC);
  const cdesc_t& D = *(const cdesc_t*)(P.desc + inst);
  cmet_t& M = *(cmet_t*)(P.sur_met + inst);
  const int pt = tid & (SUR_PTS - 1), part = tid >> 7;
  {
    Lane L;
    const bool ok = lane_setup(D, tr.local * SUR_PTS + pt, L);
    sur_features(S, M, ok ? (float)L.lx : 0.0f, ok ? (float)L.ly : 0.0f, part, act, p0s, pt, nullptr, 0, 0);
  }
  stage_dyn(P.sur_net, &dsh);
  __syncthreads();
  sur_hidden<false, true>(P.sur_net, &dsh, act, red, p0s, phis, nullptr, nullptr, 0, 0);
  Lane L;
  if (tid < SUR_PTS && lane_setup(D, tr.local * SUR_PTS + pt, L)) {
    if (L.kind == LANE_ORIGIN) {
      emit_origin(D, B.ws, 0, 0.0, D.q00);
    } else {
      const float2 ph = phis[tid];
      double2 g;
      if (S.mode == (float)SUR_MODE_SPEC) {
        g = cmk((double)ph.x, (double)ph.y);
      } else {
        const double mag = exp(-fmax((double)ph.x, 0.0));
        double sn, cs;
        sincos((double)ph.y, &sn, &cs);
        g = cmk(mag * cs, -mag * sn);
      }
      emit_mode(D, L, B.ws, 0, cmk(0.0, 0.0), cmk(g.x * D.q00, g.y * D.q00), phase_shift(D, L));
    }
  }
}

// ---------------------------------------------------------------- launchers
#define SUR_LAUNCH_CHECK()                         \
  do {                                             \
    hipError_t e_ = hipGetLastError();             \
    if (e_ != hipSuccess) return -(int)e_ - 1000;  \
  } while (0)

int bldfm_launch_sur_plan(const PlanDev& P, const ExecBufs& B, int grid_cap, hipStream_t stream) {
  (void)grid_cap;  // one workgroup per tile
  if (P.tot_sur <= 0) return 0;
  k_sur_plan<<<(unsigned)P.tot_sur, SUR_THREADS, 0, stream>>>(P, B);
  SUR_LAUNCH_CHECK();
  return 0;
}

int bldfm_launch_sur_forward(const uint8_t* net, const SurMet* mets, int nmet, const int32_t* blk_met,
                             const double2* kxy, const double2* gt, int64_t N, __bf16* xs, __bf16* zs,
                             float2* dphi, float2* phi, float* loss, int train, hipStream_t stream) {
  const unsigned nb = (unsigned)(N / SUR_PTS);
  if (train)
    k_sur_fwd<true><<<nb, SUR_THREADS, 0, stream>>>(net, mets, nmet, blk_met, kxy, gt, N, xs, zs, dphi, phi, loss);
  else
    k_sur_fwd<false><<<nb, SUR_THREADS, 0, stream>>>(net, mets, nmet, blk_met, kxy, nullptr, N, nullptr,
                                                     nullptr, nullptr, phi, nullptr);
  SUR_LAUNCH_CHECK();
  return 0;
}

int bldfm_launch_sur_backward(const uint8_t* net, int64_t N, const __bf16* xs, const __bf16* zs,
                              const float2* dphi, __bf16* dzs, float* grad, int chunk, hipStream_t stream) {
  const unsigned nb = (unsigned)(N / SUR_PTS);
  k_sur_bwd<<<nb, SUR_THREADS, 0, stream>>>(net, N, xs, zs, dphi, dzs, grad);
  SUR_LAUNCH_CHECK();
  const int nch = (int)(N / chunk);
  k_sur_wgrad<<<(unsigned)(2 * nch + (SUR_NL - 1) * 4 * nch), SUR_THREADS, 0, stream>>>(N, chunk, xs, dzs, grad);
  SUR_LAUNCH_CHECK();
  return 0;
}

int bldfm_launch_sur_adam(float* p, float* g, float* m, float* v, uint8_t* net, int update, float lr,
                          float b1, float b2, float eps, float c1, float c2, hipStream_t stream) {
  k_sur_adam<<<(SUR_NPARAM + 255) / 256, 256, 0, stream>>>(p, g, m, v, net, update, lr, b1, b2, eps, c1, c2);
  SUR_LAUNCH_CHECK();
  return 0;
}
