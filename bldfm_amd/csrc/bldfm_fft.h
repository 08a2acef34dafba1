// LDS-resident mixed-radix Stockham FFT (FP64 complex) for one workgroup.
//
// The transform of length n lives in an LDS buffer of n double2.  Each stage
// of radix R with current sub-transform length Ns reads butterfly inputs
// buf[j + r*n/R] (j in [0, n/R)), applies the twiddles W_n^{r*k*n/(Ns*R)}
// with k = j mod Ns, performs a length-R DFT and writes the outputs to
// buf[(j - k)*R + k + r*Ns] (self-sorting, natural order at the end).  The
// stage is done in place: every thread first gathers its butterflies into
// registers, then the workgroup synchronises, then all threads scatter.
//
// W_n^t = e^{-2 pi i t/n} comes from a precomputed table; the e^{+i} transform
// conjugates it.  A thread handles at most MAXE complex values per stage, so
// the launch must use T >= ceil(n / MAXE) threads.  Two configurations are
// compiled: up to 256 threads with MAXE = 8 (n <= 2048, no register spills),
// and up to 1024 threads with MAXE = BLDFM_FFT_MAXE for longer transforms.
#pragma once
#include "bldfm_math.h"



namespace bldfm_fft {

// sin/cos constants of the small butterflies (exact values to 20 digits)
#define C_SQRT1_2 0.70710678118654752440
#define C_SIN_60 0.86602540378443864676
#define C_COS_72 0.30901699437494742410
#define C_SIN_72 0.95105651629515357212
#define C_COS_144 -0.80901699437494742410
#define C_SIN_144 0.58778525229247312917

// multiply by s*i, s = +1 (inv) or -1
__device__ __forceinline__ double2 mul_si(double2 a, bool inv) {
  return inv ? cmk(-a.y, a.x) : cmk(a.y, -a.x);
}

template <int R>
__device__ __forceinline__ void butterfly(double2* x, bool inv);

template <>
__device__ __forceinline__ void butterfly<2>(double2* x, bool) {
  const double2 a = x[0], b = x[1];
  x[0] = cadd(a, b);
  x[1] = csub(a, b);
}

template <>
__device__ __forceinline__ void butterfly<3>(double2* x, bool inv) {
  const double2 t1 = cadd(x[1], x[2]), t2 = csub(x[1], x[2]);
  const double2 m = csub(x[0], cscale(0.5, t1));
  const double2 r = mul_si(cscale(C_SIN_60, t2), inv);
  x[0] = cadd(x[0], t1);
  x[1] = cadd(m, r);
  x[2] = csub(m, r);
}

template <>
__device__ __forceinline__ void butterfly<4>(double2* x, bool inv) {
  const double2 a = cadd(x[0], x[2]), b = csub(x[0], x[2]);
  const double2 c = cadd(x[1], x[3]), d = mul_si(csub(x[1], x[3]), inv);
  x[0] = cadd(a, c);
  x[2] = csub(a, c);
  x[1] = cadd(b, d);
  x[3] = csub(b, d);
}

template <>
__device__ __forceinline__ void butterfly<5>(double2* x, bool inv) {
  const double2 t1 = cadd(x[1], x[4]), t2 = cadd(x[2], x[3]);
  const double2 t3 = csub(x[1], x[4]), t4 = csub(x[2], x[3]);
  const double2 a1 = cadd(x[0], cadd(cscale(C_COS_72, t1), cscale(C_COS_144, t2)));
  const double2 a2 = cadd(x[0], cadd(cscale(C_COS_144, t1), cscale(C_COS_72, t2)));
  const double2 b1 = mul_si(cadd(cscale(C_SIN_72, t3), cscale(C_SIN_144, t4)), inv);
  const double2 b2 = mul_si(csub(cscale(C_SIN_144, t3), cscale(C_SIN_72, t4)), inv);
  x[0] = cadd(x[0], cadd(t1, t2));
  x[1] = cadd(a1, b1);
  x[4] = csub(a1, b1);
  x[2] = cadd(a2, b2);
  x[3] = csub(a2, b2);
}

template <>
__device__ __forceinline__ void butterfly<8>(double2* x, bool inv) {
  // even/odd split into two length-4 DFTs, then combine with W_8^r
  double2 e[4] = {x[0], x[2], x[4], x[6]};
  double2 o[4] = {x[1], x[3], x[5], x[7]};
  butterfly<4>(e, inv);
  butterfly<4>(o, inv);
  const double s = inv ? 1.0 : -1.0;
  // W_8^1 = sqrt(1/2) (1 + s i), W_8^2 = s i, W_8^3 = sqrt(1/2) (-1 + s i)
  const double2 o1 = cscale(C_SQRT1_2, cmk(o[1].x - s * o[1].y, o[1].y + s * o[1].x));
  const double2 o2 = mul_si(o[2], inv);
  const double2 o3 = cscale(C_SQRT1_2, cmk(-o[3].x - s * o[3].y, -o[3].y + s * o[3].x));
  x[0] = cadd(e[0], o[0]);
  x[4] = csub(e[0], o[0]);
  x[1] = cadd(e[1], o1);
  x[5] = csub(e[1], o1);
  x[2] = cadd(e[2], o2);
  x[6] = csub(e[2], o2);
  x[3] = cadd(e[3], o3);
  x[7] = csub(e[3], o3);
}

// j / d for 0 <= j, d < 2^16 by multiply-high: with m = floor(2^32/d) + 1 the
// error of j*m/2^32 against j/d is below 2^-16 < 1/d, so the floor is exact.
__device__ __forceinline__ uint64_t div_magic(int d) { return (1ull << 32) / (uint64_t)d + 1ull; }
__device__ __forceinline__ int fdiv(int j, uint64_t m) { return (int)(((uint64_t)(uint32_t)j * m) >> 32); }

template <int R, int MAXB>
__device__ __forceinline__ void stage_r(double2* buf, int n, int Ns, const double2* __restrict__ tw,
                                        bool inv, int tid, int T) {
  const int nb = n / R;
  const int stride = n / (Ns * R);
  const uint64_t mNs = div_magic(Ns);
  double2 x[MAXB][R];
#pragma unroll
  for (int b = 0; b < MAXB; ++b) {
    const int j = tid + b * T;
    if (j < nb) {
      const int k = j - fdiv(j, mNs) * Ns;
#pragma unroll
      for (int r = 0; r < R; ++r) x[b][r] = buf[j + r * nb];
      if (k != 0) {
        // W^{r k stride} for r = 1..R-1 from one table entry by repeated
        // multiplication (error grows like r*eps, not amplified downstream)
        double2 w1 = tw[k * stride];
        if (inv) w1.y = -w1.y;
        double2 w = w1;
#pragma unroll
        for (int r = 1; r < R; ++r) {
          x[b][r] = cmul(x[b][r], w);
          if (r + 1 < R) w = cmul(w, w1);
        }
      }
      butterfly<R>(x[b], inv);
    }
  }
  __syncthreads();
#pragma unroll
  for (int b = 0; b < MAXB; ++b) {
    const int j = tid + b * T;
    if (j < nb) {
      const int q = fdiv(j, mNs);
      const int k = j - q * Ns;
      const int base = q * Ns * R + k;
#pragma unroll
      for (int r = 0; r < R; ++r) buf[base + r * Ns] = x[b][r];
    }
  }
  __syncthreads();
}

// Any radix: one output per (thread, slot), out[o] = sum_m x_m W_n^{m*stride*(k + r*Ns)}.
template <int MAXE>
__device__ __forceinline__ void stage_generic(double2* buf, int n, int R, int Ns,
                                              const double2* __restrict__ tw, bool inv, int tid,
                                              int T) {
  const int nb = n / R;
  const int stride = n / (Ns * R);
  const uint64_t mNs = div_magic(Ns), mR = div_magic(R);
  double2 y[MAXE];
#pragma unroll
  for (int e = 0; e < MAXE; ++e) {
    const int o = tid + e * T;
    if (o < n) {
      const int q = fdiv(o, mNs);      // o / Ns
      const int k = o - q * Ns;        // o mod Ns
      const int g = fdiv(q, mR);       // o / (Ns R)
      const int r = q - g * R;         // (o / Ns) mod R
      const int j = g * Ns + k;
      const int step = stride * (k + r * Ns);
      double2 acc = cmk(0.0, 0.0);
      int t = 0;
      for (int m = 0; m < R; ++m) {
        double2 w = tw[t];
        if (inv) w.y = -w.y;
        acc = cadd(acc, cmul(buf[j + m * nb], w));
        t += step;
        if (t >= n) t -= n;
      }
      y[e] = acc;
    }
  }
  __syncthreads();
#pragma unroll
  for (int e = 0; e < MAXE; ++e) {
    const int o = tid + e * T;
    if (o < n) buf[o] = y[e];
  }
  __syncthreads();
}

// Full transform on buf (n points) with the plan at Pg (device memory,
// read through the constant address space: the plan is uniform);
// inv selects e^{+i}.  A thread handles at most MAXE values per stage: the
// launch uses T >= ceil(n/MAXE).
template <int MAXE>
__device__ __forceinline__ void lds_fft(double2* buf, const FftPlan* Pg, const double2* __restrict__ tw,
                                        bool inv, int tid, int T) {
  typedef const __attribute__((address_space(4))) int32_t cint32;
  const cint32* pc = (const cint32*)Pg;
  const int n = pc[0];
  const int nstage = pc[1];
  int Ns = 1;
  for (int s = 0; s < nstage; ++s) {
    const int R = pc[2 + s];
    switch (R) {
      case 2: stage_r<2, (MAXE + 1) / 2>(buf, n, Ns, tw, inv, tid, T); break;
      case 3: stage_r<3, (MAXE + 2) / 3>(buf, n, Ns, tw, inv, tid, T); break;
      case 4: stage_r<4, (MAXE + 3) / 4>(buf, n, Ns, tw, inv, tid, T); break;
      case 5: stage_r<5, (MAXE + 4) / 5>(buf, n, Ns, tw, inv, tid, T); break;
      case 8: stage_r<8, (MAXE + 7) / 8>(buf, n, Ns, tw, inv, tid, T); break;
      default: stage_generic<MAXE>(buf, n, R, Ns, tw, inv, tid, T); break;
    }
    Ns *= R;
  }
}

}  // namespace bldfm_fft
