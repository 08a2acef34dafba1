// LDS-resident mixed-radix Stockham FFT (FP64 complex) for one workgroup.
//
// Each stage of radix R with current sub-transform length Ns reads butterfly
// inputs x[j + r*n/R] (j in [0, n/R)), applies the twiddles
// W_n^{r*k*n/(Ns*R)} with k = j mod Ns, performs a length-R DFT and writes the
// outputs to x[(j - k)*R + k + r*Ns] (self-sorting, natural order at the
// end).  The first stage gathers from the caller's source (global memory:
// band scatter, zero padding and strides folded into the index map) and the
// last stage scatters to the caller's sink (cropping, scaling, output
// format); the stages between live in an LDS buffer of n double2, in place:
// every thread first gathers its butterflies into registers, then the
// workgroup synchronises, then all threads scatter.
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

// Element i of an LDS transform buffer lives at i ^ ((i >> 3) & 7) when the
// length is a multiple of 8 (a bijection within every aligned block of 8),
// else at i.  ds_read/write_b128 serve 8 lanes (128 bytes, all 32 banks) per
// cycle; the first Stockham stages write with element strides R*Ns (8, 64,
// ...) which, unswizzled, put all 8 lanes of a group on the same 4 banks.
// Swizzled, every aligned group of 8 lanes hits distinct banks in every stage
// of the factorisations used here.  The map only moves bits 3-5 into bits
// 0-2, so phys(i + 64 m) = phys(i) + 64 m: an access pattern with a stride
// that is a multiple of 64 needs one swizzle per thread, not per element.
struct Lds {
  double2* p;
  int swz;
  __device__ __forceinline__ int phys(int i) const { return i ^ ((i >> 3) & swz); }
  __device__ __forceinline__ double2& operator[](int i) const { return p[phys(i)]; }
};
__device__ __forceinline__ Lds lds_view(double2* p, int n) { return Lds{p, (n & 7) ? 0 : 7}; }

// Source and sink of the intermediate stages: the LDS buffer.
struct LdsIn {
  Lds b;
  __device__ __forceinline__ double2 operator()(int i) const { return b[i]; }
};
struct LdsOut {
  Lds b;
  __device__ __forceinline__ void operator()(int i, double2 v) const { b[i] = v; }
};

// Twiddles W_n^t from two short tables in LDS: with t = hi*L + lo (L = 2^lb,
// L*L >= n), W^t = W^{hi*L} * W^{lo}.  Both factors are correctly rounded
// table entries, so W^t is within a few ulp; the tables (L + ceil(n/L)
// entries, 88 for n = 1536) sit behind the transform buffer and replace a
// dependent global load per stage by two LDS reads (tw_split_log2 and
// tw_table_len live with the plan types).
struct TwLds {
  const double2* lo;  // W^lo, lo < L
  const double2* hi;  // W^{hi*L}, hi < ceil(n/L)
  int lb;
  __device__ __forceinline__ double2 operator()(int t) const {
    return cmul(hi[t >> lb], lo[t & ((1 << lb) - 1)]);
  }
};
// Fill the tables at dst from the n-entry global table; the caller
// synchronises the workgroup before the first lookup.
__device__ __forceinline__ TwLds tw_tables(double2* dst, const double2* __restrict__ tw, int n, int tid, int T) {
  const int lb = tw_split_log2(n), L = 1 << lb, nhi = (n + L - 1) >> lb;
  for (int i = tid; i < L + nhi; i += T) dst[i] = i < L ? tw[i < n ? i : 0] : tw[(i - L) << lb];
  return TwLds{dst, dst + L, lb};
}

template <class A, class B> struct same_t { static constexpr bool value = false; };
template <class A> struct same_t<A, A> { static constexpr bool value = true; };

// Stage constants: radix R, Ns, twiddle index step n/(Ns*R), and the
// multiply-high divisor of Ns (unused in the first stage, where Ns = 1).
struct StageK {
  int Ns, stride;
  uint32_t magic;
};

// One radix-R stage: gather the butterfly inputs in(j + r*nb), twiddle,
// length-R DFT, scatter to out(base + r*Ns).  When the sink is the LDS buffer
// the workgroup synchronises before the scatter (the stage works in place)
// and after it (the next stage reads it).  FIRST: Ns = 1 (no twiddles,
// k = 0, q = j).
template <int R, int MAXB, bool INV, bool FIRST, class In, class Out>
__device__ __forceinline__ void stage_r(In in, Out out, int n, StageK K, TwLds tw,
                                        int tid, int T) {
  constexpr bool to_lds = same_t<Out, LdsOut>::value;
  const int nb = n / R;
  double2 x[MAXB][R];
#pragma unroll
  for (int b = 0; b < MAXB; ++b) {
    const int j = tid + b * T;
    if (j < nb) {
      if constexpr (same_t<In, LdsIn>::value) {
        if ((nb & 63) == 0) {  // phys(j + r*nb) = phys(j) + r*nb
          const int pj = in.b.phys(j);
#pragma unroll
          for (int r = 0; r < R; ++r) x[b][r] = in.b.p[pj + r * nb];
        } else {
#pragma unroll
          for (int r = 0; r < R; ++r) x[b][r] = in(j + r * nb);
        }
      } else {
#pragma unroll
        for (int r = 0; r < R; ++r) x[b][r] = in(j + r * nb);
      }
      if (!FIRST) {
        const int k = j - (int)__umulhi((unsigned)j, K.magic) * K.Ns;
        if (k != 0) {
          // W^{r k stride} for r = 1..R-1 from one table entry by repeated
          // multiplication (error grows like r*eps, not amplified downstream)
          double2 w1 = tw(k * K.stride);
          if (INV) w1.y = -w1.y;
          double2 w = w1;
#pragma unroll
          for (int r = 1; r < R; ++r) {
            x[b][r] = cmul(x[b][r], w);
            if (r + 1 < R) w = cmul(w, w1);
          }
        }
      }
      butterfly<R>(x[b], INV);
    }
  }
  if (to_lds) __syncthreads();
#pragma unroll
  for (int b = 0; b < MAXB; ++b) {
    const int j = tid + b * T;
    if (j < nb) {
      int base;
      if (FIRST) {
        base = j * R;
      } else {
        const int q = (int)__umulhi((unsigned)j, K.magic);
        base = q * K.Ns * (R - 1) + j;  // q*Ns*R + k with k = j - q*Ns
      }
      if constexpr (same_t<Out, LdsOut>::value) {
        const Lds& L = out.b;
        if (FIRST && R == 8) {  // element 8j + r: bits 3-5 are j & 7
          const int jm = j & L.swz;
#pragma unroll
          for (int r = 0; r < R; ++r) L.p[base + (r ^ jm)] = x[b][r];
        } else if (!FIRST && (K.Ns & 63) == 0) {
          const int pb = L.phys(base);
#pragma unroll
          for (int r = 0; r < R; ++r) L.p[pb + r * K.Ns] = x[b][r];
        } else {
#pragma unroll
          for (int r = 0; r < R; ++r) L[base + r * (FIRST ? 1 : K.Ns)] = x[b][r];
        }
      } else {
#pragma unroll
        for (int r = 0; r < R; ++r) out(base + r * (FIRST ? 1 : K.Ns), x[b][r]);
      }
    }
  }
  if (to_lds) __syncthreads();
}

// Any radix: one output per (thread, slot), out[o] = sum_m x_m W_n^{m*stride*(k + r*Ns)}.
template <int MAXE, bool INV, bool FIRST, class In, class Out>
__device__ __forceinline__ void stage_generic(In in, Out out, int n, int R, StageK K,
                                              TwLds tw, int tid, int T) {
  constexpr bool to_lds = same_t<Out, LdsOut>::value;
  const int nb = n / R;
  const int Ns = FIRST ? 1 : K.Ns;
  const uint64_t mR = div_magic(R);
  double2 y[MAXE];
#pragma unroll
  for (int e = 0; e < MAXE; ++e) {
    const int o = tid + e * T;
    if (o < n) {
      const int q = FIRST ? o : (int)__umulhi((unsigned)o, K.magic);  // o / Ns
      const int k = o - q * Ns;        // o mod Ns
      const int g = fdiv(q, mR);       // o / (Ns R)
      const int r = q - g * R;         // (o / Ns) mod R
      const int j = g * Ns + k;
      const int step = K.stride * (k + r * Ns);
      double2 acc = cmk(0.0, 0.0);
      int t = 0;
      for (int m = 0; m < R; ++m) {
        double2 w = tw(t);
        if (INV) w.y = -w.y;
        acc = cadd(acc, cmul(in(j + m * nb), w));
        t += step;
        if (t >= n) t -= n;
      }
      y[e] = acc;
    }
  }
  if (to_lds) __syncthreads();
#pragma unroll
  for (int e = 0; e < MAXE; ++e) {
    const int o = tid + e * T;
    if (o < n) out(o, y[e]);
  }
  if (to_lds) __syncthreads();
}

template <int MAXE, bool INV, bool FIRST, class In, class Out>
__device__ __forceinline__ void run_stage(int R, In in, Out out, int n, StageK K, TwLds tw,
                                          int tid, int T) {
  switch (R) {
    case 2: stage_r<2, (MAXE + 1) / 2, INV, FIRST>(in, out, n, K, tw, tid, T); break;
    case 3: stage_r<3, (MAXE + 2) / 3, INV, FIRST>(in, out, n, K, tw, tid, T); break;
    case 4: stage_r<4, (MAXE + 3) / 4, INV, FIRST>(in, out, n, K, tw, tid, T); break;
    case 5: stage_r<5, (MAXE + 4) / 5, INV, FIRST>(in, out, n, K, tw, tid, T); break;
    case 8: stage_r<8, (MAXE + 7) / 8, INV, FIRST>(in, out, n, K, tw, tid, T); break;
    default: stage_generic<MAXE, INV, FIRST>(in, out, n, R, K, tw, tid, T); break;
  }
}

// Full transform of the n values in(0..n-1) into out(0..n-1), with the plan
// at Pg (device memory, read through the constant address space: the plan is
// uniform) and its n-entry twiddle table tw (global); INV selects e^{+i}.
// The first stage gathers straight from `in` and the last scatters straight
// to `out`; buf holds the stages between (n double2) followed by the two
// twiddle tables (tw_table_len(n) double2).  A thread handles at most MAXE
// values per stage: the launch uses T >= ceil(n/MAXE).  `reuse`: this
// workgroup transformed on buf before (its last stage may still be reading
// the tables, so the workgroup synchronises before refilling them); the
// transform buffer itself needs no such care, since every stage that writes
// buf synchronises the workgroup before its scatter.
template <int MAXE, bool INV, class In, class Out>
__device__ __forceinline__ void lds_fft(Lds buf, const FftPlan* Pg, const double2* __restrict__ twg, int tid0,
                                        int T0, In in, Out out, bool reuse) {
  typedef const __attribute__((address_space(4))) FftPlan cplan;
  const cplan& P = *(const cplan*)Pg;
  const int n = P.n;
  const int nstage = P.nstage;
  if (nstage == 0) {  // n == 1
    if (tid0 == 0) out(0, in(0));
    return;
  }
  if (reuse) __syncthreads();
  const TwLds tw = tw_tables(buf.p + n, twg, n, tid0, T0);
  {  // the radix-2/3/4/5/8 first stage reads no twiddle before its barrier
    const int R0 = P.radix[0];
    if (R0 != 2 && R0 != 3 && R0 != 4 && R0 != 5 && R0 != 8) __syncthreads();
  }
  const LdsIn li{buf};
  const LdsOut lo{buf};
  // Each stage takes opaque copies of tid and T: without them the compiler
  // hoists every stage variant's index arithmetic out of the stage loop and
  // keeps it live across all of them (doubling the VGPR count).
  const StageK K0{1, P.stride[0], 0u};
  {
    int tid = tid0, T = T0;
    asm volatile("" : "+v"(tid), "+s"(T));
    if (nstage == 1) {
      run_stage<MAXE, INV, true>(P.radix[0], in, out, n, K0, tw, tid, T);
      return;
    }
    run_stage<MAXE, INV, true>(P.radix[0], in, lo, n, K0, tw, tid, T);
  }
  int Ns = P.radix[0];
  for (int s = 1; s < nstage; ++s) {
    const int R = P.radix[s];
    const StageK K{Ns, P.stride[s], P.magic[s]};
    int tid = tid0, T = T0;
    asm volatile("" : "+v"(tid), "+s"(T));
    if (s == nstage - 1) run_stage<MAXE, INV, false>(R, li, out, n, K, tw, tid, T);
    else run_stage<MAXE, INV, false>(R, li, lo, n, K, tw, tid, T);
    Ns *= R;
  }
}

}  // namespace bldfm_fft
