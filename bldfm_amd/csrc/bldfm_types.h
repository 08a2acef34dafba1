// Shared host/device descriptors and index maps for the BLDFM MI355X solver.
//
// Spectral layout conventions (all integer index maps live here so the CPU
// bounds-check harness exercises exactly the code the kernels run):
//
//  * An axis with nl retained Fourier modes and ne padded grid points carries
//    integer frequencies f in numpy fftfreq order: truncated index j holds
//    f = j for j < ceil(nl/2), else j - nl  (src/bldfm/solver.py:148-149).
//    Frequency f sits at padded index f mod ne (solver.py:139-145, 265-278).
//  * The packed spectrum Z that feeds the inverse transform lives on a
//    compact band of nb = nl + 1 points when nl is even and nl < ne (the
//    extra point holds frequency +nl/2, the mirror of the unpaired Nyquist
//    frequency -nl/2), else nb = nl.  Compact index c = f mod nb, so the
//    mirror of c is (nb - c) mod nb.
//  * Z = Hp + i Hq, where H = (X(k) + conj X(-k))/2 is the Hermitian part of
//    the truncated spectrum X over the padded grid.  Re(DFT(X)) = DFT(H), so
//    one complex transform of Z yields conc (real part) and flux (imag part).
#pragma once
#include <stdint.h>

#if defined(__HIPCC__)
#define BLDFM_HD __host__ __device__ inline
#else
#define BLDFM_HD inline
#endif

#define BLDFM_MAX_RADIX_STAGES 16
#define BLDFM_LEVEL_STRIDE 16  // doubles per vertical step in the level table
#define BLDFM_FFT_MAXE 10      // values per thread per FFT stage, 1024-thread kernels (n <= 10033: 160 KB of LDS)
#define BLDFM_FFT_SMALL_N 2048 // longest transform of the 256-thread kernels (8 values per thread)

// Mixed-radix Stockham plan for one transform length.
// Stage s has radix radix[s] and current sub-transform length
// Ns = radix[0]*...*radix[s-1]; stride[s] = n/(Ns*radix[s]) is its twiddle
// index step and magic[s] = floor(2^32/Ns) + 1 divides by Ns with one
// multiply-high (j/Ns = (j*magic) >> 32 for j < 2^16, Ns >= 2; 0 at s = 0).
//
// A length too long for one workgroup's LDS is split, n = n1 * n2 (the
// four-step method): n2 transforms of length n1 over the strided inputs
// x[j1 n2 + j2], the twiddle W_n^{j2 k1}, then n1 transforms of length n2
// giving X[k1 + n1 k2].  Such a plan has nstage = 0, n1 > 0 and the plan ids
// (or, for the plain 2-D FFT, unused) of its two factor lengths.
struct FftPlan {
  int32_t n;
  int32_t nstage;
  int32_t radix[BLDFM_MAX_RADIX_STAGES];
  int32_t stride[BLDFM_MAX_RADIX_STAGES];
  uint32_t magic[BLDFM_MAX_RADIX_STAGES];
  int64_t tw_off;  // offset (in double2) of the n twiddles e^{-2 pi i t/n}
  int32_t n1, n2;        // four-step factors (n1 = 0: one LDS transform)
  int32_t plan1, plan2;  // plan ids of the factor lengths
};

// Factor n1 of a four-step split of a length n > limit: the smallest divisor
// of n at or above sqrt(n) with n1 <= limit (then n2 = n/n1 <= sqrt(n) <=
// limit); 0 when n fits one workgroup, -1 when no such split exists (n has
// no divisor in [n/limit, limit], e.g. a prime).
BLDFM_HD int fft_split(int n, int limit) {
  if (n <= limit) return 0;
  int r = 1;
  while ((int64_t)r * r < n) ++r;
  for (int d = r; d <= limit; ++d)
    if (n % d == 0 && n / d <= limit) return d;
  return -1;
}

// Twiddle tables of an LDS transform of length n: W^lo (lo < L = 2^lb,
// L*L >= n) and W^{hi*L} (hi < ceil(n/L)), behind the n-point buffer.
BLDFM_HD int tw_split_log2(int n) {
  int lb = 0;
  while ((1 << (2 * lb)) < n) ++lb;
  return lb;
}
BLDFM_HD int tw_table_len(int n) {
  const int lb = tw_split_log2(n);
  return (1 << lb) + ((n + (1 << lb) - 1) >> lb);
}
// Dynamic LDS of one transform: buffer and twiddle tables.
BLDFM_HD int fft_lds_bytes(int n) { return (n + tw_table_len(n)) * 16; }

// One tile of a kernel's work: the instance and the tile's index inside it.
struct TileRec {
  int32_t inst;
  int32_t local;
};

// Mode kinds produced by the lane decoder.
enum LaneKind : int32_t {
  LANE_IDLE = 0,
  LANE_HALF = 1,    // paired mode in the upper half plane; writes k and -k
  LANE_ORIGIN = 2,  // the degenerate (0,0) mode
  LANE_NYQ = 3,     // unpaired mode on a Nyquist row/column
};

// Per-instance descriptor.  All offsets are element offsets into the
// buffers named in the comments.
struct InstDesc {
  // geometry
  int32_t nx, ny;      // output grid (nslot, ny, nx); srf_flx is (ny, nx) (the same except for
                       // a footprint whose padded length and mode count differ in parity)
  int32_t nlx, nly;    // retained modes (after clamping)
  int32_t nxe, nye;    // padded grid
  int32_t px, py;      // halo pad widths
  int32_t nxb, nyb;    // compact band lengths of Z
  int32_t nz;          // vertical points
  int32_t nslot;       // output levels
  int32_t nfilled;     // slots filled by the numerical branch
  int32_t max_level;   // highest filled level index (numerical branch)
  int32_t footprint, analytic, single, has_shift, out_f32;
  int32_t plan_x, plan_y;  // plan ids for nxe and nye
  int32_t lanes;           // shooting lanes for this instance
  int32_t variant;         // shooting kernel: 0 one register snapshot, 1 up to four,
                           // 2 two-pass march, 3 analytic
  int32_t exact;           // 1: reference rounding (no fused multiply-add) in the march
  int32_t snap_level[4];   // level of output slot s < 4 of the numerical branch, or -1
  double dx, dy;
  double sx, sy;        // wavenumber scales ((2 pi)/dx)/nxe, ((2 pi)/dy)/nye
  double vx, vy;        // fftfreq value factors 1/(nl*(1/nl))
  double shx, shy;      // phase-shift distances
  double p000;          // surface background concentration
  double q00;           // footprint-mode delta spectrum (1/nxe)(1/nye), rounded as numpy forms it
  // offsets
  int64_t off_prof;     // prof buffer: z,u,v,Kx,Ky,Kz (6*nz doubles)
  int64_t off_tab;      // level table: (nz-1)*BLDFM_LEVEL_STRIDE doubles
  int64_t off_slot;     // int buffer: level->slot map (nz) then level index per slot (nslot)
  int64_t off_src;      // source buffer: srf_flx (ny*nx doubles), dispersion only
  int64_t off_aw;       // workspace double2: forward pass-A rows (ny*nlx)
  int64_t off_q0;       // workspace double2: q^0 band, x-major (nlx*nly)
  int64_t off_z;        // workspace double2: packed Z, (nslot, nxb, nyb)
  int64_t off_nyq;      // workspace double2: Nyquist modes, (nslot, nlx+nly, 2)
  int64_t off_v;        // workspace double2: after y-pass, (nslot, nxb, ny), or (nslot, ny, nxb)
                        // when the plan's x-pass layout is 1
  int64_t off_vt;       // workspace double2: (nslot, ny, nxb) transpose of v (x-pass layout 2)
  int64_t off_out;      // output element offset: (nslot, ny, nx)
  int64_t off_long;     // workspace double2: four-step scratch (lines x n of the longest split pass)
  int32_t long_x, long_y;  // 1: the padded x / y length is split (four-step passes)
};

// ---------------------------------------------------------------- frequencies

// Frequency held by truncated index j of an axis with nl modes (fftfreq order).
BLDFM_HD int freq_of_trunc(int j, int nl) { return j < (nl + 1) / 2 ? j : j - nl; }

// Truncated index of frequency f (inverse of freq_of_trunc).
BLDFM_HD int trunc_of_freq(int f, int nl) { return f >= 0 ? f : f + nl; }

// Padded-grid index of frequency f.
BLDFM_HD int padded_of_freq(int f, int ne) { return f >= 0 ? f : f + ne; }

// Compact band length.
BLDFM_HD int band_len(int nl, int ne) { return (nl % 2 == 0 && nl < ne) ? nl + 1 : nl; }

// Compact index of frequency f (f within [-nl/2, nl/2]).
BLDFM_HD int compact_of_freq(int f, int nb) { return f >= 0 ? f : f + nb; }

// Frequency of compact index c.
BLDFM_HD int freq_of_compact(int c, int nb) { return c < (nb + 1) / 2 ? c : c - nb; }

// Mirror (negated frequency) of compact index c.
BLDFM_HD int compact_mirror(int c, int nb) { return c == 0 ? 0 : nb - c; }

// Largest |f| that has its negation among the retained frequencies.
BLDFM_HD int paired_max(int nl) { return (nl - 1) / 2; }

// Number of unpaired (Nyquist) modes: row f_y=-nly/2 plus column f_x=-nlx/2.
BLDFM_HD int nyq_count(int nlx, int nly) {
  int n = 0;
  if (nly % 2 == 0) n += nlx;
  if (nlx % 2 == 0) n += 2 * paired_max(nly) + 1;
  return n;
}

// Shooting lanes: upper half plane + positive x axis + origin + Nyquist.
BLDFM_HD int shoot_lanes(int nlx, int nly) {
  const int Px = paired_max(nlx), Py = paired_max(nly);
  return Py * (2 * Px + 1) + Px + 1 + nyq_count(nlx, nly);
}

// Decode a shooting lane into (kind, fx, fy, nyquist index).  Lanes are
// ordered with fy fastest inside the half plane so that consecutive lanes
// write consecutive addresses of the x-major packed spectrum.
BLDFM_HD int decode_lane(int lane, int nlx, int nly, int* fx, int* fy, int* nidx) {
  const int Px = paired_max(nlx), Py = paired_max(nly);
  const int nA = Py * (2 * Px + 1);
  *nidx = -1;
  if (lane < 0) return LANE_IDLE;
  if (lane < nA) {
    *fx = lane / Py - Px;
    *fy = lane % Py + 1;
    return LANE_HALF;
  }
  lane -= nA;
  if (lane < Px) {
    *fx = lane + 1;
    *fy = 0;
    return LANE_HALF;
  }
  lane -= Px;
  if (lane == 0) {
    *fx = 0;
    *fy = 0;
    return LANE_ORIGIN;
  }
  lane -= 1;
  if (nly % 2 == 0) {
    if (lane < nlx) {  // Nyquist row f_y = -nly/2, every f_x
      *fx = freq_of_trunc(lane, nlx);
      *fy = -nly / 2;
      *nidx = lane;
      return LANE_NYQ;
    }
    lane -= nlx;
  }
  if (nlx % 2 == 0) {
    if (lane < 2 * Py + 1) {  // Nyquist column f_x = -nlx/2, paired f_y
      *fx = -nlx / 2;
      *fy = lane - Py;
      *nidx = (nly % 2 == 0 ? nlx : 0) + lane;
      return LANE_NYQ;
    }
  }
  return LANE_IDLE;
}

// Nyquist index of the unpaired mode at compact position (cx, cy), or -1.
BLDFM_HD int nyq_index_at(int cx, int cy, int nlx, int nly, int nxb, int nyb) {
  const int fx = freq_of_compact(cx, nxb), fy = freq_of_compact(cy, nyb);
  const int Py = paired_max(nly);
  if (nly % 2 == 0 && fy == -nly / 2) {
    if (fx >= -(nlx / 2) && fx <= (nlx - 1) / 2 && !(nlx % 2 == 0 && nxb > nlx && fx == nlx / 2))
      return trunc_of_freq(fx, nlx);
    return -1;
  }
  if (nlx % 2 == 0 && fx == -nlx / 2) {
    if (fy >= -Py && fy <= Py) return (nly % 2 == 0 ? nlx : 0) + fy + Py;
  }
  return -1;
}

// Positions of Z filled by the Nyquist fix-up: the rows with |f_y| = nly/2
// and the columns with |f_x| = nlx/2 (only when that axis has a Nyquist
// frequency).  Returns the count and decodes position p (row-positions
// first, then column-positions not already on such a row).
BLDFM_HD int nyq_fix_rows(int nly, int nyb) {
  if (nly % 2) return 0;
  return nyb > nly ? 2 : 1;  // compact rows of -nly/2 and (if distinct) +nly/2
}
BLDFM_HD int nyq_fix_cols(int nlx, int nxb) {
  if (nlx % 2) return 0;
  return nxb > nlx ? 2 : 1;
}
BLDFM_HD int nyq_fix_count(int nlx, int nly, int nxb, int nyb) {
  const int nr = nyq_fix_rows(nly, nyb), nc = nyq_fix_cols(nlx, nxb);
  return nr * nxb + nc * (nyb - nr);
}
// Row/column compact indices of the fix-up lines.
BLDFM_HD int nyq_row_c(int r, int nly, int nyb) {
  return r == 0 ? compact_of_freq(-nly / 2, nyb) : compact_of_freq(nly / 2, nyb);
}
BLDFM_HD int nyq_col_c(int r, int nlx, int nxb) {
  return r == 0 ? compact_of_freq(-nlx / 2, nxb) : compact_of_freq(nlx / 2, nxb);
}
BLDFM_HD void nyq_fix_decode(int p, int nlx, int nly, int nxb, int nyb, int* cx, int* cy) {
  const int nr = nyq_fix_rows(nly, nyb);
  if (p < nr * nxb) {
    *cy = nyq_row_c(p / nxb, nly, nyb);
    *cx = p % nxb;
    return;
  }
  p -= nr * nxb;
  const int per = nyb - nr;  // column positions not on a fix-up row
  const int col = p / per;
  // k-th compact row that is not a fix-up row: skip the (sorted) fix-up rows
  int c = p % per;
  if (nr > 0) {
    int lo = nyq_row_c(0, nly, nyb), hi = nr > 1 ? nyq_row_c(1, nly, nyb) : nyb + 1;
    if (lo > hi) { const int t = lo; lo = hi; hi = t; }
    if (c >= lo) ++c;
    if (c >= hi) ++c;
  }
  *cx = nyq_col_c(col, nlx, nxb);
  *cy = c;
}

// Padded index of compact index c on an axis with compact length nb.
BLDFM_HD int padded_of_compact(int c, int nb, int ne) { return padded_of_freq(freq_of_compact(c, nb), ne); }
