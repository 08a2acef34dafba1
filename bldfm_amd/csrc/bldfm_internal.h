// Internal launch interface between the plan/ABI layer and the kernels.
#pragma once
#include <hip/hip_runtime.h>
#include "bldfm_types.h"
#include "bldfm_math.h"
#include "bldfm_sur.h"

// Device-side metadata of one plan (all pointers are device pointers owned
// by the plan).
struct PlanDev {
  const InstDesc* desc;       // ninst descriptors
  const ProfSetup* met;       // ninst met setups (prof_source == 1), else null
  const int32_t* met_flag;    // ninst: 1 = profiles from met, 0 = caller arrays
  const int32_t* ints;        // slot maps and level indices
  const FftPlan* plans;       // FFT plans
  const double2* tw;          // twiddle tables
  // per-kernel tile records {instance, local tile index}
  const TileRec* t_tab;
  const TileRec* t_fwd_rows;
  const TileRec* t_fwd_cols;
  const TileRec* t_shoot;
  const TileRec* t_nyq;
  const TileRec* t_inv_cols;
  const TileRec* t_inv_rows;
  // four-step lines of split lengths: [2*pass + step] for the PassId order
  // (forward rows, forward columns, inverse y-pass, inverse x-pass), step A/B
  const TileRec* t_long[8];
  int64_t tot_long[8];
  int32_t ninst;
  int32_t any_dispersion;
  int32_t variant_used[4][2]; // shooting kernels this plan launches [variant][exact]
  int32_t fft_small;       // all transforms <= BLDFM_FFT_SMALL_N: 256-thread kernels
  int32_t fft_threads;     // threads per FFT workgroup
  int32_t fft_lds_bytes;   // dynamic LDS per FFT workgroup
  int32_t shoot_block;     // modes per workgroup of the vertical solve (64, 128, 256)
  int32_t xlayout;         // x-pass input: 0 strided gather, 1 y-pass writes transposed,
                           // 2 tiled transpose kernel between the passes
  const TileRec* t_tr;     // 32x32 tiles of the transpose (xlayout 2)
  int64_t tot_tr;
  int64_t tot_tab, tot_fwd_rows, tot_fwd_cols, tot_shoot, tot_nyq, tot_inv_cols, tot_inv_rows;
  // learned surrogate of footprint instances (unused while sur_net is null)
  const TileRec* t_sur;    // SUR_PTS-lane tiles of surrogate-flagged instances
  const TileRec* t_spot;   // 256-sample tiles of the spot-check error gate
  const SurMet* sur_met;   // ninst network inputs
  const int32_t* sur_ok;   // ninst: 1 = flagged and inside the network's training hull
  const uint8_t* sur_net;  // network buffer (SUR_NET_BYTES)
  double sur_tol;          // largest accepted spot-check relative L2 error
  int32_t sur_stride;      // one spot-checked lane in sur_stride
  int64_t tot_sur, tot_spot;
};

struct ExecBufs {
  double* prof;        // 6*nz doubles per instance
  const double* src;   // surface flux (dispersion instances)
  void* conc;          // outputs
  void* flx;
  double* tab;         // level tables (workspace)
  double2* ws;         // complex workspace
  double2* gate;       // ninst spot-check sums (|surrogate - solver|^2, |solver|^2)
  int32_t* mode;       // ninst: 0 surrogate accepted, 1 FP64 solver; null: every instance solved
};

// ev (optional): BLDFM_PHASES + 1 events recorded at the phase boundaries
// [level table | forward transform | surrogate and gate | shooting | Nyquist fix-up | y-pass | x-pass]
#define BLDFM_PHASES 7
int bldfm_launch_all(const PlanDev& P, const ExecBufs& B, int grid_cap, hipStream_t stream,
                     hipEvent_t* ev = nullptr);
// Surrogate inference over the lanes of accepted instances (packed spectrum).
int bldfm_launch_sur_plan(const PlanDev& P, const ExecBufs& B, int grid_cap, hipStream_t stream);
// Surrogate training and evaluation on point lists (N a multiple of SUR_PTS).
int bldfm_launch_sur_forward(const uint8_t* net, const SurMet* mets, int nmet, const int32_t* blk_met,
                             const double2* kxy, const double2* gt, int64_t N, __bf16* xs, __bf16* zs,
                             float2* dphi, float2* phi, float* loss, int train, hipStream_t stream);
int bldfm_launch_sur_backward(const uint8_t* net, int64_t N, const __bf16* xs, const __bf16* zs,
                              const float2* dphi, __bf16* dzs, float* grad, int chunk, hipStream_t stream);
int bldfm_launch_sur_adam(float* p, float* g, float* m, float* v, uint8_t* net, int update, float lr,
                          float b1, float b2, float eps, float c1, float c2, hipStream_t stream);
// FP64 point solver: level tables from met setups, then G at the measurement level.
int bldfm_launch_tables(const ProfSetup* S, int nmet, int rows, double* tab, hipStream_t stream);
int bldfm_launch_transfer(const double* tab, int rows, const int32_t* nzs, const int32_t* lvls, int nmet,
                          const int32_t* blk_met, int nblk, const double2* kxy, double2* g, int exact,
                          hipStream_t stream);
int bldfm_launch_profiles(const ProfSetup& S, double* out, hipStream_t stream);
// Kormann & Meixner footprint grids: per footprint the cell geometry and the
// scalars of the closed form, f = scale x^e1 exp(-Xi/x - (gA y x^e2)^2 / 2).
struct Km01Dev {
  double xmin, ymax, res, mx, my, wd_rad;
  double scale, e1, e2, Xi, gA;
  int64_t off;
  int32_t nx, ny, has_wd, valid;
};
int bldfm_launch_km01(const Km01Dev* P, int n, int64_t max_cells, double* out, hipStream_t stream);
// Device-resident plan of one axis of the plain 2-D FFT: F (n entries of
// twiddles at tw) and, for a split length (F1 non-null), the plans of its
// factors n1, n2 with their twiddles.
struct FftSplit {
  const FftPlan* F;
  const double2* tw;
  const FftPlan* F1;
  const double2* tw1;
  const FftPlan* F2;
  const double2* tw2;
  int32_t n1, n2;
};
// A0/A1: the axis lengths n0 (columns) and n1 (rows); scr: batch*n0*n1
// values of four-step scratch when an axis is split (else unused)
int bldfm_launch_fft2(const FftSplit& A0, int32_t n0, const FftSplit& A1, int32_t n1, int32_t batch,
                      const double2* in, double2* mid, double2* scr, double2* out, int32_t sign, double scale,
                      int threads, int lds_bytes, int grid_cap, hipStream_t stream);
void bldfm_set_fft_lds_attr(int bytes);
int bldfm_launch_stability(const double* x, int64_t n, double* psi, double* phi, hipStream_t stream);
int bldfm_launch_wind(const double* spd, const double* dir, int64_t n, double* u, double* v, hipStream_t stream);
int bldfm_launch_math_probe(int op, const double* a, const double* b, int64_t n, double* out, hipStream_t stream);
