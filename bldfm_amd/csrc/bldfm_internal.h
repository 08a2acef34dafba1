// Internal launch interface between the plan/ABI layer and the kernels.
#pragma once
#include <hip/hip_runtime.h>
#include "bldfm_types.h"
#include "bldfm_math.h"

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
  int32_t ninst;
  int32_t any_dispersion;
  int32_t variant_used[4][2]; // shooting kernels this plan launches [variant][exact]
  int32_t fft_small;       // all transforms <= BLDFM_FFT_SMALL_N: 256-thread kernels
  int32_t fft_threads;     // threads per FFT workgroup
  int32_t fft_lds_bytes;   // dynamic LDS per FFT workgroup
  int64_t tot_tab, tot_fwd_rows, tot_fwd_cols, tot_shoot, tot_nyq, tot_inv_cols, tot_inv_rows;
};

struct ExecBufs {
  double* prof;        // 6*nz doubles per instance
  const double* src;   // surface flux (dispersion instances)
  void* conc;          // outputs
  void* flx;
  double* tab;         // level tables (workspace)
  double2* ws;         // complex workspace
};

// ev (optional): 7 events recorded at the phase boundaries
// [level table | forward transform | shooting | Nyquist fix-up | y-pass | x-pass]
int bldfm_launch_all(const PlanDev& P, const ExecBufs& B, int grid_cap, hipStream_t stream,
                     hipEvent_t* ev = nullptr);
int bldfm_launch_profiles(const ProfSetup& S, double* out, hipStream_t stream);
// F0/F1: device-resident plans of the two axis lengths n0 (columns) and n1 (rows)
int bldfm_launch_fft2(const FftPlan* F0, int32_t n0, const double2* tw0, const FftPlan* F1, int32_t n1,
                      const double2* tw1, int32_t batch, const double2* in, double2* mid, double2* out,
                      int32_t sign, double scale, int threads, int lds_bytes, int grid_cap,
                      hipStream_t stream);
void bldfm_set_fft_lds_attr(int bytes);
int bldfm_launch_stability(const double* x, int64_t n, double* psi, double* phi, hipStream_t stream);
int bldfm_launch_wind(const double* spd, const double* dir, int64_t n, double* u, double* v, hipStream_t stream);
int bldfm_launch_math_probe(int op, const double* a, const double* b, int64_t n, double* out, hipStream_t stream);
