/*
 * bldfm_amd.h -- C ABI of the MI355X-native BLDFM steady-state solver.
 *
 * The reference (SchlutowSM2Group/BLDFM) is pure Python; its hot path is
 *   steady_state_transport_solver   src/bldfm/solver.py:16-304
 *   ivp_solver (numba JIT)          src/bldfm/solver.py:307-374
 *   vertical_profiles / psi / phi   src/bldfm/pbl_model.py:8-250
 *   FFT backend (pyFFTW -> FFTW3)   src/bldfm/fft_manager.py:78-157
 * The Python drop-ins in bldfm_amd/ keep those signatures and call the
 * functions below through ctypes.  No torch types cross this boundary:
 * every buffer is a plain device pointer (the Python layer allocates them
 * as torch tensors) and every stream is a hipStream_t.
 *
 * Conventions: functions return 0 on success and a negative code on error;
 * bldfm_last_error() returns the calling thread's last message.  Nothing
 * here frees or retains caller memory.  A plan may be executed on any
 * stream of the device it was created on; executions of one plan that
 * overlap in time must use distinct workspaces.
 */
#ifndef BLDFM_AMD_H
#define BLDFM_AMD_H

#include <stddef.h>
#include <stdint.h>
#include <hip/hip_runtime_api.h>

#ifdef __cplusplus
extern "C" {
#endif

#define BLDFM_ABI_VERSION 2

/* Error codes */
#define BLDFM_OK 0
#define BLDFM_EINVAL -1   /* invalid argument (message says which) */
#define BLDFM_ENOMEM -2   /* device allocation failed */
#define BLDFM_EHIP -3     /* HIP runtime error */
#define BLDFM_ELIMIT -4   /* size beyond what the kernels support */

/* Closure ids (src/bldfm/pbl_model.py:65-182). */
#define BLDFM_CLOSURE_MOST 0
#define BLDFM_CLOSURE_CONSTANT 1
#define BLDFM_CLOSURE_MOSTM 2
#define BLDFM_CLOSURE_OAAHOC 3

/*
 * Meteorology / profile parameters: the arguments of
 * vertical_profiles(n, meas_height, wind, ustar, z0, mol, prsc, closure,
 * domain_height, stretch, z0_min, z0_max, tke)   src/bldfm/pbl_model.py:8-22.
 * Exactly one of has_ustar/has_z0 for MOST, MOSTM and CONSTANT, unless
 * has_prologue is set.
 *
 * The scalar prologue (src/bldfm/pbl_model.py:65-120: the derived z0 or u*,
 * the stretched-grid constants aa = bb e^{-z0/h}, bb = zm/(e^{-z0/h} - e^{-zm/h})
 * and zeta_max = aa - bb e^{-zmx/h}) is a handful of scalar exp/log calls.
 * When has_prologue is 0 the library evaluates it with correctly rounded
 * functions.  A caller that must reproduce a particular host's libm bit for
 * bit (the Python drop-in passes numpy's own values: numpy's exp is not
 * correctly rounded on AVX-512 hosts) sets has_prologue = 1 and both of
 * ustar and z0 plus grid_aa, grid_bb and grid_zetamax; the library then
 * uses them as given and evaluates only the per-level profiles.
 */
typedef struct bldfm_met {
  int32_t n;                 /* grid points between z0 and meas_height */
  int32_t closure;           /* BLDFM_CLOSURE_* */
  int32_t has_ustar, has_z0, has_domain_height, has_stretch;
  double meas_height;        /* zm [m] */
  double um, vm;             /* wind components at zm [m/s] */
  double ustar, z0, mol, prsc, domain_height, stretch, tke;
  int32_t has_prologue;      /* 1: ustar, z0 and grid_* below are the caller's prologue */
  int32_t reserved;
  double grid_aa, grid_bb, grid_zetamax;
} bldfm_met;

/*
 * One solve: the arguments of steady_state_transport_solver after the
 * reference's validation and halo/mode arithmetic (src/bldfm/solver.py:82-130).
 */
typedef struct bldfm_instance {
  int32_t nx, ny;            /* srf_flx is (ny, nx), C order */
  int32_t nlx, nly;          /* modes after clamping to the padded grid */
  int32_t px, py;            /* halo pad widths int(halo/dx), int(halo/dy) */
  int32_t nz;                /* len(z); ignored when prof_from_met (derived) */
  int32_t nslot;             /* len(levels) */
  int32_t footprint;         /* 1: Green's function / footprint */
  int32_t analytic;          /* 1: constant-profile closed form */
  int32_t single;            /* 1: precision="single" (complex64 spectra) */
  int32_t out_f32;           /* 1: write float32 outputs (single, no shift) */
  int32_t has_shift;         /* 1: apply the phase shift below */
  int32_t prof_from_met;     /* 1: evaluate profiles from met on the device */
  int32_t fast;              /* 0: reproduce the reference's rounding in the vertical
                                march (default); 1: fused multiply-adds (faster,
                                agrees to the roundoff of the shooting cancellation) */
  int32_t surrogate;         /* 1: may be served by the plan's surrogate network
                                (bldfm_plan_set_surrogate); 0: always solved */
  double dx, dy;             /* xmax/nx, ymax/ny */
  double shift_x, shift_y;   /* xm+halo, ym+halo (footprint) or xm-xmax/2, ym-ymax/2 */
  double srf_bg_conc;        /* p000 */
  bldfm_met met;             /* used when prof_from_met */
} bldfm_instance;

typedef struct bldfm_plan bldfm_plan;

/* Version of this ABI (BLDFM_ABI_VERSION). */
int bldfm_abi_version(void);

/* Last error message of the calling thread. */
const char* bldfm_last_error(void);

/* Device memory: bytes the library itself holds on `device` (the FFT plan
 * and twiddle blocks of bldfm_fft2, an LRU cache capped at 16 MiB; plans are
 * the caller's, see bldfm_plan_destroy) and hipMemGetInfo's free/total.
 * No counterpart in the reference; it serves the memory checks that follow
 * tests/test_memory.py:64-119 of the reference.  Null pointers are skipped. */
int bldfm_mem_info(int device, size_t* cache_bytes, size_t* free_bytes, size_t* total_bytes);

/*
 * Replaces the scalar prologue of vertical_profiles
 * (src/bldfm/pbl_model.py:58-129): number of grid points and the derived
 * roughness length / friction velocity.
 */
int bldfm_profile_query(const bldfm_met* met, int32_t* nlev, double* z0, double* ustar);

/*
 * Replaces the array part of vertical_profiles (src/bldfm/pbl_model.py:127-182):
 * writes z,u,v,Kx,Ky,Kz (6*nlev doubles, each array contiguous) to prof6_dev
 * on the device.
 */
int bldfm_profiles(const bldfm_met* met, double* prof6_dev, hipStream_t stream);

/*
 * MOST stability functions on n device values x = z/L: psi (momentum
 * correction) and phi (diffusivity), either output may be null.
 * Replaces psi / phi (src/bldfm/pbl_model.py:207-250).
 */
int bldfm_stability(const double* x, int64_t n, double* psi, double* phi, hipStream_t stream);

/*
 * Met wind speed and direction (degrees, blowing FROM, clockwise from north)
 * to (u, v) for n device values.  Replaces compute_wind_fields
 * (src/bldfm/utils.py:7-27).
 */
int bldfm_wind_fields(const double* speed, const double* dir, int64_t n, double* u, double* v,
                      hipStream_t stream);

/*
 * Builds the device metadata (descriptors, tile tables, FFT twiddles) for a
 * batch of ninst solves on `device`.  `levels` holds, per instance in order,
 * nz level->slot entries (slot index or -1; the reference's `i in levels`
 * membership in ascending i, src/bldfm/solver.py:244,352) followed by nslot
 * level indices in the caller's order (used by the analytic branch,
 * src/bldfm/solver.py:197).  For prof_from_met instances nz is the value
 * bldfm_profile_query returns.  All instances must share out_f32.
 * Replaces the per-call setup of steady_state_transport_solver
 * (src/bldfm/solver.py:82-191) and the FFTW plan cache
 * (src/bldfm/fft_manager.py:23-44).
 */
int bldfm_plan_create(int device, const bldfm_instance* inst, int32_t ninst,
                      const int32_t* levels, bldfm_plan** plan);

/*
 * Launch shape of a plan.  No counterpart in the reference, whose numba and
 * FFTW back ends choose their own; bldfm_amd.autotune picks it by timing the
 * candidates on the device (bldfm_plan_create uses the defaults).
 *   shoot_block   modes per workgroup of the vertical solve: 64, 128 or 256
 *                 (0 selects 256)
 *   xpass_layout  how the x-pass reads the y-pass output: 0 gathers each row
 *                 at stride ny; 1 the y-pass stores its columns transposed;
 *                 2 a coalesced LDS-tiled transpose runs between the passes
 *                 (one more y-pass-sized workspace block)
 *   fft_threads   threads per workgroup of the FFT passes (one transform per
 *                 workgroup).  0 selects ceil(n/8) (n <= 2048) or ceil(n/10),
 *                 rounded up to a multiple of 64, for the plan's longest
 *                 transform n; otherwise a multiple of 64 between that value
 *                 and 256 (n <= 2048) or 1024.
 * Every choice gives the same results (bit for bit).
 */
typedef struct {
  int32_t shoot_block;
  int32_t xpass_layout;
  int32_t fft_threads;
  int32_t reserved[5];
} bldfm_tuning;

int bldfm_plan_create_tuned(int device, const bldfm_instance* inst, int32_t ninst,
                            const int32_t* levels, const bldfm_tuning* tuning, bldfm_plan** plan);

int bldfm_plan_destroy(bldfm_plan* plan);

/*
 * New meteorology for the instances of an existing plan (no counterpart in
 * the reference, which rebuilds everything per call): mets[i] replaces the
 * met of instance i when it is a prof_from_met instance (ignored otherwise).
 * Every instance must keep its number of vertical points, so the tile
 * tables and buffer layout stay valid; otherwise BLDFM_EINVAL and the plan
 * is unchanged.  The upload is ordered on `stream` after earlier executions
 * and the call returns once it is done; graphs captured from the plan replay
 * with the new values (serving loops reuse one plan and one graph).  The
 * surrogate's hull flags are re-evaluated against the hull last set.
 */
int bldfm_plan_update_met(bldfm_plan* plan, const bldfm_met* mets, int32_t n, hipStream_t stream);

/* Buffer sizes of a plan: workspace bytes, profile doubles (6*nz summed),
 * source doubles (ny*nx summed over dispersion instances), output elements
 * (nslot*ny*nx summed). */
int bldfm_plan_sizes(const bldfm_plan* plan, size_t* workspace_bytes, int64_t* prof_doubles,
                     int64_t* src_doubles, int64_t* out_elems);

/* Element offsets of instance i inside the prof, src and output buffers. */
int bldfm_plan_offsets(const bldfm_plan* plan, int32_t i, int64_t* off_prof, int64_t* off_src,
                       int64_t* off_out, int32_t* nz);

/*
 * Runs the whole path on `stream`: level tables (and profiles for
 * prof_from_met instances), forward transform of srf_flx (dispersion
 * instances), vertical solve with shooting, inverse transforms, crop.
 * Replaces src/bldfm/solver.py:132-290 including ivp_solver and the two
 * fft2/ifft2 calls.  conc/flx receive (nslot, ny, nx) per instance as float64,
 * or float32 when out_f32.  Asynchronous; capturable into a hipGraph.
 */
int bldfm_plan_execute(bldfm_plan* plan, double* prof, const double* src, void* conc, void* flx,
                       void* workspace, hipStream_t stream);

/*
 * No counterpart in the reference.  For callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc): bldfm_plan_execute with its workspace taken from a
 * library-owned static device array of BLDFM_STATIC_SCRATCH_BYTES on each GPU
 * the library runs on (held until the process ends).  Calls no hipMalloc,
 * hipFree or allocating routine.  Plans whose workspace (bldfm_plan_sizes)
 * exceeds the array are refused with BLDFM_ELIMIT.  Every call on one GPU
 * shares that GPU's array and it keeps the last call's contents (what must
 * start at zero is cleared by a kernel of each execution): executions that
 * overlap on one GPU -- other streams, other host threads -- race and must be
 * serialised by the caller.
 */
#define BLDFM_STATIC_SCRATCH_BYTES (64u << 20)
int bldfm_plan_execute_static(bldfm_plan* plan, double* prof, const double* src, void* conc, void* flx,
                              hipStream_t stream);

/*
 * Phase timing (no counterpart in the reference; for benchmarks): after
 * bldfm_plan_timing_begin, the next max_execs executions record HIP events
 * on their stream at the phase boundaries; bldfm_plan_timing_end waits for
 * them and returns the summed milliseconds of the seven phases
 * [level table, forward transform, surrogate and its gate, shooting,
 * Nyquist fix-up, y-pass, x-pass] and the number of timed executions.
 * Not for use under graph capture.
 */
int bldfm_plan_timing_begin(bldfm_plan* plan, int32_t max_execs);
int bldfm_plan_timing_end(bldfm_plan* plan, double* phase_ms, int32_t* nexecs);

/*
 * 2-D complex FFT of `batch` contiguous (n0, n1) complex128 arrays with
 * numpy semantics: sign -1 (fft2) or +1 (ifft2), every output multiplied by
 * `scale`.  Replaces FFTManager.fft2/ifft2 (src/bldfm/fft_manager.py:78-102).
 * `workspace` needs bldfm_fft2_workspace_bytes().
 */
size_t bldfm_fft2_workspace_bytes(int32_t n0, int32_t n1, int32_t batch);
int bldfm_fft2(int device, int32_t n0, int32_t n1, int32_t batch, const double* in, double* out,
               int32_t sign, double scale, void* workspace, hipStream_t stream);

/*
 * Diagnostics (no counterpart in the reference): evaluates the device
 * arithmetic helpers on n device inputs so tests can compare their rounding
 * with the host's numpy/libm bit for bit.  op: 0 hypot(a,b); 1 complex sqrt
 * of a+ib (2n outputs); 2 complex divide, 3 numpy-rounded complex multiply,
 * 4 plain complex multiply (a, b hold n complex pairs, 2n outputs); 5 sqrt(a).
 */
int bldfm_math_probe(int32_t op, const double* a, const double* b, int64_t n, double* out,
                     hipStream_t stream);


/* ------------------------------------------------------------------------
 * Learned spectral surrogate of the footprint transfer function.
 * No counterpart in the reference.  A coordinate MLP maps one Fourier mode
 * of one footprint solve (wavenumber and met scalars) to the flux transfer
 * G(k) = q^(zm; k)/q^(0; k) that steady_state_transport_solver computes in
 * footprint mode (src/bldfm/solver.py:132-134, 220-235, 307-374).  The FP64
 * solver generates its training data, checks it on spot-check modes, and
 * stays its fallback.
 */
#define BLDFM_SUR_SPOT_STRIDE 100   /* one spot-checked mode in 100 (1 %) */
#define BLDFM_SUR_MIN_SPOT 16       /* instances with fewer spot-checked modes are always solved */

/* Box of met conditions and the vertical-grid convention a network was
 * trained on.  Instances outside it are always solved. */
typedef struct bldfm_sur_hull {
  double ustar_lo, ustar_hi;   /* friction velocity [m/s] */
  double invl_lo, invl_hi;     /* 1/L [1/m] */
  double z0_lo, z0_hi;         /* roughness length [m] */
  double z0_zm_max;            /* largest z0/zm */
  double zm_lo, zm_hi;         /* measurement height [m] */
  double kmax;                 /* largest |kx|, |ky| [rad/m] */
  double top_ratio;            /* domain_height / meas_height */
  int32_t n;                   /* grid points below zm */
  int32_t closure;             /* BLDFM_CLOSURE_* */
} bldfm_sur_hull;

/* Sizes: fp32 parameter count, network buffer bytes, training workspace
 * bytes per point, points per training block (npts must be a multiple). */
int bldfm_sur_sizes(int64_t* nparam, int64_t* net_bytes, int64_t* train_bytes_per_point,
                    int32_t* block_points);

/* Network inputs of nmet met samples, 76 floats each: u*, 1/L, ln z0, zm,
 * |U|, wind direction [rad], zm/L, ln(z0/zm) (with the derived z0 / u* of
 * vertical_profiles, src/bldfm/pbl_model.py:65-77), the unit wind vector,
 * zm, the number nw of vertical intervals below zm, then 32 pairs
 * (h_j, U_j/K_j) of those intervals (the WKB prior of the network). */
int bldfm_sur_met(const bldfm_met* met, int32_t nmet, float* surmet);

/* Level tables of nmet met samples on the device (rows doubles*16 each,
 * rows >= every sample's level count); nz (host, nmet) receives the counts. */
int bldfm_transfer_tables(int device, const bldfm_met* met, int32_t nmet, int32_t rows, double* tab,
                          int32_t* nz, hipStream_t stream);

/* FP64 flux transfer of single modes: block b (256 points, kxy as (lx, ly)
 * pairs) uses met sample blk_met[b] and its measurement level lvl[m].
 * exact = 1 reproduces the reference's rounding of the march. */
int bldfm_transfer_solve(const double* tab, int32_t rows, const int32_t* nz, const int32_t* lvl,
                         int32_t nmet, const int32_t* blk_met, int32_t nblk, const double* kxy,
                         double* g, int32_t exact, hipStream_t stream);

/* Packs fp32 parameters (bldfm_sur_sizes' nparam) into a network buffer
 * whose first 1024 bytes (features and normalisation) the caller wrote. */
int bldfm_sur_pack(const float* params, void* net, hipStream_t stream);

/* Phi = -log G (two floats per point) of npts points; block b of
 * block_points points uses surmet[blk_met[b]]. */
int bldfm_sur_forward(const void* net, const float* surmet, int32_t nmet, const int32_t* blk_met,
                      const double* kxy, int64_t npts, float* phi, hipStream_t stream);

/* Loss against solver transfers g and its gradient: loss (one float,
 * accumulated) and grad (nparam floats: weights accumulated, biases and
 * output layer overwritten).  ws: npts * train_bytes_per_point bytes. */
int bldfm_sur_train_grad(const void* net, const float* surmet, int32_t nmet, const int32_t* blk_met,
                         const double* kxy, const double* g, int64_t npts, void* ws, float* grad,
                         float* loss, float* phi, hipStream_t stream);

/* Adam step t (1-based) on fp32 master parameters with moments m, v; clears
 * grad and repacks the network buffer. */
int bldfm_sur_adam(float* params, float* grad, float* m, float* v, void* net, float lr, float beta1,
                   float beta2, float eps, int32_t step, hipStream_t stream);

/* Lets a plan serve its surrogate-flagged instances with the network in
 * `net` (null: none): each execution evaluates the network on every mode of
 * those inside `hull`, solves one mode in BLDFM_SUR_SPOT_STRIDE in FP64,
 * and solves the whole instance instead when the spot-check relative L2
 * error exceeds tol or no spot-checked mode carried a nonzero reference.
 * Instances with fewer than BLDFM_SUR_MIN_SPOT spot-checked modes are
 * outside every hull.  Served instances get the flux footprint; their conc
 * output is 0. */
int bldfm_plan_set_surrogate(bldfm_plan* plan, const void* net, const bldfm_sur_hull* hull, double tol);

/* After an execution (synchronises `stream`): per instance the spot-check
 * relative L2 error (NaN when not served) and the path taken
 * (0 surrogate, 1 solver). */
int bldfm_plan_sur_stats(const bldfm_plan* plan, const void* workspace, double* err, int32_t* mode,
                         hipStream_t stream);

/* ------------------------------------------------------------------------
 * Kormann & Meixner (2001) analytic footprint, the reference's comparison
 * model (src/bldfm/ffm_kormann_meixner.py:110-330).  One record per
 * footprint; the grid is numpy's arange(xmin + res/2, xmax, res) by
 * arange(ymax - res/2, ymin, -res) (rows from the top).
 */
typedef struct bldfm_km01 {
  double zm, z0, ws, ustar, mo_len, sigma_v;   /* estimateFootprint arguments */
  double xmin, xmax, ymin, ymax, res;          /* grid_domain, grid_res */
  double mx, my;                               /* mxy */
  double wd;                                   /* wind direction [deg], used when has_wd */
  int32_t has_wd, reserved;
} bldfm_km01;

/* Grid shape (rows ny, columns nx) of one record. */
int bldfm_km01_grid(const bldfm_km01* p, int32_t* nx, int32_t* ny);

/* Footprints of n records into ffm (device, records' ny*nx grids one after
 * another).  status[i] (host) is 0, or 1 when Eq. (11) gives U < 0 (the
 * reference then returns an all-zero footprint, as here). */
int bldfm_km01_footprint(int device, const bldfm_km01* p, int32_t n, double* ffm, int32_t* status,
                         hipStream_t stream);

#ifdef __cplusplus
}
#endif
#endif /* BLDFM_AMD_H */
