// Learned spectral surrogate: shared layout of the network, its parameter
// vector and its device buffer (host and device).
//
// The network maps one Fourier mode of one footprint solve to the flux
// transfer G(k) = q^(zm; k) / q^(0; k) of the solver's footprint mode.  It
// predicts Phi = -log G (complex), so G = exp(-Phi):
//
//   x0  = [RFF(c) (64) | c (4) | met (8) | Phi0/8 (2) | 0 (18)]   96 inputs
//   z_l = W_l x_l + b_l,  x_{l+1} = silu(z_l),   l = 0..4         width 256
//   Phi = Phi0 + Wo x5 + bo                                      2 outputs
//
// c are wind-frame coordinates of the wavenumber scaled by the measurement
// height: kpar = k.w zm, kperp = |k x w| zm, and their signed square roots;
// RFF(c) = [sin(2 pi B c), cos(2 pi B c)] with a fixed 32x4 matrix B.
// Phi0 is a WKB estimate of -log G from the profiles (SurMet below), so the
// network only learns the correction.
// Hidden activations and weights are bf16 with fp32 accumulation (MFMA);
// the output layer runs in fp32 on the last hidden layer's fp32 values.
#pragma once
#include <stdint.h>

#include "bldfm_types.h"

#define SUR_IN 96        // padded input width
#define SUR_W 256        // hidden width
#define SUR_NL 5         // hidden layers
#define SUR_NRFF 32      // random Fourier frequencies (64 features)
#define SUR_NC 4         // coordinate inputs
#define SUR_NMET 8       // met inputs
#define SUR_F_PHI0 (2 * SUR_NRFF + SUR_NC + SUR_NMET)  // input index of Phi0/8
#define SUR_PHI0_SCALE 0.125f
#define SUR_PTS 128      // points per workgroup
#define SUR_THREADS 256  // four waves

// Flat fp32 parameter vector (master weights, gradients, Adam moments).
// W_l is row-major [out][in].
#define SUR_P_W0 0
#define SUR_P_B0 (SUR_W * SUR_IN)
#define SUR_P_L1 (SUR_P_B0 + SUR_W)  // start of layers 1..4
#define SUR_P_LSTRIDE (SUR_W * SUR_W + SUR_W)
#define SUR_P_W(l) ((l) == 0 ? SUR_P_W0 : SUR_P_L1 + ((l) - 1) * SUR_P_LSTRIDE)
#define SUR_P_B(l) ((l) == 0 ? SUR_P_B0 : SUR_P_W(l) + SUR_W * SUR_W)
#define SUR_P_WO (SUR_P_L1 + (SUR_NL - 1) * SUR_P_LSTRIDE)
#define SUR_P_BO (SUR_P_WO + 2 * SUR_W)
#define SUR_NPARAM (SUR_P_BO + 2)
#define SUR_NPARAM_PAD ((SUR_NPARAM + 255) / 256 * 256)

// Fixed (trained-with) constants of a network, set once by the host.
//
// mode 0 is the design above.  mode 1 is the network exactly as the
// surrogate's specification words it: x0 = [RFF(kx, ky) (64) | 6 met scalars
// (u*, 1/L, ln z0, zm, |U|, direction; normalised) | 0], no prior, and the two
// outputs are Re G and Im G themselves (rff[j][2..3] unused).
#define SUR_MODE_SPEC 1
#define SUR_NMET_SPEC 6
struct SurStatic {
  float rff[SUR_NRFF][SUR_NC];  // B
  float cscale[SUR_NC];         // coordinate scales
  float met_mu[SUR_NMET];       // met normalisation: (m - mu) / sd
  float met_sd[SUR_NMET];
  float mode;                   // 0 or SUR_MODE_SPEC
  float pad[64 - 2 * SUR_NMET - SUR_NC - 1];
};

// Parameters in the layout the kernels read, written by the packing kernel.
struct SurDyn {
  float b[SUR_NL][SUR_W];
  float wo[2][SUR_W];
  float bo[4];
};

// Device buffer of one network (byte offsets).
#define SUR_OFF_STATIC 0
#define SUR_OFF_DYN 1024
#define SUR_OFF_W0P (SUR_OFF_DYN + 8192)                    // bf16, 256 x 96 packed
#define SUR_OFF_WP(l) (SUR_OFF_W0P + SUR_W * SUR_IN * 2 + ((l) - 1) * SUR_W * SUR_W * 2)   // l = 1..4
#define SUR_OFF_WT(l) (SUR_OFF_WP(SUR_NL) + ((l) - 1) * SUR_W * SUR_W * 2)                 // l = 1..4
#define SUR_NET_BYTES SUR_OFF_WT(SUR_NL)

#define SUR_NWKB 32      // most vertical intervals below zm in the WKB prior

// Per-footprint inputs of the network: 8 raw met values
// (u*, 1/L, ln z0, zm, |U|, wind direction, zm/L, ln(z0/zm)), the unit wind
// vector, zm, and the WKB prior table: for each of the nw intervals between
// the surface and zm its height h_j and w_j = U_j / K_j (along-wind speed
// over diffusivity at the interval's midpoint), so that
//   Phi0(k) = sum_j h_j sqrt(|k|^2 + i kpar w_j)
// approximates -log G (the network adds the rest).
struct SurMet {
  float met[SUR_NMET];
  float wx, wy, zm, nw;
  float wkb[SUR_NWKB][2];
};

// Packed bf16 fragment index of W[o][k] for the forward product
// (A operand of v_mfma_f32_32x32x16_bf16: lane = o%32 + 32*((k%16)/8), 8 values).
BLDFM_HD int sur_pack_fwd(int o, int k, int kin) {
  const int kb16 = kin / 16;
  const int blk = (o / 32) * kb16 + k / 16;
  const int lane = (o % 32) + 32 * ((k % 16) / 8);
  return (blk * 64 + lane) * 8 + (k % 8);
}

// Packed index of W[o][k] for the transposed product W^T dZ (rows k, reduction over o).
BLDFM_HD int sur_pack_bwd(int o, int k) {
  const int blk = (k / 32) * (SUR_W / 16) + o / 16;
  const int lane = (k % 32) + 32 * ((o % 16) / 8);
  return (blk * 64 + lane) * 8 + (o % 8);
}

// Bf16 element offset of (point, feature) in a workgroup's LDS activation
// tile: rows of 256 bf16, 16-byte chunks XOR-swizzled by the row so that the
// 32 rows read by one MFMA operand load fall in distinct banks.
BLDFM_HD int sur_act_off(int pt, int f) { return pt * SUR_W + (((f >> 3) ^ (pt & 31)) << 3) + (f & 7); }
