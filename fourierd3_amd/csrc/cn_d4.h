// D4 coordination numbers in reciprocal space (cn_d4.hip).
#pragma once

#include "fd3_internal.h"
#include "fft_plan.h"

namespace fd3 {

// Arguments of the D4 CN k-space kernel.  spec: the species spectra after the
// z and y passes (channel X = species, r = 1), transformed in x and
// overwritten with the convolution; rho_hat: the x-transformed density
// spectrum (written by the forward variant, read by the backward one).
struct CnKsArgs {
  float2* spec;
  float2* rho_hat;
  const float2* twx;
  const float* scx; const float* scy; const float* scz;
  const float* r0;     // (nS, nS) Bohr
  const float* delta;  // (nS, nS)
  float k0;
  float4* tab;         // (pairs X <= Y, kGtabN, 2) per-pair tables of G^CN
  float tkinv;         // 1 / (table interval width), Bohr
  double* partials;    // (blocks, 8): strain-derivative sums of the backward variant
  double inv_vol;      // 1 / Omega (Bohr^-3)
  float recip[9];      // rows b_i of 2 pi cell^-T (Bohr^-1)
  int nx, ny, nz, M1, nS, NK, KT, use_sinc;
  const CellDyn* dyn = nullptr;  // when set, recip and inv_vol are read from it
  float2* sink = nullptr;        // 256 values written by idle lanes (split x kernels), never read
};

struct CnKsPlan {
  int NK, KT, blocks, maxb;
  bool reg;             // species count <= 4: the register-resident forward kernel
  bool xs;              // nx = ny = 256, species <= 4: split x FFT (k_kspace_cnp_xs), the
                        // y passes run the x FFT's first stage
  size_t lds, lds_bwd;  // dynamic LDS of the forward and backward variants
};

int cn_ks_plan(const FftPlan& p, int nS, CnKsPlan* kp);
// Builds a.tab (interval width dk) from a.r0, a.delta, a.k0.
int cn_tab_init(const CnKsArgs& a, double dk, hipStream_t s);
int cn_ks_run(const CnKsArgs& a, const CnKsPlan& kp, bool bwd, hipStream_t s);

// Per-atom steps (slot order).
struct CnAtomArgs {
  int64_t n;
  int nx, ny, nz;
  double inv[9];      // cell^-1 (Bohr^-1)
  double inv_vol;     // 1 / Omega (Bohr^-3)
  const int32_t* spec_s;
  const float* selfcont;  // (nS) f_XX(0)
  const float* accA;  // (n) sum_nodes W phi (forward gather)
  const float* accG;  // (n, 3) sum_nodes dW/du of the gathered field
  double* cn;         // (n) D4 CN (written by cn_d4_fin)
  float* g1;          // (n, 3) forward gather's sum_nodes dW/du phi, kept for the force
  const double* dEdcn; // (n) dE/dCN (backward)
  double* gpos;       // (n, 3) dE/dr (Bohr), added to
  double* partials;   // (blocks, 8) strain-derivative sums
  const CellDyn* dyn = nullptr;  // when set, inv and inv_vol are read from it
};

int cn_d4_fin(const CnAtomArgs& a, hipStream_t s);
int cn_d4_force(const CnAtomArgs& a, hipStream_t s);
int fill_value(float* p, int64_t n, float v, hipStream_t s);

}  // namespace fd3
