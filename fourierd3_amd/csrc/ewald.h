// method='ewald' kernels (ewald.hip): direct structure-factor sum over the
// full k grid instead of the mesh path.
#pragma once

#include "fd3_internal.h"
#include "fft_plan.h"

namespace fd3 {

struct EwaldArgs {
  int64_t n;
  int nS, r;
  int nx, ny, nz;          // k grid: m_d in fftfreq(n_d) * n_d
  double inv[9];           // cell^-1 (Bohr^-1)
  double recip[9];         // rows b_i of 2 pi cell^-T
  double volume;           // Bohr^3
  const double4* pos4;     // (n) slot order, Bohr
  const int32_t* spec_s;   // (n) slot order
  const float* charges;    // (n, r) slot order
  const float* eigs;       // (r)
  const KsPair* pc;        // (nS, nS) pair constants
  const float4* htab;      // h6 / h8 table
  float2* S;               // (nk, nS, r) structure factors
  float2* phi;             // (nk, nS, r)
  int32_t* perm;           // (n) atoms by species
  int32_t* seg;            // (nS + 1)
  double* partials;        // (ceil(nk / 256), 16), the k-space kernel's format
  float* accA;             // (n, r) out: dE/dc (k-space part)
  float* accG;             // (n, 3) out: dE/df (fractional coordinates)
  int need_params;
};

// The D4 CN on the Ewald k grid (cnfunc='d4', method='ewald';
// src/fourierd3/core.py:430-441): per species structure factors of weights
// w_i (1, or dE/dCN_i in the adjoint), S_X(k) = sum_{i in X} w_i e_ik,
// phi_X = sum_Y G^CN_XY(|k|) S_Y, then per atom
//   accA_i = sum_k Re(conj(e_ik) phi_{X_i}(k))                (forward only)
//   accG_i = sum_k 2 pi m Im(conj(e_ik) phi_{X_i}(k))         (d/df of the same)
// The adjoint also sums the strain derivative of sum_ij g_i raw_i:
// sum_k sum_XY dG/dk (-k k^T / |k|) Re(conj(S^g_X) S^rho_Y) / Omega.
struct EwCnArgs {
  int64_t n;
  int nS, nx, ny, nz;
  double inv[9];
  double recip[9];
  double volume;
  const double4* pos4;
  const int32_t* spec_s;
  const float* w;          // (n) atom weights
  const int32_t* perm;     // (n) atoms by species
  const int32_t* seg;      // (nS + 1)
  const float4* tab;       // per-pair G^CN tables
  float tkinv;
  float2* S;               // (nk, nS) this pass's structure factors
  float2* Srho;            // (nk, nS) the forward pass's (written by it, read by the adjoint)
  float2* phi;             // (nk, nS)
  double* partials;        // (ceil(nk / 256), 8) adjoint strain sums
  float* accA;             // (n)
  float* accG;             // (n, 3)
};
int ewald_partition(const int32_t* spec_s, int64_t n, int nS, int32_t* perm, int32_t* seg, hipStream_t s);
int ewcn_pass(const EwCnArgs& a, bool bwd, hipStream_t s);

// A plan holding the k grid and channel counts (no FFT tiling); ks_blocks is
// the number of partial rows k_ewald_sf writes.
int ewald_plan_make(const int32_t ns[3], int nS, int r, FftPlan* p);
size_t ewald_buffer_elems(const FftPlan& p);  // complex elements of S (and of phi)
int ewald_run(const EwaldArgs& a, bool grad, hipStream_t s);

}  // namespace fd3
