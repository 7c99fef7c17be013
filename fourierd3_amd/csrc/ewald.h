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

// A plan holding the k grid and channel counts (no FFT tiling); ks_blocks is
// the number of partial rows k_ewald_sf writes.
int ewald_plan_make(const int32_t ns[3], int nS, int r, FftPlan* p);
size_t ewald_buffer_elems(const FftPlan& p);  // complex elements of S (and of phi)
int ewald_run(const EwaldArgs& a, bool grad, hipStream_t s);

}  // namespace fd3
