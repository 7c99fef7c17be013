// Tiling plan and tables of the multi-channel 3-D real FFT (fft.hip).
#pragma once

#include <algorithm>

#include "fd3_internal.h"

namespace fd3 {

// Radix-4 butterflies held per thread per stage by the in-LDS line FFT, and
// the resulting cap on complex values per 256-thread block (32 KB of LDS).
constexpr int kFftMaxB = 4;
constexpr int kFftLines = kFftMaxB * 4 * 256;

struct FftPlan {
  int nx, ny, nz, M, M1;  // M = nz/2, M1 = M + 1 (half-spectrum length)
  int nS, r, C;           // species, rank, channels = nS * r
  int TY;                 // y lines per block of the z passes
  int TX;                 // x lines per block of the y passes
  int LPB;                // lines per block of the plain x pass
  int RT, NK;             // ranks and ky rows per block of the k-space kernel
  int ks_reg;             // k-space variant: nS (<= kKsRegMaxS) with S and phi in registers, 0 = generic
  int ks_blocks;          // grid size of the k-space kernel
  size_t ks_lds;          // dynamic LDS bytes of the k-space kernel
  size_t mesh_complex() const { return (size_t)C * nx * M1 * ny; }
  size_t mesh_real() const { return (size_t)C * nx * ny * nz; }
};

struct FftTables {
  float2 *twx, *twy, *twz;
  float *scx, *scy, *scz;
  float4* htab;  // 3 * (kHtabN + 1) entries: polynomial table of h6, h8
};

// Species counts for which the k-space kernel keeps a k point's spectra in registers.
constexpr int kKsRegMaxS = 4;

int fft_plan_make(const int32_t ns[3], int nS, int r, FftPlan* p);
void fft_tables_carve(const FftPlan& p, Arena& ar, FftTables* t);
int fft_tables_init(const FftPlan& p, const FftTables& t, int order, int use_sinc, hipStream_t s);
// rho (real) -> A -> spec; spec holds the y/z-transformed data (x not yet transformed)
int fft_forward_zy(const FftPlan& p, const FftTables& t, const float* rho, float2* A, float2* spec, hipStream_t s);
// y pass alone: dir -1 A -> spec, dir +1 spec -> A
int fft_y(const FftPlan& p, const FftTables& t, const float2* in, float2* out, int dir, hipStream_t s);
// plain in-place x transform of spec (dir -1 forward, +1 inverse)
int fft_x(const FftPlan& p, const FftTables& t, float2* spec, int dir, hipStream_t s);
// spec (x already inverse-transformed) -> A -> out (real)
int fft_inverse_yz(const FftPlan& p, const FftTables& t, const float2* spec, float2* A, float* out, hipStream_t s);
// fused forward-x / k-space / inverse-x; partials has p.ks_blocks * 16 doubles
int kspace_run(const FftPlan& p, const FftTables& t, const Geom& g, float2* spec, const float* eigs, const float* qq,
               double s6, double s8, double a1, double a2, int need_grad, int need_params, double* partials,
               hipStream_t s);

}  // namespace fd3
