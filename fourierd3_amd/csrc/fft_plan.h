// Tiling plan and tables of the multi-channel 3-D real FFT (fft.hip).
#pragma once

#include <algorithm>

#include "fd3_internal.h"

namespace fd3 {

// Radix-4 butterflies held per thread per stage by the in-LDS line FFT, and
// the resulting cap on complex values per 256-thread block (32 KB of LDS).
constexpr int kFftMaxB = 4;
constexpr int kFftLines = kFftMaxB * 4 * 256;
// Complex values per block of the y passes (1024 threads, 64 KB of LDS):
// 32 x lines of 256, so the transposed side moves 256-byte rows.
constexpr int kYLines = 2 * kFftLines;
// Complex values per block of the z passes (256 threads, 16 y lines of 128
// at nz = 256): small tiles keep eight blocks in flight per CU, so one
// block's atom loads and barriers overlap the others' FFT work.
constexpr int kZLines = kFftLines / 2;

struct FftPlan {
  int nx, ny, nz, M, M1;  // M = nz/2, M1 = M + 1 (half-spectrum length)
  int nS, r, C;           // species, rank, channels = nS * r
  int TY;                 // y lines per block of the z passes
  int TX;                 // x lines per block of the y passes
  int LPB;                // lines per block of the plain x pass
  int RT, NK;             // ranks and ky rows per tile of the k-space kernel
  int KT;                 // tiles (of NK ky rows) per k-space block
  int ks_reg;             // k-space variant: nS (<= kKsRegMaxS) with S and phi in registers, 0 = generic
  int ks2;                // register variant with the two-pass (16 x 16) x FFT: nx = 256, NK = 2
  int xs;                 // ks2 with ny = 256: the y passes run the x FFT's first stage (split x FFT)
  int ks_blocks;          // grid size of the k-space kernel
  size_t ks_lds;          // dynamic LDS bytes of the k-space kernel
  size_t mesh_complex() const { return (size_t)C * nx * M1 * ny; }
  size_t mesh_real() const { return (size_t)C * nx * ny * nz; }
};

struct FftTables {
  float2 *twx, *twy, *twz;
  float *scx, *scy, *scz;
  float4* htab;  // 3 * (kHtabN + 1) entries: polynomial table of h6, h8
  float4* gtab;  // (pairs X <= Y, kGtabN, 2): per-pair G tables of the register k-space kernel
  KsPair* pc;    // (nS, nS) per-pair constants of the k-space kernel
  float2* sink;  // 256: written, never read (stores of the split x k-space kernel's idle lanes)
};

// Inputs of the per-pair constants (device table qq = sqrt(Q_X) sqrt(Q_Y)).
struct PairInit {
  const float* qq;
  int nS;
  float s6, s8, a1, a2;
  double gdk;  // interval width of the per-pair G tables (Bohr^-1), 0: none
};

// Butterflies per thread held by the register k-space variant's line FFTs:
// its tiles hold nS * RT * NK * nx <= 1024 * ks_maxb(nS * RT) values.
constexpr int ks_maxb(int lines_per_row) { return lines_per_row <= 8 ? 2 : 4; }

// Line stride (complex) of the two-pass x FFT of the k-space kernel: 2 * 272
// = 32 (mod 64) banks between lines.
constexpr int kKs2Ls = 272;
// Dynamic LDS of the split x FFT variant (kspace_xs_body): 2 ky rows of
// nl = nS * RT lines of 256 points (at most 32 KB, so five blocks fit the
// 160 KB of a gfx950 CU), then with parameter gradients the h table and the
// pair constants.
constexpr size_t ksx_lds(int nl, bool params) {
  return (size_t)2 * nl * 256 * sizeof(float2) +
         (params ? (size_t)3 * (kHtabN + 1) * sizeof(float4) + (size_t)nl * nl * sizeof(KsPair) : 0);
}
// Dynamic LDS of the pipelined split x kernel without parameter gradients
// (kspace_xsp_body): the lines of nl spectra, the per-pair G tables of nS
// species, the rank eigenvalues, sinc_x and sinc_y (256 each) and sinc_z (M1).
inline size_t ksxp_lds(int nS, int nl, int M1) {
  return (size_t)2 * nl * 256 * sizeof(float2) + (size_t)nS * (nS + 1) * kGtabN * sizeof(float4) +
         (size_t)(kMaxRank + 512 + M1) * sizeof(float);
}
// float4 entries of its table region: the h table with parameter gradients,
// none without (the per-pair G tables are read from global memory).
constexpr int ks2_table_f4(bool params) { return params ? 3 * (kHtabN + 1) : 0; }

// Species counts for which the k-space kernel keeps a k point's spectra in registers.
constexpr int kKsRegMaxS = 4;
// float4 entries of the table region of the register k-space kernel: the h
// table or the per-pair G tables, whichever is larger (one layout serves both
// variants).
constexpr int ks_table_f4(int nS) {
  return 3 * (kHtabN + 1) > nS * (nS + 1) * kGtabN ? 3 * (kHtabN + 1) : nS * (nS + 1) * kGtabN;
}
// The generic k-space kernel: LDS budget (of the 160 KB a gfx950 workgroup may
// use; 2 KB of static LDS come on top) and butterflies per thread of its line
// FFTs (tiles of nS * RT * NK * nx <= 1024 * kKsGenericMaxB values).
constexpr int kKsGenericLds = 150 * 1024;
constexpr int kKsGenericMaxB = 8;

int fft_plan_make(const int32_t ns[3], int nS, int r, FftPlan* p);
// Interval width of the per-pair G tables: kmax / kGtabN, kmax the largest
// |k| of the half spectrum (a corner of the index box).
double ks_gtab_dk(const FftPlan& p, const Geom& g);
void fft_tables_carve(const FftPlan& p, Arena& ar, FftTables* t);
int fft_tables_init(const FftPlan& p, const FftTables& t, int order, int use_sinc, hipStream_t s,
                    const PairInit* pairs = nullptr);
// rho (real) -> A -> spec; spec holds the y/z-transformed data (x not yet transformed)
int fft_forward_zy(const FftPlan& p, const FftTables& t, const float* rho, float2* A, float2* spec, hipStream_t s);
// y pass alone: dir -1 A -> spec, dir +1 spec -> A.  xs (needs p.xs): the
// pass also runs the first stage of the x FFT, for kspace_run(.., xs = true)
int fft_y(const FftPlan& p, const FftTables& t, const float2* in, float2* out, int dir, hipStream_t s,
          bool xs = false);
// plain in-place x transform of spec (dir -1 forward, +1 inverse)
int fft_x(const FftPlan& p, const FftTables& t, float2* spec, int dir, hipStream_t s);
// spec (x already inverse-transformed) -> A -> out (real)
int fft_inverse_yz(const FftPlan& p, const FftTables& t, const float2* spec, float2* A, float* out, hipStream_t s);
// fused forward-x / k-space / inverse-x; partials has p.ks_blocks * 16 doubles
// (the pair constants come from fft_tables_init with a PairInit)
// k0 (C): the k = 0 coefficients of dE/drho, zeroed in the spectrum.
// xs: spec comes from (and goes back to) the y passes with the split x FFT.
int kspace_run(const FftPlan& p, const FftTables& t, const Geom& g, float2* spec, const float* eigs, int need_grad,
               int need_params, double* partials, float2* k0, hipStream_t s, bool xs = false);

}  // namespace fd3
