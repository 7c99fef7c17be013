// 3-D real FFTs of multi-channel meshes and the fused reciprocal-space kernel.
//
// Mesh layouts (float32, c = species * r + rank):
//   rho  [c][x][y][z]          real, z fastest (spreading output / gather input)
//   A    [c][x][kz][y]         complex, after the z pass (kz = 0..nz/2)
//   spec [c][kz][ky][kx]       complex, after the y and x passes
// Every pass reads and writes the mesh once, transforms the axis that is
// contiguous in its input and writes its output transposed so the next pass
// again reads contiguous lines; the x pass of the forward transform, the
// k-space product and the x pass of the inverse transform run in one kernel
// (k_kspace), so the spectrum is read and written once between them.
//
// Line FFTs are in-LDS Stockham radix-4 (+ one radix-2 stage for odd log2 n)
// with twiddles W_n^k = exp(-2 pi i k / n) from a table computed in double.
// The real-to-complex z transform packs a real line of length nz into nz/2
// complex values, transforms them and separates even/odd halves.

#include <array>
#include <map>
#include <mutex>
#include <tuple>

#include "fd3_internal.h"
#include "fft_plan.h"
#include "launch.h"
#include "fft_lds.h"
#include "fft_reg.h"

namespace fd3 {

// ------------------------------------------------------------ tables ---
// tw_axis[k] = exp(-2 pi i k / n) for each axis; sinc_axis[m] = sinc(m/n)^(2p)
// with m the fftfreq integer of index m (rfftfreq for z).  Also the h table
// and, when pi.qq is set, the per-pair constants and (pi.gdk > 0) the per-pair
// G tables.  The tables take one thread per Chebyshev node (the double
// exponentials and cosines dominate) and one per interval for the fit:
// blocks [0, kInitHB) build the h table (42 intervals x 6 nodes each), the
// next ngb blocks the G tables (32 intervals x 8 nodes), the rest the
// twiddles, sinc factors and pair constants.
constexpr int kInitHB = (kHtabN + 1 + 41) / 42;
constexpr int kInitGI = 256 / (kGtabDeg + 1);  // G-table intervals per block

inline int init_gtab_blocks(const PairInit& pi) {
  if (!pi.qq || !(pi.gdk > 0.0) || pi.nS > kKsRegMaxS) return 0;
  return (pi.nS * (pi.nS + 1) / 2 * kGtabN + kInitGI - 1) / kInitGI;
}

struct InitTablesArgs {
  float2 *twx, *twy, *twz;
  float *scx, *scy, *scz;
  float4* htab;
  float4* gtab;
  KsPair* pc;
  PairInit pi;
  int ngb, nx, ny, nz, order, use_sinc;
};

__device__ __forceinline__ void init_tables_body(const BlockInfo& bk, const InitTablesArgs& ia) {
  float2 *twx = ia.twx, *twy = ia.twy, *twz = ia.twz;
  float *scx = ia.scx, *scy = ia.scy, *scz = ia.scz;
  float4* htab = ia.htab;
  float4* gtab = ia.gtab;
  KsPair* pc = ia.pc;
  const PairInit& pi = ia.pi;
  const int ngb = ia.ngb, nx = ia.nx, ny = ia.ny, nz = ia.nz, order = ia.order, use_sinc = ia.use_sinc;
  const unsigned bx = bk.id.x;
  __shared__ double f[2][256];
  const int tid = threadIdx.x;
  if (bx < kInitHB) {  // h table
    const int i = bx * 42 + tid / 6, j = tid % 6;
    if (tid < 252 && i < kHtabN) h_exact(htab_node(i, j), &f[0][tid], &f[1][tid]);
    __syncthreads();
    const int ii = bx * 42 + tid;
    if (tid < 42 && ii <= kHtabN) htab_fit(ii, &f[0][6 * tid], &f[1][6 * tid], htab + 3 * ii);
    return;
  }
  if ((int)bx < kInitHB + ngb) {  // per-pair G tables
    const int np = pi.nS * (pi.nS + 1) / 2;
    const int it = (bx - kInitHB) * kInitGI + tid / (kGtabDeg + 1), j = tid % (kGtabDeg + 1);
    if (it < np * kGtabN) {
      const int pt = it / kGtabN, i = it - pt * kGtabN;  // pair pt = (X, Y >= X) in row order
      int X = 0, r0 = 0;
      while (pt >= r0 + pi.nS - X) { r0 += pi.nS - X; ++X; }
      const KsPair pk = ks_pair(pi.qq[X * pi.nS + X + (pt - r0)], pi.s6, pi.s8, pi.a1, pi.a2);
      f[0][tid] = g_exact(gtab_node(pi.gdk, i, j), (double)pk.Rs * kHtabDx, pk.p6, pk.p8);
    }
    __syncthreads();
    const int io = (bx - kInitHB) * kInitGI + tid;
    if (tid < kInitGI && io < np * kGtabN) gtab_fit(&f[0][(kGtabDeg + 1) * tid], gtab + 2 * io);
    return;
  }
  const int t = (bx - kInitHB - ngb) * blockDim.x + tid;
  if (pi.qq && t < pi.nS * pi.nS) pc[t] = ks_pair(pi.qq[t], pi.s6, pi.s8, pi.a1, pi.a2);
  auto tw = [](float2* dst, int k, int n) {
    double s, c;
    sincospi(-2.0 * (double)k / (double)n, &s, &c);
    dst[k] = make_float2((float)c, (float)s);
  };
  auto sincp = [&](int m, int n) -> float {
    if (!use_sinc) return 1.0f;
    if (m == 0) return 1.0f;
    const double a = (double)m / (double)n;
    const double s = sinpi(a) / (kPi * a);
    return (float)pow(s, 2.0 * order);
  };
  if (t < nx) { tw(twx, t, nx); scx[t] = sincp(t < (nx + 1) / 2 ? t : t - nx, nx); }
  if (t < ny) { tw(twy, t, ny); scy[t] = sincp(t < (ny + 1) / 2 ? t : t - ny, ny); }
  if (t < nz) tw(twz, t, nz);
  if (t <= nz / 2) scz[t] = sincp(t, nz);
}

// ------------------------------------------------------------ z passes ---
// rho[c][x][y][0..nz) real -> A[c][x][kz][y], kz = 0..M (M = nz/2).
// grid (ny/TY, nx, C); LDS TY lines of M+1 complex.
__global__ void __launch_bounds__(256) k_r2c_z(const float* __restrict__ rho, float2* __restrict__ A, int nx, int ny,
                                               int nz, int TY, const float2* __restrict__ twz) {
  extern __shared__ float2 lds[];
  const int M = nz >> 1, ls = M + 1;
  const int y0 = blockIdx.x * TY, x = blockIdx.y, c = blockIdx.z;
  float2* tws = lds + TY * ls;
  stage_table(tws, twz, nz);
  const float2* src = (const float2*)rho + ((size_t)(c * nx + x) * ny + y0) * M;
  const int lM = ilog2(M), lT = ilog2(TY);
  tile2_to_lds<kCopyU>(lds, src, TileShape{TY, lM, 1, ls, 1}, StridedLines{M});
  __syncthreads();
  lds_fft<-1, kFftMaxB, 256>(lds, M, TY, ls, tws, 2);
  z_pairs<true>(lds, TY, ls, M, tws);  // X[k] = E[k] + W_nz^k O[k]
  __syncthreads();
  float2* dst = A + (size_t)(c * nx + x) * (M + 1) * ny + y0;
  tile2_from_lds<kCopyU>(dst, lds, TileShape{M + 1, lT, 1, 1, ls}, StridedLines{ny});
}

// A[c][x][kz][y0..y0+TY) -> out[c][x][y][0..nz) real (unnormalised inverse;
// imaginary parts of the kz = 0 and kz = nz/2 terms are ignored, as C2R does).
__global__ void __launch_bounds__(256) k_c2r_z(const float2* __restrict__ A, float* __restrict__ out, int nx, int ny,
                                               int nz, int TY, const float2* __restrict__ twz) {
  extern __shared__ float2 lds[];
  const int M = nz >> 1, ls = M + 1;
  const int y0 = blockIdx.x * TY, x = blockIdx.y, c = blockIdx.z;
  float2* tws = lds + TY * ls;
  stage_table(tws, twz, nz);
  const float2* src = A + (size_t)(c * nx + x) * (M + 1) * ny + y0;
  const int lM = ilog2(M), lT = ilog2(TY);
  tile2_to_lds<kCopyU>(lds, src, TileShape{M + 1, lT, 1, 1, ls}, StridedLines{ny});
  __syncthreads();
  z_pairs<false>(lds, TY, ls, M, tws);
  __syncthreads();
  lds_fft<1, kFftMaxB, 256>(lds, M, TY, ls, tws, 2);
  float2* dst = (float2*)out + ((size_t)(c * nx + x) * ny + y0) * M;
  tile2_from_lds<kCopyU>(dst, lds, TileShape{TY, lM, 1, ls, 1}, StridedLines{M});
}

// ------------------------------------------------------------ y passes ---
// forward: A[c][x][kz][0..ny) (x in tile) -> B[c][kz][ky][x]; grid (nx/TX, M+1, C);
// blocks of 2^LOG threads.
struct YPassArgs {
  const float2* in;
  float2* out;
  int nx, ny, M1, TX;
  const float2* twy;
};

template <int DIR, int LOG>
__device__ __forceinline__ void c2c_y_body(const BlockInfo& bk, const YPassArgs& ya) {
  constexpr int U = kYLines >> LOG;     // a whole tile in one batch of loads
  constexpr int MAXB = kYLines >> (LOG + 2);
  extern __shared__ float2 lds[];
  const float2* __restrict__ in = ya.in;
  float2* __restrict__ out = ya.out;
  const float2* __restrict__ twy = ya.twy;
  const int nx = ya.nx, ny = ya.ny, M1 = ya.M1, TX = ya.TX;
  const int ls = ny + 1;
  const int x0 = bk.id.x * TX, kz = bk.id.y, c = bk.id.z;
  float2* tws = lds + TX * ls;
  stage_table(tws, twy, ny);
  const int lTX = ilog2(TX);
  // the y lines (stride M1 ny apart, y contiguous) go straight between HBM and
  // the FFT's first (forward) or last (inverse) stage; the transposed side
  // (x contiguous) passes through LDS
  long long* lb = (long long*)(tws + ny);  // line bases of the TX y lines
  for (int l = threadIdx.x; l < TX; l += blockDim.x) lb[l] = ((long long)(c * nx + x0 + l) * M1 + kz) * ny;
  if (DIR < 0) {
    __syncthreads();
    lds_fft<DIR, MAXB, (1 << LOG)>(lds, ny, TX, ls, tws, 1, GlobalLines{const_cast<float2*>(in), lb});
    float2* dst = out + ((size_t)(c * M1 + kz) * ny) * nx + x0;
    tile2_from_lds<U, LOG>(dst, lds, TileShape{ny, lTX, 1, 1, ls}, StridedLines{nx});
  } else {
    const float2* src = in + ((size_t)(c * M1 + kz) * ny) * nx + x0;
    tile2_to_lds<U, LOG>(lds, src, TileShape{ny, lTX, 1, 1, ls}, StridedLines{nx});
    __syncthreads();
    lds_fft<DIR, MAXB, (1 << LOG)>(lds, ny, TX, ls, tws, 1, LdsLines{}, GlobalLines{out, lb});
  }
}

// Exchange slot of (k1, n2) in a 256-point line of the two-pass FFTs (rows
// of 16 padded to 17).
__device__ __forceinline__ int ks2_slot(int k1, int n2) { return 17 * k1 + n2; }

// W256^(DIR m j) for j = 1..15 from four table reads (tw[m] = exp(-2 pi i m / 256)).
template <int DIR>
__device__ __forceinline__ void ks2_twiddle(float2 (&v)[16], const float2* tw, int m) {
  float2 w1 = tw[m & 255], w2 = tw[(2 * m) & 255], w4 = tw[(4 * m) & 255], w8 = tw[(8 * m) & 255];
  if (DIR > 0) { w1 = conjf2(w1); w2 = conjf2(w2); w4 = conjf2(w4); w8 = conjf2(w8); }
  const float2 w3 = cmul(w1, w2);
  float2 lo[8];
  lo[0] = make_float2(1.f, 0.f); lo[1] = w1; lo[2] = w2; lo[3] = w3;
  lo[4] = w4; lo[5] = cmul(w4, w1); lo[6] = cmul(w4, w2); lo[7] = cmul(w4, w3);
#pragma unroll
  for (int j = 1; j < 8; ++j) v[j] = cmul(v[j], lo[j]);
  v[8] = cmul(v[8], w8);
#pragma unroll
  for (int j = 9; j < 16; ++j) v[j] = cmul(v[j], cmul(w8, lo[j - 8]));
}

// The y passes for ny = 256: NL = 16 or 32 y lines per block of 16 NL
// threads, the two-pass (16 x 16) line FFT of the k-space kernel (pass A
// reads the y lines from global memory, one LDS exchange, pass B), and the
// transposed side (x contiguous, NL x values per ky row) through an LDS tile
// T of rows of NL, rotated by the row index,
//   T(ky, l) = NL ky + ((l + ky) mod NL):
// LDS serves a 64-bit access 16 lanes at a time over 32 banks, so 16 lanes
// are conflict-free when their float2 indices differ modulo 16; the pass-B
// stores (16 lanes over k1) and the row-wise copies (16 lanes over l) both
// are.  T takes the place of the lines.
//
// XS (split x FFT, nx = ny = 256): the block also runs the first stage of
// the 256-point x FFT (256 = 16 x 16, x = 16 n1 + n2) on its tile.  Block b
// takes the x lines n2 in [H b, H b + H), H = NL / 16, all n1 (tile column
// l = 16 (n2 - H b) + n1), and the rows are stored in the order 16 n2 + k1
// (tile column 16 (n2 - H b) + k1 at row position NL b + column):
//   forward  Y[k1, n2] = W256^(n2 k1) sum_n1 W16^(n1 k1) x[16 n1 + n2],
//   inverse  x[16 n1 + n2] = sum_k1 W16^-(n1 k1) W256^-(n2 k1) Y[k1, n2],
// and the k-space kernel finishes (or starts) the transform with one
// 16-point FFT over n2 per (row, k1), reading 16 coalesced values.  The
// stage runs in place on T, one thread per (row, n2): 16 lanes over
// consecutive rows hit distinct banks.
template <int NL>
__device__ __forceinline__ int y2_tile(int ky, int l) { return NL * ky + ((l + ky) & (NL - 1)); }

template <int DIR, int NL>
__device__ __forceinline__ void y2_x_stage(float2* T, const float2* tws, int b) {
  constexpr int NH = NL / 16;  // n2 values per block
  const int t = threadIdx.x, ky = t & 255, h = t >> 8;
  float2 v[16];
#pragma unroll
  for (int i = 0; i < 16; ++i) v[i] = T[y2_tile<NL>(ky, 16 * h + i)];
  if (DIR < 0) {
    fft16r<-1>(v);                      // over n1 -> k1
    ks2_twiddle<-1>(v, tws, NH * b + h);  // W256^(n2 k1)
  } else {
    ks2_twiddle<1>(v, tws, NH * b + h);   // W256^-(n2 k1)
    fft16r<1>(v);                       // over k1 -> n1
  }
#pragma unroll
  for (int i = 0; i < 16; ++i) T[y2_tile<NL>(ky, 16 * h + i)] = v[i];
}

template <int DIR, int NL, bool XS = false>
__device__ __forceinline__ void c2c_y2_body(const BlockInfo& bk, const YPassArgs& ya) {
  static_assert(NL == 16 || NL == 32, "y2: 16 or 32 lines");
  constexpr int NT = 16 * NL, LR = NL == 16 ? 4 : 5;
  extern __shared__ float2 lds[];
  const float2* __restrict__ in = ya.in;
  float2* __restrict__ out = ya.out;
  const int nx = XS ? 256 : ya.nx, M1 = ya.M1;
  constexpr int ny = 256;
  const int x0 = bk.id.x * NL, kz = bk.id.y, c = bk.id.z;
  float2* S = lds;  // NL lines of kKs2Ls, or the tile T
  float2* tws = lds + NL * kKs2Ls;
  stage_table(tws, ya.twy, ny);
  const int t = threadIdx.x, l = t >> 4, j = t & 15;
  // x of this thread's y line: x0 + l, or with XS 16 n1 + n2 (l = 16 (n2 - 2b) + n1)
  const int xl = XS ? 16 * (l & 15) + (NL / 16) * bk.id.x + (l >> 4) : x0 + l;
  const long long yline = ((long long)(c * nx + xl) * M1 + kz) * ny;  // this thread's y line
  // rows ky = (t >> LR) + 16 it of the transposed side, x = x0 + (t mod NL)
  const size_t row0 = ((size_t)(c * M1 + kz) * ny + (t >> LR)) * nx + x0 + (t & (NL - 1));
  const int ty = t >> LR, tl = t & (NL - 1);  // tile row ty + 16 it, column tl
  static_assert(NT == 16 * NL, "");
  if (DIR < 0) {
    __syncthreads();  // twiddles staged
    float2 v[16];
    const float2* src = in + yline + j;
#pragma unroll
    for (int n1 = 0; n1 < 16; ++n1) v[n1] = FD3_LD_NT(src + 16 * n1);  // streaming: read once
    fft16r<-1>(v);  // n2 = j
    ks2_twiddle<-1>(v, tws, j);
    float2* L = S + l * kKs2Ls;
#pragma unroll
    for (int k1 = 0; k1 < 16; ++k1) L[ks2_slot(k1, j)] = v[k1];
    __syncthreads();
    float2 z[16];
#pragma unroll
    for (int n2 = 0; n2 < 16; ++n2) z[n2] = L[ks2_slot(j, n2)];  // k1 = j
    fft16r<-1>(z);  // X[j + 16 k2]
    __syncthreads();  // the lines are read: T takes their place
#pragma unroll
    for (int k2 = 0; k2 < 16; ++k2) S[y2_tile<NL>(j + 16 * k2, l)] = z[k2];
    __syncthreads();
    if constexpr (XS) {
      y2_x_stage<-1, NL>(S, tws, bk.id.x);
      __syncthreads();
    }
    float2 w[16];
#pragma unroll
    for (int it = 0; it < 16; ++it) w[it] = S[y2_tile<NL>(ty + 16 * it, tl)];
#pragma unroll
    for (int it = 0; it < 16; ++it) out[row0 + (size_t)16 * it * nx] = w[it];
  } else {
    float2 w[16];
#pragma unroll
    for (int it = 0; it < 16; ++it) w[it] = FD3_LD_NT(in + row0 + (size_t)16 * it * nx);  // streaming
#pragma unroll
    for (int it = 0; it < 16; ++it) S[y2_tile<NL>(ty + 16 * it, tl)] = w[it];
    __syncthreads();  // T filled, twiddles staged
    if constexpr (XS) {
      y2_x_stage<1, NL>(S, tws, bk.id.x);
      __syncthreads();
    }
    float2 z[16];
#pragma unroll
    for (int k2 = 0; k2 < 16; ++k2) z[k2] = S[y2_tile<NL>(j + 16 * k2, l)];  // X[j + 16 k2] of line l
    fft16r<1>(z);  // over k2 -> n2
    ks2_twiddle<1>(z, tws, j);
    __syncthreads();  // T is read: the lines take its place
    float2* L = S + l * kKs2Ls;
#pragma unroll
    for (int n2 = 0; n2 < 16; ++n2) L[ks2_slot(j, n2)] = z[n2];
    __syncthreads();
    float2 v[16];
#pragma unroll
    for (int k1 = 0; k1 < 16; ++k1) v[k1] = L[ks2_slot(k1, j)];  // n2 = j
    fft16r<1>(v);  // x[j + 16 n1]
    float2* dst = out + yline + j;
#pragma unroll
    for (int n1 = 0; n1 < 16; ++n1) FD3_ST_NT(dst + 16 * n1, v[n1]);
  }
}

// ------------------------------------------------------------ x pass ---
// In place on contiguous lines of length nx (spec layout), LPB lines per block.
template <int DIR>
__global__ void __launch_bounds__(256) k_c2c_x(float2* __restrict__ spec, int nx, int nlines, int LPB,
                                               const float2* __restrict__ twx) {
  extern __shared__ float2 lds[];
  const int ls = nx + 1;
  const size_t l0 = (size_t)blockIdx.x * LPB;
  const int nl = min(LPB, nlines - (int)l0);
  float2* base = spec + l0 * nx;
  float2* tws = lds + LPB * ls;
  stage_table(tws, twx, nx);
  const int lnx = ilog2(nx);
  const TileShape ts{nl, lnx, 1, ls, 1};
  tile2_to_lds<kCopyU>(lds, base, ts, StridedLines{nx});
  __syncthreads();
  lds_fft<DIR, kFftMaxB, 256>(lds, nx, nl, ls, tws, 1);
  tile2_from_lds<kCopyU>(base, lds, ts, StridedLines{nx});
}

// ------------------------------------------------- fused k-space kernel ---
// For one kz plane, NK consecutive ky rows and RT ranks [q0, q0+RT) of every
// species (a tile): forward x FFT of the nS*RT*NK lines, then per k-point
//   phi_X^q = sum_Y G~_XY S_Y^q,  G~ = G(|k|) / sinc^2p (P3M only)
//   e      += w sum_X sum_q lambda_q Re(S_X^q conj(phi_X^q))
//   dsum   += w sum_XY dG~_XY/d|k| sum_q lambda_q Re(S_X^q conj(S_Y^q))
// (w = 1 on the kz = 0 and nz/2 planes, 2 elsewhere: the half-spectrum weights
// of src/fourierd3/pair_pot.py:209-218), writes -(lambda_q/Omega) phi back and
// inverse-transforms along x, leaving dE/drho's spectrum in place.  A block
// runs KT tiles (consecutive ky rows) and reduces once.
// Per-block partial sums (double): [0] sum e, [1..6] sum w dsum k_a k_b/|k|
// (xx,yy,zz,xy,xz,yz), [7..10] dE/d(s6,s8,a1,a2) spectral parts (unscaled).
// G~ is evaluated once per unordered species pair and shared by all ranks.
struct KspaceArgs {
  float2* spec;
  const float* eigs;
  const KsPair* pc;     // (nS, nS)
  const float2* twx;
  const float4* htab;
  const float4* gtab;   // per-pair G tables (register kernel, no parameter gradients)
  float2* sink;         // 256 values written by idle lanes (kspace_xsp_body), never read
  float gkinv;          // 1 / (interval width) of gtab, Bohr
  const float* scx; const float* scy; const float* scz;
  double* partials;     // (gridDim, 16)
  // dE/drho's k = 0 coefficient per channel.  It is a constant on the mesh:
  // it adds exactly itself to every atom's sum_nodes W dE/drho and nothing to
  // the force sums (sum_nodes dW/du = 0), so it is passed to the C6 adjoint
  // directly and zeroed in the spectrum, which keeps its float rounding out of
  // the forces.
  float2* k0;           // (C)
  float recip[9];       // rows b_i of 2 pi cell^-T (Bohr^-1)
  float neg_inv_vol;    // -1/Omega
  const CellDyn* dyn = nullptr;  // when set, recip and neg_inv_vol are read from it
  int nx, ny, nz, M1, nS, r, RT, NK, KT, use_sinc, need_grad, need_params;
};

// Reciprocal vectors and volume from the device cell when there is one
// (wave-uniform: kept in scalar registers like the by-value arguments).
__device__ inline float uniform_f(float v) {
  return __int_as_float(__builtin_amdgcn_readfirstlane(__float_as_int(v)));
}

struct KsCell {
  float recip[9];
  float neg_inv_vol;
};

// Only the cell-dependent values are copied out of the arguments (the rest
// stay in the argument block and load as scalars where they are used).
__device__ inline KsCell ks_cell(const KspaceArgs& a) {
  KsCell c;
  for (int i = 0; i < 9; ++i) c.recip[i] = a.dyn ? uniform_f(a.dyn->recip_f[i]) : a.recip[i];
  c.neg_inv_vol = a.dyn ? uniform_f(a.dyn->neg_inv_vol_f) : a.neg_inv_vol;
  return c;
}

// Dynamic LDS of both k-space kernels: [h table | x twiddles | lines ... |
// line bases], the table first so that its float4 entries are 16-byte aligned.
__device__ inline void kspace_stage_consts(const KspaceArgs& a, float4* htab, float2* tws) {
  for (int i = threadIdx.x; i < 3 * (kHtabN + 1); i += blockDim.x) htab[i] = a.htab[i];
  stage_table(tws, a.twx, a.nx);
}

// Element offset in spec of every tile line l = (X*RT + qi)*NK + kyi, or -1
// for the lines of ranks past the last (qi >= rt).
__device__ inline void kspace_line_bases(const KspaceArgs& a, long long* lbase, int nlines, int RT, int rt, int ky0,
                                         int kz, int q0) {
  for (int l = threadIdx.x; l < nlines; l += blockDim.x) {
    const int kyi = l % a.NK, xq = l / a.NK, qi = xq % RT, X = xq / RT;
    lbase[l] = qi < rt ? (((long long)(X * a.r + q0 + qi) * a.M1 + kz) * a.ny + ky0 + kyi) * a.nx : -1LL;
  }
}

// k vector, |k| and 1/sinc^2p of column (kx, ky).
struct KPoint {
  float kv0, kv1, kv2, kn, isc;
};
__device__ inline KPoint kspace_point(const KspaceArgs& a, const KsCell& c, int kx, int ky, int kz) {
  const float mx = (float)(kx < (a.nx + 1) / 2 ? kx : kx - a.nx);
  const float my = (float)(ky < (a.ny + 1) / 2 ? ky : ky - a.ny);
  const float mz = (float)kz;
  KPoint p;
  p.kv0 = mx * c.recip[0] + my * c.recip[3] + mz * c.recip[6];
  p.kv1 = mx * c.recip[1] + my * c.recip[4] + mz * c.recip[7];
  p.kv2 = mx * c.recip[2] + my * c.recip[5] + mz * c.recip[8];
  p.kn = __builtin_amdgcn_sqrtf(p.kv0 * p.kv0 + p.kv1 * p.kv1 + p.kv2 * p.kv2);
  p.isc = 1.0f;
  if (a.use_sinc) p.isc = __builtin_amdgcn_rcpf(fmaxf(a.scx[kx] * a.scy[ky] * a.scz[kz], 1e-10f));
  return p;
}

// Derivative of G with respect to (s6, s8, a1, a2) at one pair, times t.
__device__ inline void kspace_param_terms(const KsPair& pk, float kn, float h6, float d6, float h8, float d8, float t,
                                          float* p_acc) {
  const float R = pk.Rs * kHtabDx;
  const float h6p = d6 * kHtabInvDx, h8p = d8 * kHtabInvDx;
  const float dgdR = pk.p6 * (-3.0f * h6 / R + kn * h6p) + pk.p8 * (-5.0f * h8 / R + kn * h8p);
  p_acc[0] += pk.pf6 * h6 * t;
  p_acc[1] += pk.pf8 * h8 * t;
  p_acc[2] += dgdR * pk.dR * t;
  p_acc[3] += dgdR * t;
}

// e_acc, d_acc and p_acc are sums over G, not G~: the 1/sinc^2p factor of
// the k point is applied here.
__device__ inline void kspace_accumulate(double (&acc)[11], const KPoint& p, float w, float e_acc, float d_acc,
                                         const float* p_acc, bool params) {
  acc[0] += (double)(w * e_acc * p.isc);
  if (p.kn > 0.f) {
    const double s = (double)(w * d_acc * p.isc * __builtin_amdgcn_rcpf(p.kn));
    acc[1] += s * p.kv0 * p.kv0; acc[2] += s * p.kv1 * p.kv1; acc[3] += s * p.kv2 * p.kv2;
    acc[4] += s * p.kv0 * p.kv1; acc[5] += s * p.kv0 * p.kv2; acc[6] += s * p.kv1 * p.kv2;
  }
  if (params) {
#pragma unroll
    for (int i = 0; i < 4; ++i) acc[7 + i] += (double)(w * p_acc[i] * p.isc);
  }
}

// The same in float, for a thread's own partial sums over its few k points
// (a handful of terms; the block sums are double).
__device__ inline void kspace_accumulate_f(float (&acc)[11], const KPoint& p, float w, float e_acc, float d_acc,
                                           const float* p_acc, bool params) {
  acc[0] += w * e_acc * p.isc;
  if (p.kn > 0.f) {
    const float s = w * d_acc * p.isc * __builtin_amdgcn_rcpf(p.kn);
    acc[1] += s * p.kv0 * p.kv0; acc[2] += s * p.kv1 * p.kv1; acc[3] += s * p.kv2 * p.kv2;
    acc[4] += s * p.kv0 * p.kv1; acc[5] += s * p.kv0 * p.kv2; acc[6] += s * p.kv1 * p.kv2;
  }
  if (params) {
#pragma unroll
    for (int i = 0; i < 4; ++i) acc[7 + i] += w * p_acc[i] * p.isc;
  }
}

// Block sums of the 7 (or, with parameter gradients, 11) partials; the
// unused parameter slots are written as zeros.  Ends with a barrier.
__device__ inline void kspace_reduce(const BlockInfo& bk, double (&acc)[11], double* red, const KspaceArgs& a,
                                     bool params) {
  const int bid = (bk.id.z * bk.n.y + bk.id.y) * bk.n.x + bk.id.x;
  double* out = a.partials + (size_t)bid * 16;
  if (params) {
    block_sum_store<11>(acc, red, out);
  } else {
    double v[7];
#pragma unroll
    for (int i = 0; i < 7; ++i) v[i] = acc[i];
    block_sum_store<7>(v, red, out);
    if (threadIdx.x >= 7 && threadIdx.x < 11) out[threadIdx.x] = 0.0;
  }
}

// Two-pass x FFT of the k-space kernel for nx = 256 (256 = 16 x 16), 16
// lines of stride kKs2Ls in LDS, 256 threads, one (line, column) task each.
// Pass A: thread (l, n2) reads x[16 n1 + n2] (n1 < 16) of its line from
// global memory, does a 16-point FFT in registers and multiplies by
// W256^(n2 k1); pass B: thread (l, k1) gathers the 16 values of its k1 from
// the exchange slots 17 k1 + n2 and does the second 16-point FFT, giving
// X[k1 + 16 k2], stored in natural order.  The inverse mirrors it and writes
// the lines back to global memory.  One LDS exchange per direction instead
// of one per radix-4 stage.  LDS serves a 64-bit access 16 lanes at a time
// over 32 banks (conflict-free when the float2 indices differ modulo 16);
// the 16 lanes of a group share a line and differ in j, and with the slot
// rows padded to 17 (slot 17 k1 + n2) both passes' accesses are
// conflict-free, each thread's 16 addresses one base plus immediate offsets.
// lbase[l] < 0: a dead line (read as zeros, not written).  Lines >= nlines
// idle but take part in the barriers.

// Forward: global lines -> spectra in natural order in LDS.  Ends with a barrier.
__device__ __forceinline__ void ks2_forward(float2* S, const float2* tw, const float2* __restrict__ g,
                                            const long long* lbase, int nlines) {
  const int t = threadIdx.x, l = t >> 4, j = t & 15;
  float2* L = S + l * kKs2Ls;
  const bool on = l < nlines;
  if (on) {
    const long long b = lbase[l];
    const float2* src = g + (b >= 0 ? b : 0) + j;
    float2 v[16];
#pragma unroll
    for (int n1 = 0; n1 < 16; ++n1) v[n1] = src[16 * n1];
    if (b < 0)
#pragma unroll
      for (int n1 = 0; n1 < 16; ++n1) v[n1] = make_float2(0.f, 0.f);
    fft16r<-1>(v);  // over n1 -> k1 (n2 = j)
    ks2_twiddle<-1>(v, tw, j);
#pragma unroll
    for (int k1 = 0; k1 < 16; ++k1) L[ks2_slot(k1, j)] = v[k1];
  }
  __syncthreads();
  float2 z[16];
  if (on) {
#pragma unroll
    for (int n2 = 0; n2 < 16; ++n2) z[n2] = L[ks2_slot(j, n2)];  // k1 = j
    fft16r<-1>(z);  // over n2 -> k2: X[j + 16 k2]
  }
  __syncthreads();  // every slot read is done
  if (on)
#pragma unroll
    for (int k2 = 0; k2 < 16; ++k2) L[j + 16 * k2] = z[k2];
  __syncthreads();
}

// Inverse (unnormalised): spectra in natural order in LDS -> global lines.
__device__ __forceinline__ void ks2_inverse(float2* S, const float2* tw, float2* __restrict__ g,
                                            const long long* lbase, int nlines) {
  const int t = threadIdx.x, l = t >> 4, j = t & 15;
  float2* L = S + l * kKs2Ls;
  const bool on = l < nlines;
  float2 z[16];
  if (on) {
#pragma unroll
    for (int k2 = 0; k2 < 16; ++k2) z[k2] = L[j + 16 * k2];  // k1 = j
    fft16r<1>(z);  // over k2 -> n2
    ks2_twiddle<1>(z, tw, j);  // W256^-(k1 n2)
  }
  __syncthreads();  // every natural-order read is done
  if (on)
#pragma unroll
    for (int n2 = 0; n2 < 16; ++n2) L[ks2_slot(j, n2)] = z[n2];
  __syncthreads();
  if (on) {
    float2 v[16];
#pragma unroll
    for (int k1 = 0; k1 < 16; ++k1) v[k1] = L[ks2_slot(k1, j)];  // n2 = j
    fft16r<1>(v);  // over k1 -> n1: x[j + 16 n1]
    const long long b = lbase[l];
    if (b >= 0) {
      float2* dst = g + b + j;
#pragma unroll
      for (int n1 = 0; n1 < 16; ++n1) dst[16 * n1] = v[n1];
    }
  }
}

// One k point (column kx of tile row kyi, element kxe of a line) of the
// register k-space kernels: G~ of every species pair, phi_X = sum_Y G~ S_Y
// over the NS*RT spectra of the point (read from the tile lines in S,
// overwritten with -(lambda_q/Omega) phi), and the energy, virial and
// parameter partials added to acc (float: a handful of terms per thread).
// Without parameter gradients G comes from the per-pair tables gt (one index
// per k point, one Horner pass per pair, the coefficients of the next PD
// pairs in flight while a pair is evaluated); with them from the h table
// tab and the pair constants pcl, which also give h6 and h8 separately.
template <int NS, int RT, bool PARAMS>
__device__ __forceinline__ void kspace_kpoint(const KspaceArgs& a, const KsCell& kc, float2* S, int ls, int NK,
                                              int kyi, int kx, int kxe, int ky, int kz, const float (&lam)[RT],
                                              float w, const KsPair* pcl, const float4* tab, const float4* gt,
                                              float (&accf)[11]) {
  constexpr int NP = NS * (NS + 1) / 2;  // unordered species pairs
  const KPoint kp = kspace_point(a, kc, kx, ky, kz);
  // G and dG/d|k| (or dG/dw) of every pair first, then the spectra
  float gp[NP], dgp[NP], p_acc[4] = {0.f, 0.f, 0.f, 0.f};
  float h6v[PARAMS ? NP : 1], d6v[PARAMS ? NP : 1], h8v[PARAMS ? NP : 1], d8v[PARAMS ? NP : 1];
  if constexpr (PARAMS) {
#pragma unroll
    for (int pi = 0; pi < NP; ++pi) {
      float h6, d6, h8, d8;
      green_eval_tab(pcl[pi], tab, kp.kn, &gp[pi], &dgp[pi], &h6, &d6, &h8, &d8);
      h6v[pi] = h6; d6v[pi] = d6; h8v[pi] = h8; d8v[pi] = d8;
      asm volatile("" : "+v"(h6v[pi]), "+v"(d6v[pi]), "+v"(h8v[pi]), "+v"(d8v[pi]));
      // pin g and dG/dk here: left to itself the compiler defers the
      // derivative and keeps every pair's Horner intermediates live
      asm volatile("" : "+v"(gp[pi]), "+v"(dgp[pi]));
    }
  } else {
    const float fk = kp.kn * a.gkinv;
    const int i = min((int)fk, kGtabN - 1);
    const float wk = fk - (float)i;
    const float4* e = gt + 2 * i;
    constexpr int PD = 2;
    float4 ca[PD + 1], cb[PD + 1];
    static_for<PD>([&](auto dc) {
      constexpr int d = decltype(dc)::value;
      if constexpr (d < NP) { ca[d] = e[2 * kGtabN * d]; cb[d] = e[2 * kGtabN * d + 1]; }
    });
    static_for<NP>([&](auto pc) {
      constexpr int pi = decltype(pc)::value, nxt = pi + PD;
      if constexpr (nxt < NP) {
        ca[nxt % (PD + 1)] = e[2 * kGtabN * nxt];
        cb[nxt % (PD + 1)] = e[2 * kGtabN * nxt + 1];
      }
      gtab_eval(ca[pi % (PD + 1)], cb[pi % (PD + 1)], wk, &gp[pi], &dgp[pi]);
      asm volatile("" : "+v"(gp[pi]), "+v"(dgp[pi]));
    });
  }
  float2 sv[NS][RT], ph[NS][RT];
#pragma unroll
  for (int X = 0; X < NS; ++X)
#pragma unroll
    for (int q = 0; q < RT; ++q) {
      sv[X][q] = S[((X * RT + q) * NK + kyi) * ls + kxe];
      ph[X][q] = make_float2(0.f, 0.f);
    }
  float d_acc = 0.f;
#pragma unroll
  for (int X = 0, pi = 0; X < NS; ++X) {
#pragma unroll
    for (int Y = X; Y < NS; ++Y, ++pi) {
      const float gs = gp[pi];
      float tre = 0.f;  // sum_q lambda_q Re(S_X conj S_Y)
#pragma unroll
      for (int q = 0; q < RT; ++q) {
        const float2 sx = sv[X][q], sy = sv[Y][q];
        tre += lam[q] * (sx.x * sy.x + sx.y * sy.y);
        ph[X][q].x += gs * sy.x;
        ph[X][q].y += gs * sy.y;
        if (Y != X) {
          ph[Y][q].x += gs * sx.x;
          ph[Y][q].y += gs * sx.y;
        }
      }
      if (Y != X) tre *= 2.0f;
      d_acc += dgp[pi] * tre;
      if (PARAMS) kspace_param_terms(pcl[pi], kp.kn, h6v[pi], d6v[pi], h8v[pi], d8v[pi], tre, p_acc);
    }
  }
  if (!PARAMS) d_acc *= a.gkinv;  // dG/dw -> dG/d|k|
  // ph holds sum_Y G S_Y; the 1/sinc^2p of G~ goes into f and the accumulation
  float e_acc = 0.f, f[RT];
#pragma unroll
  for (int q = 0; q < RT; ++q) f[q] = lam[q] * kc.neg_inv_vol * kp.isc;
#pragma unroll
  for (int X = 0; X < NS; ++X)
#pragma unroll
    for (int q = 0; q < RT; ++q) {
      e_acc += lam[q] * (sv[X][q].x * ph[X][q].x + sv[X][q].y * ph[X][q].y);
      S[((X * RT + q) * NK + kyi) * ls + kxe] = make_float2(f[q] * ph[X][q].x, f[q] * ph[X][q].y);
    }
  kspace_accumulate_f(accf, kp, w, e_acc, d_acc, p_acc, PARAMS);
}

// The k = 0 coefficient of each spectrum of the tile (column 0 of row 0)
// goes to k0 and is zeroed in the mesh (thread 0 of the tile holding it).
template <int NS, int RT>
__device__ __forceinline__ void kspace_take_k0(const KspaceArgs& a, float2* S, int ls, int NK, int q0, int rt) {
  for (int l = 0; l < NS * RT; ++l) {
    float2* sp = &S[l * NK * ls];
    const int X = l / RT, q = l - X * RT;
    if (q < rt) a.k0[X * a.r + q0 + q] = *sp;
    *sp = make_float2(0.f, 0.f);
  }
}

// Pair constants of the NP unordered species pairs (X, Y >= X) in row order.
template <int NS>
__device__ __forceinline__ void kspace_stage_pairs(const KspaceArgs& a, KsPair* pcl) {
  constexpr int NP = NS * (NS + 1) / 2;
  for (int t = threadIdx.x; t < NP; t += blockDim.x) {
    int X = 0, r0 = 0;
    while (t >= r0 + NS - X) { r0 += NS - X; ++X; }
    pcl[t] = a.pc[X * NS + X + (t - r0)];
  }
}

// nS = NS <= kKsRegMaxS species and RT ranks per tile: each thread keeps the
// NS*RT spectra of its k point in registers and writes phi back over them,
// so the block needs a single LDS buffer.  Lines l = (X*RT + qi)*NK + kyi.
// LDS: [tables | x twiddles | lines | line bases]; the block sums reuse the
// lines at the end.
// TWO: nx = 256 with the two-pass x FFT (ks2_forward / ks2_inverse), NK = 2.
template <int NS, int RT, bool PARAMS, bool TWO = false>
__device__ __forceinline__ void kspace_reg_body(const BlockInfo& bk, const KspaceArgs& a) {
  const KsCell kc = ks_cell(a);
  constexpr int MAXB = ks_maxb(NS * RT);
  constexpr int NP = NS * (NS + 1) / 2;  // unordered species pairs
  extern __shared__ float4 lds4[];
  __shared__ KsPair pcl[PARAMS ? NP : 1];
  const int nx = TWO ? 256 : a.nx, NK = TWO ? 2 : a.NK;
  const int kz = bk.id.y, q0 = bk.id.z * RT;
  const int rt = min(RT, a.r - q0);
  const int nlines = NS * RT * NK;
  // padded lines when lds_fft takes its unrolled variant (same rule on the host)
  const bool pad = !TWO && fft_fixed_tile(nx, nlines, MAXB, 256);
  const int ls = TWO ? kKs2Ls : pad ? padded_stride(nx) : nx + 1;
  const int lnx = ilog2(nx);  // nx a power of two
  // GT: the two-pass variant reads the per-pair G tables straight from
  // global memory (L1-resident) instead of staging them: its lines need 35 KB
  // of LDS, and without the tables four blocks fit a CU
  constexpr bool GT = TWO && !PARAMS;
  float4* tab = lds4;
  const int ntab = PARAMS ? 3 * (kHtabN + 1) : 2 * NP * kGtabN;
  float2* tws = (float2*)(tab + (TWO ? ks2_table_f4(PARAMS) : ks_table_f4(NS)));
  float2* S = tws + nx;
  long long* lbase = (long long*)(S + nlines * ls);
  const float4* src_tab = PARAMS ? a.htab : a.gtab;
  if (!GT)
    for (int i = threadIdx.x; i < ntab; i += blockDim.x) tab[i] = src_tab[i];
  const float4* gt = GT ? a.gtab : tab;
  stage_table(tws, a.twx, a.nx);
  if (PARAMS) kspace_stage_pairs<NS>(a, pcl);

  float lam[RT];
#pragma unroll
  for (int q = 0; q < RT; ++q) lam[q] = q < rt ? a.eigs[q0 + q] : 0.0f;
  // each thread's partial sums over its k points: without parameter
  // gradients in float per tile (NK * nx / 256 terms) and in double over the
  // tiles; with them (11 sums) in float over the block's tiles, at most
  // KT * NK * nx / 256 <= 8 terms, which keeps the kernel at three waves per
  // SIMD; the block sums in double
  double acc[11];
#pragma unroll
  for (int i = 0; i < 11; ++i) acc[i] = 0.0;
  float accf[11];
#pragma unroll
  for (int i = 0; i < 11; ++i) accf[i] = 0.0f;
  const float w = (kz == 0 || kz == a.M1 - 1) ? 1.0f : 2.0f;
  for (int it = 0; it < a.KT; ++it) {
    if constexpr (!PARAMS) {
#pragma unroll
      for (int i = 0; i < 7; ++i) accf[i] = 0.0f;
    }
    const int ky0 = (bk.id.x * a.KT + it) * NK;
    if (it) __syncthreads();  // the previous tile is done with S and lbase
    kspace_line_bases(a, lbase, nlines, RT, rt, ky0, kz, q0);
    __syncthreads();
    if constexpr (TWO) ks2_forward(S, tws, a.spec, lbase, nlines);
    else lds_fft<-1, MAXB, 256>(S, nx, nlines, ls, tws, 1, GlobalLines{a.spec, lbase});
    for (int t = threadIdx.x; t < NK * nx; t += blockDim.x) {
      const int kyi = t >> lnx, kx = t & (nx - 1);
      const int kxe = pad ? lpad(kx) : kx;  // element kx of a line
      kspace_kpoint<NS, RT, PARAMS>(a, kc, S, ls, NK, kyi, kx, kxe, ky0 + kyi, kz, lam, w, pcl, tab, gt, accf);
    }
    if constexpr (!PARAMS) {
#pragma unroll
      for (int i = 0; i < 7; ++i) acc[i] += (double)accf[i];
    }
    if (kz == 0 && ky0 == 0 && threadIdx.x == 0) kspace_take_k0<NS, RT>(a, S, ls, NK, q0, rt);
    if (!a.need_grad) continue;
    __syncthreads();
    if constexpr (TWO) ks2_inverse(S, tws, a.spec, lbase, nlines);
    else lds_fft<1, MAXB, 256>(S, nx, nlines, ls, tws, 1, LdsLines{}, GlobalLines{a.spec, lbase});
  }
  if constexpr (PARAMS) {
#pragma unroll
    for (int i = 0; i < 11; ++i) acc[i] = (double)accf[i];
  }
  __syncthreads();  // the lines are free: reuse them for the block sums
  kspace_reduce(bk, acc, (double*)S, a, PARAMS);
}

// The spectra part of one k point (column kx of tile row kyi) given G and
// dG/dw of its species pairs: phi_X = sum_Y G S_Y rank by rank (the NS
// spectra of one rank in registers at a time), S <- -(lambda_q/Omega) phi /
// sinc^2p, and the energy and virial partials added to acc.
template <int NS, int RT, int NK>
__device__ __forceinline__ void ks_kpoint_spectra(const KspaceArgs& a, const KsCell& kc, float2* S, int kyi, int kx,
                                                  const KPoint& kp, const float (&gp)[NS * (NS + 1) / 2],
                                                  const float (&dgp)[NS * (NS + 1) / 2], const float (&lam)[RT],
                                                  float w, float (&accf)[7]) {
  constexpr int ls = 256;
  float e_acc = 0.f, d_acc = 0.f;
#pragma unroll
  for (int q = 0; q < RT; ++q) {
    float2 sv[NS], ph[NS];
#pragma unroll
    for (int X = 0; X < NS; ++X) {
      sv[X] = S[((X * RT + q) * NK + kyi) * ls + kx];
      ph[X] = make_float2(0.f, 0.f);
    }
    float dq = 0.f;  // sum over pairs of dG/dw Re(S_X conj S_Y) (both orders)
#pragma unroll
    for (int X = 0, pi = 0; X < NS; ++X) {
#pragma unroll
      for (int Y = X; Y < NS; ++Y, ++pi) {
        const float gs = gp[pi];
        const float2 sx = sv[X], sy = sv[Y];
        const float tre = sx.x * sy.x + sx.y * sy.y;
        ph[X].x += gs * sy.x;
        ph[X].y += gs * sy.y;
        if (Y != X) {
          ph[Y].x += gs * sx.x;
          ph[Y].y += gs * sx.y;
          dq += (2.0f * dgp[pi]) * tre;
        } else {
          dq += dgp[pi] * tre;
        }
      }
    }
    float eq = 0.f;
    const float f = lam[q] * kc.neg_inv_vol * kp.isc;
#pragma unroll
    for (int X = 0; X < NS; ++X) {
      eq += sv[X].x * ph[X].x + sv[X].y * ph[X].y;
      S[((X * RT + q) * NK + kyi) * ls + kx] = make_float2(f * ph[X].x, f * ph[X].y);
    }
    e_acc += lam[q] * eq;
    d_acc += lam[q] * dq;
    asm volatile("" ::: "memory");  // one rank's spectra in registers at a time
  }
  d_acc *= a.gkinv;  // dG/dw -> dG/d|k|
  accf[0] += w * e_acc * kp.isc;
  if (kp.kn > 0.f) {
    const float sv = w * d_acc * kp.isc * __builtin_amdgcn_rcpf(kp.kn);
    accf[1] += sv * kp.kv0 * kp.kv0; accf[2] += sv * kp.kv1 * kp.kv1; accf[3] += sv * kp.kv2 * kp.kv2;
    accf[4] += sv * kp.kv0 * kp.kv1; accf[5] += sv * kp.kv0 * kp.kv2; accf[6] += sv * kp.kv1 * kp.kv2;
  }
}

// The register k-space kernel with the split x FFT (nx = ny = 256, the y
// passes run the first stage): blocks of 256 threads over a 1-D grid, block b
// taking the KT consecutive tiles b KT .. b KT + KT - 1 of the
// (ny/2) x M1 x ceil(r/RT) tiles of two ky rows x NS*RT spectra, ky pair
// fastest.  Thread (l, j) owns column j of line l = (X*RT + qi)*2 + kyi of
// every tile (the 16-point FFT of ksx_load / ksx_forward_regs /
// ksx_inverse); its load is issued before the barrier that frees the lines.
// Dynamic LDS: the 16 lines of 256 points (32 KB), then with parameter
// gradients the h table and the pair constants.  The block sums reuse the
// lines at the end.  Each thread adds its tile's two k points in float and
// the tile sums in double.  It runs with parameter gradients;
// kspace_xsp_body (below) without them.  Measured on the 100k-atom box
// (256^3, 4 species, rank 2, without parameter gradients): 0.267 ms against
// 0.274 ms for the same tiles with padded lines, twiddle and line-base
// tables in LDS (35 KB), 0.297 ms for persistent blocks striding over the
// tiles, 0.29 ms for one ky row per tile; streaming (nontemporal) loads of
// the spectrum cost 40 %.
template <int NS, int RT, bool PARAMS>
__device__ __forceinline__ void kspace_xs_body(const BlockInfo& bk, const KspaceArgs& a) {
  static_assert(NS * RT <= 8, "split x k-space: at most 8 spectra per tile");
  constexpr int NP = NS * (NS + 1) / 2, NK = 2, ls = 256, nlines = NS * RT * NK;
  const KsCell kc = ks_cell(a);
  extern __shared__ float4 lds4[];
  float2* S = (float2*)lds4;
  float4* tab = lds4 + nlines * ls / 2;
  KsPair* pcl = (KsPair*)(tab + (PARAMS ? 3 * (kHtabN + 1) : 0));
  if constexpr (PARAMS) {
    for (int i = threadIdx.x; i < 3 * (kHtabN + 1); i += blockDim.x) tab[i] = a.htab[i];
    kspace_stage_pairs<NS>(a, pcl);
  }
  const int nyp = a.ny >> 1, zb = (a.r + RT - 1) / RT;
  const int ntiles = nyp * a.M1 * zb;
  const int xl = threadIdx.x >> 4, xj = threadIdx.x & 15;
  const int xqi = (xl >> 1) % RT, xX = (xl >> 1) / RT;
  constexpr int NA = PARAMS ? 11 : 7;  // partial sums carried over the tiles
  double acc[11];
#pragma unroll
  for (int i = 0; i < 11; ++i) acc[i] = 0.0;

  const int u0 = bk.id.x * a.KT, u1 = min(ntiles, u0 + a.KT);
  for (int u = u0; u < u1; ++u) {
    const int kyp = u % nyp, rest = u / nyp, kz = rest % a.M1, q0 = rest / a.M1 * RT;
    const int rt = min(RT, a.r - q0), ky0 = 2 * kyp;
    const bool live = xl < nlines && xqi < rt;
    const long long e0 = (((long long)(xX * a.r + q0 + xqi) * a.M1 + kz) * a.ny + ky0 + (xl & 1)) * 256 + xj;
    float lam[RT];
#pragma unroll
    for (int q = 0; q < RT; ++q) lam[q] = q < rt ? a.eigs[q0 + q] : 0.0f;
    const float w = (kz == 0 || kz == a.M1 - 1) ? 1.0f : 2.0f;
    float accf[11];  // this tile's two k points in float, the tile sums in double
#pragma unroll
    for (int i = 0; i < 11; ++i) accf[i] = 0.0f;
    float2 v[16];
    ksx_load(v, a.spec, live, e0);
    __syncthreads();  // the previous tile (and the staged tables) are done
    ksx_forward_regs(S, ls, v, nlines);
    for (int t = threadIdx.x; t < NK * 256; t += blockDim.x) {
      const int kyi = t >> 8, kx = t & 255;
      kspace_kpoint<NS, RT, PARAMS>(a, kc, S, ls, NK, kyi, kx, kx, ky0 + kyi, kz, lam, w, pcl, tab, a.gtab, accf);
    }
#pragma unroll
    for (int i = 0; i < NA; ++i) acc[i] += (double)accf[i];
    if (kz == 0 && ky0 == 0 && threadIdx.x == 0) kspace_take_k0<NS, RT>(a, S, ls, NK, q0, rt);
    if (!a.need_grad) continue;
    __syncthreads();
    ksx_inverse(S, ls, a.spec, live, e0);
  }
  __syncthreads();  // the lines are free: reuse them for the block sums
  kspace_reduce(bk, acc, (double*)S, a, PARAMS);
}

// k vector and |k| of column (kx, ky, kz) with a given 1/sinc^2p.
__device__ inline KPoint kspace_point_isc(const KspaceArgs& a, const KsCell& c, int kx, int ky, int kz, float isc) {
  const float mx = (float)(kx < (a.nx + 1) / 2 ? kx : kx - a.nx);
  const float my = (float)(ky < (a.ny + 1) / 2 ? ky : ky - a.ny);
  const float mz = (float)kz;
  KPoint p;
  p.kv0 = mx * c.recip[0] + my * c.recip[3] + mz * c.recip[6];
  p.kv1 = mx * c.recip[1] + my * c.recip[4] + mz * c.recip[7];
  p.kv2 = mx * c.recip[2] + my * c.recip[5] + mz * c.recip[8];
  p.kn = __builtin_amdgcn_sqrtf(p.kv0 * p.kv0 + p.kv1 * p.kv1 + p.kv2 * p.kv2);
  p.isc = isc;
  return p;
}

// The split-x k-space kernel without parameter gradients (nx = ny = 256,
// the y passes run the first stage of the x FFT), software-pipelined over
// the block's KT tiles: each thread loads its 16 values of the next tile
// into registers right after the forward pass has moved the current tile
// into LDS, so that load is in flight while the current tile is filtered and
// transformed back.  The vector-memory counter retires in issue order, so
// nothing else in the loop may wait on it, and the compiler places its
// waits from what every path to them has issued:
//  * the per-pair G tables, the rank eigenvalues and the sinc factors are
//    staged in LDS once per block, so the loop's only vector-memory
//    operations are that load and the tile's stores;
//  * the loop is rotated (k points, inverse and stores of tile u, then the
//    forward pass of tile u + 1 and the load of tile u + 2), and every lane
//    stores (GRAD) on every path, lanes of idle lines into a 256-value sink,
//    so the wait before a forward pass leaves the 16 stores in flight.
// Tiles, lines and thread roles as kspace_xs_body; each thread's two k
// points (column kx of both tile rows) are taken one after the other, rank
// by rank, with float sums per tile and double sums over the tiles.  LDS:
// the lines (32 KB), the G tables, the eigenvalues and the sinc factors
// (ksxp_lds); three blocks fit a CU.
template <int NS, int RT, bool GRAD>
__device__ __forceinline__ void kspace_xsp_body(const BlockInfo& bk, const KspaceArgs& a) {
  static_assert(NS * RT <= 8, "split x k-space: at most 8 spectra per tile");
  constexpr int NP = NS * (NS + 1) / 2, NK = 2, ls = 256, nlines = NS * RT * NK;
  const KsCell kc = ks_cell(a);
  extern __shared__ float4 lds4[];
  float2* S = (float2*)lds4;
  float4* tab = lds4 + nlines * ls / 2;
  float* eig_l = (float*)(tab + 2 * NP * kGtabN);  // kMaxRank
  float* scx_l = eig_l + kMaxRank;                  // 256
  float* scy_l = scx_l + 256;                       // 256
  float* scz_l = scy_l + 256;                       // M1
  for (int i = threadIdx.x; i < 2 * NP * kGtabN; i += blockDim.x) tab[i] = a.gtab[i];
  for (int i = threadIdx.x; i < a.r; i += blockDim.x) eig_l[i] = a.eigs[i];
  if (a.use_sinc) {
    for (int i = threadIdx.x; i < 256; i += blockDim.x) {
      scx_l[i] = a.scx[i];
      scy_l[i] = a.scy[i];
    }
    for (int i = threadIdx.x; i < a.M1; i += blockDim.x) scz_l[i] = a.scz[i];
  }
  const int nyp = 128, zb = (a.r + RT - 1) / RT;
  const int ntiles = nyp * a.M1 * zb;
  const int xl = threadIdx.x >> 4, xj = threadIdx.x & 15;
  const int xqi = (xl >> 1) % RT, xX = (xl >> 1) / RT;
  const int kx = threadIdx.x;  // column of this thread's two k points
  // spectrum offset of this thread's line column in tile u, or -1 (idle:
  // a line past the tile's spectra or of a rank past the last)
  auto tile_elem = [&](int u) -> long long {
    const int kyp = u % nyp, rest = u / nyp, kz = rest % a.M1, q0 = rest / a.M1 * RT;
    if (xl >= nlines || xqi >= min(RT, a.r - q0)) return -1LL;
    return (((long long)(xX * a.r + q0 + xqi) * a.M1 + kz) * 256 + 2 * kyp + (xl & 1)) * 256 + xj;
  };
  double acc[11];  // tile sums (7 used; the parameter slots stay zero)
#pragma unroll
  for (int i = 0; i < 11; ++i) acc[i] = 0.0;
  const int u0 = bk.id.x * a.KT, u1 = min(ntiles, u0 + a.KT);
  float2 v[16];
  // idle lanes load the spectrum's first 256 values (ignored) and store
  // into the sink
  auto load_tile = [&](long long e) {
    const float2* src = a.spec + (e >= 0 ? e : xj);
#pragma unroll
    for (int n2 = 0; n2 < 16; ++n2) v[n2] = src[16 * n2];
  };
  const float2* Lr = S + (xl < nlines ? xl : 0) * ls;  // the line this lane transforms (any, when idle)
  auto forward = [&](bool live) {  // the tile's column of this lane into LDS, then a barrier
    if (xl < nlines) {
      float2 z[16];
#pragma unroll
      for (int n2 = 0; n2 < 16; ++n2) z[n2] = v[n2];
      fft16r<-1>(z);  // over n2 -> k2: X[xj + 16 k2]
      float2* L = S + xl * ls;
#pragma unroll
      for (int k2 = 0; k2 < 16; ++k2) L[xj + 16 * k2] = live ? z[k2] : make_float2(0.f, 0.f);
    }
    __syncthreads();
  };
  if (u0 < u1) {
    long long e0 = tile_elem(u0);
    load_tile(e0);
    __syncthreads();  // the staged tables
    forward(e0 >= 0);
    long long e1 = u0 + 1 < u1 ? tile_elem(u0 + 1) : -1LL;
    if (u0 + 1 < u1) load_tile(e1);
    for (int u = u0; u < u1; ++u) {
      asm volatile("" ::: "memory");  // the load is issued before the k points
      const int kyp = u % nyp, rest = u / nyp, kz = rest % a.M1, q0 = rest / a.M1 * RT;
      const int rt = min(RT, a.r - q0), ky0 = 2 * kyp;
      float lam[RT];
#pragma unroll
      for (int q = 0; q < RT; ++q) lam[q] = q < rt ? eig_l[q0 + q] : 0.0f;
      const float w = (kz == 0 || kz == a.M1 - 1) ? 1.0f : 2.0f;
      const float scxz = a.use_sinc ? scx_l[kx] * scz_l[kz] : 1.0f;
      float accb[7];
#pragma unroll
      for (int i = 0; i < 7; ++i) accb[i] = 0.0f;
#pragma unroll 1
      for (int kyi = 0; kyi < 2; ++kyi) {
        float isc = 1.0f;
        if (a.use_sinc) isc = __builtin_amdgcn_rcpf(fmaxf(scxz * scy_l[ky0 + kyi], 1e-10f));
        const KPoint kp = kspace_point_isc(a, kc, kx, ky0 + kyi, kz, isc);
        float gp[NP], dgp[NP];
        ks_gtab_lds<NP>(tab, a.gkinv, kp.kn, gp, dgp);
        ks_kpoint_spectra<NS, RT, NK>(a, kc, S, kyi, kx, kp, gp, dgp, lam, w, accb);
        asm volatile("" ::: "memory");  // one k point at a time
      }
#pragma unroll
      for (int i = 0; i < 7; ++i) acc[i] += (double)accb[i];
      if (kz == 0 && ky0 == 0 && threadIdx.x == 0) kspace_take_k0<NS, RT>(a, S, ls, NK, q0, rt);
      if constexpr (GRAD) {
        __syncthreads();
        float2 z[16];
#pragma unroll
        for (int k2 = 0; k2 < 16; ++k2) z[k2] = Lr[xj + 16 * k2];  // k1 = xj
        fft16r<1>(z);  // over k2 -> n2: Y[xj, n2]
        float2* dst = e0 >= 0 ? a.spec + e0 : a.sink + xj;
#pragma unroll
        for (int n2 = 0; n2 < 16; ++n2) FD3_ST_NT(dst + 16 * n2, z[n2]);  // streaming
      }
      if (u + 1 < u1) {
        __syncthreads();  // every lane is done with the lines
        forward(e1 >= 0);
        e0 = e1;
        e1 = u + 2 < u1 ? tile_elem(u + 2) : -1LL;
        if (u + 2 < u1) load_tile(e1);
      }
    }
  }
  __syncthreads();  // the lines are free: reuse them for the block sums
  kspace_reduce(bk, acc, (double*)S, a, false);
}

// Any number of species: phi accumulates in a second LDS buffer P.
__device__ __forceinline__ void kspace_body(const BlockInfo& bk, const KspaceArgs& a) {
  const KsCell kc = ks_cell(a);
  extern __shared__ float4 lds4[];
  __shared__ double red[16 * 16];
  const int nx = a.nx, nS = a.nS, RT = a.RT, NK = a.NK;
  const int kz = bk.id.y, q0 = bk.id.z * RT;
  const int rt = min(RT, a.r - q0);
  const int nlines = nS * RT * NK;
  const int ls = nx + 1;
  const int lnx = ilog2(nx);  // nx a power of two
  float4* htab = lds4;
  float2* tws = (float2*)(htab + 3 * (kHtabN + 1));
  float2* S = tws + nx;
  float2* P = S + nlines * ls;
  long long* lbase = (long long*)(P + nlines * ls);
  kspace_stage_consts(a, htab, tws);

  const bool params = a.need_params;
  double acc[11];
#pragma unroll
  for (int i = 0; i < 11; ++i) acc[i] = 0.0;
  const float w = (kz == 0 || kz == a.M1 - 1) ? 1.0f : 2.0f;
  for (int it = 0; it < a.KT; ++it) {
    const int ky0 = (bk.id.x * a.KT + it) * NK;
    if (it) __syncthreads();
    kspace_line_bases(a, lbase, nlines, RT, rt, ky0, kz, q0);
    __syncthreads();
    lds_fft<-1, kKsGenericMaxB, 256>(S, nx, nlines, ls, tws, 1, GlobalLines{a.spec, lbase});
    for (int t = threadIdx.x; t < NK * nx; t += blockDim.x) {
      const int kyi = t >> lnx, kx = t & (nx - 1);
      const KPoint kp = kspace_point(a, kc, kx, ky0 + kyi, kz);
      float d_acc = 0.f, p_acc[4] = {0.f, 0.f, 0.f, 0.f};
      // phi accumulates in P (this thread owns column (kyi, kx) of every line)
      for (int l = 0; l < nS * RT; ++l) P[(l * NK + kyi) * ls + kx] = make_float2(0.f, 0.f);
      for (int X = 0; X < nS; ++X) {
        for (int Y = X; Y < nS; ++Y) {
          const KsPair pk = a.pc[X * nS + Y];
          float g, dg, h6, d6, h8, d8;
          green_eval_tab(pk, htab, kp.kn, &g, &dg, &h6, &d6, &h8, &d8);
          const float gs = g;  // G, not G~: 1/sinc^2p is applied to phi and the sums below
          float tre = 0.f;  // sum_q lambda_q Re(S_X conj S_Y)
          for (int qi = 0; qi < rt; ++qi) {
            const int ix = ((X * RT + qi) * NK + kyi) * ls + kx;
            const int iy = ((Y * RT + qi) * NK + kyi) * ls + kx;
            const float2 sx = S[ix], sy = S[iy];
            float2 px = P[ix];
            px.x += gs * sy.x;
            px.y += gs * sy.y;
            P[ix] = px;
            if (Y != X) {
              float2 py = P[iy];
              py.x += gs * sx.x;
              py.y += gs * sx.y;
              P[iy] = py;
            }
            tre += a.eigs[q0 + qi] * (sx.x * sy.x + sx.y * sy.y);
          }
          if (Y != X) tre *= 2.0f;
          d_acc += dg * tre;
          if (params) kspace_param_terms(pk, kp.kn, h6, d6, h8, d8, tre, p_acc);
        }
      }
      // energy sum_X sum_q lambda_q Re(S conj phi); P <- -(lambda_q / Omega) phi / sinc^2p
      float e_acc = 0.f;
      for (int X = 0; X < nS; ++X) {
        for (int qi = 0; qi < rt; ++qi) {
          const int ix = ((X * RT + qi) * NK + kyi) * ls + kx;
          const float lam = a.eigs[q0 + qi];
          const float2 sx = S[ix], ph = P[ix];
          e_acc += lam * (sx.x * ph.x + sx.y * ph.y);
          const float f = lam * kc.neg_inv_vol * kp.isc;
          float2 v = make_float2(f * ph.x, f * ph.y);
          if (kz == 0 && ky0 + kyi == 0 && kx == 0) {  // k = 0: kept out of the mesh (see KspaceArgs::k0)
            a.k0[X * a.r + q0 + qi] = v;
            v = make_float2(0.f, 0.f);
          }
          P[ix] = v;
        }
      }
      kspace_accumulate(acc, kp, w, e_acc, d_acc, p_acc, params);
    }
    if (!a.need_grad) continue;
    __syncthreads();
    lds_fft<1, kKsGenericMaxB, 256>(P, nx, nlines, ls, tws, 1, LdsLines{}, GlobalLines{a.spec, lbase});
  }
  kspace_reduce(bk, acc, red, a, params);
}

// --------------------------------------------------------------- host ---
constexpr int kYLog = 10;  // y passes: 1024-thread blocks
constexpr int kBatchKT = 8;  // k-space tiles per block of a merged batch launch (meshes other than 256 x 256 x/y)

// Tile overrides measured by the autotuner, per (mesh, species, rank):
// z-pass lines TY, y-pass lines TX, k-space ranks per tile RT and tiles per
// block KT (0 = the plan's own choice).  An override that does not fit the
// kernels' limits for its plan is ignored.
using TuneKey = std::tuple<int, int, int, int, int>;
static std::mutex g_tune_mu;
static std::map<TuneKey, std::array<int, 4>> g_tune;

static bool tune_lookup(const FftPlan& p, std::array<int, 4>* t) {
  std::lock_guard<std::mutex> lock(g_tune_mu);
  const auto it = g_tune.find(TuneKey{p.nx, p.ny, p.nz, p.nS, p.r});
  if (it == g_tune.end()) return false;
  *t = it->second;
  return true;
}

int fft_plan_make(const int32_t ns[3], int nS, int r, FftPlan* p) {
  for (int i = 0; i < 3; ++i) {
    const int n = ns[i];
    if (n < 2 || n > 1024 || (n & (n - 1))) {
      set_error("mesh size %d on axis %d is not a power of two in [2, 1024]", n, i);
      return FD3_ERR_UNSUPPORTED;
    }
  }
  p->nx = ns[0]; p->ny = ns[1]; p->nz = ns[2];
  p->M = p->nz / 2;
  p->M1 = p->M + 1;
  p->nS = nS; p->r = r; p->C = nS * r;
  const int budget = kFftLines;  // complex values per block for the pass kernels
  // lines per block: the register budget caps lines * length at kFftLines
  // (kZLines for the z passes)
  p->TY = std::max(1, std::min(p->ny, kZLines / std::max(p->M, 1)));
  p->TY = 1 << (31 - __builtin_clz(p->TY));
  p->TX = std::max(1, std::min(p->nx, kYLines / p->ny));
  p->TX = 1 << (31 - __builtin_clz(p->TX));
  p->LPB = std::max(1, budget / (p->nx + 1));
  std::array<int, 4> tune{0, 0, 0, 0};
  const bool tuned = tune_lookup(*p, &tune);
  auto pow2 = [](int v) { return v > 0 && (v & (v - 1)) == 0; };
  if (tuned && pow2(tune[0]) && tune[0] <= p->ny && (size_t)tune[0] * p->M <= (size_t)kFftLines) p->TY = tune[0];
  if (tuned && pow2(tune[1]) && tune[1] <= p->nx && (size_t)tune[1] * p->ny <= (size_t)kYLines) p->TX = tune[1];
  // k-space: the h table, the x twiddles, one (nS <= kKsRegMaxS) or two LDS
  // buffers of nS*RT*NK lines of nx+1 complex, and the line bases, within
  // 62 KB (the block also holds 2 KB of static LDS)
  p->ks_reg = nS <= kKsRegMaxS ? nS : 0;
  const int nbuf = p->ks_reg ? 1 : 2;
  // the generic variant may take most of gfx950's 160 KB of LDS per workgroup
  const size_t ks_budget = p->ks_reg ? (size_t)62 * 1024 : (size_t)kKsGenericLds;
  // (the register variant also reuses its lines for the block sums: >= 2 KB)
  auto ks_bytes = [&](int rt_, int nk_) {
    const size_t tab = p->ks_reg ? (size_t)ks_table_f4(nS) * sizeof(float4) : (size_t)3 * (kHtabN + 1) * sizeof(float4);
    // the register variant pads its lines when lds_fft unrolls (fft_fixed_tile)
    const bool pad = p->ks_reg && fft_fixed_tile(p->nx, nS * rt_ * nk_, ks_maxb(nS * rt_), 256);
    const int ls = pad ? padded_stride(p->nx) : p->nx + 1;
    const size_t lines =
        (size_t)nbuf * nS * rt_ * nk_ * ls * sizeof(float2) + (size_t)nS * rt_ * nk_ * sizeof(long long);
    return tab + (size_t)p->nx * sizeof(float2) + std::max(lines, (size_t)16 * 16 * sizeof(double));
  };
  // RT: ranks per tile; the register variant takes any RT <= 4, the generic
  // one a power of two <= 8 (a last partial tile is padded with zero lines)
  int RT = 1;
  if (p->ks_reg) RT = std::min(r, std::max(1, 8 / nS));  // nS * RT <= 8 spectra per thread: no register spills
  else
    while (RT < std::min(r, 8)) RT <<= 1;
  int NK = std::max(1, std::min(p->ny, 256 / p->nx));
  auto ks_fits = [&](int rt_, int nk_) {
    const int maxb = p->ks_reg ? ks_maxb(nS * rt_) : kKsGenericMaxB;
    return ks_bytes(rt_, nk_) <= ks_budget && (size_t)nS * rt_ * nk_ * p->nx <= (size_t)1024 * maxb;
  };
  if (tuned && tune[2] >= 1 && tune[2] <= std::min(r, p->ks_reg ? 4 : 8) && (p->ks_reg || pow2(tune[2])) &&
      ks_fits(tune[2], 1))
    RT = tune[2];
  while (RT > 1 && !ks_fits(RT, NK)) RT = p->ks_reg ? RT - 1 : RT >> 1;
  while (NK > 1 && !ks_fits(RT, NK)) NK >>= 1;
  if (!ks_fits(RT, NK)) {
    set_error("k-space tile of %d species x %d points exceeds the LDS of a workgroup", nS, p->nx);
    return FD3_ERR_UNSUPPORTED;
  }
  p->RT = RT; p->NK = NK;
  p->ks_lds = ks_bytes(RT, NK);
  // nx = 256 with at most 8 lines per ky row: the two-pass x FFT over two ky
  // rows (16 lines of kKs2Ls complex, no padding, 256 (line, column) tasks)
  p->ks2 = 0;
  if (p->ks_reg && p->nx == 256 && p->ny % 2 == 0 && nS * RT <= 8) {
    const size_t nl = (size_t)nS * RT * 2;
    const size_t b = (size_t)ks2_table_f4(true) * sizeof(float4) + (size_t)p->nx * sizeof(float2) +
                     std::max(nl * kKs2Ls * sizeof(float2) + nl * sizeof(long long), (size_t)16 * 16 * sizeof(double));
    if (b <= ks_budget) {
      p->ks2 = 1;
      p->NK = NK = 2;
      p->ks_lds = b;
    }
  }
  // ny = 256 as well: the y passes run the first stage of the x FFT and the
  // k-space kernel the second (kspace_xs_body, 32 KB of LDS per block)
  p->xs = p->ks2 && p->ny == 256;
  // tiles per block: up to 4 ky tiles, keeping at least ~8k blocks in flight
  const int zb = (r + RT - 1) / RT;
  int KT = 1;
  while (KT < 4 && (p->ny / NK) % (2 * KT) == 0 && (size_t)(p->ny / NK / (2 * KT)) * p->M1 * zb >= 8192) KT <<= 1;
  // planned for a merged batch launch (a recorder is installed): the blocks
  // of all boxes share one grid, so small meshes take up to kBatchKT tiles
  // per block, which spreads each block's table staging and partial-sum row
  // over more tiles (the 256-box batching workload: k-space 1.71 -> 1.30 ms,
  // partial-sum reduction 0.26 -> 0.06 ms at 4 tiles, 1.26 / 0.04 ms at 8)
  if (current_recorder())
    while (KT < kBatchKT && (p->ny / NK) % (2 * KT) == 0) KT <<= 1;
  // the split x kernel takes any tile count per block (1-D grid): five
  // tiles (measured with the nontemporal y passes beside it: 210 us against
  // 229 / 223 / 217 / 211 us for 2 / 3 / 4 / 6 at the 100k-atom box, 256^3,
  // 4 species, rank 2), fewer when that would leave fewer than ~3k blocks
  if (p->xs) KT = std::max(1, std::min(5, (int)((size_t)p->ny / 2 * p->M1 * zb / 3072)));
  if (tuned && p->xs && tune[3] >= 1 && tune[3] <= 64) KT = tune[3];
  else if (tuned && pow2(tune[3]) && tune[3] <= 16 && (p->ny / NK) % tune[3] == 0) KT = tune[3];
  p->KT = KT;
  p->ks_blocks = (p->ny / NK / KT) * p->M1 * zb;
  if (p->xs) p->ks_blocks = (p->ny / 2 * p->M1 * zb + KT - 1) / KT;  // 1-D grid of KT-tile blocks
  // register budget of lds_fft<.., kFftMaxB>: lines * n <= kFftLines
  const size_t lim = kFftLines;
  if ((size_t)p->TY * p->M > lim || (size_t)p->TX * p->ny > (size_t)kYLines || (size_t)p->LPB * p->nx > lim) {
    set_error("internal FFT tiling exceeds the per-thread register budget");
    return FD3_ERR_UNSUPPORTED;
  }
  return FD3_OK;
}

double ks_gtab_dk(const FftPlan& p, const Geom& g) {
  double kmax = 0.0;
  for (int sx = -1; sx <= 1; sx += 2)
    for (int sy = -1; sy <= 1; sy += 2) {
      const double m[3] = {0.5 * sx * p.nx, 0.5 * sy * p.ny, 0.5 * p.nz};
      double k2 = 0.0;
      for (int c = 0; c < 3; ++c) {
        const double kc = m[0] * g.recip[c] + m[1] * g.recip[3 + c] + m[2] * g.recip[6 + c];
        k2 += kc * kc;
      }
      kmax = std::max(kmax, std::sqrt(k2));
    }
  return std::max(kmax, 1e-30) * (1.0 + 1e-6) * g.kreach / kGtabN;
}

void fft_tables_carve(const FftPlan& p, Arena& ar, FftTables* t) {
  t->twx = ar.take<float2>(p.nx);
  t->twy = ar.take<float2>(p.ny);
  t->twz = ar.take<float2>(p.nz);
  t->scx = ar.take<float>(p.nx);
  t->scy = ar.take<float>(p.ny);
  t->scz = ar.take<float>(p.M1);
  t->htab = ar.take<float4>(3 * (kHtabN + 1));
  t->gtab = ar.take<float4>((size_t)ks_table_f4(kKsRegMaxS));
  t->pc = ar.take<KsPair>((size_t)p.nS * p.nS);
  t->sink = ar.take<float2>(256);
}

int fft_tables_init(const FftPlan& p, const FftTables& t, int order, int use_sinc, hipStream_t s,
                    const PairInit* pairs) {
  PairInit pi{};
  if (pairs) pi = *pairs;
  const int n = std::max(std::max(std::max(p.nx, p.ny), p.nz), pi.nS * pi.nS);
  const int ngb = init_gtab_blocks(pi);
  ProfScope _ps("k_init_tables", s);
  const InitTablesArgs ia{t.twx, t.twy, t.twz, t.scx, t.scy, t.scz, t.htab, t.gtab, t.pc, pi,
                          ngb, p.nx, p.ny, p.nz, order, use_sinc};
  return launch<InitTablesArgs, init_tables_body, 256>("k_init_tables", dim3(kInitHB + ngb + (n + 255) / 256), 0, s,
                                                       ia);
}

// One y pass (dir -1: A -> spec, +1: spec -> A); more than 64 KB of dynamic
// LDS is allowed per kernel by launch().
static int launch_y(const FftPlan& p, const FftTables& t, const float2* in, float2* out, int dir, hipStream_t s,
                    bool xs) {
  if (xs && !(p.nx == 256 && p.ny == 256)) {
    set_error("split x FFT planned for a mesh that is not 256 x 256 in x, y");
    return FD3_ERR_UNSUPPORTED;
  }
  if (p.ny == 256 && p.nx % 16 == 0) {  // two-pass y lines, 32 per block when nx allows
    const int nl = p.nx % 32 == 0 ? 32 : 16;
    const size_t sh2 = (size_t)nl * kKs2Ls * sizeof(float2) + (size_t)p.ny * sizeof(float2);
    const dim3 grid2(p.nx / nl, p.M1, p.C);
    const YPassArgs ya{in, out, p.nx, p.ny, p.M1, nl, t.twy};
    if (dir < 0) {
      ProfScope _ps("k_c2c_y_fwd", s);
      // split x FFT: 16 lines per forward block, 32 per inverse block (the
      // faster of the two for each direction: 0.226 / 0.230 ms forward,
      // 0.230 / 0.220 ms inverse on the 100k-atom box)
      if (xs)
        return launch<YPassArgs, c2c_y2_body<-1, 16, true>, 256>("k_c2c_y_fwd", dim3(16, p.M1, p.C),
                                                                 (size_t)16 * kKs2Ls * sizeof(float2) +
                                                                     (size_t)p.ny * sizeof(float2),
                                                                 s, ya);
      if (nl == 32) return launch<YPassArgs, c2c_y2_body<-1, 32>, 512>("k_c2c_y_fwd", grid2, sh2, s, ya);
      return launch<YPassArgs, c2c_y2_body<-1, 16>, 256>("k_c2c_y_fwd", grid2, sh2, s, ya);
    }
    ProfScope _ps("k_c2c_y_inv", s);
    if (xs) return launch<YPassArgs, c2c_y2_body<1, 32, true>, 512>("k_c2c_y_inv", grid2, sh2, s, ya);
    if (nl == 32) return launch<YPassArgs, c2c_y2_body<1, 32>, 512>("k_c2c_y_inv", grid2, sh2, s, ya);
    return launch<YPassArgs, c2c_y2_body<1, 16>, 256>("k_c2c_y_inv", grid2, sh2, s, ya);
  }
  const size_t sh = (size_t)p.TX * (p.ny + 1) * sizeof(float2) + (size_t)p.ny * sizeof(float2) +
                    (size_t)p.TX * sizeof(long long);
  const dim3 grid(p.nx / p.TX, p.M1, p.C);
  const YPassArgs ya{in, out, p.nx, p.ny, p.M1, p.TX, t.twy};
  constexpr int NT = 1 << kYLog;
  if (dir < 0) {
    ProfScope _ps("k_c2c_y_fwd", s);
    return launch<YPassArgs, c2c_y_body<-1, kYLog>, NT>("k_c2c_y_fwd", grid, sh, s, ya);
  }
  ProfScope _ps("k_c2c_y_inv", s);
  return launch<YPassArgs, c2c_y_body<1, kYLog>, NT>("k_c2c_y_inv", grid, sh, s, ya);
}

int fft_forward_zy(const FftPlan& p, const FftTables& t, const float* rho, float2* A, float2* spec, hipStream_t s) {
  int rc;
  {
    const size_t sh = (size_t)p.TY * (p.M + 1) * sizeof(float2) + (size_t)p.nz * sizeof(float2);
    ProfScope _ps("k_r2c_z", s);
    hipLaunchKernelGGL(k_r2c_z, dim3(p.ny / p.TY, p.nx, p.C), dim3(256), sh, s, rho, A, p.nx, p.ny, p.nz, p.TY,
                       t.twz);
    FD3_LAUNCH_CHECK("k_r2c_z");
  }
  return launch_y(p, t, A, spec, -1, s, false);
}

int fft_x(const FftPlan& p, const FftTables& t, float2* spec, int dir, hipStream_t s) {
  const int nlines = p.C * p.M1 * p.ny;
  const size_t sh = (size_t)p.LPB * (p.nx + 1) * sizeof(float2) + (size_t)p.nx * sizeof(float2);
  const dim3 grid((nlines + p.LPB - 1) / p.LPB);
  ProfScope _ps("k_c2c_x", s);
  if (dir < 0)
    hipLaunchKernelGGL(k_c2c_x<-1>, grid, dim3(256), sh, s, spec, p.nx, nlines, p.LPB, t.twx);
  else
    hipLaunchKernelGGL(k_c2c_x<1>, grid, dim3(256), sh, s, spec, p.nx, nlines, p.LPB, t.twx);
  FD3_LAUNCH_CHECK("k_c2c_x");
  return FD3_OK;
}

int fft_inverse_yz(const FftPlan& p, const FftTables& t, const float2* spec, float2* A, float* out, hipStream_t s) {
  int rc;
  if ((rc = launch_y(p, t, spec, A, 1, s, false))) return rc;
  {
    const size_t sh = (size_t)p.TY * (p.M + 1) * sizeof(float2) + (size_t)p.nz * sizeof(float2);
    ProfScope _ps("k_c2r_z", s);
    hipLaunchKernelGGL(k_c2r_z, dim3(p.ny / p.TY, p.nx, p.C), dim3(256), sh, s, A, out, p.nx, p.ny, p.nz, p.TY,
                       t.twz);
    FD3_LAUNCH_CHECK("k_c2r_z");
  }
  return FD3_OK;
}

int fft_y(const FftPlan& p, const FftTables& t, const float2* in, float2* out, int dir, hipStream_t s, bool xs) {
  return launch_y(p, t, in, out, dir, s, xs);
}

template <int NS, int RT>
static int ks_launch(bool two, bool xs, bool params, dim3 grid, size_t sh, hipStream_t s, const KspaceArgs& a) {
  if constexpr (NS * RT <= 8) {
    // the plan sizes the two-pass variant's LDS with the h table, which only
    // the parameter-gradient kernel stages
    const size_t sh_g = sh - (size_t)(ks2_table_f4(true) - ks2_table_f4(false)) * sizeof(float4);
    if (xs && params) return launch<KspaceArgs, kspace_xs_body<NS, RT, true>, 256>("k_kspace", grid, sh, s, a);
    if (xs && a.need_grad) return launch<KspaceArgs, kspace_xsp_body<NS, RT, true>, 256>("k_kspace", grid, sh, s, a);
    if (xs) return launch<KspaceArgs, kspace_xsp_body<NS, RT, false>, 256>("k_kspace", grid, sh, s, a);
    if (two && params) return launch<KspaceArgs, kspace_reg_body<NS, RT, true, true>, 256>("k_kspace", grid, sh, s, a);
    if (two) return launch<KspaceArgs, kspace_reg_body<NS, RT, false, true>, 256>("k_kspace", grid, sh_g, s, a);
  }
  if (params) return launch<KspaceArgs, kspace_reg_body<NS, RT, true>, 256>("k_kspace", grid, sh, s, a);
  return launch<KspaceArgs, kspace_reg_body<NS, RT, false>, 256>("k_kspace", grid, sh, s, a);
}

int kspace_run(const FftPlan& p, const FftTables& t, const Geom& g, float2* spec, const float* eigs, int need_grad,
               int need_params, double* partials, float2* k0, hipStream_t s, bool xs) {
  if (xs && !(p.xs && p.ks2)) {
    set_error("split x FFT requested for a k-space plan without it");
    return FD3_ERR_UNSUPPORTED;
  }
  KspaceArgs a;
  a.k0 = k0;
  a.spec = spec; a.eigs = eigs; a.pc = t.pc; a.twx = t.twx; a.htab = t.htab; a.gtab = t.gtab; a.sink = t.sink;
  a.gkinv = (float)(1.0 / ks_gtab_dk(p, g));
  a.scx = t.scx; a.scy = t.scy; a.scz = t.scz; a.partials = partials;
  for (int i = 0; i < 9; ++i) a.recip[i] = (float)g.recip[i];
  a.neg_inv_vol = (float)(-1.0 / g.volume);
  a.dyn = g.dyn;
  a.nx = p.nx; a.ny = p.ny; a.nz = p.nz; a.M1 = p.M1; a.nS = p.nS; a.r = p.r; a.RT = p.RT; a.NK = p.NK; a.KT = p.KT;
  a.use_sinc = g.interp == FD3_INTERP_P3M;
  a.need_grad = need_grad; a.need_params = need_params;
  const size_t sh = !xs ? p.ks_lds : need_params ? ksx_lds(p.nS * p.RT, true) : ksxp_lds(p.nS, p.nS * p.RT, p.M1);
  const dim3 grid = xs ? dim3(p.ks_blocks) : dim3(p.ny / p.NK / p.KT, p.M1, (p.r + p.RT - 1) / p.RT);
  ProfScope _ps("k_kspace", s);
#define FD3_KS(NS_, RT_)                                                                              \
  case NS_ * 8 + RT_:                                                                                 \
    return ks_launch<NS_, RT_>(p.ks2 != 0, xs, need_params != 0, grid, sh, s, a);
  switch (p.ks_reg * 8 + p.RT) {
    FD3_KS(1, 1) FD3_KS(1, 2) FD3_KS(1, 3) FD3_KS(1, 4)
    FD3_KS(2, 1) FD3_KS(2, 2) FD3_KS(2, 3) FD3_KS(2, 4)
    FD3_KS(3, 1) FD3_KS(3, 2) FD3_KS(3, 3) FD3_KS(3, 4)
    FD3_KS(4, 1) FD3_KS(4, 2) FD3_KS(4, 3) FD3_KS(4, 4)
    default:
      return launch<KspaceArgs, kspace_body, 256>("k_kspace", grid, sh, s, a);
  }
#undef FD3_KS
}

}  // namespace fd3

using namespace fd3;

extern "C" {

int fd3_tune_set(const int32_t ns[3], int32_t n_species, int32_t n_rank, const int32_t tiles[4]) {
  if (!ns || !tiles || n_species < 1 || n_rank < 1) {
    set_error("bad tuning entry");
    return FD3_ERR_ARG;
  }
  std::lock_guard<std::mutex> lock(g_tune_mu);
  g_tune[TuneKey{ns[0], ns[1], ns[2], n_species, n_rank}] = {tiles[0], tiles[1], tiles[2], tiles[3]};
  return FD3_OK;
}

int fd3_tune_clear(void) {
  std::lock_guard<std::mutex> lock(g_tune_mu);
  g_tune.clear();
  return FD3_OK;
}

int fd3_plan_tiles(const int32_t ns[3], int32_t n_species, int32_t n_rank, int32_t tiles[4]) {
  if (!ns || !tiles || n_species < 1 || n_rank < 1) {
    set_error("plan tiles: NULL argument or no species / ranks");
    return FD3_ERR_ARG;
  }
  FftPlan p;
  const int rc = fft_plan_make(ns, n_species, n_rank, &p);
  if (rc) return rc;
  tiles[0] = p.TY; tiles[1] = p.TX; tiles[2] = p.RT; tiles[3] = p.KT;
  return FD3_OK;
}

}  // extern "C"
