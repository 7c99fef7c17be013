// Real-space lattice sum of the BJ-damped D3 pair energy, in double.
//
// The reciprocal-space energy of src/fourierd3/core.py:553-581 is the Fourier
// representation of
//   E = -1/2 sum_{i,j,n}' C6_ij [s6 / (r^6 + R0^6) + 3 s8 QQ_ij / (r^8 + R0^8)]
// with C6_ij = sum_q eigs_q c_iq c_jq the low-rank C6 of the decomposition and
// R0_ij = a1 sqrt(3 QQ_ij) + a2 (src/fourierd3/pair_pot.py:13-21,64).  This
// file sums it directly over lattice images up to rmax, plus the continuum
// tail beyond rmax, with forces, the strain derivative and the parameter
// gradient: a check of the mesh path and an evaluator for small boxes.  The
// reference has no real-space pair sum; the mesh energy does not include
// this one (adding both would count every pair twice).
//
// One workgroup per atom i (slot order).  Its 256 lanes split the pairs
// (j, image): with n >= 256 each lane takes whole j rows of images, with
// fewer atoms the image cube of each j is split into slices so that every
// lane has work.  A lane keeps the image sums of one (i, j) in registers and
// applies C6_ij once per row; the block then reduces its lanes with wavefront
// shuffles and one LDS pass (fixed order: deterministic).

#include <cmath>

#include "direct.h"
#include "launch.h"

namespace fd3 {

namespace {

constexpr int kDirNT = 256;

// 8-point Gauss-Legendre rule on [-1, 1] (Abramowitz & Stegun, table 25.4).
__device__ constexpr double kGlX[4] = {0.1834346424956498, 0.5255324099163290, 0.7966664774136267,
                                       0.9602898564975363};
__device__ constexpr double kGlW[4] = {0.3626837833783620, 0.3137066458778873, 0.2223810344533745,
                                       0.1012285362903763};

// Tail integrals beyond rmax, after r = rmax / u (u in (0, 1], a smooth
// integrand with poles at |u| = rmax / R0):
//   I6 = int_rmax^inf r^2 / (r^6 + R^6) dr = int_0^1 rmax^3 u^2 / (rmax^6 + R^6 u^6) du
//   I8 = int_rmax^inf r^2 / (r^8 + R^8) dr = int_0^1 rmax^3 u^4 / (rmax^8 + R^8 u^8) du
// and their derivatives in R; composite 8-point Gauss-Legendre on 16 panels.
__device__ void tail_integrals(double rm, double R, double* out) {
  const double rm3 = rm * rm * rm, rm6 = rm3 * rm3, rm8 = rm6 * rm * rm;
  const double R2 = R * R, R6 = R2 * R2 * R2, R8 = R6 * R2;
  double i6 = 0.0, i8 = 0.0, d6 = 0.0, d8 = 0.0;
  constexpr int kPanels = 16;
  for (int pnl = 0; pnl < kPanels; ++pnl) {
    const double mid = (pnl + 0.5) / kPanels, half = 0.5 / kPanels;
    for (int k = 0; k < 8; ++k) {
      const double u = mid + (k < 4 ? -kGlX[k] : kGlX[k - 4]) * half;
      const double w = (k < 4 ? kGlW[k] : kGlW[k - 4]) * half;
      const double u2 = u * u, u4 = u2 * u2, u6 = u4 * u2, u8 = u4 * u4;
      const double den6 = 1.0 / (rm6 + R6 * u6), den8 = 1.0 / (rm8 + R8 * u8);
      i6 += w * rm3 * u2 * den6;
      i8 += w * rm3 * u4 * den8;
      d6 -= w * 6.0 * (R6 / R) * rm3 * u2 * u6 * den6 * den6;
      d8 -= w * 8.0 * (R8 / R) * rm3 * u4 * u8 * den8 * den8;
    }
  }
  out[0] = i6; out[1] = i8; out[2] = d6; out[3] = d8;
}

// Blocks [0, nS r): the species charge sums qsum[X, q]; the blocks after
// them: the tail integrals of every species pair, one lane each.
__device__ __forceinline__ void direct_tables_body(const BlockInfo& bk, const DirectArgs& a) {
  __shared__ double red[16];
  const int nq = a.nS * a.r;
  const int b = bk.id.x;
  if (b < nq) {
    const int X = b / a.r, q = b % a.r;
    double v[1] = {0.0};
    for (int64_t i = threadIdx.x; i < a.n; i += blockDim.x)
      if (a.spec_s[i] == X) v[0] += (double)a.charges[i * a.r + q];
    block_sum_store<1>(v, red, a.qsum + b);
    return;
  }
  const int p = (b - nq) * blockDim.x + threadIdx.x;
  if (p >= a.nS * a.nS) return;
  const double qq = (double)a.qq[p];
  tail_integrals(a.rmax, a.a1 * sqrt(3.0 * qq) + a.a2, a.tab + 4 * p);
}

// Pair constants of (X, Y) for a row of species X: R0, R0^6, R0^8, 3 QQ, sqrt(3 QQ).
constexpr int kPairC = 5;

template <int RM>
__device__ __forceinline__ void direct_pairs_body(const BlockInfo& bk, const DirectArgs& a) {
  constexpr int NV = 14 + RM;  // E, dE/dr (3), strain (6), params (4), dE/dc (RM)
  extern __shared__ double dsh[];
  double* pc = dsh;                        // (nS, kPairC)
  double* red = dsh + kPairC * a.nS;       // (16, NV) block reduction
  double* res = red + 16 * NV;             // (NV)
  const int64_t i = bk.id.x;
  const int nS = a.nS, r = a.r;
  const int X = a.spec_s[i];
  for (int Y = threadIdx.x; Y < nS; Y += blockDim.x) {
    const double qq = (double)a.qq[X * nS + Y];
    const double sq = sqrt(3.0 * qq);
    const double R = a.a1 * sq + a.a2;
    const double R2 = R * R, R6 = R2 * R2 * R2;
    double* p = pc + kPairC * Y;
    p[0] = R; p[1] = R6; p[2] = R6 * R2; p[3] = 3.0 * qq; p[4] = sq;
  }
  __syncthreads();
  const double4 xi = a.pos4[i];
  double ci[RM], lam[RM];
#pragma unroll
  for (int q = 0; q < RM; ++q) {
    ci[q] = q < r ? (double)a.charges[i * r + q] : 0.0;
    lam[q] = q < r ? (double)a.eigs[q] : 0.0;
  }
  const int n0 = a.nimg[0], n1 = a.nimg[1], n2 = a.nimg[2];
  const int dy = 2 * n1 + 1, dz = 2 * n2 + 1;
  const int64_t ncube = (int64_t)(2 * n0 + 1) * dy * dz;
  const int64_t per = kDirNT / a.n;
  const int64_t slices = a.n >= kDirNT ? 1 : (per < ncube ? per : ncube);
  const int64_t units = a.n * slices;
  const double rmax2 = a.rmax * a.rmax;
  double acc[NV];
#pragma unroll
  for (int k = 0; k < NV; ++k) acc[k] = 0.0;
  for (int64_t u = threadIdx.x; u < units; u += blockDim.x) {
    const int64_t j = u % a.n, sl = u / a.n;
    const double4 xj = a.pos4[j];
    const int Y = a.spec_s[j];
    double cj[RM];
    double c6 = 0.0;
#pragma unroll
    for (int q = 0; q < RM; ++q) {
      cj[q] = q < r ? (double)a.charges[j * r + q] : 0.0;
      c6 += lam[q] * ci[q] * cj[q];
    }
    const double* p = pc + kPairC * Y;
    const double R = p[0], R6 = p[1], R8 = p[2], qq3 = p[3], sq = p[4];
    const double w8 = a.s8 * qq3;
    const double d0 = xj.x - xi.x, d1 = xj.y - xi.y, d2 = xj.z - xi.z;
    // image sums of this (i, j): f, g6, g8, g6^2, g8^2, df/dr^2 v, df/dr^2 v v
    double sf = 0.0, s6 = 0.0, s8 = 0.0, h6 = 0.0, h8 = 0.0;
    double gg[3] = {0.0, 0.0, 0.0}, vv[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
    for (int64_t img = sl; img < ncube; img += slices) {
      const int iz = (int)(img % dz) - n2;
      const int64_t t = img / dz;
      const int iy = (int)(t % dy) - n1, ix = (int)(t / dy) - n0;
      const double v0 = d0 + ix * a.cell[0] + iy * a.cell[3] + iz * a.cell[6];
      const double v1 = d1 + ix * a.cell[1] + iy * a.cell[4] + iz * a.cell[7];
      const double v2 = d2 + ix * a.cell[2] + iy * a.cell[5] + iz * a.cell[8];
      const double r2 = v0 * v0 + v1 * v1 + v2 * v2;
      if (r2 >= rmax2 || (j == i && ix == 0 && iy == 0 && iz == 0)) continue;
      const double r4 = r2 * r2, r6 = r4 * r2, r8 = r4 * r4;
      const double g6 = 1.0 / (r6 + R6), g8 = 1.0 / (r8 + R8);
      sf += a.s6 * g6 + w8 * g8;
      s6 += g6; s8 += g8;
      h6 += g6 * g6; h8 += g8 * g8;
      // d f / d(r^2) with f = s6 g6 + 3 s8 QQ g8
      const double dfr = -(3.0 * a.s6 * r4 * g6 * g6 + 4.0 * w8 * r6 * g8 * g8);
      gg[0] += dfr * v0; gg[1] += dfr * v1; gg[2] += dfr * v2;
      vv[0] += dfr * v0 * v0; vv[1] += dfr * v1 * v1; vv[2] += dfr * v2 * v2;
      vv[3] += dfr * v0 * v1; vv[4] += dfr * v0 * v2; vv[5] += dfr * v1 * v2;
    }
    // phi = -C6 f, v = r_j + n - r_i: row i holds 1/2 of E and of dE/dp,
    // all of dE/dr_i = sum dphi/dr^2 (-2 v), and sum dphi/dr^2 v v of the strain
    acc[0] -= 0.5 * c6 * sf;
    acc[1] += 2.0 * c6 * gg[0]; acc[2] += 2.0 * c6 * gg[1]; acc[3] += 2.0 * c6 * gg[2];
#pragma unroll
    for (int k = 0; k < 6; ++k) acc[4 + k] -= c6 * vv[k];
    const double R5 = R6 / R, R7 = R8 / R;
    const double dEdR = -0.5 * c6 * (-6.0 * a.s6 * R5 * h6 - 8.0 * w8 * R7 * h8);
    acc[10] -= 0.5 * c6 * s6;
    acc[11] -= 0.5 * c6 * qq3 * s8;
    acc[12] += dEdR * sq;
    acc[13] += dEdR;
#pragma unroll
    for (int q = 0; q < RM; ++q) acc[14 + q] -= lam[q] * cj[q] * sf;
  }
  block_sum_store<NV>(acc, red, res);
  if (threadIdx.x != 0) return;
  // continuum tail: dE/dc_iq = -eigs_q sum_Y T_XY qsum_Yq with
  // T_XY = 4 pi / Omega (s6 I6 + 3 s8 QQ I8); E_i = 1/2 sum_q c_iq dE/dc_iq
  const double pref = 4.0 * kPi / a.volume;
  double et = 0.0, pt[4] = {0.0, 0.0, 0.0, 0.0};
  for (int q = 0; q < r; ++q) {
    double sT = 0.0, s6T = 0.0, s8T = 0.0, sa1 = 0.0, sa2 = 0.0;
    for (int Y = 0; Y < nS; ++Y) {
      const double* tb = a.tab + 4 * (X * nS + Y);
      const double* p = pc + kPairC * Y;
      const double Q = a.qsum[Y * r + q];
      sT += pref * (a.s6 * tb[0] + a.s8 * p[3] * tb[1]) * Q;
      s6T += pref * tb[0] * Q;
      s8T += pref * p[3] * tb[1] * Q;
      const double dR = pref * (a.s6 * tb[2] + a.s8 * p[3] * tb[3]) * Q;
      sa1 += dR * p[4];
      sa2 += dR;
    }
    const double l = (double)a.eigs[q], c = (double)a.charges[i * r + q];
    const double dq = -l * sT;
    res[14 + q] += dq;
    et += 0.5 * c * dq;
    pt[0] -= 0.5 * c * l * s6T;
    pt[1] -= 0.5 * c * l * s8T;
    pt[2] -= 0.5 * c * l * sa1;
    pt[3] -= 0.5 * c * l * sa2;
  }
  double* row = a.rows + 16 * i;
  row[0] = res[0] + et;
  // the tail scales as 1/Omega: its strain derivative is -E_tail on the diagonal
  for (int k = 0; k < 6; ++k) row[1 + k] = res[4 + k] - (k < 3 ? et : 0.0);
  for (int k = 0; k < 4; ++k) row[7 + k] = res[10 + k] + pt[k];
  for (int d = 0; d < 3; ++d) a.gpos[3 * i + d] = res[1 + d];
  for (int q = 0; q < r; ++q) a.dEdq[i * r + q] = res[14 + q];
}

// One workgroup: sums of the per-atom rows and of the CN strain partials.
__device__ __forceinline__ void direct_finalize_body(const BlockInfo& bk, const DirectArgs& a) {
  (void)bk;
  constexpr int NV = 17;
  __shared__ double red[16 * NV];
  __shared__ double tot[NV];
  double v[NV];
#pragma unroll
  for (int k = 0; k < NV; ++k) v[k] = 0.0;
  for (int64_t i = threadIdx.x; i < a.n; i += blockDim.x)
#pragma unroll
    for (int k = 0; k < 11; ++k) v[k] += a.rows[16 * i + k];
  if (a.cnv)
    for (int b = threadIdx.x; b < a.ncnv; b += blockDim.x)
#pragma unroll
      for (int k = 0; k < 6; ++k) v[11 + k] += a.cnv[(size_t)8 * b + k];
  block_sum_store<NV>(v, red, tot);
  if (threadIdx.x != 0) return;
  a.energy[0] = tot[0];
  if (a.virial) {
    // symmetric 3x3 from (xx, yy, zz, xy, xz, yz)
    const int map[9] = {0, 3, 4, 3, 1, 5, 4, 5, 2};
    for (int k = 0; k < 9; ++k) a.virial[k] = tot[1 + map[k]] + tot[11 + map[k]];
  }
  if (a.grad_params)
    for (int k = 0; k < 4; ++k) a.grad_params[k] = tot[7 + k];
}

template <int RM>
int launch_pairs(const DirectArgs& a, hipStream_t s) {
  const size_t sh = (size_t)(kPairC * a.nS + 16 * (14 + RM) + (14 + RM)) * sizeof(double);
  return launch<DirectArgs, direct_pairs_body<RM>, kDirNT>("k_direct_pairs", dim3((unsigned)a.n), sh, s, a);
}

}  // namespace

void direct_images(const double cell[9], const double inv[9], double rmax, int nimg[3]) {
  (void)cell;
  for (int ax = 0; ax < 3; ++ax) {
    // plane spacing along axis ax: 1 / |column ax of cell^-1|
    const double b0 = inv[0 * 3 + ax], b1 = inv[1 * 3 + ax], b2 = inv[2 * 3 + ax];
    const double d = 1.0 / std::sqrt(b0 * b0 + b1 * b1 + b2 * b2);
    nimg[ax] = (int)std::ceil(rmax / d) + 1;
  }
}

int launch_direct(const DirectArgs& a, hipStream_t s) {
  int rc;
  const int nq = a.nS * a.r, np = (a.nS * a.nS + kDirNT - 1) / kDirNT;
  {
    ProfScope _ps("k_direct_tables", s);
    if ((rc = launch<DirectArgs, direct_tables_body, kDirNT>("k_direct_tables", dim3(nq + np), 0, s, a))) return rc;
  }
  if (a.n == 0) return FD3_OK;
  ProfScope _ps("k_direct_pairs", s);
  if (a.r <= 4) return launch_pairs<4>(a, s);
  if (a.r <= 8) return launch_pairs<8>(a, s);
  if (a.r <= 16) return launch_pairs<16>(a, s);
  return launch_pairs<32>(a, s);
}

int launch_direct_finalize(const DirectArgs& a, hipStream_t s) {
  ProfScope _ps("k_direct_finalize", s);
  return launch<DirectArgs, direct_finalize_body, kDirNT>("k_direct_finalize", dim3(1), 0, s, a);
}

}  // namespace fd3
