// B-spline spreading fused into the forward z pass, gathering fused into the
// inverse z pass, and the per-atom adjoint of the CN -> C6 chain.
//
// A block of the z pass owns the lines (c, x, y0..y0+TY-1, all z) of one
// channel c = species X * r + rank q.  Instead of reading a spread mesh it
// builds those lines in LDS from the atoms of species X whose p x p x p
// stencil touches them: atoms are sorted into mesh columns keyed by (X, x
// node 0, y node 0), so the contributing atoms are the columns with x node 0
// in {x - p + 1 .. x} and y node 0 in {y0 - p + 1 .. y0 + TY - 1}: at most two
// contiguous runs per x node.  Each atom adds c_q Wx Wy Wz into the LDS lines
// (compare-and-swap adds); the block then runs the in-LDS z FFT and writes the half
// spectrum, so the real mesh is never written to or read from HBM.
//
// The inverse z pass mirrors it: after the C2R transform the block holds the
// real lines of dE/drho in LDS and the same atoms accumulate
//   sum_nodes W dE/drho            (-> dE/dc, rank q)
//   c_q sum_nodes dW/du_d dE/drho  (-> dE/du)
// over the nodes in the block, stored as the atom's partial sums for this
// (x node, y tile, rank) and summed per atom by k_gather_sum in a fixed order.
// (points_to_mesh / mesh_to_points: src/fourierd3/mesh_interpolator_d3.py:99-189.)

#include "fd3_internal.h"
#include "fft_lds.h"
#include "fft_reg.h"
#include "fused.h"
#include "launch.h"
#include "mesh_kernels.h"
#include "neigh_kernels.h"

namespace fd3 {

void col_carve(const FftPlan& p, int64_t n, int order, Arena& ar, ColBuffers* b) {
  const size_t ncol = (size_t)p.nS * p.nx * p.ny;
  const int64_t nn = std::max<int64_t>(n, 1);
  b->key = ar.take<int32_t>(nn);
  b->tmp_i = ar.take<int4>(nn);
  b->tmp_f = ar.take<float4>(nn);
  b->colcount = ar.take<int32_t>(2 * ncol);  // colcount and colcur: cleared by one fill
  b->colcur = offset_or_null(b->colcount, ncol);
  b->colstart = ar.take<int32_t>(ncol + 1);
  b->scantmp = ar.take<int32_t>(scan_tmp_elems((int64_t)ncol));
  b->rec_i = ar.take<int4>(nn);
  b->rec_f = ar.take<float4>(nn);
  b->accA = ar.take<float>(nn * p.r);
  b->accG = ar.take<float>(nn * 3);
  b->part = ar.take<float4>((size_t)nn * std::max(order, 1) * 2 * p.r);
}

// ------------------------------------------------------------ column sort ---
__device__ __forceinline__ void col_prep_body(const BlockInfo& bk, const FusedArgs& a) {
  const int64_t s = (int64_t)bk.id.x * blockDim.x + threadIdx.x;
  if (s >= a.n) return;
  double inv[9];
  for (int i = 0; i < 9; ++i) inv[i] = a.dyn ? a.dyn->inv[i] : a.inv[i];
  const double4 r = a.pos4[s];
  const int X = a.spec_s[s];
  const int ns[3] = {a.nx, a.ny, a.nz};
  int b0[3];
  float xo[3];
  for (int d = 0; d < 3; ++d) {
    const double f = r.x * inv[0 * 3 + d] + r.y * inv[1 * 3 + d] + r.z * inv[2 * 3 + d];
    double x;
    b0[d] = wrap(interp_base((double)ns[d] * f, a.order, &x), ns[d]);
    xo[d] = (float)x;
  }
  const int key = (X * a.nx + b0[0]) * a.ny + b0[1];
  a.cb.key[s] = key;
  a.cb.tmp_i[s] = make_int4(b0[0], b0[1], b0[2], (int)s);
  a.cb.tmp_f[s] = make_float4(xo[0], xo[1], xo[2], 0.f);
  atomicAdd(a.cb.colcount + key, 1);
}

__device__ __forceinline__ void col_scatter_body(const BlockInfo& bk, const FusedArgs& a) {
  const int64_t s = (int64_t)bk.id.x * blockDim.x + threadIdx.x;
  if (s >= a.n) return;
  const int key = a.cb.key[s];
  const int pos = a.cb.colstart[key] + atomicAdd(a.cb.colcur + key, 1);
  a.cb.rec_i[pos] = a.cb.tmp_i[s];
  a.cb.rec_f[pos] = a.cb.tmp_f[s];
}


// Deterministic mode: the records of each column (placed in arrival order
// by col_scatter_body) sorted by slot, one thread per column (insertion sort:
// columns hold a few atoms).
__device__ __forceinline__ void col_order_body(const BlockInfo& bk, const FusedArgs& a) {
  const int64_t key = (int64_t)bk.id.x * blockDim.x + threadIdx.x;
  if (key >= (int64_t)a.nS * a.nx * a.ny) return;
  const int b = a.cb.colstart[key], m = a.cb.colcount[key];
  int4* ri = a.cb.rec_i + b;
  float4* rf = a.cb.rec_f + b;
  for (int i = 1; i < m; ++i) {
    const int4 ki = ri[i];
    const float4 kf = rf[i];
    int j = i - 1;
    for (; j >= 0 && ri[j].w > ki.w; --j) {
      ri[j + 1] = ri[j];
      rf[j + 1] = rf[j];
    }
    ri[j + 1] = ki;
    rf[j + 1] = kf;
  }
}

int col_sort(const FusedArgs& a, hipStream_t s) {
  const size_t ncol = (size_t)a.nS * a.nx * a.ny;
  int rc;
  ProfScope _ps("k_col_sort", s);
  if ((rc = launch_fill_i32(a.cb.colcount, 2 * ncol, s))) return rc;  // and colcur
  if (a.n > 0 && (rc = launch<FusedArgs, col_prep_body, 256>("k_col_prep", dim3(cdiv(a.n, 256)), 0, s, a)))
    return rc;
  if ((rc = scan_exclusive(a.cb.colcount, (int64_t)ncol, a.cb.colstart, a.cb.scantmp, nullptr, s))) return rc;
  if (a.n > 0 && (rc = launch<FusedArgs, col_scatter_body, 256>("k_col_scatter", dim3(cdiv(a.n, 256)), 0, s, a)))
    return rc;
  if (a.det && a.n > 0 &&
      (rc = launch<FusedArgs, col_order_body, 256>("k_col_order", dim3(cdiv((int64_t)ncol, 256)), 0, s, a)))
    return rc;
  return FD3_OK;
}

// Runs of the column-sorted atom list whose stencils touch plane x and rows
// [y0, y0 + TY) of species X: for each of the p x nodes j, the columns with
// y node 0 in [y0 - p + 1, y0 + TY - 1] (mod ny) form at most two contiguous
// runs.  Built in two steps so that the up to 2p colstart lookups are issued
// by 2p threads at once (segs_lookup, before a barrier) and compacted by one
// thread (segs_compact, before the next barrier).
struct Segs {
  int ca0[2 * kMaxOrder], cn[2 * kMaxOrder];  // candidate runs (start, length)
  int a0[2 * kMaxOrder];
  int jx[2 * kMaxOrder];
  int off[2 * kMaxOrder + 1];
  int n;
};

__device__ inline void segs_lookup(Segs* sg, const int32_t* __restrict__ colstart, int X, int x, int y0, int TY,
                                   int nx, int ny, int p) {
  const int t = threadIdx.x;
  if (t >= 2 * p) return;
  const int j = t >> 1, k = t & 1;
  const int bx0 = wrap(x - j, nx);
  const int base = (X * nx + bx0) * ny;
  int lo[2], hi[2], nr;
  const int l = y0 - (p - 1), h = y0 + TY - 1;
  if (TY + p - 1 >= ny) {
    lo[0] = 0; hi[0] = ny - 1; nr = 1;
  } else if (l < 0) {
    lo[0] = l + ny; hi[0] = ny - 1; lo[1] = 0; hi[1] = h; nr = 2;
  } else if (h >= ny) {
    lo[0] = l; hi[0] = ny - 1; lo[1] = 0; hi[1] = h - ny; nr = 2;
  } else {
    lo[0] = l; hi[0] = h; nr = 1;
  }
  int a0 = 0, cnt = 0;
  if (k < nr) {
    a0 = colstart[base + lo[k]];
    cnt = colstart[base + hi[k] + 1] - a0;
  }
  sg->ca0[t] = a0;
  sg->cn[t] = cnt;
}

__device__ inline void segs_compact(Segs* sg, int p) {
  if (threadIdx.x != 0) return;
  int ns = 0, off = 0;
  for (int t = 0; t < 2 * p; ++t) {
    if (sg->cn[t] > 0) {
      sg->a0[ns] = sg->ca0[t];
      sg->jx[ns] = t >> 1;
      sg->off[ns] = off;
      off += sg->cn[t];
      ++ns;
    }
  }
  sg->off[ns] = off;
  sg->n = ns;
}

__device__ inline float pick(const float* w, int j, int p) {
  float v = 0.f;
  for (int i = 0; i < p; ++i)
    if (i == j) v = w[i];
  return v;
}

// ------------------------------------------------------ spread + z pass ---
// line[z[j]] += v[j], j < p.  Float atomic adds to LDS run at about a third
// of a lane per clock per CU on gfx950, integer compare-and-swap at the rate
// of plain LDS traffic, so the adds are compare-and-swaps of the bit
// patterns, repeated for the lanes whose value another lane changed in
// between (the same results as atomic adds, in some order).
template <int P>
__device__ __forceinline__ void lds_add_f32_cas(float* line, const int* z, const float* v, int p) {
  float cur[kMaxOrder];
  for (int j = 0; j < p; ++j) cur[j] = line[z[j]];
  unsigned pend = (1u << p) - 1u;
  while (pend) {
    for (int j = 0; j < p; ++j) {
      if ((pend >> j) & 1u) {
        const unsigned as = __float_as_uint(cur[j]);
        const unsigned old = atomicCAS((unsigned*)(line + z[j]), as, __float_as_uint(cur[j] + v[j]));
        if (old == as) pend &= ~(1u << j);
        else cur[j] = __uint_as_float(old);
      }
    }
  }
}

// Deterministic spread (FD3_DETERMINISTIC): the block's atoms one after the
// other in list order (columns in segment order, each column ordered by slot
// by col_order_body); the first wave adds an atom's p x p stencil nodes of
// this plane, lane k taking node (k / p, k mod p), with plain LDS
// read-modify-writes (distinct nodes: no two lanes meet), or one lane doing
// all of them when the mesh is shorter than the stencil along y or z.
__device__ __forceinline__ void spread_ordered(float* ldsf, const Segs& sg, const FusedArgs& a, int q, int y0, int TY,
                                            int nz, int ls, int p) {
  if (threadIdx.x >= 64) return;
  const int lane = threadIdx.x;
  const bool serial = p > a.ny || p > nz;
  const int total = sg.off[sg.n];
  for (int i = 0; i < total; ++i) {
    int k = 0;
    while (i >= sg.off[k + 1]) ++k;
    const int ai = sg.a0[k] + (i - sg.off[k]);
    const int4 bi = a.cb.rec_i[ai];
    const float4 xf = a.cb.rec_f[ai];
    float w[kMaxOrder], dw[kMaxOrder], wy[kMaxOrder], wz[kMaxOrder];
    interp_weights<float>(p, a.interp, xf.x, w, dw);
    const float wx = pick(w, sg.jx[k], p) * a.charges[(size_t)bi.w * a.r + q];
    interp_weights<float>(p, a.interp, xf.y, wy, dw);
    interp_weights<float>(p, a.interp, xf.z, wz, dw);
    const int n0 = serial ? 0 : lane, n1 = serial ? (lane == 0 ? p * p : 0) : min(lane + 1, p * p);
    for (int node = n0; node < n1; ++node) {
      const int jy = node / p, jz = node - jy * p;
      const int t = (bi.y + jy) % a.ny - y0;
      if (t < 0 || t >= TY) continue;
      const int z = (bi.z + jz) % nz;
      ldsf[2 * t * ls + z] += wx * pick(wy, jy, p) * pick(wz, jz, p);
    }
  }
}

// ------------------------------------- two-pass z lines for nz = 256 ---
// With nz = 256 (M = 128 complex points per packed real line) and TY = 16
// lines, the z FFT runs as 128 = 8 x 16: pass A, one thread per (line, n2),
// n2 < 16, does an 8-point FFT of x[16 n1 + n2] in registers and multiplies
// by W128^(n2 k1); pass B, one thread per (line, k1), k1 < 8, does a
// 16-point FFT of those values in registers, giving X[k1 + 8 k2].  The
// values cross between the passes once through LDS, instead of once per
// radix-4 stage, and the R2C split (or C2R merge) happens in registers on
// the way to (or from) global memory.  LDS serves a 64-bit access 16 lanes
// at a time over 32 four-byte banks, so 16 lanes are conflict-free when
// their float2 indices differ modulo 16: with the lines as the fastest
// thread index and an odd line stride (kZ2Ls = 129 complex), every access
// of a pass (fixed slot, 16 lines) is.  The exchange slot of (k1, n2) is
// 16 k1 + n2.
constexpr int kZ2Ls = 129;

__device__ __forceinline__ int z2_slot(int k1, int n2) { return 16 * k1 + n2; }

// k1 of lane group kk = t / 16 in pass B: k1 and 8 - k1 (the R2C partners)
// share a wave (groups 2j and 2j + 1; k1 = 0 and 4 are their own partners).
__device__ __forceinline__ int z2_k1(int kk) { return (0x53627140 >> (4 * kk)) & 15; }

// Forward: the 16 spread lines in LDS -> half spectrum rows kz = 0..M of
// dst (row kz at dst + kz * ny, 16 contiguous y).  256 threads.
__device__ __forceinline__ void z2_forward_store(float2* lds, const float2* tws, float2* __restrict__ dst, int ny) {
  const int t = threadIdx.x;
  {
    const int l = t & 15, n2 = t >> 4;
    float2* L = lds + l * kZ2Ls;
    float2 v[8];
#pragma unroll
    for (int n1 = 0; n1 < 8; ++n1) v[n1] = L[16 * n1 + n2];
    fft8r<-1>(v);
    // W128^(n2 k1) from W128^n2, ^2n2, ^4n2 (three table reads, at most two
    // products each: fewer registers than seven reads in flight)
    {
      const float2 w1 = tws[(2 * n2) & 255], w2 = tws[(4 * n2) & 255], w4 = tws[(8 * n2) & 255];
      const float2 w3 = cmul(w1, w2);
      v[1] = cmul(v[1], w1);
      v[2] = cmul(v[2], w2);
      v[3] = cmul(v[3], w3);
      v[4] = cmul(v[4], w4);
      v[5] = cmul(v[5], cmul(w4, w1));
      v[6] = cmul(v[6], cmul(w4, w2));
      v[7] = cmul(v[7], cmul(w4, w3));
    }
    __syncthreads();  // every pass-A read is done
#pragma unroll
    for (int k1 = 0; k1 < 8; ++k1) L[z2_slot(k1, n2)] = v[k1];
  }
  __syncthreads();
  if (t < 128) {  // waves 0 and 1
    const int l = t & 15, kk = t >> 4, k1 = z2_k1(kk);
    const float2* L = lds + l * kZ2Ls;
    float2 z[16];
#pragma unroll
    for (int n2 = 0; n2 < 16; ++n2) z[n2] = L[z2_slot(k1, n2)];
    fft16r<-1>(z);  // z[k2] = Z[k1 + 8 k2]
    // Z[M - k] = Z[(8 - k1) + 8 (15 - k2)]: the partner lane's z[15 - k2]
    // (k1 = 4: its own), or z[16 - k2] for k1 = 0
    const int src = (l + 16 * (kk < 2 ? kk : (kk ^ 1))) & 63;
    const float2 w1 = tws[k1];  // W256^k1
    static_for<16>([&](auto kc) {
      constexpr int k2 = decltype(kc)::value;
      float2 zm = make_float2(__shfl(z[15 - k2].x, src), __shfl(z[15 - k2].y, src));
      if (k1 == 0) zm = z[(16 - k2) & 15];
      const float2 zk = z[k2];
      const int k = k1 + 8 * k2;
      if (k2 == 0 && k1 == 0) {
        // streaming stores (the y pass reads each value once)
        FD3_ST_NT(dst + l, make_float2(zk.x + zk.y, 0.0f));
        FD3_ST_NT(dst + (size_t)128 * ny + l, make_float2(zk.x - zk.y, 0.0f));
        return;
      }
      // X_k = E_k + W256^k O_k, E_k = (Z_k + conj Z_{M-k}) / 2, O_k = -i (Z_k - conj Z_{M-k}) / 2
      const float2 E = make_float2(0.5f * (zk.x + zm.x), 0.5f * (zk.y - zm.y));
      const float2 O = make_float2(0.5f * (zk.y + zm.y), -0.5f * (zk.x - zm.x));
      const float2 T = cmul(tw_mul<32, k2, -1>(w1), O);
      FD3_ST_NT(dst + (size_t)k * ny + l, make_float2(E.x + T.x, E.y + T.y));
    });
  }
}

// Inverse: half spectrum rows kz = 0..M of src -> the 16 real lines in LDS
// (unnormalised, as c2r_merge + lds_fft<1>).  256 threads; ends with a barrier.
__device__ __forceinline__ void z2_load_inverse(float2* lds, const float2* tws, const float2* __restrict__ src,
                                                int ny) {
  const int t = threadIdx.x;
  if (t < 128) {
    const int l = t & 15, k1 = t >> 4;
    const float2 w1 = conjf2(tws[k1]);  // W256^-k1 (inverse sign)
    float2 z[16];
    // four rounds of 8 loads (X_k and X_{M-k} of four k): all 32 at once
    // would hold 64 VGPRs beside z and halve the occupancy
    static_for<4>([&](auto hc) {
      constexpr int h = decltype(hc)::value;
      float2 xa[4], xb[4];
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        const int k = k1 + 8 * (4 * h + u);
        xa[u] = src[(size_t)k * ny + l];
        xb[u] = src[(size_t)(128 - k) * ny + l];
      }
      static_for<4>([&](auto uc) {
        constexpr int u = decltype(uc)::value, k2 = 4 * h + u;
        const float2 xk = xa[u], xm = xb[u];
        // Z_k = (X_k + conj X_{M-k}) + i (X_k - conj X_{M-k}) w^k, w = exp(2 pi i / 256)
        const float2 Ak = make_float2(xk.x + xm.x, xk.y - xm.y);
        const float2 Bk = cmul(make_float2(xk.x - xm.x, xk.y + xm.y), tw_mul<32, k2, 1>(w1));
        z[k2] = make_float2(Ak.x - Bk.y, Ak.y + Bk.x);
        if (k2 == 0 && k1 == 0) z[0] = make_float2(xk.x + xm.x, xk.x - xm.x);  // X_0, X_M real
      });
      asm volatile("" ::: "memory");  // keep the rounds apart
    });
    fft16r<1>(z);  // over k2 -> n2
    // W128^-(k1 n2) from W128^k1, ^2k1, ^4k1, ^8k1 (four table reads, at
    // most three products each)
    {
      const float2 w1 = conjf2(tws[(2 * k1) & 255]), w2 = conjf2(tws[(4 * k1) & 255]);
      const float2 w4 = conjf2(tws[(8 * k1) & 255]), w8 = conjf2(tws[(16 * k1) & 255]);
      const float2 w3 = cmul(w1, w2);
      float2 lo[8];
      lo[0] = make_float2(1.f, 0.f); lo[1] = w1; lo[2] = w2; lo[3] = w3;
      lo[4] = w4; lo[5] = cmul(w4, w1); lo[6] = cmul(w4, w2); lo[7] = cmul(w4, w3);
#pragma unroll
      for (int n2 = 1; n2 < 8; ++n2) z[n2] = cmul(z[n2], lo[n2]);
      z[8] = cmul(z[8], w8);
#pragma unroll
      for (int n2 = 9; n2 < 16; ++n2) z[n2] = cmul(z[n2], cmul(w8, lo[n2 - 8]));
    }
    float2* L = lds + l * kZ2Ls;
#pragma unroll
    for (int n2 = 0; n2 < 16; ++n2) L[z2_slot(k1, n2)] = z[n2];
  }
  __syncthreads();
  {
    const int l = t & 15, n2 = t >> 4;
    float2* L = lds + l * kZ2Ls;
    float2 v[8];
#pragma unroll
    for (int k1 = 0; k1 < 8; ++k1) v[k1] = L[z2_slot(k1, n2)];
    fft8r<1>(v);  // over k1 -> n1: z[n2 + 16 n1]
    __syncthreads();  // every read of the exchange slots is done
#pragma unroll
    for (int n1 = 0; n1 < 8; ++n1) L[n2 + 16 * n1] = v[n1];
  }
  __syncthreads();
}

// fused z passes: 256-thread blocks of kZLines complex values (measured
// against 512 threads x 32 lines and 128 x 8); z_tile relies on the two
// being equal
constexpr int kZLogSpread = 8, kZLogGather = 8;
static_assert(kZLogSpread == kZLogGather, "one z tile for both passes");
// grid (ny/TY, nx, C); LDS: TY lines of M+1 complex (= nz + 2 floats).
// MAXB: butterflies per thread of lds_fft (TY * M <= 4 MAXB 2^LOG).
struct ZPassArgs {
  FusedArgs a;
  float2* A;  // z-pass spectrum A[c][x][kz][y] (written by the spread, read by the gather)
  int TY;
  const float2* twz;
};

// Z2: nz = 256 and TY = 16, two-pass z FFT (z2_forward_store)
template <int LOG, int P, int MAXB, bool Z2 = false>  // P > 0: compile-time interpolation order
__device__ __forceinline__ void spread_r2c_z_body(const BlockInfo& bk, const ZPassArgs& za) {
  extern __shared__ float2 lds[];
  __shared__ Segs sg;
  const FusedArgs& a = za.a;
  float2* __restrict__ A = za.A;
  const float2* __restrict__ twz = za.twz;
  const int TY = Z2 ? 16 : za.TY;
  const int nz = Z2 ? 256 : a.nz, M = nz >> 1, ls = Z2 ? kZ2Ls : M + 1, p = P > 0 ? P : a.order;
  const int y0 = bk.id.x * TY, x = bk.id.y, c = bk.id.z;
  const int X = c / a.r, q = c - X * a.r;
  float2* tws = lds + TY * ls;
  stage_table(tws, twz, nz);
  segs_lookup(&sg, a.cb.colstart, X, x, y0, TY, a.nx, a.ny, p);
  for (int e = threadIdx.x; e < TY * ls; e += blockDim.x) lds[e] = make_float2(0.f, 0.f);
  __syncthreads();
  segs_compact(&sg, p);
  __syncthreads();
  float* ldsf = (float*)lds;
  const int total = sg.off[sg.n];
  if (a.det) {
    spread_ordered(ldsf, sg, a, q, y0, TY, nz, ls, p);
    __syncthreads();
  }
  for (int i = a.det ? total : threadIdx.x; i < total; i += blockDim.x) {
    int k = 0;
    while (i >= sg.off[k + 1]) ++k;
    const int ai = sg.a0[k] + (i - sg.off[k]);
    const int4 bi = a.cb.rec_i[ai];
    const float4 xf = a.cb.rec_f[ai];
    float w[kMaxOrder], dw[kMaxOrder], wy[kMaxOrder], wz[kMaxOrder];
    interp_weights<float>(p, a.interp, xf.x, w, dw);
    const float wx = pick(w, sg.jx[k], p) * a.charges[(size_t)bi.w * a.r + q];
    interp_weights<float>(p, a.interp, xf.y, wy, dw);
    interp_weights<float>(p, a.interp, xf.z, wz, dw);
    for (int jy = 0; jy < p; ++jy) {
      int y = bi.y + jy;
      if (y >= a.ny) y -= a.ny;
      const int t = y - y0;
      if (t < 0 || t >= TY) continue;
      float* line = ldsf + 2 * t * ls;
      const float v = wx * wy[jy];
      int zz[kMaxOrder];
      float add[kMaxOrder];
      for (int jz = 0; jz < p; ++jz) {
        int z = bi.z + jz;
        if (z >= nz) z -= nz;
        zz[jz] = z;
        add[jz] = v * wz[jz];
      }
      lds_add_f32_cas<P>(line, zz, add, p);
    }
  }
  __syncthreads();
  float2* dst = A + (size_t)(c * a.nx + x) * (M + 1) * a.ny + y0;
  if constexpr (Z2) {
    z2_forward_store(lds, tws, dst, a.ny);
  } else {
    lds_fft<-1, MAXB, (1 << LOG)>(lds, M, TY, ls, tws, 2);
    z_pairs<true>(lds, TY, ls, M, tws);
    __syncthreads();
    const int lT = ilog2(TY), ny = a.ny;
    tile2_from_lds<4 * MAXB + 1, LOG>(dst, lds, TileShape{M + 1, lT, 1, 1, ls}, StridedLines{ny});
  }
}

// ------------------------------------------------------ z pass + gather ---
// Z2: nz = 256 and TY = 16, two-pass z FFT (z2_load_inverse)
template <int LOG, int P, int MAXB, bool Z2 = false>  // P > 0: compile-time interpolation order
__device__ __forceinline__ void c2r_z_gather_body(const BlockInfo& bk, const ZPassArgs& za) {
  extern __shared__ float2 lds[];
  __shared__ Segs sg;
  const FusedArgs& a = za.a;
  const float2* __restrict__ A = za.A;
  const float2* __restrict__ twz = za.twz;
  const int TY = Z2 ? 16 : za.TY;
  const int nz = Z2 ? 256 : a.nz, M = nz >> 1, ls = Z2 ? kZ2Ls : M + 1, p = P > 0 ? P : a.order;
  const int y0 = bk.id.x * TY, x = bk.id.y, c = bk.id.z;
  const int X = c / a.r, q = c - X * a.r;
  float2* tws = lds + TY * ls;
  stage_table(tws, twz, nz);
  const float2* src = A + (size_t)(c * a.nx + x) * (M + 1) * a.ny + y0;
  const int lT = ilog2(TY), ny = a.ny;
  segs_lookup(&sg, a.cb.colstart, X, x, y0, TY, a.nx, a.ny, p);
  if constexpr (Z2) {
    __syncthreads();  // twiddles staged, segment lookups written
    segs_compact(&sg, p);
    z2_load_inverse(lds, tws, src, ny);  // its barriers publish sg
  } else {
    tile2_to_lds<4 * MAXB + 1, LOG>(lds, src, TileShape{M + 1, lT, 1, 1, ls}, StridedLines{ny});
    __syncthreads();
    segs_compact(&sg, p);
    z_pairs<false>(lds, TY, ls, M, tws);
    __syncthreads();
    lds_fft<1, MAXB, (1 << LOG)>(lds, M, TY, ls, tws, 2);
  }
  const float* ldsf = (const float*)lds;
  const int total = sg.off[sg.n];
  for (int i = threadIdx.x; i < total; i += blockDim.x) {
    int k = 0;
    while (i >= sg.off[k + 1]) ++k;
    const int ai = sg.a0[k] + (i - sg.off[k]);
    const int4 bi = a.cb.rec_i[ai];
    const float4 xf = a.cb.rec_f[ai];
    float w[kMaxOrder], dw[kMaxOrder], wy[kMaxOrder], dwy[kMaxOrder], wz[kMaxOrder], dwz[kMaxOrder];
    interp_weights<float>(p, a.interp, xf.x, w, dw);
    const float wx = pick(w, sg.jx[k], p), dwx = pick(dw, sg.jx[k], p);
    interp_weights<float>(p, a.interp, xf.y, wy, dwy);
    interp_weights<float>(p, a.interp, xf.z, wz, dwz);
    float acc = 0.f, g0 = 0.f, g1 = 0.f, g2 = 0.f;
    for (int jy = 0; jy < p; ++jy) {
      int y = bi.y + jy;
      if (y >= a.ny) y -= a.ny;
      const int t = y - y0;
      if (t < 0 || t >= TY) continue;
      const float* line = ldsf + 2 * t * ls;
      float sz = 0.f, sdz = 0.f;
      for (int jz = 0; jz < p; ++jz) {
        int z = bi.z + jz;
        if (z >= nz) z -= nz;
        const float v = line[z];
        sz += wz[jz] * v;
        sdz += dwz[jz] * v;
      }
      acc += wy[jy] * sz;
      g1 += dwy[jy] * sz;
      g2 += wy[jy] * sdz;
    }
    const float cq = a.charges[(size_t)bi.w * a.r + q];
    g0 = dwx * acc;
    // slot h = 0: the tile holding the atom's first y row, h = 1: the next
    // tile (rows that run past the first tile); the h = 0 block zeroes the
    // h = 1 slot when the footprint stays in its tile
    const bool first = bi.y >= y0 && bi.y < y0 + TY;
    const size_t slot = ((size_t)(sg.jx[k] * 2 + (first ? 0 : 1)) * a.r + q) * a.n + ai;
    a.cb.part[slot] = make_float4(wx * acc, cq * g0, cq * wx * g1, cq * wx * g2);
    if (first && (TY == a.ny || bi.y - y0 + p - 1 < TY))
      a.cb.part[slot + (size_t)a.r * a.n] = make_float4(0.f, 0.f, 0.f, 0.f);
  }
}

// y lines per block of the fused z passes: the plan's TY, raised (within the
// register budget of MAXB = 4) until an atom's p y nodes span at most two
// tiles, which the gather's two partial-sum slots per x node assume.
// The two-pass z lines: nz = 256, 16 lines per block of 256 threads.
static bool z2_fits(const FftPlan& p, int TY) { return p.nz == 256 && TY == 16 && kZLogSpread == 8; }

static int z_tile(const FusedArgs& a, const FftPlan& p, int* maxb) {
  int TY = p.TY;
  while (TY + 1 < a.order && TY < p.ny) TY <<= 1;
  const size_t vals = (size_t)TY * p.M;
  *maxb = vals <= (size_t)kZLines ? 2 : 4;
  if (vals > (size_t)4 * 4 * (1 << kZLogSpread)) {
    set_error("z tile of %d lines of %d exceeds the FFT register budget", TY, p.M);
    return -1;
  }
  return TY;
}

int spread_r2c_z(const FusedArgs& a, const FftPlan& p, const FftTables& t, float2* A, hipStream_t s) {
  int maxb;
  const int TY = z_tile(a, p, &maxb);
  if (TY < 0) return FD3_ERR_UNSUPPORTED;
  const bool z2 = z2_fits(p, TY);
  const size_t sh = (size_t)TY * (z2 ? kZ2Ls : p.M + 1) * sizeof(float2) + (size_t)p.nz * sizeof(float2);
  ProfScope _ps("k_spread_r2c_z", s);
  const dim3 grid(p.ny / TY, p.nx, p.C);
  constexpr int NT = 1 << kZLogSpread;
  const ZPassArgs za{a, A, TY, t.twz};
  if (z2 && a.order == 5)
    return launch<ZPassArgs, spread_r2c_z_body<kZLogSpread, 5, 2, true>, NT>("k_spread_r2c_z", grid, sh, s, za);
  if (z2) return launch<ZPassArgs, spread_r2c_z_body<kZLogSpread, 0, 2, true>, NT>("k_spread_r2c_z", grid, sh, s, za);
  if (maxb == 4) return launch<ZPassArgs, spread_r2c_z_body<kZLogSpread, 0, 4>, NT>("k_spread_r2c_z", grid, sh, s, za);
  if (a.order == 5)
    return launch<ZPassArgs, spread_r2c_z_body<kZLogSpread, 5, 2>, NT>("k_spread_r2c_z", grid, sh, s, za);
  return launch<ZPassArgs, spread_r2c_z_body<kZLogSpread, 0, 2>, NT>("k_spread_r2c_z", grid, sh, s, za);
}

// accA / accG of each atom from the gather's partial sums: p x nodes, two
// y slots, r channels (column order in, slot order out).  A fixed order of
// additions, no atomics.  ORD > 0: p = ORD, and the 2 ORD partials of a
// channel are loaded together (one memory wait per channel, not per partial).
template <int ORD>
__device__ __forceinline__ void gather_sum_body(const BlockInfo& bk, const FusedArgs& a) {
  const int64_t ai = (int64_t)bk.id.x * blockDim.x + threadIdx.x;
  if (ai >= a.n) return;
  const int s = a.cb.rec_i[ai].w, r = a.r, nk = 2 * a.order;
  const float4* pp = a.cb.part + ai;
  float g0 = 0.f, g1 = 0.f, g2 = 0.f;
  for (int q = 0; q < r; ++q) {
    float A = 0.f;
    if constexpr (ORD > 0) {
      float4 v[2 * ORD];
#pragma unroll
      for (int k = 0; k < 2 * ORD; ++k) v[k] = pp[((size_t)k * r + q) * a.n];
#pragma unroll
      for (int k = 0; k < 2 * ORD; ++k) {
        A += v[k].x; g0 += v[k].y; g1 += v[k].z; g2 += v[k].w;
      }
    } else {
      for (int k = 0; k < nk; ++k) {
        const float4 v = pp[((size_t)k * r + q) * a.n];
        A += v.x; g0 += v.y; g1 += v.z; g2 += v.w;
      }
    }
    a.cb.accA[(size_t)s * r + q] = A;
  }
  a.cb.accG[3 * (size_t)s + 0] = g0;
  a.cb.accG[3 * (size_t)s + 1] = g1;
  a.cb.accG[3 * (size_t)s + 2] = g2;
}

int c2r_z_gather(const FusedArgs& a, const FftPlan& p, const FftTables& t, const float2* A, hipStream_t s) {
  int maxb;
  const int TY = z_tile(a, p, &maxb);
  if (TY < 0) return FD3_ERR_UNSUPPORTED;
  const bool z2 = z2_fits(p, TY);
  const size_t sh = (size_t)TY * (z2 ? kZ2Ls : p.M + 1) * sizeof(float2) + (size_t)p.nz * sizeof(float2);
  ProfScope _ps("k_c2r_z_gather", s);
  const dim3 grid(p.ny / TY, p.nx, p.C);
  constexpr int NT = 1 << kZLogGather;
  const ZPassArgs za{a, const_cast<float2*>(A), TY, t.twz};
  int rc;
  if (z2 && a.order == 5)
    rc = launch<ZPassArgs, c2r_z_gather_body<kZLogGather, 5, 2, true>, NT>("k_c2r_z_gather", grid, sh, s, za);
  else if (z2)
    rc = launch<ZPassArgs, c2r_z_gather_body<kZLogGather, 0, 2, true>, NT>("k_c2r_z_gather", grid, sh, s, za);
  else if (maxb == 4) rc = launch<ZPassArgs, c2r_z_gather_body<kZLogGather, 0, 4>, NT>("k_c2r_z_gather", grid, sh, s, za);
  else if (a.order == 5)
    rc = launch<ZPassArgs, c2r_z_gather_body<kZLogGather, 5, 2>, NT>("k_c2r_z_gather", grid, sh, s, za);
  else rc = launch<ZPassArgs, c2r_z_gather_body<kZLogGather, 0, 2>, NT>("k_c2r_z_gather", grid, sh, s, za);
  if (rc) return rc;
  if (a.n > 0 && a.order == 5)
    return launch<FusedArgs, gather_sum_body<5>, 256>("k_gather_sum", dim3(cdiv(a.n, 256)), 0, s, a);
  if (a.n > 0) return launch<FusedArgs, gather_sum_body<0>, 256>("k_gather_sum", dim3(cdiv(a.n, 256)), 0, s, a);
  return FD3_OK;
}

// ------------------------------------------------------ C6 chain adjoint ---
// Per atom, from the gathered sums (in double, like the forward chain):
//   dE/dc_q  = accA_q + lambda_q c_q phi_XX(0)       (self term, src/fourierd3/core.py:547-550)
//   dE/dr_b  = sum_d accG_d n_d cell^-1[b][d]
//   dE/dCN   = sum_p L_p (g_p - sum_p' L_p' g_p') (-8)(CN' - ref_p) a_X,   g_p = sum_q dE/dc_q V[X,p,q]
// (adjoint of src/fourierd3/core.py:469-478, 530-532 and of cncorr, :383-384).
// Block partials (16 doubles): sum_i f_i (x) dE/dr_i, f = r cell^-1.
__device__ __forceinline__ void c6_bwd_body(const BlockInfo& bk, const C6BwdArgs& a) {
  __shared__ double red[16 * 9];
  double inv[9];
  for (int i = 0; i < 9; ++i) inv[i] = a.dyn ? a.dyn->inv[i] : a.inv[i];
  const int64_t s = (int64_t)bk.id.x * blockDim.x + threadIdx.x;
  double part[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
  if (s < a.n) {
    const int X = a.spec_s[s];
    const double ns[3] = {(double)a.nx, (double)a.ny, (double)a.nz};
    if (!a.dEdq) {
      const double4 r = a.pos4[s];
      double gr[3], f[3];
      for (int d = 0; d < 3; ++d) f[d] = r.x * inv[0 * 3 + d] + r.y * inv[1 * 3 + d] + r.z * inv[2 * 3 + d];
      for (int b = 0; b < 3; ++b) {
        gr[b] = 0.0;
        for (int d = 0; d < 3; ++d) gr[b] += (double)a.accG[3 * s + d] * ns[d] * inv[b * 3 + d];
        a.gpos[3 * s + b] = gr[b];
      }
      for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) part[i * 3 + j] = f[i] * gr[j];
    }
    const float qqx = a.qq[X * a.nS + X];
    const double R = a.a1 * sqrt(3.0 * (double)qqx) + a.a2;
    const double R2 = R * R, R6 = R2 * R2 * R2;
    const double phi0 = a.s6 / R6 + 3.0 * a.s8 * (double)qqx / (R6 * R2);
    double cn = a.cn_raw[s];
    double ca = 1.0;
    if (a.cncorr) {
      ca = (double)a.cncorr[2 * X + 1];
      cn = ca * cn + (double)a.cncorr[2 * X];
    }
    double L[kNRef];
    c6_softmax(a.cnref, X, cn, L);
    double gp[kNRef] = {0, 0, 0, 0, 0, 0, 0};
    for (int q = 0; q < a.r; ++q) {
      // with the charges the mesh and the self term used (float-rounded)
      double dc;
      if (a.dEdq) {
        dc = a.dEdq[s * a.r + q];
      } else {
        dc = (double)a.accA[s * a.r + q] + (double)a.eigs[q] * (double)a.charges[s * a.r + q] * phi0;
        if (a.k0) dc += (double)a.k0[X * a.r + q].x;
      }
      for (int pp = 0; pp < kNRef; ++pp) gp[pp] += dc * (double)a.vq[(X * kNRef + pp) * a.r + q];
    }
    double gbar = 0.0;
    for (int pp = 0; pp < kNRef; ++pp) gbar += L[pp] * gp[pp];
    double dcn = 0.0;
    for (int pp = 0; pp < kNRef; ++pp) {
      const float ref = a.cnref[X * kNRef + pp];
      if (ref != -1.0f) dcn += L[pp] * (gp[pp] - gbar) * (-8.0) * (cn - (double)ref);
    }
    a.dEdcn[s] = dcn * ca;
  }
  if (a.partials) block_sum_store<9>(part, red, a.partials + (size_t)bk.id.x * 16);
}

int launch_c6_bwd(const C6BwdArgs& a, hipStream_t s) {
  ProfScope _ps("k_c6_bwd", s);
  return launch<C6BwdArgs, c6_bwd_body, 256>("k_c6_bwd", dim3(c6_blocks(a.n)), 0, s, a);
}

}  // namespace fd3
