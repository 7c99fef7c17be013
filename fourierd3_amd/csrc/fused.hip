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
  b->colcur = b->colcount + ncol;
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

template <int LOG, int P, int MAXB>  // P > 0: compile-time interpolation order
__device__ __forceinline__ void spread_r2c_z_body(const BlockInfo& bk, const ZPassArgs& za) {
  extern __shared__ float2 lds[];
  __shared__ Segs sg;
  const FusedArgs& a = za.a;
  float2* __restrict__ A = za.A;
  const float2* __restrict__ twz = za.twz;
  const int TY = za.TY;
  const int nz = a.nz, M = nz >> 1, ls = M + 1, p = P > 0 ? P : a.order;
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
  for (int i = threadIdx.x; i < total; i += blockDim.x) {
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
  lds_fft<-1, MAXB, (1 << LOG)>(lds, M, TY, ls, tws, 2);
  z_pairs<true>(lds, TY, ls, M, tws);
  __syncthreads();
  float2* dst = A + (size_t)(c * a.nx + x) * (M + 1) * a.ny + y0;
  const int lT = ilog2(TY), ny = a.ny;
  tile2_from_lds<4 * MAXB + 1, LOG>(dst, lds, TileShape{M + 1, lT, 1, 1, ls}, StridedLines{ny});
}

// ------------------------------------------------------ z pass + gather ---
template <int LOG, int P, int MAXB>  // P > 0: compile-time interpolation order
__device__ __forceinline__ void c2r_z_gather_body(const BlockInfo& bk, const ZPassArgs& za) {
  extern __shared__ float2 lds[];
  __shared__ Segs sg;
  const FusedArgs& a = za.a;
  const float2* __restrict__ A = za.A;
  const float2* __restrict__ twz = za.twz;
  const int TY = za.TY;
  const int nz = a.nz, M = nz >> 1, ls = M + 1, p = P > 0 ? P : a.order;
  const int y0 = bk.id.x * TY, x = bk.id.y, c = bk.id.z;
  const int X = c / a.r, q = c - X * a.r;
  float2* tws = lds + TY * ls;
  stage_table(tws, twz, nz);
  const float2* src = A + (size_t)(c * a.nx + x) * (M + 1) * a.ny + y0;
  const int lT = ilog2(TY), ny = a.ny;
  segs_lookup(&sg, a.cb.colstart, X, x, y0, TY, a.nx, a.ny, p);
  tile2_to_lds<4 * MAXB + 1, LOG>(lds, src, TileShape{M + 1, lT, 1, 1, ls}, StridedLines{ny});
  __syncthreads();
  segs_compact(&sg, p);
  z_pairs<false>(lds, TY, ls, M, tws);
  __syncthreads();
  lds_fft<1, MAXB, (1 << LOG)>(lds, M, TY, ls, tws, 2);
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
  const size_t sh = (size_t)TY * (p.M + 1) * sizeof(float2) + (size_t)p.nz * sizeof(float2);
  ProfScope _ps("k_spread_r2c_z", s);
  const dim3 grid(p.ny / TY, p.nx, p.C);
  constexpr int NT = 1 << kZLogSpread;
  const ZPassArgs za{a, A, TY, t.twz};
  if (maxb == 4) return launch<ZPassArgs, spread_r2c_z_body<kZLogSpread, 0, 4>, NT>("k_spread_r2c_z", grid, sh, s, za);
  if (a.order == 5)
    return launch<ZPassArgs, spread_r2c_z_body<kZLogSpread, 5, 2>, NT>("k_spread_r2c_z", grid, sh, s, za);
  return launch<ZPassArgs, spread_r2c_z_body<kZLogSpread, 0, 2>, NT>("k_spread_r2c_z", grid, sh, s, za);
}

// accA / accG of each atom from the gather's partial sums: p x nodes, two
// y slots, r channels (column order in, slot order out).  A fixed order of
// additions, no atomics.
__device__ __forceinline__ void gather_sum_body(const BlockInfo& bk, const FusedArgs& a) {
  const int64_t ai = (int64_t)bk.id.x * blockDim.x + threadIdx.x;
  if (ai >= a.n) return;
  const int s = a.cb.rec_i[ai].w, r = a.r, nk = 2 * a.order;
  const float4* pp = a.cb.part + ai;
  float g0 = 0.f, g1 = 0.f, g2 = 0.f;
  for (int q = 0; q < r; ++q) {
    float A = 0.f;
    for (int k = 0; k < nk; ++k) {
      const float4 v = pp[((size_t)k * r + q) * a.n];
      A += v.x; g0 += v.y; g1 += v.z; g2 += v.w;
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
  const size_t sh = (size_t)TY * (p.M + 1) * sizeof(float2) + (size_t)p.nz * sizeof(float2);
  ProfScope _ps("k_c2r_z_gather", s);
  const dim3 grid(p.ny / TY, p.nx, p.C);
  constexpr int NT = 1 << kZLogGather;
  const ZPassArgs za{a, const_cast<float2*>(A), TY, t.twz};
  int rc;
  if (maxb == 4) rc = launch<ZPassArgs, c2r_z_gather_body<kZLogGather, 0, 4>, NT>("k_c2r_z_gather", grid, sh, s, za);
  else if (a.order == 5)
    rc = launch<ZPassArgs, c2r_z_gather_body<kZLogGather, 5, 2>, NT>("k_c2r_z_gather", grid, sh, s, za);
  else rc = launch<ZPassArgs, c2r_z_gather_body<kZLogGather, 0, 2>, NT>("k_c2r_z_gather", grid, sh, s, za);
  if (rc) return rc;
  if (a.n > 0) return launch<FusedArgs, gather_sum_body, 256>("k_gather_sum", dim3(cdiv(a.n, 256)), 0, s, a);
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
