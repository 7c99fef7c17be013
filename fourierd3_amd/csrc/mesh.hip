// Per-atom kernels around the mesh: position staging, the CN -> C6 charge
// chain and its adjoint, B-spline spreading and gathering.
//
// Per-atom arrays are kept in a "working order" of slots: the cell-list order
// for the calculator path (atoms sorted by bin, so neighbouring atoms touch
// neighbouring mesh lines) and the caller's order for the edge-list path.
// slot_atom[s] maps a slot back to the caller's atom index.

#include "fd3_internal.h"
#include "launch.h"
#include "mesh_kernels.h"

namespace fd3 {

// Stage positions (Angstrom, f32/f64) as Bohr doubles in caller order.
__global__ void k_prep_pos(const void* pos, int dtype, const int32_t* __restrict__ species, int64_t n,
                           double4* __restrict__ pos4, int32_t* __restrict__ spec_s) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  double r[3];
  load_pos(pos, dtype, i, r);
  pos4[i] = make_double4(r[0], r[1], r[2], 0.0);
  spec_s[i] = species[i];
}

// CN -> L_p (softmax over valid references) -> c_q, and the per-atom self
// term.  src/fourierd3/core.py:383-384 (cncorr), 469-478 (weights),
// 530-532 (charges), 547-550 (V_self).  The chain runs in double (the
// softmax of -4 (CN - ref)^2 is steep, and its float rounding shows in the
// forces of heavy atoms); the charges are stored as float for the mesh, and
// the self term uses those same float-rounded charges so that it cancels the
// mesh self-interaction of exactly the charges that were spread.
// partials per block (8 doubles): [0] sum_i sum_q lambda_q c_iq^2 phi_XX(0),
// [1..4] the same sum with phi_XX(0) replaced by d phi_XX(0)/d(s6, s8, a1, a2).
__device__ __forceinline__ void c6_fwd_body(const BlockInfo& bk, const C6Args& a) {
  __shared__ double red[16 * 8];
  const int64_t s = (int64_t)bk.id.x * blockDim.x + threadIdx.x;
  double acc[5] = {0, 0, 0, 0, 0};
  if (s < a.n) {
    const int X = a.spec_s[s];
    double cn = a.cn_raw[s];
    if (a.cncorr) cn = (double)a.cncorr[2 * X + 1] * cn + (double)a.cncorr[2 * X];
    if (a.cn_out) a.cn_out[a.slot_atom ? a.slot_atom[s] : s] = (float)cn;
    double L[kNRef];
    c6_softmax(a.cnref, X, cn, L);
    const float qqx = a.qq[X * a.nS + X];
    const double R = a.a1 * sqrt(3.0 * (double)qqx) + a.a2;
    const double R2 = R * R, R6 = R2 * R2 * R2, R8 = R6 * R2;
    const double phi0 = a.s6 / R6 + 3.0 * a.s8 * qqx / R8;
    double t = 0.0;
    for (int q = 0; q < a.r; ++q) {
      double c = 0.0;
      for (int p = 0; p < kNRef; ++p) c += L[p] * (double)a.vq[(X * kNRef + p) * a.r + q];
      const float cf = (float)c;
      a.charges[s * a.r + q] = cf;
      t += (double)a.eigs[q] * (double)cf * (double)cf;
    }
    acc[0] = t * phi0;
    acc[1] = t / R6;
    acc[2] = t * 3.0 * qqx / R8;
    const double dphidR = -6.0 * a.s6 / (R6 * R) - 24.0 * a.s8 * qqx / (R8 * R);
    acc[3] = t * dphidR * sqrt(3.0 * (double)qqx);
    acc[4] = t * dphidR;
  }
  block_sum_store<5>(acc, red, a.partials + (size_t)bk.id.x * 8);
}

// Mesh coordinate setup of one atom: base node and 1-D weights per axis.
struct AtomStencil {
  int b[3];
  float w[3][kMaxOrder];
  float dw[3][kMaxOrder];
  double u[3];
};

__device__ inline void atom_stencil(const MeshGeom& g, const double4 r, AtomStencil* st) {
  const double ns[3] = {(double)g.nx, (double)g.ny, (double)g.nz};
  for (int d = 0; d < 3; ++d) {
    const double f = r.x * g.inv[0 * 3 + d] + r.y * g.inv[1 * 3 + d] + r.z * g.inv[2 * 3 + d];
    const double u = ns[d] * f;
    st->u[d] = f;
    double x;
    st->b[d] = interp_base(u, g.order, &x);
    interp_weights<float>(g.order, g.interp, (float)x, st->w[d], st->dw[d]);
  }
}

__device__ inline double4 atom_pos(const double4* pos4, const void* raw, int dtype, int64_t s) {
  if (pos4) return pos4[s];
  double r[3];
  load_pos(raw, dtype, s, r);
  return make_double4(r[0], r[1], r[2], 0.0);
}

// Spreading (points_to_mesh, src/fourierd3/mesh_interpolator_d3.py:99-133):
// rho[X*r + q][gx][gy][gz] += c_iq Wx Wy Wz over the p^3 nodes of atom i.
// LPA lanes per atom; lane (ia, ib) handles one x/y node pair and all z nodes.
template <int LPA>
__global__ void __launch_bounds__(256) k_spread(SpreadArgs a) {
  const int64_t gid = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t s = gid / LPA;
  const int lane = (int)(gid % LPA);
  const int p = a.g.order;
  if (s >= a.n || lane >= p * p) return;
  AtomStencil st;
  atom_stencil(a.g, atom_pos(a.pos4, a.raw_pos, a.raw_dtype, s), &st);
  const int ia = lane / p, ib = lane - ia * p;
  const int X = a.spec_s[s];
  const int gx = wrap(st.b[0] + ia, a.g.nx), gy = wrap(st.b[1] + ib, a.g.ny);
  const float wxy = st.w[0][ia] * st.w[1][ib];
  const size_t M = (size_t)a.g.nx * a.g.ny * a.g.nz;
  float* base = a.rho + (size_t)X * a.r * M + ((size_t)gx * a.g.ny + gy) * a.g.nz;
  const float* c = a.charges + s * a.r;
  for (int cz = 0; cz < p; ++cz) {
    const int gz = wrap(st.b[2] + cz, a.g.nz);
    const float v = wxy * st.w[2][cz];
    for (int q = 0; q < a.r; ++q) atomicAdd(base + q * M + gz, c[q] * v);
  }
}

template <int LPA>
__device__ inline float lane_sum(float v) {
#pragma unroll
  for (int off = LPA / 2; off > 0; off >>= 1) v += __shfl_xor(v, off, LPA);
  return v;
}

// Plain gather (mesh_to_points): out[i, q] = sum_nodes W mesh[X_i r + q].
template <int LPA, int RMAX>
__global__ void __launch_bounds__(256) k_gather_plain(GatherArgs a, float* out) {
  const int64_t gid = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t s = gid / LPA;
  const int lane = (int)(gid % LPA);
  const int p = a.g.order;
  float A[RMAX];
#pragma unroll
  for (int q = 0; q < RMAX; ++q) A[q] = 0.f;
  if (s < a.n && lane < p * p) {
    AtomStencil st;
    atom_stencil(a.g, atom_pos(a.pos4, a.raw_pos, a.raw_dtype, s), &st);
    const int X = a.spec_s[s];
    const int ia = lane / p, ib = lane - ia * p;
    const int gx = wrap(st.b[0] + ia, a.g.nx), gy = wrap(st.b[1] + ib, a.g.ny);
    const size_t M = (size_t)a.g.nx * a.g.ny * a.g.nz;
    const float* base = a.mesh + (size_t)X * a.r * M + ((size_t)gx * a.g.ny + gy) * a.g.nz;
    const float wxy = st.w[0][ia] * st.w[1][ib];
    for (int cz = 0; cz < p; ++cz) {
      const int gz = wrap(st.b[2] + cz, a.g.nz);
      const float w = wxy * st.w[2][cz];
#pragma unroll
      for (int q = 0; q < RMAX; ++q)
        if (q < a.r) A[q] += w * base[q * M + gz];
    }
  }
#pragma unroll
  for (int q = 0; q < RMAX; ++q) A[q] = lane_sum<LPA>(A[q]);
  if (s < a.n && lane == 0)
    for (int q = 0; q < a.r; ++q) out[s * a.r + q] = A[q];
}

__global__ void k_to_f32(const double* __restrict__ src, float* __restrict__ dst, int64_t n) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) dst[i] = (float)src[i];
}

int launch_to_f32(const double* src, float* dst, int64_t n, hipStream_t s) {
  if (n <= 0) return FD3_OK;
  hipLaunchKernelGGL(k_to_f32, dim3(cdiv(n, 256)), dim3(256), 0, s, src, dst, n);
  FD3_LAUNCH_CHECK("k_to_f32");
  return FD3_OK;
}

// grad_pos[slot_atom[s]] = gpos[s] * (Bohr per Angstrom)
__global__ void k_unpermute(const double* __restrict__ gpos, const int32_t* __restrict__ slot_atom, int64_t n,
                            double* __restrict__ out) {
  const int64_t s = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= n) return;
  const int64_t i = slot_atom ? slot_atom[s] : s;
  for (int b = 0; b < 3; ++b) out[3 * i + b] = gpos[3 * s + b] * kA2B;
}

template <typename T>
struct FillArgs {
  T* p;
  size_t n;
  T v;
};

template <typename T>
__device__ __forceinline__ void fill_body(const BlockInfo& bk, const FillArgs<T>& a) {
  for (size_t i = (size_t)bk.id.x * blockDim.x + threadIdx.x; i < a.n; i += (size_t)bk.n.x * blockDim.x) a.p[i] = a.v;
}

template <typename T>
static int launch_fill(T* p, size_t n, size_t max_blocks, hipStream_t s) {
  if (!n) return FD3_OK;
  const size_t blocks = std::min<size_t>((n + 255) / 256, max_blocks);
  return launch<FillArgs<T>, fill_body<T>, 256>("k_fill", dim3((unsigned)blocks), 0, s, FillArgs<T>{p, n, T(0)});
}

// ------------------------------------------------------------------ host ---
int launch_fill_f32(float* p, size_t n, hipStream_t s) {
  ProfScope _ps("k_fill", s);
  return launch_fill(p, n, 65536, s);
}
int launch_fill_f64(double* p, size_t n, hipStream_t s) { return launch_fill(p, n, 65536, s); }
int launch_fill_i32(int32_t* p, size_t n, hipStream_t s) { return launch_fill(p, n, 65536, s); }
int launch_fill_i64(int64_t* p, size_t n, hipStream_t s) { return launch_fill(p, n, 1, s); }

int launch_prep_pos(const void* pos, int dtype, const int32_t* species, int64_t n, double4* pos4, int32_t* spec_s,
                    hipStream_t s) {
  if (!n) return FD3_OK;
  ProfScope _ps("k_prep_pos", s);
  hipLaunchKernelGGL(k_prep_pos, dim3(cdiv(n, 256)), dim3(256), 0, s, pos, dtype, species, n, pos4, spec_s);
  FD3_LAUNCH_CHECK("k_prep_pos");
  return FD3_OK;
}

int c6_blocks(int64_t n) { return (int)std::max<int64_t>(1, cdiv(n, 256)); }

int launch_c6_fwd(const C6Args& a, hipStream_t s) {
  ProfScope _ps("k_c6_fwd", s);
  return launch<C6Args, c6_fwd_body, 256>("k_c6_fwd", dim3(c6_blocks(a.n)), 0, s, a);
}

static int lanes_per_atom(int p) { return p * p <= 32 ? 32 : 64; }

int gather_blocks(int64_t n, int order) { return (int)std::max<int64_t>(1, cdiv(n * lanes_per_atom(order), 256)); }

int launch_spread(const SpreadArgs& a, hipStream_t s) {
  if (!a.n) return FD3_OK;
  const int lpa = lanes_per_atom(a.g.order);
  const dim3 grid(cdiv(a.n * lpa, 256));
  ProfScope _ps("k_spread", s);
  if (lpa == 32)
    hipLaunchKernelGGL(k_spread<32>, grid, dim3(256), 0, s, a);
  else
    hipLaunchKernelGGL(k_spread<64>, grid, dim3(256), 0, s, a);
  FD3_LAUNCH_CHECK("k_spread");
  return FD3_OK;
}

int launch_gather_plain(const GatherArgs& a, float* out, hipStream_t s) {
  if (!a.n) return FD3_OK;
  const int lpa = lanes_per_atom(a.g.order);
  const dim3 grid(gather_blocks(a.n, a.g.order));
  if (a.r > 32) {
    set_error("rank %d exceeds the gather kernel's limit of 32", a.r);
    return FD3_ERR_UNSUPPORTED;
  }
  if (lpa == 32) {
    if (a.r <= 8) hipLaunchKernelGGL((k_gather_plain<32, 8>), grid, dim3(256), 0, s, a, out);
    else hipLaunchKernelGGL((k_gather_plain<32, 32>), grid, dim3(256), 0, s, a, out);
  } else {
    if (a.r <= 8) hipLaunchKernelGGL((k_gather_plain<64, 8>), grid, dim3(256), 0, s, a, out);
    else hipLaunchKernelGGL((k_gather_plain<64, 32>), grid, dim3(256), 0, s, a, out);
  }
  FD3_LAUNCH_CHECK("k_gather_plain");
  return FD3_OK;
}

int launch_unpermute(const double* gpos, const int32_t* slot_atom, int64_t n, double* out, hipStream_t s) {
  if (!n) return FD3_OK;
  ProfScope _ps("k_unpermute", s);
  hipLaunchKernelGGL(k_unpermute, dim3(cdiv(n, 256)), dim3(256), 0, s, gpos, slot_atom, n, out);
  FD3_LAUNCH_CHECK("k_unpermute");
  return FD3_OK;
}

}  // namespace fd3
