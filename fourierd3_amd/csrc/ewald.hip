// method='ewald': the structure-factor sum over the full k grid
// (src/fourierd3/core.py:225-242, 563-577).
//
// With m_d in fftfreq(ns_d) * ns_d (k = 0 included, no half-spectrum and no
// mesh), k = sum_d m_d b_d and fractional coordinates f_i = r_i cell^-1,
//   S_X^q(k)   = sum_{i in X} c_iq exp(2 pi i m.f_i)
//   phi_X^q(k) = sum_Y G_XY(|k|) S_Y^q(k)
//   E_k        = sum_k sum_X sum_q |lambda_q| Re(S_X^q conj phi_X^q)
// (the reference weights the bilinear form with sqrt|lambda| on both sides),
// and E = -E_k / (2 Omega) + V_self / 2 as for SPME.  The adjoint is
//   dE/dc_iq = -(|lambda_q| / Omega) sum_k Re(e_ik conj phi_X^q)
//   dE/df_id = (2 pi / Omega) sum_k m_d sum_q |lambda_q| c_iq Im(e_ik conj phi_X^q)
// with e_ik = exp(2 pi i m.f_i).  The phases m.f are formed in double and
// reduced to [-1/2, 1/2] before the float sine/cosine, so the exponentials
// keep float accuracy for any m.
//
// k_ewald_sf: one thread per k point: S (atoms of each species staged in LDS
// in turn), then phi, the energy, virial and dE/dtheta partials in the
// k-space kernel's format.  k_ewald_adj: a block per (256 atoms, k range),
// partial sums added atomically into the per-atom accumulators that the C6
// adjoint reads.  Atoms are first partitioned by species (one block, stable).

#include "ewald.h"
#include "fd3_internal.h"
#include "mesh_kernels.h"

namespace fd3 {

constexpr int kEwRMax = 32;  // ranks held per thread
constexpr int kEwKChunk = 32;  // k points per LDS stage of the adjoint

__device__ inline float2 phase_exp(double t) {  // exp(2 pi i t)
  const double f = t - rint(t);
  float s, c;
  sincospif(2.0f * (float)f, &s, &c);
  return make_float2(c, s);
}

__device__ inline int fftfreq_m(int i, int n) { return i < (n + 1) / 2 ? i : i - n; }

// Stable partition of the atoms by species: perm lists the atoms of species
// 0, then 1, ...; seg[X] .. seg[X+1] is species X.  One block of 1024 threads.
__global__ void __launch_bounds__(1024) k_ewald_partition(const int32_t* __restrict__ spec, int64_t n, int nS,
                                                          int32_t* __restrict__ perm, int32_t* __restrict__ seg) {
  __shared__ int wtot[16];
  __shared__ int base;
  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
  if (tid == 0) base = 0;
  __syncthreads();
  for (int X = 0; X < nS; ++X) {
    if (tid == 0) seg[X] = base;
    for (int64_t c0 = 0; c0 < n; c0 += 1024) {
      const int64_t i = c0 + tid;
      const bool f = i < n && spec[i] == X;
      const unsigned long long bal = __ballot(f);
      const int pre = __popcll(bal & ((1ull << lane) - 1ull));
      if (lane == 0) wtot[w] = __popcll(bal);
      __syncthreads();
      int off = base;
      for (int v = 0; v < w; ++v) off += wtot[v];
      if (f) perm[off + pre] = (int32_t)i;
      __syncthreads();
      if (tid == 0) {
        int t = 0;
        for (int v = 0; v < 16; ++v) t += wtot[v];
        base += t;
      }
      __syncthreads();
    }
  }
  if (tid == 0) seg[nS] = (int)n;
}

__device__ inline void frac_of(const double4* __restrict__ pos4, int64_t i, const double* inv, double* f) {
  const double4 r = pos4[i];
  for (int d = 0; d < 3; ++d) f[d] = r.x * inv[0 * 3 + d] + r.y * inv[1 * 3 + d] + r.z * inv[2 * 3 + d];
}

__global__ void __launch_bounds__(256) k_ewald_sf(EwaldArgs a) {
  __shared__ double fr[256][3];
  __shared__ float ch[256][kEwRMax];
  __shared__ float4 htab[3 * (kHtabN + 1)];
  __shared__ double red[16 * 16];
  for (int i = threadIdx.x; i < 3 * (kHtabN + 1); i += blockDim.x) htab[i] = a.htab[i];
  const int r = a.r, nS = a.nS;
  const int64_t nk = (int64_t)a.nx * a.ny * a.nz;
  const int64_t kk = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const bool live = kk < nk;
  const int64_t kc = live ? kk : 0;
  const int iz = (int)(kc % a.nz), iy = (int)((kc / a.nz) % a.ny), ix = (int)(kc / ((int64_t)a.nz * a.ny));
  const double m0 = fftfreq_m(ix, a.nx), m1 = fftfreq_m(iy, a.ny), m2 = fftfreq_m(iz, a.nz);
  // structure factors, species by species
  for (int X = 0; X < nS; ++X) {
    float2 acc[kEwRMax];
#pragma unroll
    for (int q = 0; q < kEwRMax; ++q) acc[q] = make_float2(0.f, 0.f);
    const int s0 = a.seg[X], s1 = a.seg[X + 1];
    for (int c0 = s0; c0 < s1; c0 += 256) {
      __syncthreads();
      const int j = c0 + threadIdx.x;
      if (j < s1) {
        const int i = a.perm[j];
        double f[3];
        frac_of(a.pos4, i, a.inv, f);
        fr[threadIdx.x][0] = f[0]; fr[threadIdx.x][1] = f[1]; fr[threadIdx.x][2] = f[2];
        for (int q = 0; q < r; ++q) ch[threadIdx.x][q] = a.charges[(size_t)i * r + q];
      }
      __syncthreads();
      const int nc = min(256, s1 - c0);
      for (int jj = 0; jj < nc; ++jj) {
        const float2 e = phase_exp(m0 * fr[jj][0] + m1 * fr[jj][1] + m2 * fr[jj][2]);
#pragma unroll
        for (int q = 0; q < kEwRMax; ++q)
          if (q < r) {
            acc[q].x += ch[jj][q] * e.x;
            acc[q].y += ch[jj][q] * e.y;
          }
      }
    }
    if (live)
#pragma unroll
      for (int q = 0; q < kEwRMax; ++q)
        if (q < r) a.S[(kk * nS + X) * r + q] = acc[q];
  }
  __syncthreads();  // htab staged
  // phi, energy, virial and parameter partials
  double acc[11];
#pragma unroll
  for (int i = 0; i < 11; ++i) acc[i] = 0.0;
  if (live) {
    const float kv0 = (float)(m0 * a.recip[0] + m1 * a.recip[3] + m2 * a.recip[6]);
    const float kv1 = (float)(m0 * a.recip[1] + m1 * a.recip[4] + m2 * a.recip[7]);
    const float kv2 = (float)(m0 * a.recip[2] + m1 * a.recip[5] + m2 * a.recip[8]);
    const float kn = sqrtf(kv0 * kv0 + kv1 * kv1 + kv2 * kv2);
    double e_acc = 0.0, d_acc = 0.0, p_acc[4] = {0.0, 0.0, 0.0, 0.0};
    const float2* Sk = a.S + kk * nS * r;
    for (int X = 0; X < nS; ++X) {
      float2 ph[kEwRMax], sx[kEwRMax];
#pragma unroll
      for (int q = 0; q < kEwRMax; ++q) {
        ph[q] = make_float2(0.f, 0.f);
        sx[q] = q < r ? Sk[X * r + q] : make_float2(0.f, 0.f);
      }
      for (int Y = 0; Y < nS; ++Y) {
        const KsPair pk = a.pc[X * nS + Y];
        float g, dg, h6, d6, h8, d8;
        green_eval_tab(pk, htab, kn, &g, &dg, &h6, &d6, &h8, &d8);
        float tre = 0.f;  // sum_q |lambda_q| Re(S_X conj S_Y)
#pragma unroll
        for (int q = 0; q < kEwRMax; ++q)
          if (q < r) {
            const float2 sy = Sk[Y * r + q];
            ph[q].x += g * sy.x;
            ph[q].y += g * sy.y;
            tre += fabsf(a.eigs[q]) * (sx[q].x * sy.x + sx[q].y * sy.y);
          }
        d_acc += (double)(dg * tre);
        if (a.need_params) {
          const float R = pk.Rs * kHtabDx;
          const float h6p = d6 * kHtabInvDx, h8p = d8 * kHtabInvDx;
          const float dgdR = pk.p6 * (-3.0f * h6 / R + kn * h6p) + pk.p8 * (-5.0f * h8 / R + kn * h8p);
          p_acc[0] += (double)(pk.pf6 * h6 * tre);
          p_acc[1] += (double)(pk.pf8 * h8 * tre);
          p_acc[2] += (double)(dgdR * pk.dR * tre);
          p_acc[3] += (double)(dgdR * tre);
        }
      }
#pragma unroll
      for (int q = 0; q < kEwRMax; ++q)
        if (q < r) {
          e_acc += (double)(fabsf(a.eigs[q]) * (sx[q].x * ph[q].x + sx[q].y * ph[q].y));
          a.phi[(kk * nS + X) * r + q] = ph[q];
        }
    }
    acc[0] = e_acc;
    if (kn > 0.f) {
      const double s = d_acc / (double)kn;
      acc[1] = s * kv0 * kv0; acc[2] = s * kv1 * kv1; acc[3] = s * kv2 * kv2;
      acc[4] = s * kv0 * kv1; acc[5] = s * kv0 * kv2; acc[6] = s * kv1 * kv2;
    }
    for (int i = 0; i < 4; ++i) acc[7 + i] = p_acc[i];
  }
  block_sum_store<11>(acc, red, a.partials + (size_t)blockIdx.x * 16);
}

// grid (ceil(n / 256), ksplit): atoms of slot s = blockIdx.x * 256 + t, k in
// this block's range.
__global__ void __launch_bounds__(256) k_ewald_adj(EwaldArgs a, int64_t kper) {
  __shared__ float2 phs[kEwKChunk * 64];  // kEwKChunk k points x C <= 64 channels
  const int r = a.r, nS = a.nS, C = nS * r;
  const int64_t nk = (int64_t)a.nx * a.ny * a.nz;
  const int64_t s = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const bool live = s < a.n;
  const int64_t k0 = (int64_t)blockIdx.y * kper, k1 = min(nk, k0 + kper);
  int X = 0;
  double f[3] = {0.0, 0.0, 0.0};
  float cq[kEwRMax], lam[kEwRMax];
  if (live) {
    X = a.spec_s[s];
    frac_of(a.pos4, s, a.inv, f);
  }
#pragma unroll
  for (int q = 0; q < kEwRMax; ++q) {
    cq[q] = (live && q < r) ? a.charges[(size_t)s * r + q] : 0.f;
    lam[q] = q < r ? fabsf(a.eigs[q]) : 0.f;
  }
  double A[kEwRMax], B[3] = {0.0, 0.0, 0.0};
#pragma unroll
  for (int q = 0; q < kEwRMax; ++q) A[q] = 0.0;
  for (int64_t kc = k0; kc < k1; kc += kEwKChunk) {
    const int nc = (int)min((int64_t)kEwKChunk, k1 - kc);
    __syncthreads();
    for (int e = threadIdx.x; e < nc * C; e += blockDim.x) phs[e] = a.phi[kc * C + e];
    __syncthreads();
    if (!live) continue;
    for (int j = 0; j < nc; ++j) {
      const int64_t kk = kc + j;
      const int iz = (int)(kk % a.nz), iy = (int)((kk / a.nz) % a.ny), ix = (int)(kk / ((int64_t)a.nz * a.ny));
      const double m0 = fftfreq_m(ix, a.nx), m1 = fftfreq_m(iy, a.ny), m2 = fftfreq_m(iz, a.nz);
      const float2 e = phase_exp(m0 * f[0] + m1 * f[1] + m2 * f[2]);
      float bim = 0.f;
#pragma unroll
      for (int q = 0; q < kEwRMax; ++q)
        if (q < r) {
          const float2 p = phs[j * C + X * r + q];
          A[q] += (double)(e.x * p.x + e.y * p.y);
          bim += lam[q] * cq[q] * (e.y * p.x - e.x * p.y);
        }
      B[0] += m0 * bim; B[1] += m1 * bim; B[2] += m2 * bim;
    }
  }
  if (!live) return;
  const double io = 1.0 / a.volume;
#pragma unroll
  for (int q = 0; q < kEwRMax; ++q)
    if (q < r) atomicAdd(a.accA + s * r + q, (float)(-(double)lam[q] * io * A[q]));
  for (int d = 0; d < 3; ++d) atomicAdd(a.accG + 3 * s + d, (float)(2.0 * kPi * io * B[d]));
}

// ----------------------------------------------------------- D4 CN ---
template <bool BWD>
__global__ void __launch_bounds__(256) k_ewcn_sf(EwCnArgs a) {
  __shared__ double fr[256][3];
  __shared__ float wt[256];
  __shared__ double red[16 * 6];
  const int nS = a.nS;
  const int64_t nk = (int64_t)a.nx * a.ny * a.nz;
  const int64_t kk = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const bool live = kk < nk;
  const int64_t kc = live ? kk : 0;
  const int iz = (int)(kc % a.nz), iy = (int)((kc / a.nz) % a.ny), ix = (int)(kc / ((int64_t)a.nz * a.ny));
  const double m0 = fftfreq_m(ix, a.nx), m1 = fftfreq_m(iy, a.ny), m2 = fftfreq_m(iz, a.nz);
  for (int X = 0; X < nS; ++X) {
    float2 acc = make_float2(0.f, 0.f);
    const int s0 = a.seg[X], s1 = a.seg[X + 1];
    for (int c0 = s0; c0 < s1; c0 += 256) {
      __syncthreads();
      const int j = c0 + threadIdx.x;
      if (j < s1) {
        const int i = a.perm[j];
        double f[3];
        frac_of(a.pos4, i, a.inv, f);
        fr[threadIdx.x][0] = f[0]; fr[threadIdx.x][1] = f[1]; fr[threadIdx.x][2] = f[2];
        wt[threadIdx.x] = a.w[i];
      }
      __syncthreads();
      const int nc = min(256, s1 - c0);
      for (int jj = 0; jj < nc; ++jj) {
        const float2 e = phase_exp(m0 * fr[jj][0] + m1 * fr[jj][1] + m2 * fr[jj][2]);
        acc.x += wt[jj] * e.x;
        acc.y += wt[jj] * e.y;
      }
    }
    if (live) {
      a.S[kk * nS + X] = acc;
      if (!BWD) a.Srho[kk * nS + X] = acc;
    }
  }
  double vir[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
  if (live) {
    const float kv0 = (float)(m0 * a.recip[0] + m1 * a.recip[3] + m2 * a.recip[6]);
    const float kv1 = (float)(m0 * a.recip[1] + m1 * a.recip[4] + m2 * a.recip[7]);
    const float kv2 = (float)(m0 * a.recip[2] + m1 * a.recip[5] + m2 * a.recip[8]);
    const float kn = sqrtf(kv0 * kv0 + kv1 * kv1 + kv2 * kv2);
    const float fk = kn * a.tkinv;
    const int ti = min((int)fk, kGtabN - 1);
    const float tw = fk - (float)ti;
    const float4* te = a.tab + 2 * ti;
    const float2* Sk = a.S + kk * nS;
    const float2* Rk = a.Srho + kk * nS;
    float vs = 0.f;
    for (int X = 0; X < nS; ++X) {
      float2 ph = make_float2(0.f, 0.f);
      const float2 sx = Sk[X];
      for (int Y = 0; Y < nS; ++Y) {
        const int lo = min(X, Y), hi = max(X, Y);
        const int pt = lo * nS - lo * (lo - 1) / 2 + (hi - lo);  // pair (lo, hi) in row order
        float g, dg;
        gtab_eval(te[2 * kGtabN * pt], te[2 * kGtabN * pt + 1], tw, &g, &dg);
        const float2 sy = Sk[Y];
        ph.x += g * sy.x;
        ph.y += g * sy.y;
        if (BWD) {
          const float2 ry = Rk[Y];
          vs += dg * a.tkinv * (sx.x * ry.x + sx.y * ry.y);
        }
      }
      a.phi[kk * nS + X] = ph;
    }
    if (BWD && kn > 0.f) {
      const double f = -(double)vs / kn / a.volume;
      vir[0] = f * kv0 * kv0; vir[1] = f * kv1 * kv1; vir[2] = f * kv2 * kv2;
      vir[3] = f * kv0 * kv1; vir[4] = f * kv0 * kv2; vir[5] = f * kv1 * kv2;
    }
  }
  if (BWD) block_sum_store<6>(vir, red, a.partials + (size_t)blockIdx.x * 8);
}

// grid (ceil(n / 256), k ranges): per atom, this block's k range.
template <bool BWD>
__global__ void __launch_bounds__(256) k_ewcn_atom(EwCnArgs a, int64_t kper) {
  __shared__ float2 phs[kEwKChunk * 64];  // kEwKChunk k points x nS <= 64 species
  const int nS = a.nS;
  const int64_t nk = (int64_t)a.nx * a.ny * a.nz;
  const int64_t s = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const bool live = s < a.n;
  const int64_t k0 = (int64_t)blockIdx.y * kper, k1 = min(nk, k0 + kper);
  int X = 0;
  double f[3] = {0.0, 0.0, 0.0};
  if (live) {
    X = a.spec_s[s];
    frac_of(a.pos4, s, a.inv, f);
  }
  double A = 0.0, B[3] = {0.0, 0.0, 0.0};
  for (int64_t kc = k0; kc < k1; kc += kEwKChunk) {
    const int nc = (int)min((int64_t)kEwKChunk, k1 - kc);
    __syncthreads();
    for (int e = threadIdx.x; e < nc * nS; e += blockDim.x) phs[e] = a.phi[kc * nS + e];
    __syncthreads();
    if (!live) continue;
    for (int j = 0; j < nc; ++j) {
      const int64_t kk = kc + j;
      const int iz = (int)(kk % a.nz), iy = (int)((kk / a.nz) % a.ny), ix = (int)(kk / ((int64_t)a.nz * a.ny));
      const double m0 = fftfreq_m(ix, a.nx), m1 = fftfreq_m(iy, a.ny), m2 = fftfreq_m(iz, a.nz);
      const float2 e = phase_exp(m0 * f[0] + m1 * f[1] + m2 * f[2]);
      const float2 p = phs[j * nS + X];
      if (!BWD) A += (double)(e.x * p.x + e.y * p.y);
      const float im = e.x * p.y - e.y * p.x;  // Im(conj(e) phi)
      B[0] += m0 * im; B[1] += m1 * im; B[2] += m2 * im;
    }
  }
  if (!live) return;
  if (!BWD) atomicAdd(a.accA + s, (float)A);
  for (int d = 0; d < 3; ++d) atomicAdd(a.accG + 3 * s + d, (float)(2.0 * kPi * B[d]));
}

int ewald_partition(const int32_t* spec_s, int64_t n, int nS, int32_t* perm, int32_t* seg, hipStream_t s) {
  ProfScope _ps("k_ewald_partition", s);
  hipLaunchKernelGGL(k_ewald_partition, dim3(1), dim3(1024), 0, s, spec_s, n, nS, perm, seg);
  FD3_LAUNCH_CHECK("k_ewald_partition");
  return FD3_OK;
}

int ewcn_pass(const EwCnArgs& a, bool bwd, hipStream_t s) {
  if (a.nS > 64) {
    set_error("the Ewald D4 CN supports up to 64 species");
    return FD3_ERR_UNSUPPORTED;
  }
  const int64_t nk = (int64_t)a.nx * a.ny * a.nz;
  {
    ProfScope _ps(bwd ? "k_ewcn_sf_bwd" : "k_ewcn_sf", s);
    if (bwd) hipLaunchKernelGGL(k_ewcn_sf<true>, dim3((unsigned)((nk + 255) / 256)), dim3(256), 0, s, a);
    else hipLaunchKernelGGL(k_ewcn_sf<false>, dim3((unsigned)((nk + 255) / 256)), dim3(256), 0, s, a);
    FD3_LAUNCH_CHECK("k_ewcn_sf");
  }
  if (a.n == 0) return FD3_OK;
  int rc;
  if ((rc = launch_fill_f32(a.accA, (size_t)a.n, s))) return rc;
  if ((rc = launch_fill_f32(a.accG, (size_t)a.n * 3, s))) return rc;
  const unsigned nbx = (unsigned)((a.n + 255) / 256);
  int64_t ksplit = std::max<int64_t>(1, std::min<int64_t>((nk + kEwKChunk - 1) / kEwKChunk, 1024 / nbx + 1));
  int64_t kper = (nk + ksplit - 1) / ksplit;
  kper = (kper + kEwKChunk - 1) / kEwKChunk * kEwKChunk;
  ksplit = (nk + kper - 1) / kper;
  ProfScope _ps(bwd ? "k_ewcn_atom_bwd" : "k_ewcn_atom", s);
  if (bwd) hipLaunchKernelGGL(k_ewcn_atom<true>, dim3(nbx, (unsigned)ksplit), dim3(256), 0, s, a, kper);
  else hipLaunchKernelGGL(k_ewcn_atom<false>, dim3(nbx, (unsigned)ksplit), dim3(256), 0, s, a, kper);
  FD3_LAUNCH_CHECK("k_ewcn_atom");
  return FD3_OK;
}

// ------------------------------------------------------------------ host ---
int ewald_plan_make(const int32_t ns[3], int nS, int r, FftPlan* p) {
  for (int i = 0; i < 3; ++i)
    if (ns[i] < 1 || ns[i] > 4096) {
      set_error("Ewald k grid size %d on axis %d is outside [1, 4096]", ns[i], i);
      return FD3_ERR_UNSUPPORTED;
    }
  if (r > kEwRMax || nS * r > 64) {
    set_error("the Ewald kernels support rank <= %d and species x rank <= 64 (got %d x %d)", kEwRMax, nS, r);
    return FD3_ERR_UNSUPPORTED;
  }
  *p = FftPlan{};
  p->nx = ns[0]; p->ny = ns[1]; p->nz = ns[2];
  p->M = p->nz / 2;
  p->M1 = p->M + 1;
  p->nS = nS; p->r = r; p->C = nS * r;
  const int64_t nk = (int64_t)p->nx * p->ny * p->nz;
  p->ks_blocks = (int)((nk + 255) / 256);
  return FD3_OK;
}

size_t ewald_buffer_elems(const FftPlan& p) { return (size_t)p.nx * p.ny * p.nz * p.C; }

int ewald_run(const EwaldArgs& a0, bool grad, hipStream_t s) {
  EwaldArgs a = a0;
  const int64_t nk = (int64_t)a.nx * a.ny * a.nz;
  {
    ProfScope _ps("k_ewald_partition", s);
    hipLaunchKernelGGL(k_ewald_partition, dim3(1), dim3(1024), 0, s, a.spec_s, a.n, a.nS, a.perm, a.seg);
    FD3_LAUNCH_CHECK("k_ewald_partition");
  }
  {
    ProfScope _ps("k_ewald_sf", s);
    hipLaunchKernelGGL(k_ewald_sf, dim3((unsigned)((nk + 255) / 256)), dim3(256), 0, s, a);
    FD3_LAUNCH_CHECK("k_ewald_sf");
  }
  if (!grad || a.n == 0) return FD3_OK;
  int rc;
  if ((rc = launch_fill_f32(a.accA, (size_t)a.n * a.r, s))) return rc;
  if ((rc = launch_fill_f32(a.accG, (size_t)a.n * 3, s))) return rc;
  const unsigned nbx = (unsigned)((a.n + 255) / 256);
  // enough k ranges for ~4 blocks per CU, each range a whole number of chunks
  int64_t ksplit = std::max<int64_t>(1, std::min<int64_t>((nk + kEwKChunk - 1) / kEwKChunk, 1024 / nbx + 1));
  int64_t kper = (nk + ksplit - 1) / ksplit;
  kper = (kper + kEwKChunk - 1) / kEwKChunk * kEwKChunk;
  ksplit = (nk + kper - 1) / kper;
  ProfScope _ps("k_ewald_adj", s);
  hipLaunchKernelGGL(k_ewald_adj, dim3(nbx, (unsigned)ksplit), dim3(256), 0, s, a, kper);
  FD3_LAUNCH_CHECK("k_ewald_adj");
  return FD3_OK;
}

}  // namespace fd3
