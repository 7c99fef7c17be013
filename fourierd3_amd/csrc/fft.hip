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

#include "fd3_internal.h"
#include "fft_plan.h"
#include "fft_lds.h"

namespace fd3 {

// ------------------------------------------------------------ tables ---
// tw_axis[k] = exp(-2 pi i k / n) for each axis; sinc_axis[m] = sinc(m/n)^(2p)
// with m the fftfreq integer of index m (rfftfreq for z).
__global__ void k_init_tables(float2* twx, float2* twy, float2* twz, float* scx, float* scy, float* scz, int nx,
                              int ny, int nz, int order, int use_sinc) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
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
  for (int e = threadIdx.x; e < TY * M; e += blockDim.x) lds[(e >> lM) * ls + (e & (M - 1))] = src[e];
  __syncthreads();
  lds_fft<-1, kFftMaxB>(lds, M, TY, ls, tws, 2);
  z_pairs<true>(lds, TY, ls, M, tws);  // X[k] = E[k] + W_nz^k O[k]
  __syncthreads();
  float2* dst = A + (size_t)(c * nx + x) * (M + 1) * ny + y0;
  for (int e = threadIdx.x; e < TY * (M + 1); e += blockDim.x) {
    const int kz = e >> lT, t = e & (TY - 1);
    dst[(size_t)kz * ny + t] = lds[t * ls + kz];
  }
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
  for (int e = threadIdx.x; e < TY * (M + 1); e += blockDim.x) {
    const int kz = e >> lT, t = e & (TY - 1);
    lds[t * ls + kz] = src[(size_t)kz * ny + t];
  }
  __syncthreads();
  z_pairs<false>(lds, TY, ls, M, tws);
  __syncthreads();
  lds_fft<1, kFftMaxB>(lds, M, TY, ls, tws, 2);
  float2* dst = (float2*)out + ((size_t)(c * nx + x) * ny + y0) * M;
  for (int e = threadIdx.x; e < TY * M; e += blockDim.x) dst[e] = lds[(e >> lM) * ls + (e & (M - 1))];
}

// ------------------------------------------------------------ y passes ---
// forward: A[c][x][kz][0..ny) (x in tile) -> B[c][kz][ky][x]; grid (nx/TX, M+1, C)
template <int DIR>
__global__ void __launch_bounds__(256) k_c2c_y(const float2* __restrict__ in, float2* __restrict__ out, int nx,
                                               int ny, int M1, int TX, const float2* __restrict__ twy) {
  extern __shared__ float2 lds[];
  const int ls = ny + 1;
  const int x0 = blockIdx.x * TX, kz = blockIdx.y, c = blockIdx.z;
  float2* tws = lds + TX * ls;
  stage_table(tws, twy, ny);
  const int lny = ilog2(ny), lTX = ilog2(TX);
  if (DIR < 0) {
    const float2* src = in + ((size_t)(c * nx + x0) * M1 + kz) * ny;
    for (int e = threadIdx.x; e < TX * ny; e += blockDim.x) {
      const int l = e >> lny, y = e & (ny - 1);
      lds[l * ls + y] = src[(size_t)l * M1 * ny + y];
    }
  } else {
    const float2* src = in + ((size_t)(c * M1 + kz) * ny) * nx + x0;
    for (int e = threadIdx.x; e < TX * ny; e += blockDim.x) {
      const int ky = e >> lTX, l = e & (TX - 1);
      lds[l * ls + ky] = src[(size_t)ky * nx + l];
    }
  }
  __syncthreads();
  lds_fft<DIR, kFftMaxB>(lds, ny, TX, ls, tws, 1);
  if (DIR < 0) {
    float2* dst = out + ((size_t)(c * M1 + kz) * ny) * nx + x0;
    for (int e = threadIdx.x; e < TX * ny; e += blockDim.x) {
      const int ky = e >> lTX, l = e & (TX - 1);
      dst[(size_t)ky * nx + l] = lds[l * ls + ky];
    }
  } else {
    float2* dst = out + ((size_t)(c * nx + x0) * M1 + kz) * ny;
    for (int e = threadIdx.x; e < TX * ny; e += blockDim.x) {
      const int l = e >> lny, y = e & (ny - 1);
      dst[(size_t)l * M1 * ny + y] = lds[l * ls + y];
    }
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
  for (int e = threadIdx.x; e < nl * nx; e += blockDim.x) lds[(e >> lnx) * ls + (e & (nx - 1))] = base[e];
  __syncthreads();
  lds_fft<DIR, kFftMaxB>(lds, nx, nl, ls, tws, 1);
  for (int e = threadIdx.x; e < nl * nx; e += blockDim.x) base[e] = lds[(e >> lnx) * ls + (e & (nx - 1))];
}

// ------------------------------------------------- fused k-space kernel ---
// For one kz plane, NK consecutive ky rows and RT ranks [q0, q0+RT) of every
// species: forward x FFT of the nS*RT*NK lines, then per k-point
//   phi_X^q = sum_Y G~_XY S_Y^q,  G~ = G(|k|) / sinc^2p (P3M only)
//   e      += w sum_X sum_q lambda_q Re(S_X^q conj(phi_X^q))
//   dsum   += w sum_XY dG~_XY/d|k| sum_q lambda_q Re(S_X^q conj(S_Y^q))
// (w = 1 on the kz = 0 and nz/2 planes, 2 elsewhere: the half-spectrum weights
// of src/fourierd3/pair_pot.py:209-218), writes -(lambda_q/Omega) phi back and
// inverse-transforms along x, leaving dE/drho's spectrum in place.
// Per-block partial sums (double): [0] sum e, [1..6] sum w dsum k_a k_b/|k|
// (xx,yy,zz,xy,xz,yz), [7..10] dE/d(s6,s8,a1,a2) spectral parts (unscaled).
struct KspaceArgs {
  float2* spec;
  const float* eigs;
  const float* qq;      // (nS, nS)
  const float2* twx;
  const float* scx; const float* scy; const float* scz;
  double* partials;     // (gridDim, 16)
  float recip[9];       // rows b_i of 2 pi cell^-T (Bohr^-1)
  float s6, s8, a1, a2;
  float neg_inv_vol;    // -1/Omega
  int nx, ny, nz, M1, nS, r, RT, NK, use_sinc, need_grad, need_params;
};

template <bool PARAMS>
__global__ void __launch_bounds__(256) k_kspace(KspaceArgs a) {
  extern __shared__ float2 lds[];
  __shared__ double red[16 * 16];
  const int nx = a.nx, nS = a.nS, RT = a.RT, NK = a.NK;
  const int ky0 = blockIdx.x * NK, kz = blockIdx.y, q0 = blockIdx.z * RT;
  const int rt = min(RT, a.r - q0);
  const int nlines = nS * RT * NK;
  const int ls = nx + 1;
  const int lnx = ilog2(nx), lNK = ilog2(NK), lRT = ilog2(RT);  // RT, NK, nx are powers of two
  float2* S = lds;
  float2* P = lds + nlines * ls;
  float2* tws = P + nlines * ls;              // twiddles of the x lines
  PairK* pairs = (PairK*)(tws + nx);          // (nS, nS)
  float* dR = (float*)(pairs + nS * nS);      // dR0_XY/da1 = sqrt(3 QQ)
  stage_table(tws, a.twx, nx);
  // per-pair constants
  for (int t = threadIdx.x; t < nS * nS; t += blockDim.x) {
    const float qq = a.qq[t];
    const float sq = sqrtf(3.0f * qq);
    const float R = a.a1 * sq + a.a2;
    PairK pk;
    pk.R = R;
    pk.p6 = a.s6 * float(2.0 * kPi * kPi / 3.0) / (R * R * R);
    pk.p8 = 3.0f * a.s8 * qq * float(-kPi * kPi) / (R * R * R * R * R);
    pairs[t] = pk;
    dR[t] = sq;
  }
  // load lines: line l = (X*RT + qi)*NK + kyi
  for (int e = threadIdx.x; e < nlines * nx; e += blockDim.x) {
    const int l = e >> lnx, kx = e & (nx - 1);
    const int kyi = l & (NK - 1), xq = l >> lNK, qi = xq & (RT - 1), X = xq >> lRT;
    float2 v = make_float2(0.f, 0.f);
    if (qi < rt) {
      const int c = X * a.r + q0 + qi;
      v = a.spec[(((size_t)c * a.M1 + kz) * a.ny + ky0 + kyi) * nx + kx];
    }
    S[l * ls + kx] = v;
  }
  __syncthreads();
  lds_fft<-1, kFftMaxB>(S, nx, nlines, ls, tws, 1);

  double acc[11];
#pragma unroll
  for (int i = 0; i < 11; ++i) acc[i] = 0.0;
  const float w = (kz == 0 || kz == a.M1 - 1) ? 1.0f : 2.0f;
  const float mz = (float)kz;
  for (int t = threadIdx.x; t < NK * nx; t += blockDim.x) {
    const int kyi = t >> lnx, kx = t & (nx - 1);
    const int ky = ky0 + kyi;
    const float mx = (float)(kx < (nx + 1) / 2 ? kx : kx - nx);
    const float my = (float)(ky < (a.ny + 1) / 2 ? ky : ky - a.ny);
    const float kv0 = mx * a.recip[0] + my * a.recip[3] + mz * a.recip[6];
    const float kv1 = mx * a.recip[1] + my * a.recip[4] + mz * a.recip[7];
    const float kv2 = mx * a.recip[2] + my * a.recip[5] + mz * a.recip[8];
    const float kn = sqrtf(kv0 * kv0 + kv1 * kv1 + kv2 * kv2);
    float sc = 1.0f;
    if (a.use_sinc) sc = fmaxf(a.scx[kx] * a.scy[ky] * a.scz[kz], 1e-10f);
    const float isc = 1.0f / sc;
    float e_acc = 0.f, d_acc = 0.f, p_acc[4] = {0.f, 0.f, 0.f, 0.f};
    // phi accumulates in P (this thread owns column (kyi, kx) of every line)
    for (int l = 0; l < nS * RT; ++l) P[(l * NK + kyi) * ls + kx] = make_float2(0.f, 0.f);
    // one Green's function evaluation per unordered species pair (G~ is symmetric)
    for (int X = 0; X < nS; ++X) {
      for (int Y = X; Y < nS; ++Y) {
        const PairK pk = pairs[X * nS + Y];
        float g, dg;
        green_eval(pk, kn, &g, &dg);
        const float gs = g * isc;
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
        d_acc += dg * isc * tre;
        if (PARAMS) {
          // dG/ds6 = f6, dG/ds8 = 3 QQ f8, dG/dR0 for a1 (x sqrt(3QQ)) and a2
          float h6, h6p, h8, h8p;
          const float R = pk.R;
          green_h(kn * R, &h6, &h6p, &h8, &h8p);
          const float f6 = float(2.0 * kPi * kPi / 3.0) / (R * R * R) * h6;
          const float qq = a.qq[X * nS + Y];
          const float f8q = 3.0f * qq * float(-kPi * kPi) / (R * R * R * R * R) * h8;
          const float dgdR = pk.p6 * (-3.0f * h6 / R + kn * h6p) + pk.p8 * (-5.0f * h8 / R + kn * h8p);
          p_acc[0] += f6 * isc * tre;
          p_acc[1] += f8q * isc * tre;
          p_acc[2] += dgdR * dR[X * nS + Y] * isc * tre;
          p_acc[3] += dgdR * isc * tre;
        }
      }
    }
    // energy sum_X sum_q lambda_q Re(S conj phi); P <- -(lambda_q / Omega) phi
    for (int X = 0; X < nS; ++X) {
      for (int qi = 0; qi < rt; ++qi) {
        const int ix = ((X * RT + qi) * NK + kyi) * ls + kx;
        const float lam = a.eigs[q0 + qi];
        const float2 sx = S[ix], ph = P[ix];
        e_acc += lam * (sx.x * ph.x + sx.y * ph.y);
        const float f = lam * a.neg_inv_vol;
        P[ix] = make_float2(f * ph.x, f * ph.y);
      }
    }
    acc[0] += (double)(w * e_acc);
    if (kn > 0.f) {
      const double s = (double)(w * d_acc) / (double)kn;
      acc[1] += s * kv0 * kv0; acc[2] += s * kv1 * kv1; acc[3] += s * kv2 * kv2;
      acc[4] += s * kv0 * kv1; acc[5] += s * kv0 * kv2; acc[6] += s * kv1 * kv2;
    }
    if (PARAMS) {
#pragma unroll
      for (int i = 0; i < 4; ++i) acc[7 + i] += (double)(w * p_acc[i]);
    }
  }
  const int bid = (blockIdx.z * gridDim.y + blockIdx.y) * gridDim.x + blockIdx.x;
  block_sum_store<11>(acc, red, a.partials + (size_t)bid * 16);
  if (!a.need_grad) return;
  lds_fft<1, kFftMaxB>(P, nx, nlines, ls, tws, 1);
  for (int e = threadIdx.x; e < nlines * nx; e += blockDim.x) {
    const int l = e >> lnx, kx = e & (nx - 1);
    const int kyi = l & (NK - 1), xq = l >> lNK, qi = xq & (RT - 1), X = xq >> lRT;
    if (qi < rt) {
      const int c = X * a.r + q0 + qi;
      a.spec[(((size_t)c * a.M1 + kz) * a.ny + ky0 + kyi) * nx + kx] = P[l * ls + kx];
    }
  }
}

// --------------------------------------------------------------- host ---
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
  p->TY = std::max(1, std::min(p->ny, budget / std::max(p->M, 1)));
  p->TY = 1 << (31 - __builtin_clz(p->TY));
  p->TX = std::max(1, std::min(p->nx, budget / p->ny));
  p->TX = 1 << (31 - __builtin_clz(p->TX));
  p->LPB = std::max(1, budget / (p->nx + 1));
  // k-space: two LDS buffers of nS*RT*NK lines of nx+1 complex plus the
  // per-pair table, within 62 KB (the block also holds 2 KB of static LDS)
  if (nS > 32) {
    set_error("the k-space kernel supports up to 32 species, got %d", nS);
    return FD3_ERR_UNSUPPORTED;
  }
  auto ks_bytes = [&](int rt_, int nk_) {
    return (size_t)2 * nS * rt_ * nk_ * (p->nx + 1) * sizeof(float2) + (size_t)p->nx * sizeof(float2) + (size_t)nS * nS * 16;
  };
  const size_t ks_budget = 62 * 1024;
  // RT: ranks per block, a power of two (a last partial tile is padded with zero lines)
  int RT = 1;
  while (RT < std::min(r, 8)) RT <<= 1;
  int NK = std::max(1, std::min(p->ny, 256 / p->nx));
  auto ks_fits = [&](int rt_, int nk_) {
    return ks_bytes(rt_, nk_) <= ks_budget && (size_t)nS * rt_ * nk_ * p->nx <= (size_t)kFftLines;
  };
  while (RT > 1 && !ks_fits(RT, NK)) RT >>= 1;
  while (NK > 1 && !ks_fits(RT, NK)) NK >>= 1;
  if (!ks_fits(RT, NK)) {
    set_error("k-space tile of %d species x %d points does not fit in LDS", nS, p->nx);
    return FD3_ERR_UNSUPPORTED;
  }
  p->RT = RT; p->NK = NK;
  p->ks_lds = ks_bytes(RT, NK);
  p->ks_blocks = (p->ny / NK) * p->M1 * ((r + RT - 1) / RT);
  // register budget of lds_fft<.., kFftMaxB>: lines * n <= kFftLines
  const size_t lim = kFftLines;
  if ((size_t)p->TY * p->M > lim || (size_t)p->TX * p->ny > lim || (size_t)p->LPB * p->nx > lim ||
      (size_t)nS * RT * NK * p->nx > lim) {
    set_error("internal FFT tiling exceeds the per-thread register budget");
    return FD3_ERR_UNSUPPORTED;
  }
  return FD3_OK;
}

void fft_tables_carve(const FftPlan& p, Arena& ar, FftTables* t) {
  t->twx = ar.take<float2>(p.nx);
  t->twy = ar.take<float2>(p.ny);
  t->twz = ar.take<float2>(p.nz);
  t->scx = ar.take<float>(p.nx);
  t->scy = ar.take<float>(p.ny);
  t->scz = ar.take<float>(p.M1);
}

int fft_tables_init(const FftPlan& p, const FftTables& t, int order, int use_sinc, hipStream_t s) {
  const int n = std::max(std::max(p.nx, p.ny), p.nz);
  ProfScope _ps("k_init_tables", s);
  hipLaunchKernelGGL(k_init_tables, dim3((n + 255) / 256), dim3(256), 0, s, t.twx, t.twy, t.twz, t.scx, t.scy, t.scz,
                     p.nx, p.ny, p.nz, order, use_sinc);
  FD3_LAUNCH_CHECK("k_init_tables");
  return FD3_OK;
}

int fft_forward_zy(const FftPlan& p, const FftTables& t, const float* rho, float2* A, float2* spec, hipStream_t s) {
  {
    const size_t sh = (size_t)p.TY * (p.M + 1) * sizeof(float2) + (size_t)p.nz * sizeof(float2);
    ProfScope _ps("k_r2c_z", s);
    hipLaunchKernelGGL(k_r2c_z, dim3(p.ny / p.TY, p.nx, p.C), dim3(256), sh, s, rho, A, p.nx, p.ny, p.nz, p.TY,
                       t.twz);
    FD3_LAUNCH_CHECK("k_r2c_z");
  }
  {
    const size_t sh = (size_t)p.TX * (p.ny + 1) * sizeof(float2) + (size_t)p.ny * sizeof(float2);
    ProfScope _ps("k_c2c_y_fwd", s);
    hipLaunchKernelGGL(k_c2c_y<-1>, dim3(p.nx / p.TX, p.M1, p.C), dim3(256), sh, s, A, spec, p.nx, p.ny, p.M1, p.TX,
                       t.twy);
    FD3_LAUNCH_CHECK("k_c2c_y fwd");
  }
  return FD3_OK;
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
  {
    const size_t sh = (size_t)p.TX * (p.ny + 1) * sizeof(float2) + (size_t)p.ny * sizeof(float2);
    ProfScope _ps("k_c2c_y_inv", s);
    hipLaunchKernelGGL(k_c2c_y<1>, dim3(p.nx / p.TX, p.M1, p.C), dim3(256), sh, s, spec, A, p.nx, p.ny, p.M1, p.TX,
                       t.twy);
    FD3_LAUNCH_CHECK("k_c2c_y inv");
  }
  {
    const size_t sh = (size_t)p.TY * (p.M + 1) * sizeof(float2) + (size_t)p.nz * sizeof(float2);
    ProfScope _ps("k_c2r_z", s);
    hipLaunchKernelGGL(k_c2r_z, dim3(p.ny / p.TY, p.nx, p.C), dim3(256), sh, s, A, out, p.nx, p.ny, p.nz, p.TY,
                       t.twz);
    FD3_LAUNCH_CHECK("k_c2r_z");
  }
  return FD3_OK;
}

int fft_y(const FftPlan& p, const FftTables& t, const float2* in, float2* out, int dir, hipStream_t s) {
  const size_t sh = (size_t)p.TX * (p.ny + 1) * sizeof(float2) + (size_t)p.ny * sizeof(float2);
  if (dir < 0) {
    ProfScope _ps("k_c2c_y_fwd", s);
    hipLaunchKernelGGL(k_c2c_y<-1>, dim3(p.nx / p.TX, p.M1, p.C), dim3(256), sh, s, in, out, p.nx, p.ny, p.M1, p.TX,
                       t.twy);
  } else {
    ProfScope _ps("k_c2c_y_inv", s);
    hipLaunchKernelGGL(k_c2c_y<1>, dim3(p.nx / p.TX, p.M1, p.C), dim3(256), sh, s, in, out, p.nx, p.ny, p.M1, p.TX,
                       t.twy);
  }
  FD3_LAUNCH_CHECK("k_c2c_y");
  return FD3_OK;
}

int kspace_run(const FftPlan& p, const FftTables& t, const Geom& g, float2* spec, const float* eigs, const float* qq,
               double s6, double s8, double a1, double a2, int need_grad, int need_params, double* partials,
               hipStream_t s) {
  KspaceArgs a;
  a.spec = spec; a.eigs = eigs; a.qq = qq; a.twx = t.twx;
  a.scx = t.scx; a.scy = t.scy; a.scz = t.scz; a.partials = partials;
  for (int i = 0; i < 9; ++i) a.recip[i] = (float)g.recip[i];
  a.s6 = (float)s6; a.s8 = (float)s8; a.a1 = (float)a1; a.a2 = (float)a2;
  a.neg_inv_vol = (float)(-1.0 / g.volume);
  a.nx = p.nx; a.ny = p.ny; a.nz = p.nz; a.M1 = p.M1; a.nS = p.nS; a.r = p.r; a.RT = p.RT; a.NK = p.NK;
  a.use_sinc = g.interp == FD3_INTERP_P3M;
  a.need_grad = need_grad; a.need_params = need_params;
  const size_t sh = p.ks_lds;
  const dim3 grid(p.ny / p.NK, p.M1, (p.r + p.RT - 1) / p.RT);
  ProfScope _ps("k_kspace", s);
  if (need_params)
    hipLaunchKernelGGL(k_kspace<true>, grid, dim3(256), sh, s, a);
  else
    hipLaunchKernelGGL(k_kspace<false>, grid, dim3(256), sh, s, a);
  FD3_LAUNCH_CHECK("k_kspace");
  return FD3_OK;
}

}  // namespace fd3
