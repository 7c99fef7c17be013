// D4 coordination numbers in reciprocal space (cnfunc = 'd4'):
//   CN_i = (1/Omega) sum_nodes W_i phi_{X_i} - f_{X_i X_i}(0),
//   phi_X = C2R( sum_Y G^CN_XY(|k|) / sinc^2p(k) * R2C(rho_Y) ),
// rho_Y the one-hot density of species Y spread with the same weights as the
// D3 path (src/fourierd3/core.py:389-447, src/fourierd3/kspace_filter_d3.py:101-158).
// G^CN is the Fourier transform of f_XY(r) = (delta_XY / 2) erfc(k0 (r - r0_XY) / r0_XY)
// (src/fourierd3/pair_pot.py:235-402).
//
// The mesh passes reuse the fused spread / z FFT, y FFT and z FFT / gather
// kernels with one rank (channels = species).  The k-space step is its own
// kernel: x FFT, the species-pair convolution with G^CN from per-pair tables
// (degree-7 pieces on [0, kmax], as for the D3 Green's function), inverse x
// FFT.  Its backward
// variant applies G^CN to the dE/dCN-weighted density and accumulates the
// strain derivative of sum_j dE/dCN_j CN_j, for which the forward variant
// stores the x-transformed density spectrum.

#include "cn_d4.h"
#include "fft_reg.h"
#include "fft_lds.h"

namespace fd3 {

// G^CN in double (the table's samples): the same closed form and series.
__host__ __device__ inline double gcn_exact(double k, double r0, double delta, double k0) {
  const double x = k * r0, a = 0.25 / (k0 * k0), b = 0.5 / (k0 * k0);
  double P, Q;
  if (x < 0.5) {
    const double x2 = x * x;
    P = 1.0 / 3.0 + x2 * (-1.0 / 30.0 + x2 * (1.0 / 840.0 + x2 * (-1.0 / 45360.0 + x2 * (1.0 / 3991680.0))));
    Q = 1.0 + x2 * (-1.0 / 6.0 + x2 * (1.0 / 120.0 + x2 * (-1.0 / 5040.0 + x2 * (1.0 / 362880.0))));
  } else {
    const double s = sin(x), c = cos(x);
    P = (s - x * c) / (x * x * x);
    Q = s / x;
  }
  return 4.0 * kPi * delta * r0 * r0 * r0 * exp(-a * x * x) * (P + b * Q);
}

// Per-pair tables of G^CN on [0, kGtabN dk): degree-7 pieces, the layout of
// the D3 per-pair tables (fit on Chebyshev nodes in double, one thread per
// node; 32 intervals per block).
__global__ void __launch_bounds__(256) k_cn_tab_init(const float* r0, const float* delta, float k0, int nS, double dk,
                                                     float4* tab) {
  __shared__ double f[256];
  constexpr int GI = 256 / (kGtabDeg + 1);
  const int tid = threadIdx.x, np = nS * (nS + 1) / 2;
  const int it = blockIdx.x * GI + tid / (kGtabDeg + 1), j = tid % (kGtabDeg + 1);
  if (it < np * kGtabN) {
    const int pt = it / kGtabN, i = it - pt * kGtabN;  // pair pt = (X, Y >= X) in row order
    int X = 0, rr = 0;
    while (pt >= rr + nS - X) { rr += nS - X; ++X; }
    const int Y = X + (pt - rr);
    f[tid] = gcn_exact(gtab_node(dk, i, j), r0[X * nS + Y], delta[X * nS + Y], k0);
  }
  __syncthreads();
  const int io = blockIdx.x * GI + tid;
  if (tid < GI && io < np * kGtabN) gtab_fit(&f[(kGtabDeg + 1) * tid], tab + 2 * io);
}

int cn_tab_init(const CnKsArgs& a, double dk, hipStream_t s) {
  const int np = a.nS * (a.nS + 1) / 2;
  const int GI = 256 / (kGtabDeg + 1);
  ProfScope _ps("k_cn_tab_init", s);
  hipLaunchKernelGGL(k_cn_tab_init, dim3(cdiv(np * kGtabN, GI)), dim3(256), 0, s, a.r0, a.delta, a.k0, a.nS, dk,
                     a.tab);
  FD3_LAUNCH_CHECK("k_cn_tab_init");
  return FD3_OK;
}

__device__ inline float sinc_prod(const CnKsArgs& a, int kx, int ky, int kz) {
  if (!a.use_sinc) return 1.0f;
  return 1.0f / fmaxf(a.scx[kx] * a.scy[ky] * a.scz[kz], 1e-10f);
}

// One block: kz plane blockIdx.y, ky rows [ky0, ky0 + NK) for KT tiles.
// LDS: [x twiddles | S lines | P lines | line bases | R lines (backward)]; S
// holds the spectra of the nS species (line l = X * NK + kyi), P the
// convolution being formed, R the stored density spectrum of the same lines.
// NS > 0 (species count known at compile time, forward only): each thread
// keeps the nS values of its k point in registers and writes the convolution
// straight back into S, so there is no P region.
template <bool BWD, int MAXB, int NS>
__global__ void __launch_bounds__(256) k_kspace_cn(CnKsArgs a) {
  static_assert(!(BWD && NS > 0), "the register form is the forward pass only");
  if (a.dyn) {
    for (int i = 0; i < 9; ++i)
      a.recip[i] = __int_as_float(__builtin_amdgcn_readfirstlane(__float_as_int(a.dyn->recip_f[i])));
    a.inv_vol = a.dyn->inv_volume;
  }
  extern __shared__ float2 lds[];
  __shared__ double red[16 * 6];
  const int nx = a.nx, nS = NS > 0 ? NS : a.nS, NK = a.NK;
  const int kz = blockIdx.y;
  const int nlines = nS * NK, ls = nx + 1, lnx = ilog2(nx);
  float2* tws = lds;
  float2* S = tws + nx;
  float2* P = S + nlines * ls;
  long long* lbase = (long long*)(P + (NS > 0 ? 0 : nlines * ls));
  float2* R = (float2*)(lbase + nlines);  // backward only: the stored density spectrum of the tile
  stage_table(tws, a.twx, nx);
  const float w = (kz == 0 || kz == a.M1 - 1) ? 1.0f : 2.0f;
  double vir[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
  for (int it = 0; it < a.KT; ++it) {
    const int ky0 = (blockIdx.x * a.KT + it) * NK;
    if (it) __syncthreads();
    for (int l = threadIdx.x; l < nlines; l += blockDim.x) {
      const int kyi = l % NK, X = l / NK;
      lbase[l] = (((long long)X * a.M1 + kz) * a.ny + ky0 + kyi) * nx;
    }
    __syncthreads();
    lds_fft<-1, MAXB, 256>(S, nx, nlines, ls, tws, 1, GlobalLines{a.spec, lbase});
    if (!BWD) {  // keep the x-transformed density for the backward pass
      tile2_from_lds<4>(a.rho_hat, S, TileShape{nlines, lnx, 1, ls, 1}, LineTable{lbase});
      __syncthreads();  // the copy reads other threads' columns before they are overwritten
    } else {
      tile2_to_lds<4>(R, a.rho_hat, TileShape{nlines, lnx, 1, ls, 1}, LineTable{lbase});
      __syncthreads();
    }
    for (int t = threadIdx.x; t < NK * nx; t += blockDim.x) {
      const int kyi = t >> lnx, kx = t & (nx - 1), ky = ky0 + kyi;
      const float mx = (float)(kx < (nx + 1) / 2 ? kx : kx - nx);
      const float my = (float)(ky < (a.ny + 1) / 2 ? ky : ky - a.ny);
      const float mz = (float)kz;
      const float k0v = mx * a.recip[0] + my * a.recip[3] + mz * a.recip[6];
      const float k1v = mx * a.recip[1] + my * a.recip[4] + mz * a.recip[7];
      const float k2v = mx * a.recip[2] + my * a.recip[5] + mz * a.recip[8];
      const float kn = sqrtf(k0v * k0v + k1v * k1v + k2v * k2v);
      const float isc = sinc_prod(a, kx, ky, kz);
      const float fk = kn * a.tkinv;
      const int ti = min((int)fk, kGtabN - 1);
      const float tw = fk - (float)ti;
      const float4* te = a.tab + 2 * ti;
      if constexpr (NS > 0) {
        float2 sx[NS], px[NS];
#pragma unroll
        for (int X = 0; X < NS; ++X) {
          sx[X] = S[(X * NK + kyi) * ls + kx];
          px[X] = make_float2(0.f, 0.f);
        }
#pragma unroll
        for (int X = 0; X < NS; ++X)
#pragma unroll
          for (int Y = X; Y < NS; ++Y) {
            const int pt = X * NS - X * (X - 1) / 2 + (Y - X);
            float g, dg;
            gtab_eval(te[2 * kGtabN * pt], te[2 * kGtabN * pt + 1], tw, &g, &dg);
            g *= isc;
            px[X].x += g * sx[Y].x; px[X].y += g * sx[Y].y;
            if (Y != X) { px[Y].x += g * sx[X].x; px[Y].y += g * sx[X].y; }
          }
#pragma unroll
        for (int X = 0; X < NS; ++X) S[(X * NK + kyi) * ls + kx] = px[X];
        continue;
      }
      for (int X = 0; X < nS; ++X) P[(X * NK + kyi) * ls + kx] = make_float2(0.f, 0.f);
      float vs = 0.f;  // sum_XY dG/dk Re(conj(S_X) rho_Y)
      for (int X = 0, pt = 0; X < nS; ++X) {
        const int ix = (X * NK + kyi) * ls + kx;
        const float2 sx = S[ix];
        float2 rx = make_float2(0.f, 0.f);
        if (BWD) rx = R[ix];
        for (int Y = X; Y < nS; ++Y, ++pt) {
          const int iy = (Y * NK + kyi) * ls + kx;
          float g, dg;
          gtab_eval(te[2 * kGtabN * pt], te[2 * kGtabN * pt + 1], tw, &g, &dg);
          dg *= a.tkinv;  // dG/dw -> dG/dk
          g *= isc;
          const float2 sy = S[iy];
          float2 px = P[ix];
          px.x += g * sy.x; px.y += g * sy.y;
          P[ix] = px;
          if (Y != X) {
            float2 py = P[iy];
            py.x += g * sx.x; py.y += g * sx.y;
            P[iy] = py;
          }
          if (BWD) {
            const float2 ry = R[iy];
            float t2 = sx.x * ry.x + sx.y * ry.y;
            if (Y != X) t2 += sy.x * rx.x + sy.y * rx.y;
            vs += dg * isc * t2;
          }
        }
      }
      for (int X = 0; X < nS; ++X) S[(X * NK + kyi) * ls + kx] = P[(X * NK + kyi) * ls + kx];
      if (BWD && kn > 0.f) {
        // d|k|/d eps_ab = -k_a k_b / |k| under the strain cell' = (1 + eps) cell
        const double f = -(double)w * vs / kn;
        vir[0] += f * k0v * k0v; vir[1] += f * k1v * k1v; vir[2] += f * k2v * k2v;
        vir[3] += f * k0v * k1v; vir[4] += f * k0v * k2v; vir[5] += f * k1v * k2v;
      }
    }
    __syncthreads();
    lds_fft<1, MAXB, 256>(S, nx, nlines, ls, tws, 1, LdsLines{}, GlobalLines{a.spec, lbase});
  }
  if (BWD) {
    const int bid = blockIdx.y * gridDim.x + blockIdx.x;
#pragma unroll
    for (int i = 0; i < 6; ++i) vir[i] *= a.inv_vol;
    block_sum_store<6>(vir, red, a.partials + (size_t)bid * 8);
  }
}

// The same convolution with the split x FFT (nx = ny = 256, NS <= 4
// species): the y passes have run the first stage of the x FFT, so a thread
// (l, k1) of a tile (kKsxNK ky rows x NS species, line l = X * kKsxNK + kyi)
// does one 16-point FFT over n2 per direction; the NS values of a k point
// sit in registers.  The forward variant also stores the x-transformed
// density (natural order, rho_hat), the backward one reads it at its k
// points for the strain derivative.  Block b takes the KT consecutive tiles
// b KT .. b KT + KT - 1, ky tile fastest.
constexpr int kKsxNK = 4;

// The kernel is software-pipelined over the block's KT tiles, as the
// energy's k-space kernel (kspace_xsp_body, fft.hip): the per-pair G^CN
// tables and the sinc factors are staged in LDS once per block, each lane's
// 16 values of the next tile are loaded into registers right after the
// forward pass has put the current tile into LDS, the loop is rotated so
// that each wait for a prefetched tile comes after the previous tile's
// stores, and every lane stores on every path (lanes of idle lines into
// a.sink).  The backward variant reads the stored density of its four k
// points straight into registers (the next tile's after this tile's k
// points), not through LDS, so three blocks fit a CU.  LDS: the S lines,
// the tables, sinc_x, sinc_y (256 each) and sinc_z (M1): cn_ksp_lds.
template <bool BWD, int NS>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(3, 8))) k_kspace_cnp_xs(CnKsArgs a) {
  if (a.dyn) {
    for (int i = 0; i < 9; ++i)
      a.recip[i] = __int_as_float(__builtin_amdgcn_readfirstlane(__float_as_int(a.dyn->recip_f[i])));
    a.inv_vol = a.dyn->inv_volume;
  }
  constexpr int NK = kKsxNK, ls = 256, nlines = NS * NK, NP = NS * (NS + 1) / 2;
  extern __shared__ float4 lds4[];
  __shared__ double red[4 * 6];
  float2* S = (float2*)lds4;
  float4* tab = lds4 + nlines * ls / 2;
  float* scx_l = (float*)(tab + 2 * NP * kGtabN);  // 256
  float* scy_l = scx_l + 256;                       // 256
  float* scz_l = scy_l + 256;                       // M1
  for (int i = threadIdx.x; i < 2 * NP * kGtabN; i += blockDim.x) tab[i] = a.tab[i];
  if (a.use_sinc) {
    for (int i = threadIdx.x; i < 256; i += blockDim.x) {
      scx_l[i] = a.scx[i];
      scy_l[i] = a.scy[i];
    }
    for (int i = threadIdx.x; i < a.M1; i += blockDim.x) scz_l[i] = a.scz[i];
  }
  const int nyp = a.ny / NK, ntiles = nyp * a.M1;
  const int xl = threadIdx.x >> 4, xj = threadIdx.x & 15;
  const int xX = xl / NK, xkyi = xl - xX * NK;
  const bool live = xl < nlines;
  const int kx = threadIdx.x;  // column of this thread's four k points (rows 0..3 of the tile)
  // first element of this lane's line in tile u (idle lanes: -1)
  auto tile_lb = [&](int u) -> long long {
    const int kyp = u % nyp, kz = u / nyp;
    return live ? (((long long)xX * a.M1 + kz) * a.ny + kyp * NK + xkyi) * 256 : -1LL;
  };
  double vir[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
  const int u0 = blockIdx.x * a.KT, u1 = min(ntiles, u0 + a.KT);
  float2 v[16];
  float2 rr[BWD ? NK * NS : 1];  // backward: the stored density at the four k points, [kyi][X]
  auto load_tile = [&](long long lb) {  // idle lanes read the spectrum's first values (ignored)
    const float2* src = a.spec + (lb >= 0 ? lb + xj : xj);
#pragma unroll
    for (int n2 = 0; n2 < 16; ++n2) v[n2] = src[16 * n2];
  };
  auto load_r = [&](int u) {
    const int kyp = u % nyp, kz = u / nyp;
#pragma unroll
    for (int kyi = 0; kyi < NK; ++kyi)
#pragma unroll
      for (int X = 0; X < NS; ++X)
        rr[BWD ? kyi * NS + X : 0] = a.rho_hat[(((long long)X * a.M1 + kz) * a.ny + kyp * NK + kyi) * 256 + kx];
  };
  auto forward = [&](long long lb) {  // into LDS (and the forward variant's density out), then a barrier
    if (live) {
      float2 z[16];
#pragma unroll
      for (int n2 = 0; n2 < 16; ++n2) z[n2] = v[n2];
      fft16r<-1>(z);  // over n2 -> k2: X[xj + 16 k2]
#pragma unroll
      for (int k2 = 0; k2 < 16; ++k2) S[xl * ls + xj + 16 * k2] = z[k2];
      if constexpr (!BWD) {
#pragma unroll
        for (int k2 = 0; k2 < 16; ++k2) FD3_ST_NT(a.rho_hat + lb + xj + 16 * k2, z[k2]);  // read back much later
      }
    } else if constexpr (!BWD) {
#pragma unroll
      for (int k2 = 0; k2 < 16; ++k2) a.sink[xj + 16 * k2] = make_float2(0.f, 0.f);
    }
    __syncthreads();
  };
  if (u0 < u1) {
    long long lb0 = tile_lb(u0);
    if constexpr (BWD) load_r(u0);
    load_tile(lb0);
    __syncthreads();  // the staged tables
    forward(lb0);
    long long lb1 = u0 + 1 < u1 ? tile_lb(u0 + 1) : -1LL;
    if (u0 + 1 < u1) load_tile(lb1);
    for (int u = u0; u < u1; ++u) {
      asm volatile("" ::: "memory");  // the load is issued before the k points
      const int kyp = u % nyp, kz = u / nyp, ky0 = kyp * NK;
      const float w = (kz == 0 || kz == a.M1 - 1) ? 1.0f : 2.0f;
      const float scxz = a.use_sinc ? scx_l[kx] * scz_l[kz] : 1.0f;
#pragma unroll
      for (int kyi = 0; kyi < NK; ++kyi) {
        const int ky = ky0 + kyi;
        const float mx = (float)(kx < 128 ? kx : kx - 256);
        const float my = (float)(ky < (a.ny + 1) / 2 ? ky : ky - a.ny);
        const float mz = (float)kz;
        const float k0v = mx * a.recip[0] + my * a.recip[3] + mz * a.recip[6];
        const float k1v = mx * a.recip[1] + my * a.recip[4] + mz * a.recip[7];
        const float k2v = mx * a.recip[2] + my * a.recip[5] + mz * a.recip[8];
        const float kn = sqrtf(k0v * k0v + k1v * k1v + k2v * k2v);
        const float isc = a.use_sinc ? 1.0f / fmaxf(scxz * scy_l[ky], 1e-10f) : 1.0f;
        float gp[NP], dgp[NP];  // G^CN and dG^CN/dw of the species pairs (row order)
        ks_gtab_lds<NP>(tab, a.tkinv, kn, gp, dgp);
        float2 sx[NS], px[NS];
#pragma unroll
        for (int X = 0; X < NS; ++X) {
          sx[X] = S[(X * NK + kyi) * ls + kx];
          px[X] = make_float2(0.f, 0.f);
        }
        float vs = 0.f;  // sum_XY dG/dk Re(conj(S_X) rho_Y)
#pragma unroll
        for (int X = 0; X < NS; ++X)
#pragma unroll
          for (int Y = X; Y < NS; ++Y) {
            const int pt = X * NS - X * (X - 1) / 2 + (Y - X);
            const float g = gp[pt] * isc, dg = dgp[pt];
            px[X].x += g * sx[Y].x; px[X].y += g * sx[Y].y;
            if (Y != X) { px[Y].x += g * sx[X].x; px[Y].y += g * sx[X].y; }
            if constexpr (BWD) {
              const float2 rX = rr[kyi * NS + X], rY = rr[kyi * NS + Y];
              float t2 = sx[X].x * rY.x + sx[X].y * rY.y;
              if (Y != X) t2 += sx[Y].x * rX.x + sx[Y].y * rX.y;
              vs += dg * t2;
            }
          }
#pragma unroll
        for (int X = 0; X < NS; ++X) S[(X * NK + kyi) * ls + kx] = px[X];
        if (BWD && kn > 0.f) {
          // d|k|/d eps_ab = -k_a k_b / |k| under the strain cell' = (1 + eps) cell
          const double f = -(double)w * (double)(vs * a.tkinv * isc) / kn;
          vir[0] += f * k0v * k0v; vir[1] += f * k1v * k1v; vir[2] += f * k2v * k2v;
          vir[3] += f * k0v * k1v; vir[4] += f * k0v * k2v; vir[5] += f * k1v * k2v;
        }
        asm volatile("" ::: "memory");  // one k point at a time
      }
      if constexpr (BWD) {
        if (u + 1 < u1) load_r(u + 1);
      }
      __syncthreads();
      {
        float2 z[16];
        const float2* L = S + (live ? xl : 0) * ls;
#pragma unroll
        for (int k2 = 0; k2 < 16; ++k2) z[k2] = L[xj + 16 * k2];  // k1 = xj
        fft16r<1>(z);  // over k2 -> n2: Y[xj, n2]
        float2* dst = live ? a.spec + lb0 + xj : a.sink + xj;
#pragma unroll
        for (int n2 = 0; n2 < 16; ++n2) dst[16 * n2] = z[n2];
      }
      if (u + 1 < u1) {
        __syncthreads();  // every lane is done with the lines
        forward(lb1);
        lb0 = lb1;
        lb1 = u + 2 < u1 ? tile_lb(u + 2) : -1LL;
        if (u + 2 < u1) load_tile(lb1);
      }
    }
  }
  if (BWD) {
#pragma unroll
    for (int i = 0; i < 6; ++i) vir[i] *= a.inv_vol;
    block_sum_store<6>(vir, red, a.partials + (size_t)blockIdx.x * 8);
  }
}

// Dynamic LDS of k_kspace_cnp_xs.
static size_t cn_ksp_lds(int nS, int M1) {
  return (size_t)nS * kKsxNK * 256 * sizeof(float2) + (size_t)nS * (nS + 1) * kGtabN * sizeof(float4) +
         (size_t)(512 + M1) * sizeof(float);
}

int cn_ks_plan(const FftPlan& p, int nS, CnKsPlan* kp) {
  // two butterflies per thread (tiles of 2048 values: with 4 species and
  // 256-point lines a full tile of the unrolled FFT) where the species fit,
  // else up to eight
  kp->maxb = (size_t)nS * p.nx <= 2048 ? 2 : kKsGenericMaxB;
  kp->reg = nS <= 4;
  const size_t cap = (size_t)1024 * kp->maxb;
  int NK = 1;
  while (NK * 2 <= p.ny && (size_t)nS * NK * 2 * p.nx <= cap) NK <<= 1;
  if ((size_t)nS * NK * p.nx > cap) {
    set_error("D4 CN k-space tile of %d species x %d points exceeds the kernel's FFT capacity", nS, p.nx);
    return FD3_ERR_UNSUPPORTED;
  }
  kp->NK = NK;
  kp->KT = 1;
  while (kp->KT < 4 && (p.ny / NK) % (2 * kp->KT) == 0 && (size_t)(p.ny / NK / (2 * kp->KT)) * p.M1 >= 4096)
    kp->KT <<= 1;
  kp->blocks = (p.ny / NK / kp->KT) * p.M1;
  kp->xs = p.nx == 256 && p.ny == 256 && nS <= 4;
  if (kp->xs) {  // tiles of kKsxNK ky rows, three per block (k_kspace_cnp_xs)
    kp->NK = kKsxNK;
    kp->KT = 3;
    kp->blocks = (p.ny / kKsxNK * p.M1 + kp->KT - 1) / kp->KT;
  }
  const size_t tile = (size_t)nS * NK * (p.nx + 1) * sizeof(float2);
  // the backward variant stays generic: holding the stored density as well
  // takes the register form to about 170 VGPRs, and it measured slower
  kp->lds = (size_t)p.nx * sizeof(float2) + (kp->reg ? 1 : 2) * tile + (size_t)nS * NK * sizeof(long long);
  kp->lds_bwd = (size_t)p.nx * sizeof(float2) + 3 * tile + (size_t)nS * NK * sizeof(long long);
  if (kp->lds_bwd > (size_t)kKsGenericLds) {
    set_error("D4 CN k-space tile of %d species x %d points exceeds the LDS of a workgroup", nS, p.nx);
    return FD3_ERR_UNSUPPORTED;
  }
  return FD3_OK;
}

template <bool BWD, int MAXB, int NS>
static int cn_ks_launch(const CnKsArgs& b, const CnKsPlan& kp, dim3 grid, size_t sh, hipStream_t s) {
  if (sh > (size_t)64 * 1024)
    FD3_HIP(hipFuncSetAttribute((const void*)k_kspace_cn<BWD, MAXB, NS>,
                                hipFuncAttributeMaxDynamicSharedMemorySize, (int)sh));
  hipLaunchKernelGGL((k_kspace_cn<BWD, MAXB, NS>), grid, dim3(256), sh, s, b);
  return FD3_OK;
}

template <bool BWD, int MAXB>
static int cn_ks_species(const CnKsArgs& b, const CnKsPlan& kp, dim3 grid, size_t sh, hipStream_t s) {
  if constexpr (BWD) {
    return cn_ks_launch<true, MAXB, 0>(b, kp, grid, sh, s);
  } else {
    if (!kp.reg) return cn_ks_launch<false, MAXB, 0>(b, kp, grid, sh, s);
    switch (b.nS) {
      case 1: return cn_ks_launch<false, MAXB, 1>(b, kp, grid, sh, s);
      case 2: return cn_ks_launch<false, MAXB, 2>(b, kp, grid, sh, s);
      case 3: return cn_ks_launch<false, MAXB, 3>(b, kp, grid, sh, s);
      default: return cn_ks_launch<false, MAXB, 4>(b, kp, grid, sh, s);
    }
  }
}

template <bool BWD>
static int cn_ks_xs(const CnKsArgs& b, const CnKsPlan& kp, hipStream_t s) {
  const size_t sh = cn_ksp_lds(b.nS, b.M1);
  const dim3 grid(kp.blocks);
  switch (b.nS) {
    case 1: hipLaunchKernelGGL((k_kspace_cnp_xs<BWD, 1>), grid, dim3(256), sh, s, b); break;
    case 2: hipLaunchKernelGGL((k_kspace_cnp_xs<BWD, 2>), grid, dim3(256), sh, s, b); break;
    case 3: hipLaunchKernelGGL((k_kspace_cnp_xs<BWD, 3>), grid, dim3(256), sh, s, b); break;
    default: hipLaunchKernelGGL((k_kspace_cnp_xs<BWD, 4>), grid, dim3(256), sh, s, b); break;
  }
  return FD3_OK;
}

int cn_ks_run(const CnKsArgs& a, const CnKsPlan& kp, bool bwd, hipStream_t s) {
  if (kp.xs) {
    if (a.nx != 256 || a.ny != 256 || a.nS < 1 || a.nS > 4) {
      set_error("split x D4 CN k-space kernel planned for a %d x %d mesh with %d species", a.nx, a.ny, a.nS);
      return FD3_ERR_UNSUPPORTED;
    }
    CnKsArgs b = a;
    b.NK = kp.NK;
    b.KT = kp.KT;
    ProfScope _ps(bwd ? "k_kspace_cn_bwd" : "k_kspace_cn", s);
    const int rc = bwd ? cn_ks_xs<true>(b, kp, s) : cn_ks_xs<false>(b, kp, s);
    if (rc) return rc;
    FD3_LAUNCH_CHECK("k_kspace_cn");
    return FD3_OK;
  }
  const size_t sh = bwd ? kp.lds_bwd : kp.lds;
  const dim3 grid(a.ny / kp.NK / kp.KT, a.M1);
  CnKsArgs b = a;
  b.NK = kp.NK;
  b.KT = kp.KT;
  ProfScope _ps(bwd ? "k_kspace_cn_bwd" : "k_kspace_cn", s);
  int rc;
  if (kp.maxb == 2)
    rc = bwd ? cn_ks_species<true, 2>(b, kp, grid, sh, s) : cn_ks_species<false, 2>(b, kp, grid, sh, s);
  else
    rc = bwd ? cn_ks_species<true, kKsGenericMaxB>(b, kp, grid, sh, s)
             : cn_ks_species<false, kKsGenericMaxB>(b, kp, grid, sh, s);
  if (rc) return rc;
  FD3_LAUNCH_CHECK("k_kspace_cn");
  return FD3_OK;
}

// CN from the forward gather (accA = sum W phi, accG = sum dW/du phi): CN =
// accA / Omega - f_XX(0); accG is kept for the force.
__device__ inline void atom_refresh(CnAtomArgs& a) {
  if (!a.dyn) return;
  for (int i = 0; i < 9; ++i) a.inv[i] = a.dyn->inv[i];
  a.inv_vol = a.dyn->inv_volume;
}

__global__ void k_cn_d4_fin(CnAtomArgs a) {
  const int64_t s = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= a.n) return;
  atom_refresh(a);
  a.cn[s] = (double)a.accA[s] * a.inv_vol - (double)a.selfcont[a.spec_s[s]];
  for (int d = 0; d < 3; ++d) a.g1[3 * s + d] = a.accG[3 * s + d];
}

int cn_d4_fin(const CnAtomArgs& a, hipStream_t s) {
  if (a.n <= 0) return FD3_OK;
  ProfScope _ps("k_cn_d4_fin", s);
  hipLaunchKernelGGL(k_cn_d4_fin, dim3(cdiv(a.n, 256)), dim3(256), 0, s, a);
  FD3_LAUNCH_CHECK("k_cn_d4_fin");
  return FD3_OK;
}

// Force and the per-atom part of the strain derivative from the backward
// gather (accG = sum dW/du psi):
//   dE/du_i = (g_i g1_i + accG_i) / Omega  (mesh coordinates u = n f),
//   dE/dr_b += sum_d dE/du_d n_d cell^-1[b][d],
//   virial_ab -= delta_ab sum_i g_i (CN_i + f_XX(0))   (the 1/Omega of every CN).
__global__ void __launch_bounds__(256) k_cn_d4_force(CnAtomArgs a) {
  __shared__ double red[16 * 6];
  atom_refresh(a);
  const int64_t s = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  double v[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
  if (s < a.n) {
    const double g = a.dEdcn[s];
    const double ns[3] = {(double)a.nx, (double)a.ny, (double)a.nz};
    double du[3];
    for (int d = 0; d < 3; ++d) du[d] = (g * a.g1[3 * s + d] + (double)a.accG[3 * s + d]) * a.inv_vol;
    for (int b = 0; b < 3; ++b) {
      double t = 0.0;
      for (int d = 0; d < 3; ++d) t += du[d] * ns[d] * a.inv[b * 3 + d];
      a.gpos[3 * s + b] += t;
    }
    const double e = -g * (a.cn[s] + (double)a.selfcont[a.spec_s[s]]);
    v[0] = v[1] = v[2] = e;
  }
  block_sum_store<6>(v, red, a.partials + (size_t)blockIdx.x * 8);
}

int cn_d4_force(const CnAtomArgs& a, hipStream_t s) {
  if (a.n <= 0) return FD3_OK;
  ProfScope _ps("k_cn_d4_force", s);
  hipLaunchKernelGGL(k_cn_d4_force, dim3(cdiv(a.n, 256)), dim3(256), 0, s, a);
  FD3_LAUNCH_CHECK("k_cn_d4_force");
  return FD3_OK;
}

// A float buffer set to one value.
__global__ void k_fill_value(float* p, int64_t n, float v) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}

int fill_value(float* p, int64_t n, float v, hipStream_t s) {
  if (n <= 0) return FD3_OK;
  hipLaunchKernelGGL(k_fill_value, dim3(cdiv(n, 256)), dim3(256), 0, s, p, n, v);
  FD3_LAUNCH_CHECK("k_fill_value");
  return FD3_OK;
}

}  // namespace fd3
