// Internal device/host helpers shared by the fd3 kernels (gfx950, wave64).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>
#include <math.h>

#include "../../include/fd3.h"

namespace fd3 {

constexpr double kA2B = 1.0 / 0.52917726;  // Angstrom -> Bohr, src/fourierd3/core.py:120
constexpr double kPi = 3.14159265358979323846;
constexpr int kNRef = 7;
constexpr int kMaxOrder = 7;
constexpr int kMaxRank = 32;  // ranks of the C6 decomposition the kernels take

// Nontemporal (streaming) 8-byte global load and store of one complex value
// (float2).  Macros: through an inline function the store lost its
// nontemporal flag.  The mesh passes of the 256^2 plans use them where
// measured faster: the spread's stores and the forward y pass's loads of the
// z-pass spectrum, the inverse y pass's loads and stores, the split x
// k-space kernel's stores, and the D4 CN forward's stores of the density
// that its backward reads (the forward y pass then reads part of its input
// from the Infinity Cache).  Nontemporal loads in the k-space kernel and the
// gather, and stores in the forward y pass, measured slower.
#define FD3_LD_NT(p) __builtin_bit_cast(float2, __builtin_nontemporal_load((const double*)(p)))
#define FD3_ST_NT(p, v) __builtin_nontemporal_store(__builtin_bit_cast(double, (v)), (double*)(p))

// ---------------------------------------------------------------- errors ---
void set_error(const char* fmt, ...);
int hip_fail(hipError_t e, const char* what);

#define FD3_HIP(call)                                   \
  do {                                                  \
    hipError_t _e = (call);                             \
    if (_e != hipSuccess) return ::fd3::hip_fail(_e, #call); \
  } while (0)

#define FD3_LAUNCH_CHECK(name)                                   \
  do {                                                           \
    hipError_t _e = hipGetLastError();                           \
    if (_e != hipSuccess) return ::fd3::hip_fail(_e, name);      \
  } while (0)

// Optional per-kernel timing (fd3_profile_*): when enabled, launch sites wrap
// their launch in a ProfScope, which records HIP events on the launch stream.
bool prof_enabled();
void prof_set(bool on);
void prof_only(const char* name);
int prof_report(int max_entries, char* names, double* total_ms, int64_t* counts, int* n_entries);
void prof_record(const char* name, hipStream_t s, bool start);
struct ProfScope {
  const char* name;
  hipStream_t s;
  bool on;
  ProfScope(const char* n, hipStream_t st) : name(n), s(st), on(prof_enabled()) {
    if (on) prof_record(name, s, true);
  }
  ~ProfScope() {
    if (on) prof_record(name, s, false);
  }
};

// -------------------------------------------------------------- geometry ---
// Cell in device memory, written on the stream by k_cell_update from a
// device-resident cell (NPT steps: the cell changes without a host round
// trip, so a captured graph replays with the current cell).  Kernels whose
// arguments carry a non-null CellDyn pointer read the cell-dependent values
// from it instead of their by-value copies.
struct CellDyn {
  double cell[9], inv[9], recip[9];  // Bohr, as in Geom
  double volume, inv_volume;
  float recip_f[9];    // recip in float, as the k-space kernels take it
  float neg_inv_vol_f; // -1 / volume in float
};

// Box geometry in Bohr, passed to kernels by value.
struct Geom {
  double cell[9];   // rows = lattice vectors
  double inv[9];    // cell^-1
  double recip[9];  // rows b_i of 2 pi cell^-T
  double volume;
  int nx, ny, nz;
  int order, interp;
  const CellDyn* dyn = nullptr;  // device cell, or nullptr: the values above are used
  double kreach = 1.0;           // |k| tables cover kreach times this cell's largest |k|
};

int make_geom(const fd3_geom* g, Geom* out);  // validates, converts to Bohr

// ------------------------------------------------- interpolation weights ---
// Mesh coordinate u -> base node and per-node weights along one axis.
// Odd p: base = round(u), x = u - base; even p: base = floor(u), x = u - base - 1/2
// (torch-pme MeshInterpolator.compute_weights).  Node j (0..p-1) sits at
// base + j + 1 - ceil(p/2).  P3M weights are the centred cardinal B-spline
// M_p(u - node) written out as polynomials in x; Lagrange weights are the
// Lagrange basis through the p nodes.  dw = d w / d u.
template <typename T>
__host__ __device__ __forceinline__ void interp_weights(int p, int interp, T x, T* w, T* dw) {
  if (interp == FD3_INTERP_P3M) {
    const T x2 = x * x, x3 = x2 * x, x4 = x3 * x;
    switch (p) {
      case 1:
        w[0] = T(1); dw[0] = T(0);
        break;
      case 2:
        w[0] = T(0.5) - x; w[1] = T(0.5) + x;
        dw[0] = T(-1); dw[1] = T(1);
        break;
      case 3: {
        const T a = T(1) - T(2) * x, b = T(1) + T(2) * x;
        w[0] = a * a * T(0.125); w[1] = T(0.75) - x2; w[2] = b * b * T(0.125);
        dw[0] = -a * T(0.5); dw[1] = T(-2) * x; dw[2] = b * T(0.5);
        break;
      }
      case 4: {
        const T a = T(1) - T(2) * x, b = T(1) + T(2) * x;
        const T c = T(1.0 / 48.0);
        w[0] = a * a * a * c;
        w[1] = (T(23) - T(30) * x - T(12) * x2 + T(24) * x3) * c;
        w[2] = (T(23) + T(30) * x - T(12) * x2 - T(24) * x3) * c;
        w[3] = b * b * b * c;
        dw[0] = T(-6) * a * a * c;
        dw[1] = (T(-30) - T(24) * x + T(72) * x2) * c;
        dw[2] = (T(30) - T(24) * x - T(72) * x2) * c;
        dw[3] = T(6) * b * b * c;
        break;
      }
      default: {  // p == 5
        const T a = T(1) - T(2) * x, b = T(1) + T(2) * x;
        const T a2 = a * a, b2 = b * b;
        w[0] = a2 * a2 * T(1.0 / 384.0);
        w[1] = (T(19) - T(44) * x + T(24) * x2 + T(16) * x3 - T(16) * x4) * T(1.0 / 96.0);
        w[2] = (T(115) - T(120) * x2 + T(48) * x4) * T(1.0 / 192.0);
        w[3] = (T(19) + T(44) * x + T(24) * x2 - T(16) * x3 - T(16) * x4) * T(1.0 / 96.0);
        w[4] = b2 * b2 * T(1.0 / 384.0);
        dw[0] = T(-8) * a2 * a * T(1.0 / 384.0);
        dw[1] = (T(-44) + T(48) * x + T(48) * x2 - T(64) * x3) * T(1.0 / 96.0);
        dw[2] = (T(-240) * x + T(192) * x3) * T(1.0 / 192.0);
        dw[3] = (T(44) + T(48) * x - T(48) * x2 - T(64) * x3) * T(1.0 / 96.0);
        dw[4] = T(8) * b2 * b * T(1.0 / 384.0);
        break;
      }
    }
  } else {
    // Lagrange basis through nodes j0..j0+p-1 (relative to base), evaluated at
    // t = u - base = x (odd p) or x + 1/2 (even p).
    const int j0 = 1 - (p + 1) / 2;
    const T t = (p & 1) ? x : x + T(0.5);
    for (int j = 0; j < p; ++j) {
      T num = T(1), den = T(1), dnum = T(0);
      for (int k = 0; k < p; ++k) {
        if (k == j) continue;
        const T f = t - T(j0 + k);
        dnum = dnum * f + num;
        num = num * f;
        den = den * T(j - k);
      }
      w[j] = num / den;
      dw[j] = dnum / den;
    }
  }
}

// Base node and offset x of mesh coordinate u (double precision input).
__host__ __device__ inline int interp_base(double u, int p, double* x) {
  double b;
  if (p & 1) {
    b = floor(u + 0.5);  // round half up; ties are measure-zero
    *x = u - b;
  } else {
    b = floor(u);
    *x = u - b - 0.5;
  }
  return (int)b + 1 - (p + 1) / 2;  // index of node 0 (before the modulo)
}

__host__ __device__ inline int wrap(int i, int n) {
  int r = i % n;
  return r < 0 ? r + n : r;
}

// --------------------------------------------------------- Green's function ---
// G_XY(k) = s6 f6^(k; R) + 3 s8 QQ f8^(k; R), R = R0_XY, with the analytic
// Fourier transforms (src/fourierd3/pair_pot.py:152-175)
//   f6^(k) = 2 pi^2 / (3 R^3) h6(kR),  h6 = [e^-x - 2 e^-x/2 cos(pi/3 + sqrt3 x/2)] / x
//   f8^(k) =  -pi^2 / R^5     h8(kR),  h8 = [e^-x sin(pi/8) cos(pi/4 + x cos(pi/8))
//                                           + e^-x cos(pi/8) cos(3pi/4 + x sin(pi/8))] / x
// (h6(0) = 1, h8(0) = -sqrt2 sin(pi/8)).
// h6 and h8 do not depend on the species pair, so the k-space kernel reads
// them from a piecewise polynomial table instead of evaluating four
// exponentials and three sine/cosine pairs per pair and k point.  Interval i
// covers x in [i dx, (i+1) dx); on it each function is the degree-5
// polynomial interpolating it at the six Chebyshev nodes of the interval
// (interpolation error < 1e-8 of h(0) at dx = 1/2), stored as monomial
// coefficients in the local coordinate w = x/dx - i in [0, 1), which the
// device gets with one fract.  Entry 3i .. 3i+2 holds {a6_0..a6_3},
// {a6_4, a6_5, a8_0, a8_1}, {a8_2..a8_5}.  Interval kHtabN is all zeros and serves every
// x >= kHtabXmax = 64, where |h6| < 1e-15 and |h8| < 4e-13 (h6(0) = 1,
// h8(0) = -0.54): 1e-6 of float32 resolution, invisible in any sum.
constexpr int kHtabN = 128;
constexpr float kHtabDx = 0.5f;
constexpr float kHtabXmax = kHtabN * kHtabDx;
constexpr float kHtabInvDx = 1.0f / kHtabDx;  // dh/dx = kHtabInvDx dh/dw

// h6(x), h8(x) in double for x > 0 (closed forms; the cancellation at small x
// costs ~1e-16/x relative, < 1e-13 at the smallest table node).
__host__ __device__ inline void h_exact(double x, double* h6, double* h8) {
  const double s8 = 0.38268343236508977, c8 = 0.92387953251128674, sq3h = 0.86602540378443865;
  const double e2 = exp(-0.5 * x), e1 = e2 * e2;
  const double n6 = e1 - 2.0 * e2 * cos(1.04719755119659775 + sq3h * x);
  const double n8 = exp(-x * s8) * cos(0.78539816339744831 + x * c8) + exp(-x * c8) * cos(2.35619449019234492 + x * s8);
  *h6 = n6 / x;
  *h8 = n8 / x;
}

// Per-pair tables of G itself over k in [0, kmax] (the register k-space
// kernel without parameter gradients): kGtabN intervals of dk = kmax/kGtabN,
// on each the degree-7 polynomial interpolating G_XY(k) = p6 h6(kR) +
// p8 h8(kR) at its eight Chebyshev nodes, as monomial coefficients in the
// local coordinate w = k/dk - i in [0, 1]: {a0..a3}, {a4..a7}.  One index and
// w per k point serve every pair, and each pair costs one Horner pass.  An
// interval spans dx = kmax R / kGtabN in x = kR; the degree-7 interpolation
// error stays below 1e-8 of G(0) up to dx = 2, i.e. kmax R <= 112 (a 0.4 A
// mesh spacing gives kmax ~ 7 / Bohr, and R < 16 Bohr).
constexpr int kGtabN = 56;
constexpr int kGtabDeg = 7;

// Monomial coefficients a[0..n) in w in [0, 1] of the polynomial interpolating
// f at the n Chebyshev nodes u_j = cos((2j+1) pi / 2n), w_j = (u_j + 1) / 2.
__host__ __device__ inline void cheb_fit_w(const double* f, int n, double* a) {
  double c[kGtabDeg + 1], m[kGtabDeg + 1], t0[kGtabDeg + 1], t1[kGtabDeg + 1], t2[kGtabDeg + 1];
  // c_k = (2 - [k = 0]) / n sum_j f_j T_k(u_j), T_k(u_j) by the recurrence
  for (int k = 0; k < n; ++k) c[k] = 0.0;
  for (int j = 0; j < n; ++j) {
    const double u = cos((2 * j + 1) * kPi / (2.0 * n));
    double tm = 1.0, tk = u;
    c[0] += f[j];
    if (n > 1) c[1] += f[j] * u;
    for (int k = 2; k < n; ++k) {
      const double tn = 2.0 * u * tk - tm;
      c[k] += f[j] * tn;
      tm = tk;
      tk = tn;
    }
  }
  for (int k = 0; k < n; ++k) c[k] *= (k ? 2.0 : 1.0) / n;
  // sum_k c_k T_k(u) in monomials of u (T_{k+1} = 2u T_k - T_{k-1})
  for (int i = 0; i < n; ++i) { m[i] = 0.0; t0[i] = 0.0; t1[i] = 0.0; }
  t0[0] = 1.0;
  if (n > 1) t1[1] = 1.0;
  m[0] = c[0];
  for (int i = 0; n > 1 && i < n; ++i) m[i] += c[1] * t1[i];
  for (int k = 2; k < n; ++k) {
    for (int i = 0; i < n; ++i) t2[i] = (i ? 2.0 * t1[i - 1] : 0.0) - t0[i];
    for (int i = 0; i < n; ++i) { m[i] += c[k] * t2[i]; t0[i] = t1[i]; t1[i] = t2[i]; }
  }
  // u = 2w - 1: a_j = 2^j sum_{k >= j} m_k C(k, j) (-1)^(k-j)
  for (int j = 0; j < n; ++j) {
    double s = 0.0, binom = 1.0;  // C(k, j), k = j, j+1, ...
    for (int k = j; k < n; ++k) {
      s += m[k] * binom * ((k - j) & 1 ? -1.0 : 1.0);
      binom = binom * (k + 1) / (k + 1 - j);
    }
    a[j] = s * (double)(1 << j);
  }
}

// x of node j (0..5) of h-table interval i: the Chebyshev nodes of the interval.
__host__ __device__ inline double htab_node(int i, int j) {
  return kHtabDx * (i + 0.5 * (cos((2 * j + 1) * kPi / 12.0) + 1.0));
}

// The three float4 table entries of interval i (0 <= i <= kHtabN) from h6, h8
// at its nodes (unused for the zero interval i = kHtabN).
__host__ __device__ inline void htab_fit(int i, const double* f6, const double* f8, float4* out) {
  if (i >= kHtabN) {
    out[0] = out[1] = out[2] = make_float4(0.f, 0.f, 0.f, 0.f);
    return;
  }
  double a6[kGtabDeg + 1], a8[kGtabDeg + 1];
  cheb_fit_w(f6, 6, a6);
  cheb_fit_w(f8, 6, a8);
  out[0] = make_float4((float)a6[0], (float)a6[1], (float)a6[2], (float)a6[3]);
  out[1] = make_float4((float)a6[4], (float)a6[5], (float)a8[0], (float)a8[1]);
  out[2] = make_float4((float)a8[2], (float)a8[3], (float)a8[4], (float)a8[5]);
}

__host__ __device__ inline void htab_interval(int i, float4* out) {
  double f6[6] = {0, 0, 0, 0, 0, 0}, f8[6] = {0, 0, 0, 0, 0, 0};
  if (i < kHtabN)
    for (int j = 0; j < 6; ++j) h_exact(htab_node(i, j), &f6[j], &f8[j]);
  htab_fit(i, f6, f8, out);
}

// h6, h8 and their derivatives with respect to the local coordinate w from
// the table at x = fx dx (fx >= 0), by Horner's scheme carrying the derivative.
__host__ __device__ inline void green_h_tab(const float4* __restrict__ tab, float fx, float* h6, float* d6, float* h8,
                                            float* d8) {
  const int i = (int)fminf(fx, (float)kHtabN);
#if defined(__HIP_DEVICE_COMPILE__)
  const float w = __builtin_amdgcn_fractf(fx);
#else
  const float w = fx - floorf(fx);
#endif
  const float4* e = (const float4*)((const char*)tab + (unsigned)i * (unsigned)(3 * sizeof(float4)));
  const float4 A = e[0], B = e[1], C = e[2];
  float p = B.y * w + B.x, d = B.y;
  d = d * w + p; p = p * w + A.w;
  d = d * w + p; p = p * w + A.z;
  d = d * w + p; p = p * w + A.y;
  d = d * w + p; p = p * w + A.x;
  *h6 = p; *d6 = d;
  p = C.w * w + C.z; d = C.w;
  d = d * w + p; p = p * w + C.y;
  d = d * w + p; p = p * w + C.x;
  d = d * w + p; p = p * w + B.w;
  d = d * w + p; p = p * w + B.z;
  *h8 = p; *d8 = d;
}

// Per species pair constants of the k-space kernel.
struct KsPair {
  float Rs;        // R0_XY / dx (R0 in Bohr): x / dx = k Rs
  float p6, p8;    // G = p6 h6(kR) + p8 h8(kR): s6 2pi^2/(3R^3), 3 s8 QQ (-pi^2/R^5)
  float q6, q8;    // dG/dk = q6 h6_w + q8 h8_w (w-derivatives of the table): p R / dx
  float pf6, pf8;  // dG/ds6 = pf6 h6, dG/ds8 = pf8 h8
  float dR;        // dR0/da1 = sqrt(3 QQ)
};

__host__ __device__ inline KsPair ks_pair(float qq, float s6, float s8, float a1, float a2) {
  const float sq = sqrtf(3.0f * qq);
  const float R = a1 * sq + a2;
  KsPair k;
  k.Rs = R * kHtabInvDx;
  k.pf6 = float(2.0 * kPi * kPi / 3.0) / (R * R * R);
  k.pf8 = 3.0f * qq * float(-kPi * kPi) / (R * R * R * R * R);
  k.p6 = s6 * k.pf6;
  k.p8 = s8 * k.pf8;
  k.q6 = k.p6 * R * kHtabInvDx;
  k.q8 = k.p8 * R * kHtabInvDx;
  k.dR = sq;
  return k;
}

// k of node j (0..7) of interval i of a per-pair G table of interval width dk.
__host__ __device__ inline double gtab_node(double dk, int i, int j) {
  return dk * (i + 0.5 * (cos((2 * j + 1) * kPi / (2.0 * (kGtabDeg + 1))) + 1.0));
}

// The two float4 entries of one interval of a per-pair G table from G at its nodes.
__host__ __device__ inline void gtab_fit(const double* f, float4* out) {
  double a[kGtabDeg + 1];
  cheb_fit_w(f, kGtabDeg + 1, a);
  out[0] = make_float4((float)a[0], (float)a[1], (float)a[2], (float)a[3]);
  out[1] = make_float4((float)a[4], (float)a[5], (float)a[6], (float)a[7]);
}

// G_XY at k: p6 h6(kR) + p8 h8(kR).
__host__ __device__ inline double g_exact(double k, double R, double p6, double p8) {
  double h6, h8;
  h_exact(k * R, &h6, &h8);
  return p6 * h6 + p8 * h8;
}

__host__ __device__ inline void gtab_interval(double dk, int i, double R, double p6, double p8, float4* out) {
  double f[kGtabDeg + 1];
  for (int j = 0; j <= kGtabDeg; ++j) f[j] = g_exact(gtab_node(dk, i, j), R, p6, p8);
  gtab_fit(f, out);
}

// G and dG/dw at w from one interval of a per-pair G table (Horner's scheme
// carrying the derivative).
__host__ __device__ __forceinline__ void gtab_eval(float4 A, float4 B, float w, float* g, float* dgdw) {
  float p = B.w * w + B.z, d = B.w;
  d = d * w + p; p = p * w + B.y;
  d = d * w + p; p = p * w + B.x;
  d = d * w + p; p = p * w + A.w;
  d = d * w + p; p = p * w + A.z;
  d = d * w + p; p = p * w + A.y;
  d = d * w + p; p = p * w + A.x;
  *g = p;
  *dgdw = d;
}

// G and dG/dk at |k| = k from the table; h6, h8 and their w-derivatives
// (dh/dx = dh/dw / dx) for the parameter derivatives.
__host__ __device__ inline void green_eval_tab(const KsPair& pk, const float4* __restrict__ tab, float k, float* g,
                                               float* dg, float* h6, float* d6, float* h8, float* d8) {
  green_h_tab(tab, k * pk.Rs, h6, d6, h8, d8);
  *g = pk.p6 * *h6 + pk.p8 * *h8;
  *dg = pk.q6 * *d6 + pk.q8 * *d8;
}

// ----------------------------------------------------------------- misc ---
__device__ inline double warp_sum(double v) {
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  return v;
}
__device__ inline float warp_sum(float v) {
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  return v;
}

// Sum `n` doubles per thread over the block; thread 0 writes them to out[0..n).
// Requires blockDim.x a multiple of 64 and <= 1024; scratch holds 16*n doubles.
template <int NV>
__device__ inline void block_sum_store(double (&v)[NV], double* scratch, double* out) {
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6, nw = blockDim.x >> 6;
#pragma unroll
  for (int i = 0; i < NV; ++i) {
    double s = warp_sum(v[i]);
    if (lane == 0) scratch[wid * NV + i] = s;
  }
  __syncthreads();
  if (threadIdx.x < NV) {
    double s = 0.0;
    for (int w = 0; w < nw; ++w) s += scratch[w * NV + threadIdx.x];
    out[threadIdx.x] = s;
  }
  __syncthreads();
}

// Position of atom i in Bohr as double, from an f32 or f64 (N,3) array in Angstrom.
__device__ inline void load_pos(const void* p, int dtype, int64_t i, double* r) {
  if (dtype == FD3_DTYPE_F64) {
    const double* q = (const double*)p + 3 * i;
    r[0] = q[0] * kA2B; r[1] = q[1] * kA2B; r[2] = q[2] * kA2B;
  } else {
    const float* q = (const float*)p + 3 * i;
    r[0] = (double)q[0] * kA2B; r[1] = (double)q[1] * kA2B; r[2] = (double)q[2] * kA2B;
  }
}

inline int64_t cdiv(int64_t a, int64_t b) { return (a + b - 1) / b; }
inline size_t align_up(size_t v, size_t a = 256) { return (v + a - 1) / a * a; }

// Simple bump allocator over a caller-provided workspace.
// p + k, or null for a null p (a workspace query carves from a null base
// and must not offset it).
template <typename T>
inline T* offset_or_null(T* p, size_t k) {
  return p ? p + k : nullptr;
}

struct Arena {
  char* base;
  size_t cap, off;
  explicit Arena(void* b = nullptr, size_t c = 0) : base((char*)b), cap(c), off(0) {}
  template <typename T>
  T* take(size_t n) {
    size_t o = align_up(off);
    off = o + n * sizeof(T);
    return base ? (T*)(base + o) : nullptr;
  }
  bool ok() const { return base == nullptr || off <= cap; }
};

}  // namespace fd3
