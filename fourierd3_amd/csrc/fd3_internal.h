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
// Box geometry in Bohr, passed to kernels by value.
struct Geom {
  double cell[9];   // rows = lattice vectors
  double inv[9];    // cell^-1
  double recip[9];  // rows b_i of 2 pi cell^-T
  double volume;
  int nx, ny, nz;
  int order, interp;
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
__host__ __device__ inline void interp_weights(int p, int interp, T x, T* w, T* dw) {
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

// Hardware exp / sin / cos on the GPU (v_exp_f32, v_sin_f32, v_cos_f32), the
// libm versions when the same code is compiled for the host checks.
__host__ __device__ inline float fd3_exp(float x) {
#if defined(__HIP_DEVICE_COMPILE__)
  return __expf(x);
#else
  return expf(x);
#endif
}
__host__ __device__ inline void fd3_sincos(float x, float* s, float* c) {
#if defined(__HIP_DEVICE_COMPILE__)
  *s = __sinf(x);
  *c = __cosf(x);
#else
  sincosf(x, s, c);
#endif
}

// --------------------------------------------------------- Green's function ---
// h6(x) = num6(x)/x and h8(x) = num8(x)/x of the analytic Fourier transforms
// (src/fourierd3/pair_pot.py:152-175):
//   f6^(k) = 2 pi^2 / (3 R^3) h6(kR),  num6 = e^-x - 2 e^-x/2 cos(pi/3 + sqrt3 x/2)
//   f8^(k) =  -pi^2 / R^5     h8(kR),  num8 = e^-x sin(pi/8) cos(pi/4 + x cos(pi/8))
//                                           + e^-x cos(pi/8) cos(3pi/4 + x sin(pi/8))
// num6 and num8 are entire with num(0) = 0; below x = 1 their Maclaurin series
// is summed (no cancellation), above it the closed form.  hp = dh/dx.
__host__ __device__ inline void green_h(float x, float* h6, float* h6p, float* h8, float* h8p) {
  if (x < 1.0f) {
    // coefficients of x^k, k = 0..16 (exact zeros dropped)
    const float c6[17] = {1.0f, 0.0f, -3.33333333333333315e-01f, 1.25e-01f, -1.66666666666666664e-02f, 0.0f,
                          1.98412698412698413e-04f, 0.0f, -5.51146384479717850e-06f, 8.26719576719576754e-07f,
                          -5.01042167708834404e-08f, 0.0f, 1.60590438368216133e-10f, 0.0f,
                          -1.52943274636396328e-12f, 1.43384319971621551e-13f, -5.62291450869104120e-15f};
    const float c8[17] = {-5.41196100146197012e-01f, 0.0f, 9.01993500243661595e-02f, 0.0f,
                          -1.08880247073031376e-02f, 2.77777777777777788e-03f, -2.59238683507217581e-04f, 0.0f,
                          1.49139136945049866e-06f, 0.0f, -1.35581033586408978e-08f, 0.0f,
                          2.09821519285173499e-10f, -2.29414911954594490e-11f, 9.99150091834159524e-13f, 0.0f,
                          -1.52154970177953047e-15f};
    float a6 = c6[16], a8 = c8[16], d6 = 16.0f * c6[16], d8 = 16.0f * c8[16];
    for (int k = 15; k >= 1; --k) {
      a6 = a6 * x + c6[k];
      a8 = a8 * x + c8[k];
      d6 = d6 * x + float(k) * c6[k];
      d8 = d8 * x + float(k) * c8[k];
    }
    *h6 = a6 * x + c6[0];
    *h8 = a8 * x + c8[0];
    *h6p = d6;
    *h8p = d8;
  } else {
    // x >= 1: every term carries a factor exp(-x sin(pi/8)) or smaller, so the
    // hardware exp/sin/cos (relative/phase error ~1e-6 here) are accurate enough.
    const float s8 = 0.38268343236508977f, c8 = 0.92387953251128674f, sq3h = 0.86602540378443865f;
    const float e2 = fd3_exp(-0.5f * x), e1 = e2 * e2;
    float sp, cp;
    fd3_sincos(1.04719755119659775f + sq3h * x, &sp, &cp);
    const float n6 = e1 - 2.0f * e2 * cp;
    const float n6p = -e1 + e2 * cp + 2.0f * sq3h * e2 * sp;
    const float ea = fd3_exp(-x * s8), eb = fd3_exp(-x * c8);
    float sa, ca, sb, cb;
    fd3_sincos(0.78539816339744831f + x * c8, &sa, &ca);
    fd3_sincos(2.35619449019234492f + x * s8, &sb, &cb);
    const float n8 = ea * ca + eb * cb;
    const float n8p = -s8 * ea * ca - c8 * ea * sa - c8 * eb * cb - s8 * eb * sb;
    const float ix = 1.0f / x;
    *h6 = n6 * ix;
    *h8 = n8 * ix;
    *h6p = (n6p - *h6) * ix;
    *h8p = (n8p - *h8) * ix;
  }
}

// Per-pair constants for G_XY(k) = s6 f6^(k; R) + 3 s8 QQ f8^(k; R).
struct PairK {
  float R;     // R0_XY, Bohr
  float p6;    // s6 * 2 pi^2 / (3 R^3)
  float p8;    // 3 s8 QQ * (-pi^2 / R^5)
};

// G and dG/dk for |k| = k.
__host__ __device__ inline void green_eval(const PairK& pk, float k, float* g, float* dg) {
  float h6, h6p, h8, h8p;
  green_h(k * pk.R, &h6, &h6p, &h8, &h8p);
  *g = pk.p6 * h6 + pk.p8 * h8;
  *dg = (pk.p6 * h6p + pk.p8 * h8p) * pk.R;
}

// h6 and h8 do not depend on the species pair, so the k-space kernel reads
// them from a piecewise polynomial table instead of evaluating four
// exponentials and three sine/cosine pairs per pair and k point.  Interval i
// covers x in [i dx, (i+1) dx); with u = 2 (x/dx - i) - 1 in [-1, 1] each
// function is the degree-5 polynomial interpolating it at the six Chebyshev
// nodes of the interval (interpolation error < 1e-8 of h(0) at dx = 1/2).
// Entry 3i .. 3i+2 holds {a6_0..a6_3}, {a6_4, a6_5, a8_0, a8_1}, {a8_2..a8_5},
// monomial coefficients in u.  Interval kHtabN is all zeros and serves every
// x >= kHtabXmax = 64, where |h6| < 1e-15 and |h8| < 4e-13 (h6(0) = 1,
// h8(0) = -0.54): 1e-6 of float32 resolution, invisible in any sum.
constexpr int kHtabN = 128;
constexpr float kHtabDx = 0.5f;
constexpr float kHtabXmax = kHtabN * kHtabDx;

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

// The three float4 table entries of interval i (0 <= i <= kHtabN).
__host__ __device__ inline void htab_interval(int i, float4* out) {
  if (i >= kHtabN) {
    out[0] = out[1] = out[2] = make_float4(0.f, 0.f, 0.f, 0.f);
    return;
  }
  double f6[6], f8[6], u[6];
  for (int j = 0; j < 6; ++j) {
    u[j] = cos((2 * j + 1) * kPi / 12.0);
    h_exact(kHtabDx * (i + 0.5 * (u[j] + 1.0)), &f6[j], &f8[j]);
  }
  double a6[6], a8[6];
  for (int f = 0; f < 2; ++f) {
    const double* v = f ? f8 : f6;
    double c[6];  // Chebyshev coefficients
    for (int k = 0; k < 6; ++k) {
      double s = 0.0;
      for (int j = 0; j < 6; ++j) s += v[j] * cos(k * (2 * j + 1) * kPi / 12.0);
      c[k] = s * (k ? 2.0 / 6.0 : 1.0 / 6.0);
    }
    // T0..T5 in monomials
    double* a = f ? a8 : a6;
    a[0] = c[0] - c[2] + c[4];
    a[1] = c[1] - 3.0 * c[3] + 5.0 * c[5];
    a[2] = 2.0 * c[2] - 8.0 * c[4];
    a[3] = 4.0 * c[3] - 20.0 * c[5];
    a[4] = 8.0 * c[4];
    a[5] = 16.0 * c[5];
  }
  out[0] = make_float4((float)a6[0], (float)a6[1], (float)a6[2], (float)a6[3]);
  out[1] = make_float4((float)a6[4], (float)a6[5], (float)a8[0], (float)a8[1]);
  out[2] = make_float4((float)a8[2], (float)a8[3], (float)a8[4], (float)a8[5]);
}

// h6, h8 and their x-derivatives from the table (x >= 0).
__host__ __device__ inline void green_h_tab(const float4* __restrict__ tab, float x, float* h6, float* h6p, float* h8,
                                            float* h8p) {
  const float fx = x * (1.0f / kHtabDx);
  const int i = min((int)fx, kHtabN);
  const float u = 2.0f * (fx - (float)i) - 1.0f;
  const float4 A = tab[3 * i], B = tab[3 * i + 1], C = tab[3 * i + 2];
  *h6 = A.x + u * (A.y + u * (A.z + u * (A.w + u * (B.x + u * B.y))));
  *h8 = B.z + u * (B.w + u * (C.x + u * (C.y + u * (C.z + u * C.w))));
  const float d6 = A.y + u * (2.0f * A.z + u * (3.0f * A.w + u * (4.0f * B.x + u * (5.0f * B.y))));
  const float d8 = B.w + u * (2.0f * C.x + u * (3.0f * C.y + u * (4.0f * C.z + u * (5.0f * C.w))));
  *h6p = d6 * (2.0f / kHtabDx);
  *h8p = d8 * (2.0f / kHtabDx);
}

// G and dG/dk for |k| = k from the table.
__host__ __device__ inline void green_eval_tab(const PairK& pk, const float4* __restrict__ tab, float k, float* g,
                                               float* dg, float* h6o = nullptr, float* h6po = nullptr,
                                               float* h8o = nullptr, float* h8po = nullptr) {
  float h6, h6p, h8, h8p;
  green_h_tab(tab, k * pk.R, &h6, &h6p, &h8, &h8p);
  *g = pk.p6 * h6 + pk.p8 * h8;
  *dg = (pk.p6 * h6p + pk.p8 * h8p) * pk.R;
  if (h6o) { *h6o = h6; *h6po = h6p; *h8o = h8; *h8po = h8p; }
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
