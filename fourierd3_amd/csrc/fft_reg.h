// Small complex FFTs held in registers (natural order in and out) and
// multiplication by compile-time twiddle factors.  Used by the two-pass line
// FFTs of the mesh passes: one such FFT per thread, then one exchange of
// values through LDS, then a second one.
#pragma once

#include <type_traits>
#include <utility>

#include "fd3_internal.h"

namespace fd3 {

// f(std::integral_constant<int, i>{}) for i = 0..N-1, unrolled at compile time.
template <typename F, int... I>
__device__ __forceinline__ void static_for_impl(F&& f, std::integer_sequence<int, I...>) {
  (f(std::integral_constant<int, I>{}), ...);
}
template <int N, typename F>
__device__ __forceinline__ void static_for(F&& f) {
  static_for_impl(f, std::make_integer_sequence<int, N>{});
}

// cos and sin of 2 pi m / 64, m = 0..63.
__device__ constexpr double kCs64[64][2] = {
    {1.0, 0.0},
    {0.9951847266721969, 0.0980171403295606},
    {0.9807852804032304, 0.19509032201612825},
    {0.9569403357322088, 0.29028467725446233},
    {0.9238795325112867, 0.3826834323650898},
    {0.881921264348355, 0.47139673682599764},
    {0.8314696123025452, 0.5555702330196022},
    {0.773010453362737, 0.6343932841636455},
    {0.7071067811865476, 0.7071067811865475},
    {0.6343932841636455, 0.773010453362737},
    {0.5555702330196023, 0.8314696123025452},
    {0.4713967368259978, 0.8819212643483549},
    {0.38268343236508984, 0.9238795325112867},
    {0.29028467725446233, 0.9569403357322089},
    {0.19509032201612833, 0.9807852804032304},
    {0.09801714032956077, 0.9951847266721968},
    {6.123233995736766e-17, 1.0},
    {-0.09801714032956065, 0.9951847266721969},
    {-0.1950903220161282, 0.9807852804032304},
    {-0.29028467725446216, 0.9569403357322089},
    {-0.3826834323650897, 0.9238795325112867},
    {-0.4713967368259977, 0.881921264348355},
    {-0.555570233019602, 0.8314696123025455},
    {-0.6343932841636454, 0.7730104533627371},
    {-0.7071067811865475, 0.7071067811865476},
    {-0.773010453362737, 0.6343932841636455},
    {-0.8314696123025453, 0.5555702330196022},
    {-0.8819212643483549, 0.47139673682599786},
    {-0.9238795325112867, 0.3826834323650899},
    {-0.9569403357322088, 0.2902846772544624},
    {-0.9807852804032304, 0.1950903220161286},
    {-0.9951847266721968, 0.09801714032956083},
    {-1.0, 1.2246467991473532e-16},
    {-0.9951847266721969, -0.09801714032956059},
    {-0.9807852804032304, -0.19509032201612836},
    {-0.9569403357322089, -0.2902846772544621},
    {-0.9238795325112868, -0.38268343236508967},
    {-0.881921264348355, -0.47139673682599764},
    {-0.8314696123025455, -0.555570233019602},
    {-0.7730104533627371, -0.6343932841636453},
    {-0.7071067811865477, -0.7071067811865475},
    {-0.6343932841636459, -0.7730104533627367},
    {-0.5555702330196022, -0.8314696123025452},
    {-0.47139673682599786, -0.8819212643483549},
    {-0.38268343236509034, -0.9238795325112865},
    {-0.29028467725446244, -0.9569403357322088},
    {-0.19509032201612866, -0.9807852804032303},
    {-0.09801714032956045, -0.9951847266721969},
    {-1.8369701987210297e-16, -1.0},
    {0.09801714032956009, -0.9951847266721969},
    {0.1950903220161283, -0.9807852804032304},
    {0.29028467725446205, -0.9569403357322089},
    {0.38268343236509, -0.9238795325112866},
    {0.4713967368259976, -0.881921264348355},
    {0.5555702330196018, -0.8314696123025455},
    {0.6343932841636456, -0.7730104533627369},
    {0.7071067811865474, -0.7071067811865477},
    {0.7730104533627367, -0.6343932841636459},
    {0.8314696123025452, -0.5555702330196022},
    {0.8819212643483548, -0.4713967368259979},
    {0.9238795325112865, -0.3826834323650904},
    {0.9569403357322088, -0.2902846772544625},
    {0.9807852804032303, -0.19509032201612872},
    {0.9951847266721969, -0.0980171403295605}
};

// a * exp(DIR 2 pi i M / N) for compile-time M, N (N divides 64): trivial
// factors (1, -1, +-i) cost no multiplications and the eighth roots two.
template <int N, int M, int DIR>
__device__ __forceinline__ float2 tw_mul(float2 a) {
  static_assert(64 % N == 0, "tw_mul: N divides 64");
  constexpr int m = ((M % N) + N) % N * (64 / N);  // angle in units of 2 pi / 64
  if constexpr (m == 0) {
    return a;
  } else if constexpr (m == 32) {
    return make_float2(-a.x, -a.y);
  } else if constexpr (m == 16 || m == 48) {
    // exp(+-i pi/2) = +-i; the sign of the angle is DIR (m = 16) or -DIR (m = 48)
    constexpr int s = (m == 16 ? 1 : -1) * DIR;
    return s > 0 ? make_float2(-a.y, a.x) : make_float2(a.y, -a.x);
  } else if constexpr (m % 8 == 0) {
    // odd multiples of pi/4: (+-1 +-i) / sqrt 2
    constexpr float h = 0.70710678118654752f;
    constexpr float c = (float)kCs64[m][0] > 0.f ? h : -h;
    constexpr float sn = ((float)kCs64[m][1] > 0.f ? h : -h) * DIR;
    return make_float2(c * a.x - sn * a.y, c * a.y + sn * a.x);
  } else {
    constexpr float c = (float)kCs64[m][0], sn = (float)kCs64[m][1] * DIR;
    return make_float2(c * a.x - sn * a.y, c * a.y + sn * a.x);
  }
}

template <int DIR>
__device__ __forceinline__ void fft4r(float2& a0, float2& a1, float2& a2, float2& a3) {
  const float2 b0 = make_float2(a0.x + a2.x, a0.y + a2.y), b1 = make_float2(a0.x - a2.x, a0.y - a2.y);
  const float2 b2 = make_float2(a1.x + a3.x, a1.y + a3.y), d = make_float2(a1.x - a3.x, a1.y - a3.y);
  // forward: b3 = -i d ; inverse: b3 = +i d
  const float2 b3 = DIR < 0 ? make_float2(d.y, -d.x) : make_float2(-d.y, d.x);
  a0 = make_float2(b0.x + b2.x, b0.y + b2.y);
  a2 = make_float2(b0.x - b2.x, b0.y - b2.y);
  a1 = make_float2(b1.x + b3.x, b1.y + b3.y);
  a3 = make_float2(b1.x - b3.x, b1.y - b3.y);
}

// X[k] = sum_n v[n] exp(DIR 2 pi i n k / 8): n = 2 n1 + n2, k = k1 + 4 k2.
template <int DIR>
__device__ __forceinline__ void fft8r(float2 (&v)[8]) {
  float2 a[2][4];
#pragma unroll
  for (int n2 = 0; n2 < 2; ++n2) {
#pragma unroll
    for (int n1 = 0; n1 < 4; ++n1) a[n2][n1] = v[2 * n1 + n2];
    fft4r<DIR>(a[n2][0], a[n2][1], a[n2][2], a[n2][3]);
  }
  a[1][1] = tw_mul<8, 1, DIR>(a[1][1]);
  a[1][2] = tw_mul<8, 2, DIR>(a[1][2]);
  a[1][3] = tw_mul<8, 3, DIR>(a[1][3]);
#pragma unroll
  for (int k1 = 0; k1 < 4; ++k1) {
    v[k1] = make_float2(a[0][k1].x + a[1][k1].x, a[0][k1].y + a[1][k1].y);
    v[k1 + 4] = make_float2(a[0][k1].x - a[1][k1].x, a[0][k1].y - a[1][k1].y);
  }
}

// X[k] = sum_n v[n] exp(DIR 2 pi i n k / 16): n = 4 n1 + n2, k = k1 + 4 k2.
template <int DIR>
__device__ __forceinline__ void fft16r(float2 (&v)[16]) {
  float2 a[4][4];
#pragma unroll
  for (int n2 = 0; n2 < 4; ++n2) {
#pragma unroll
    for (int n1 = 0; n1 < 4; ++n1) a[n2][n1] = v[4 * n1 + n2];
    fft4r<DIR>(a[n2][0], a[n2][1], a[n2][2], a[n2][3]);
  }
  // a[n2][k1] *= W16^(n2 k1)
  a[1][1] = tw_mul<16, 1, DIR>(a[1][1]);
  a[1][2] = tw_mul<16, 2, DIR>(a[1][2]);
  a[1][3] = tw_mul<16, 3, DIR>(a[1][3]);
  a[2][1] = tw_mul<16, 2, DIR>(a[2][1]);
  a[2][2] = tw_mul<16, 4, DIR>(a[2][2]);
  a[2][3] = tw_mul<16, 6, DIR>(a[2][3]);
  a[3][1] = tw_mul<16, 3, DIR>(a[3][1]);
  a[3][2] = tw_mul<16, 6, DIR>(a[3][2]);
  a[3][3] = tw_mul<16, 9, DIR>(a[3][3]);
#pragma unroll
  for (int k1 = 0; k1 < 4; ++k1) {
    fft4r<DIR>(a[0][k1], a[1][k1], a[2][k1], a[3][k1]);
#pragma unroll
    for (int k2 = 0; k2 < 4; ++k2) v[k1 + 4 * k2] = a[k2][k1];
  }
}

// The second stage of a 256-point x FFT whose first stage the y pass has
// run (split x FFT, c2c_y2_body<.., XS> in fft.hip): a line holds Y[k1, n2]
// at 16 n2 + k1, and thread (l, k1) takes its 16 values Y[k1, 0..15] (16
// lanes over k1 read 128 contiguous bytes), does the 16-point FFT over n2
// and stores X[k1 + 16 k2] in natural order; the inverse mirrors it.  No
// twiddles, no LDS exchange.  Lines past the tile or of ranks past the last
// (live = false) read as zeros and are not written.

// this thread's 16 values of one line, starting at element e (= line base + k1)
__device__ __forceinline__ void ksx_load(float2 (&v)[16], const float2* __restrict__ g, bool live, long long e) {
  if (live) {
    const float2* src = g + e;
#pragma unroll
    for (int n2 = 0; n2 < 16; ++n2) v[n2] = src[16 * n2];
  } else {
#pragma unroll
    for (int n2 = 0; n2 < 16; ++n2) v[n2] = make_float2(0.f, 0.f);
  }
}

// Forward from the loaded values into the first nlines lines (stride ls)
// of S; ends with a barrier.
__device__ __forceinline__ void ksx_forward_regs(float2* S, int ls, const float2 (&v)[16], int nlines) {
  const int t = threadIdx.x, l = t >> 4, j = t & 15;
  if (l < nlines) {
    float2 z[16];
#pragma unroll
    for (int n2 = 0; n2 < 16; ++n2) z[n2] = v[n2];
    fft16r<-1>(z);  // over n2 -> k2: X[j + 16 k2]
    float2* L = S + l * ls;
#pragma unroll
    for (int k2 = 0; k2 < 16; ++k2) L[j + 16 * k2] = z[k2];
  }
  __syncthreads();
}

__device__ __forceinline__ void ksx_inverse(const float2* S, int ls, float2* __restrict__ g, bool live,
                                           long long e) {
  const int t = threadIdx.x, l = t >> 4, j = t & 15;
  const float2* L = S + l * ls;
  if (live) {
    float2 z[16];
#pragma unroll
    for (int k2 = 0; k2 < 16; ++k2) z[k2] = L[j + 16 * k2];  // k1 = j
    fft16r<1>(z);  // over k2 -> n2: Y[j, n2]
    float2* dst = g + e;
#pragma unroll
    for (int n2 = 0; n2 < 16; ++n2) dst[16 * n2] = z[n2];
  }
}

// G and dG/dw of the NP species pairs at |k| = kn from per-pair tables in LDS.
template <int NP>
__device__ __forceinline__ void ks_gtab_lds(const float4* tab, float gkinv, float kn, float (&gp)[NP],
                                            float (&dgp)[NP]) {
  const float fk = kn * gkinv;
  const int i = min((int)fk, kGtabN - 1);
  const float wk = fk - (float)i;
  const float4* e = tab + 2 * i;
  constexpr int PD = 2;  // pairs whose coefficients are in flight
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
    asm volatile("" : "+v"(gp[pi]), "+v"(dgp[pi])::"memory");
  });
}

}  // namespace fd3
