// In-LDS line FFT and complex helpers shared by the FFT pass kernels.
#pragma once

#include <type_traits>

#include "fd3_internal.h"

namespace fd3 {

// Copy a twiddle table of n entries into LDS (all threads; the caller's next
// __syncthreads() publishes it).
__device__ inline void stage_table(float2* dst, const float2* __restrict__ src, int n) {
  for (int i = threadIdx.x; i < n; i += blockDim.x) dst[i] = src[i];
}

// Tile copies between global memory and LDS.  A tile is nl lines of
// len = 2^llen elements; element (l, c) sits at g[gl(l) + c * gcs] in global
// memory and at lds[l * lls + c * lcs].  Element number e = k * blockDim.x + t
// (row-major over (l, c)) of thread t splits into a per-thread part and a part
// that is uniform over the block for each k, so the per-element address work
// is a few adds.  A plain copy loop waits for each global load before its LDS
// store, leaving one load per wave in flight; these issue U loads per thread
// before the first store (and U LDS reads before the first global store).
// Loads are unconditional (dead elements read g[0]) so that none sits in a
// branch the compiler has to wait inside.  gl(l) < 0 marks a dead line: it is
// read as zeros and not written.
struct StridedLines {  // line l at l * gls
  long long gls;
};
struct LineTable {  // line l at base[l], -1 for a dead line (zero-filled, not written)
  const long long* base;
};
struct TileShape {
  int nl, llen;  // lines, log2 line length
  int gcs;       // global column stride
  int lls, lcs;  // LDS line and column strides
};

// Tile copies run in blocks of 2^LOG threads (256 unless stated).
constexpr int kTileLog = 8;

// Element k * 2^LOG + t of a tile (row-major over (line, column)): thread t
// owns line lt + lu(k), column ct + cu(k), where lt, ct depend on t only and
// lu, cu on k only (uniform over the block).
template <int LOG>
struct TileIter {
  bool wide;  // a line spans several k
  int lt, ct, nk, llen;
  __device__ TileIter(const TileShape& t) : wide(t.llen > LOG), llen(t.llen) {
    const int tid = threadIdx.x;
    lt = wide ? 0 : tid >> t.llen;
    ct = wide ? tid : tid & ((1 << t.llen) - 1);
    nk = ((t.nl << t.llen) + (1 << LOG) - 1) >> LOG;
  }
  __device__ int lu(int k) const { return wide ? k >> (llen - LOG) : k << (LOG - llen); }
  __device__ int cu(int k) const { return wide ? (k & ((1 << (llen - LOG)) - 1)) << LOG : 0; }
};

template <int U, int LOG = kTileLog>
__device__ __forceinline__ void tile2_to_lds(float2* lds, const float2* __restrict__ g, TileShape t, StridedLines gl) {
  const TileIter<LOG> it(t);
  // per-thread base, kept opaque so the uniform part below stays scalar
  const float2* gt = g + ((long long)it.lt * gl.gls + (long long)it.ct * t.gcs);
  asm volatile("" : "+v"(gt));
  const int lt0 = it.lt * t.lls + it.ct * t.lcs;
  for (int k0 = 0; k0 < it.nk; k0 += U) {
    float2 v[U];
    bool in[U];
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const int k = k0 + u;
      const int lu = it.lu(k), cu = it.cu(k);
      in[u] = k < it.nk && it.lt + lu < t.nl;
      const long long uo = (long long)lu * gl.gls + (long long)cu * t.gcs;
      v[u] = in[u] ? gt[uo] : g[0];
    }
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const int k = k0 + u;
      if (in[u]) lds[lt0 + it.lu(k) * t.lls + it.cu(k) * t.lcs] = v[u];
    }
  }
}

template <int U, int LOG = kTileLog>
__device__ __forceinline__ void tile2_from_lds(float2* __restrict__ g, const float2* lds, TileShape t, StridedLines gl) {
  const TileIter<LOG> it(t);
  float2* gt = g + ((long long)it.lt * gl.gls + (long long)it.ct * t.gcs);
  asm volatile("" : "+v"(gt));
  const int lt0 = it.lt * t.lls + it.ct * t.lcs;
  for (int k0 = 0; k0 < it.nk; k0 += U) {
    float2 v[U];
    bool in[U];
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const int k = k0 + u;
      in[u] = k < it.nk && it.lt + it.lu(k) < t.nl;
      v[u] = lds[in[u] ? lt0 + it.lu(k) * t.lls + it.cu(k) * t.lcs : 0];
    }
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const int k = k0 + u;
      if (in[u]) gt[(long long)it.lu(k) * gl.gls + (long long)it.cu(k) * t.gcs] = v[u];
    }
  }
}

template <int U, int LOG = kTileLog>
__device__ __forceinline__ void tile2_to_lds(float2* lds, const float2* __restrict__ g, TileShape t, LineTable gl) {
  const TileIter<LOG> it(t);
  const int lt0 = it.lt * t.lls + it.ct * t.lcs;
  for (int k0 = 0; k0 < it.nk; k0 += U) {
    float2 v[U];
    bool in[U], live[U];
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const int k = k0 + u;
      const int l = it.lt + it.lu(k);
      in[u] = k < it.nk && l < t.nl;
      const long long b = in[u] ? gl.base[l] : -1;
      live[u] = b >= 0;
      v[u] = g[live[u] ? b + (long long)(it.ct + it.cu(k)) * t.gcs : 0];
    }
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const int k = k0 + u;
      if (in[u]) lds[lt0 + it.lu(k) * t.lls + it.cu(k) * t.lcs] = live[u] ? v[u] : make_float2(0.f, 0.f);
    }
  }
}

template <int U, int LOG = kTileLog>
__device__ __forceinline__ void tile2_from_lds(float2* __restrict__ g, const float2* lds, TileShape t, LineTable gl) {
  const TileIter<LOG> it(t);
  const int lt0 = it.lt * t.lls + it.ct * t.lcs;
  for (int k0 = 0; k0 < it.nk; k0 += U) {
    float2 v[U];
    long long go[U];
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const int k = k0 + u;
      const int l = it.lt + it.lu(k);
      const bool in = k < it.nk && l < t.nl;
      const long long b = in ? gl.base[l] : -1;
      go[u] = b >= 0 ? b + (long long)(it.ct + it.cu(k)) * t.gcs : -1;
      v[u] = lds[in ? lt0 + it.lu(k) * t.lls + it.cu(k) * t.lcs : 0];
    }
#pragma unroll
    for (int u = 0; u < U; ++u)
      if (go[u] >= 0) g[go[u]] = v[u];
  }
}

constexpr int kCopyU = 16;  // loads in flight per thread (16 x 8 B)

__device__ __forceinline__ float2 cmul(float2 a, float2 b) {
  return make_float2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
__device__ __forceinline__ float2 cmulc(float2 a, float2 b) {  // a * conj(b)
  return make_float2(a.x * b.x + a.y * b.y, a.y * b.x - a.x * b.y);
}
__device__ __forceinline__ float2 cadd(float2 a, float2 b) { return make_float2(a.x + b.x, a.y + b.y); }
__device__ __forceinline__ float2 csub(float2 a, float2 b) { return make_float2(a.x - b.x, a.y - b.y); }
__device__ __forceinline__ float2 conjf2(float2 a) { return make_float2(a.x, -a.y); }

// Where lds_fft takes its input and puts its output: the LDS lines themselves
// (LdsLines), or lines in global memory through a line table (GlobalLines:
// line l at g + base[l]; base[l] < 0 is a dead line, read as zeros and not
// written).  A global input is read straight into the first stage's
// butterflies and a global output written straight from the last stage's, so
// the lines never pass through LDS on their way in or out.
struct LineIO {
  float2* g;
  const long long* base;
};
struct LdsLines : LineIO {  // members unused
  __device__ LdsLines() : LineIO{nullptr, nullptr} {}
};
struct GlobalLines : LineIO {
  __device__ GlobalLines(float2* g_, const long long* base_) : LineIO{g_, base_} {}
};

// Padded line layout: element e of a line sits at lpad(e) = e + e/32, lines
// at stride padded_stride(n).  The radix-4 stages' scattered stores (element
// 4j + m, 16 (j/4) + j%4 + 4m, ...) then spread over the 64 LDS banks instead
// of piling 8 lanes on a bank.
__host__ __device__ constexpr int lpad(int e) { return e + (e >> 5); }
__host__ __device__ constexpr int padded_stride(int n) { return n + (n >> 5); }

// Whether lds_fft has an unrolled variant for n, nlines, MAXB and a block of
// nt threads (a full tile of n = 128 or 256).  Callers that lay their lines
// out padded (stride padded_stride(n), element e at lpad(e)) do so exactly
// when this holds.
__host__ __device__ constexpr bool fft_fixed_tile(int n, int nlines, int maxb, int nt) {
  return (n == 256 && nt >= 64 && nlines == maxb * (nt / 64)) || (n == 128 && nt >= 32 && nlines == maxb * (nt / 32));
}

// lds_fft for a compile-time length n = 2^LOGN, block size NT and a full
// tile: nlines = MAXB * NT / (n/4) lines at stride n + 1 (PAD = false) or in
// the padded layout (PAD = true).  The same Stockham stages as the general
// version below, unrolled, so that every LDS offset (line c * stride,
// element m * n/4 or m * Ns) is an immediate and no stage tests its span at
// run time.
template <int DIR, int MAXB, int LOGN, int NT, bool PAD, class IN, class OUT>
__device__ __forceinline__ void lds_fft_fixed(float2* buf, const float2* __restrict__ tw, int tws, IN in, OUT out) {
  constexpr bool kGin = std::is_same<IN, GlobalLines>::value;
  constexpr bool kGout = std::is_same<OUT, GlobalLines>::value;
  constexpr int N = 1 << LOGN, NQ = N >> 2, LNQ = LOGN - 2, LS = PAD ? padded_stride(N) : N + 1;
  static_assert(!PAD || NQ >= 32, "padded FFT: n >= 128");
  constexpr int LSTEP = NT / NQ, DL = LSTEP * LS, S4 = LOGN >> 1;
  static_assert(LOGN >= 2 && NT % NQ == 0, "fixed FFT: n/4 must divide the block");
  constexpr bool kGoLast = kGout && !(LOGN & 1);  // the last radix-4 stage writes global
  // the thread index is made opaque here so that the per-thread addresses
  // below are formed in each call, not hoisted out of a caller's loop and
  // kept live (in VGPRs) across the caller's own work
  int tid = threadIdx.x;
  asm volatile("" : "+v"(tid));
  const int j = tid & (NQ - 1);
  const int l0 = tid >> LNQ;
  float2* const base = buf + l0 * LS;
  // element j + m n/4 of a line (n/4 >= 32 when padded: lpad adds m n/128)
  const int jr = PAD ? lpad(j) : j;
  constexpr int RQ = PAD ? NQ + (NQ >> 5) : NQ;
  auto stage = [&](auto s_, auto gi, auto go) __attribute__((always_inline)) {
    constexpr int S = decltype(s_)::value;
    constexpr bool GI = decltype(gi)::value, GO = decltype(go)::value;
    constexpr int Ns = 1 << (2 * S);
    const int k = j & (Ns - 1);
    float2 w1 = make_float2(1.f, 0.f), w2 = w1, w3 = w1;
    if constexpr (Ns > 1) {
      const int t = (k << (LNQ - 2 * S)) * tws;  // k * n / (4 Ns)
      w1 = tw[t]; w2 = tw[2 * t]; w3 = tw[3 * t];
      if (DIR > 0) { w1.y = -w1.y; w2.y = -w2.y; w3.y = -w3.y; }
    }
    float2 a[MAXB][4];
    bool live[MAXB];
#pragma unroll
    for (int c = 0; c < MAXB; ++c) {
      if constexpr (GI) {
        const long long b = in.base[l0 + c * LSTEP];
        live[c] = b >= 0;
        const float2* src = in.g + (live[c] ? b : 0) + j;
#pragma unroll
        for (int m = 0; m < 4; ++m) a[c][m] = src[m * NQ];
      } else {
        live[c] = true;
#pragma unroll
        for (int m = 0; m < 4; ++m) a[c][m] = base[c * DL + jr + m * RQ];
      }
    }
    float2 v[MAXB][4];
#pragma unroll
    for (int c = 0; c < MAXB; ++c) {
      float2 a0 = a[c][0], a1 = a[c][1], a2 = a[c][2], a3 = a[c][3];
      if (GI && !live[c]) a0 = a1 = a2 = a3 = make_float2(0.f, 0.f);
      if constexpr (Ns > 1) { a1 = cmul(a1, w1); a2 = cmul(a2, w2); a3 = cmul(a3, w3); }
      const float2 b0 = cadd(a0, a2), b1 = csub(a0, a2), b2 = cadd(a1, a3);
      const float2 d = csub(a1, a3);
      const float2 b3 = DIR < 0 ? make_float2(d.y, -d.x) : make_float2(-d.y, d.x);
      v[c][0] = cadd(b0, b2); v[c][2] = csub(b0, b2);
      v[c][1] = cadd(b1, b3); v[c][3] = csub(b1, b3);
    }
    if constexpr (GO) {
#pragma unroll
      for (int c = 0; c < MAXB; ++c) {
        const long long b = out.base[l0 + c * LSTEP];
        if (b >= 0) {
          float2* dst = out.g + b + j;
#pragma unroll
          for (int m = 0; m < 4; ++m) dst[m * NQ] = v[c][m];
        }
      }
    } else {
      if (!GI) __syncthreads();  // every LDS read of this stage is done
      // element dst0 + m Ns; padded: lpad(dst0 + m Ns) = lpad(dst0) + m Ns +
      // an offset that depends on m only (dst0 % 32 + 3 Ns < 32 for Ns <= 8,
      // dst0 % 32 < 16 for Ns = 16, dst0 % 32 = k % 32 for Ns >= 32)
      const int dst0 = ((j - k) << 2) + k;
      float2* D = base + (PAD ? lpad(dst0) : dst0);
#pragma unroll
      for (int c = 0; c < MAXB; ++c)
#pragma unroll
        for (int m = 0; m < 4; ++m) {
          const int mo = !PAD ? 0 : Ns >= 32 ? m * (Ns >> 5) : Ns == 16 ? (m >> 1) : 0;
          D[c * DL + m * Ns + mo] = v[c][m];
        }
    }
    __syncthreads();
  };
  using T = std::true_type;
  using F = std::false_type;
  using I = std::integral_constant<bool, kGin>;
  stage(std::integral_constant<int, 0>{}, I{}, std::integral_constant<bool, kGoLast && S4 == 1>{});
  if constexpr (S4 > 1) stage(std::integral_constant<int, 1>{}, F{}, std::integral_constant<bool, kGoLast && S4 == 2>{});
  if constexpr (S4 > 2) stage(std::integral_constant<int, 2>{}, F{}, std::integral_constant<bool, kGoLast && S4 == 3>{});
  if constexpr (S4 > 3) stage(std::integral_constant<int, 3>{}, F{}, std::integral_constant<bool, kGoLast && S4 == 4>{});
  if constexpr (S4 > 4) stage(std::integral_constant<int, 4>{}, F{}, std::integral_constant<bool, kGoLast && S4 == 5>{});
  static_assert(S4 <= 5, "fixed FFT: n <= 1024");
  if constexpr ((LOGN & 1) != 0) {  // final radix-2 stage, Ns = n/2, in place per butterfly
    constexpr int NH = N >> 1, LH = LOGN - 1;
    constexpr int NL = MAXB * LSTEP;           // lines in the tile
    constexpr int BPT = NL * NH / NT;          // butterflies per thread
    constexpr int LPI = NT / NH;               // lines per round of NT butterflies
    static_assert(NT % NH == 0 && (NL * NH) % NT == 0, "fixed FFT: radix-2 stage shape");
    const int jj = tid & (NH - 1);
    const int line0 = tid >> LH;
    float2 w = tw[jj * tws];
    if (DIR > 0) w.y = -w.y;
    float2* const L = buf + line0 * LS + (PAD ? lpad(jj) : jj);
    constexpr int RH = PAD ? NH + (NH >> 5) : NH;  // lpad(jj + n/2) - lpad(jj), n/2 >= 32
#pragma unroll
    for (int i = 0; i < BPT; ++i) {
      const float2 a0 = L[i * LPI * LS], a1 = cmul(L[i * LPI * LS + RH], w);
      if constexpr (kGout) {
        const long long gb = out.base[line0 + i * LPI];
        if (gb >= 0) {
          out.g[gb + jj] = cadd(a0, a1);
          out.g[gb + jj + NH] = csub(a0, a1);
        }
      } else {
        L[i * LPI * LS] = cadd(a0, a1);
        L[i * LPI * LS + RH] = csub(a0, a1);
      }
    }
    __syncthreads();
  }
}

// In-place FFT of `nlines` lines of length n (power of two, n <= 4 blockDim.x)
// held in LDS at buf[l * ls + e].  tw[k * tws] = exp(-2 pi i k / n).  DIR = -1
// forward (exp(-i...)), +1 inverse (exp(+i...), unnormalised).  Stockham
// radix-4 stages (+ a final radix-2 stage for odd log2 n).  Thread t owns
// butterfly column j = t % (n/4) of lines t / (n/4) + c * (blockDim.x / (n/4)),
// c < MAXB, so its twiddles and offsets are shared by its MAXB butterflies.
// Requires nlines * n / 4 <= MAXB * blockDim.x.  With a GlobalLines input the
// caller need not fill buf; with a GlobalLines output buf is left holding
// intermediate data.  Ends with __syncthreads().
//
// NT > 0 names the block size at compile time: full tiles of n = 128 or 256
// at stride n + 1, or in the padded layout, then take lds_fft_fixed (the
// padded layout is only valid there: a caller pads exactly when
// fft_fixed_tile holds).
template <int DIR, int MAXB, int NT = 0, class IN = LdsLines, class OUT = LdsLines>
__device__ __forceinline__ void lds_fft(float2* buf, int n, int nlines, int ls, const float2* __restrict__ tw, int tws,
                                        IN in = {}, OUT out = {}) {
  constexpr bool kGin = std::is_same<IN, GlobalLines>::value;
  constexpr bool kGout = std::is_same<OUT, GlobalLines>::value;
  if constexpr (NT >= 64) {
    if (fft_fixed_tile(n, nlines, MAXB, NT)) {
      if (n == 256 && ls == 257) return lds_fft_fixed<DIR, MAXB, 8, NT, false>(buf, tw, tws, in, out);
      if (n == 256 && ls == padded_stride(256)) return lds_fft_fixed<DIR, MAXB, 8, NT, true>(buf, tw, tws, in, out);
      if (n == 128 && ls == 129) return lds_fft_fixed<DIR, MAXB, 7, NT, false>(buf, tw, tws, in, out);
      if (n == 128 && ls == padded_stride(128)) return lds_fft_fixed<DIR, MAXB, 7, NT, true>(buf, tw, tws, in, out);
    }
  }
  const int log2n = __ffs(n) - 1;
  if (kGin && log2n < 2) {  // no radix-4 stage to read the input: copy it in
    for (int e = threadIdx.x; e < (nlines << log2n); e += blockDim.x) {
      const int l = e >> log2n, i = e & (n - 1);
      if constexpr (kGin) {
        const long long b = in.base[l];
        buf[l * ls + i] = b >= 0 ? in.g[b + i] : make_float2(0.f, 0.f);
      }
    }
    __syncthreads();
  }
  if (n <= 1) {
    if constexpr (kGout)
      for (int l = threadIdx.x; l < nlines; l += blockDim.x)
        if (out.base[l] >= 0) out.g[out.base[l]] = buf[l * ls];
    return;
  }
  if (log2n >= 2) {
    const int nq = n >> 2, lognq = log2n - 2;
    const int j = threadIdx.x & (nq - 1);
    const int l0 = threadIdx.x >> lognq;
    const int lstep = blockDim.x >> lognq;
    float2* const base = buf + l0 * ls;
    const int dl = lstep * ls;
    // One radix-4 stage of span Ns = 4^s; GI: read the input from global
    // (stage 0 only), GO: write the output to global (the last stage, Ns =
    // n/4, k = j, outputs j + m n/4).  Both are template constants so that
    // no stage mixes LDS and global stores.
    auto stage = [&](int Ns, int lNs, auto gi, auto go) __attribute__((always_inline)) {
      constexpr bool GI = decltype(gi)::value, GO = decltype(go)::value;
      const int k = j & (Ns - 1);
      float2 w1 = make_float2(1.f, 0.f), w2 = w1, w3 = w1;
      if (Ns > 1) {  // stage 0 has k = 0 everywhere
        const int t = (k << (lognq - lNs)) * tws;  // k * n / (4 Ns)
        w1 = tw[t]; w2 = tw[2 * t]; w3 = tw[3 * t];
        if (DIR > 0) { w1.y = -w1.y; w2.y = -w2.y; w3.y = -w3.y; }
      }
      float2 a[MAXB][4];
      bool live[MAXB];
#pragma unroll
      for (int c = 0; c < MAXB; ++c) {
        const int l = l0 + c * lstep;
        if constexpr (GI) {
          // every load issued before the first use; dead lines read line 0
          const long long b = l < nlines ? in.base[l] : -1;
          live[c] = b >= 0;
          const float2* src = in.g + (live[c] ? b : 0) + j;
#pragma unroll
          for (int m = 0; m < 4; ++m) a[c][m] = src[m * nq];
        } else {
          live[c] = l < nlines;
          const float2* L = live[c] ? base + c * dl + j : buf;  // dead: read line 0
#pragma unroll
          for (int m = 0; m < 4; ++m) a[c][m] = L[m * nq];
        }
      }
      float2 v[MAXB][4];
#pragma unroll
      for (int c = 0; c < MAXB; ++c) {
        float2 a0 = a[c][0], a1 = a[c][1], a2 = a[c][2], a3 = a[c][3];
        if (GI && !live[c]) a0 = a1 = a2 = a3 = make_float2(0.f, 0.f);
        if (Ns > 1) { a1 = cmul(a1, w1); a2 = cmul(a2, w2); a3 = cmul(a3, w3); }
        const float2 b0 = cadd(a0, a2), b1 = csub(a0, a2), b2 = cadd(a1, a3);
        const float2 d = csub(a1, a3);
        // forward: b3 = -i d ; inverse: b3 = +i d
        const float2 b3 = DIR < 0 ? make_float2(d.y, -d.x) : make_float2(-d.y, d.x);
        v[c][0] = cadd(b0, b2); v[c][2] = csub(b0, b2);
        v[c][1] = cadd(b1, b3); v[c][3] = csub(b1, b3);
      }
      if constexpr (GO) {
#pragma unroll
        for (int c = 0; c < MAXB; ++c) {
          const int l = l0 + c * lstep;
          const long long b = l < nlines ? out.base[l] : -1;
          if (b >= 0) {
            float2* dst = out.g + b + j;
#pragma unroll
            for (int m = 0; m < 4; ++m) dst[m * nq] = v[c][m];
          }
        }
      } else {
        if (!GI) __syncthreads();  // every LDS read of this stage is done
        const int dst0 = ((j - k) << 2) + k;
#pragma unroll
        for (int c = 0; c < MAXB; ++c) {
          if (l0 + c * lstep < nlines) {
            float2* D = base + c * dl + dst0;
#pragma unroll
            for (int m = 0; m < 4; ++m) D[m * Ns] = v[c][m];
          }
        }
      }
      __syncthreads();
    };
    using T = std::true_type;
    using F = std::false_type;
    const int S4 = log2n >> 1;
    const bool go_last = kGout && !(log2n & 1);  // the last radix-4 stage writes global
    const int nmid_end = go_last ? S4 - 1 : S4;  // stages [1, nmid_end) are LDS to LDS
    if (S4 == 1 && go_last)
      stage(1, 0, std::integral_constant<bool, kGin>{}, T{});
    else
      stage(1, 0, std::integral_constant<bool, kGin>{}, F{});
    int Ns = 4, lNs = 2;
    for (int s = 1; s < nmid_end; ++s, Ns <<= 2, lNs += 2) stage(Ns, lNs, F{}, F{});
    if (go_last && S4 > 1) stage(Ns, lNs, F{}, T{});
  }
  if (log2n & 1) {  // final radix-2 stage, Ns = n/2: in place per butterfly
    const int nh = n >> 1;
    const int total2 = nlines * nh;
    for (int b = threadIdx.x; b < total2; b += blockDim.x) {
      const int line = b >> (log2n - 1), j = b & (nh - 1);
      float2* L = buf + line * ls;
      float2 w = tw[j * tws];
      if (DIR > 0) w.y = -w.y;
      const float2 a0 = L[j], a1 = cmul(L[j + nh], w);
      if constexpr (kGout) {
        const long long gb = out.base[line];
        if (gb >= 0) {
          out.g[gb + j] = cadd(a0, a1);
          out.g[gb + j + nh] = csub(a0, a1);
        }
      } else {
        L[j] = cadd(a0, a1);
        L[j + nh] = csub(a0, a1);
      }
    }
    __syncthreads();
  }
}

__device__ __forceinline__ int ilog2(int n) { return __ffs(n) - 1; }  // n a power of two

// The z-pass pre/post-processing that turns a length-M complex FFT of a real
// line of length nz = 2M (packed as z[m] = x[2m] + i x[2m+1]) into its
// half spectrum X[0..M], and back.  L points at one LDS line of M+1 values;
// threads t = 0..(M/2) of the line each handle the pair (k, M-k).
__device__ inline void r2c_split(float2* L, int M, int k, const float2* __restrict__ twz) {
  if (k == 0) {
    const float2 z0 = L[0];
    L[0] = make_float2(z0.x + z0.y, 0.0f);
    L[M] = make_float2(z0.x - z0.y, 0.0f);
  } else {
    const int km = M - k;
    const float2 zk = L[k], zm = L[km];
    // E_k = (zk + conj zm)/2, O_k = -i (zk - conj zm)/2, X_k = E_k + W^k O_k.
    // E_{M-k} = conj E_k, O_{M-k} = conj O_k and W^{M-k} = -conj W^k, so
    // X_{M-k} = conj(E_k - W^k O_k).
    const float2 Ek = make_float2(0.5f * (zk.x + zm.x), 0.5f * (zk.y - zm.y));
    const float2 Ok = make_float2(0.5f * (zk.y + zm.y), -0.5f * (zk.x - zm.x));
    const float2 T = cmul(twz[k], Ok);
    L[k] = cadd(Ek, T);
    L[km] = make_float2(Ek.x - T.x, T.y - Ek.y);
  }
}

// Inverse of r2c_split (unnormalised): Z[k] = (X[k] + conj X[M-k]) + i (X[k] - conj X[M-k]) w^k,
// w = exp(+2 pi i / nz); the imaginary parts of X[0] and X[M] are ignored.
__device__ inline void c2r_merge(float2* L, int M, int k, const float2* __restrict__ twz) {
  if (k == 0) {
    const float a = L[0].x, b = L[M].x;
    L[0] = make_float2(a + b, a - b);
  } else {
    const int km = M - k;
    const float2 xk = L[k], xm = L[km];
    // A_{M-k} = conj A_k; with w^{M-k} = -conj w^k, B_{M-k} = conj B_k
    const float2 Ak = make_float2(xk.x + xm.x, xk.y - xm.y);
    const float2 Bk = cmulc(make_float2(xk.x - xm.x, xk.y + xm.y), twz[k]);
    L[k] = make_float2(Ak.x - Bk.y, Ak.y + Bk.x);
    L[km] = make_float2(Ak.x + Bk.y, Bk.x - Ak.y);
  }
}

// Apply r2c_split (R2C = true) or c2r_merge to every pair (k, M-k), k = 0..M/2,
// of nlines LDS lines of stride ls (M a power of two).
template <bool R2C>
__device__ inline void z_pairs(float2* lds, int nlines, int ls, int M, const float2* __restrict__ tw) {
  const int half = M >> 1;
  if (half > 0) {
    const int lh = ilog2(half);
    for (int e = threadIdx.x; e < (nlines << lh); e += blockDim.x) {
      const int l = e >> lh, k = e & (half - 1);
      if (R2C) r2c_split(lds + l * ls, M, k, tw);
      else c2r_merge(lds + l * ls, M, k, tw);
    }
  }
  for (int l = threadIdx.x; l < nlines; l += blockDim.x) {  // k = M/2 (k = 0 when M = 1)
    if (R2C) r2c_split(lds + l * ls, M, half, tw);
    else c2r_merge(lds + l * ls, M, half, tw);
  }
}

}  // namespace fd3
