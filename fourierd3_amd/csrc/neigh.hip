// Cell-list neighbour search and the smooth-cutoff coordination number.
//
// Atoms are binned by their fractional coordinates wrapped into [0, 1): nb_a
// bins along lattice vector a, with nb_a = max(1, floor(d_a / r_cut)) where
// d_a is the spacing of the lattice planes normal to a.  A bin is then at
// least r_cut thick, and every image within r_cut of an atom lies in the bins
// b + d, |d_a| <= s_a = ceil(r_cut nb_a / d_a), with unwrapped bin b + d
// mapping to bin (b + d) mod nb_a and lattice shift floor((b + d) / nb_a).
// Thin boxes (d_a < r_cut) get nb_a = 1 and s_a > 1, i.e. several images of
// the same atoms, including an atom's own images.  This enumerates exactly
// matscipy's neighbour_list("ijS") pairs (src/fourierd3/calculator.py:103-109):
// every (i, j, S) with |r_j + S cell - r_i| < r_cut except i == j, S == 0.
//
// The CN of src/fourierd3/core.py:348-386 sums, for each atom, over these
// pairs, so no edge list is materialised on the calculator path.

#include <climits>

#include "fd3_internal.h"
#include "launch.h"
#include "neigh_kernels.h"

namespace fd3 {

// Smooth-cutoff CN term of one pair at squared distance r2 (Bohr^2) and its
// derivative wrt r divided by r:
//   r1 = (Rcov_ij + r_cut)/2, r' = max(r, r1)
//   k  = 16 + (r' - r1)^2 / ((r_cut - r')^2 + 1e-6)
//   f  = sigmoid(k (4/3 Rcov_ij / r - 1)),  f = 0 for r <= 1e-6
// (src/fourierd3/core.py:352-378), evaluated in double: with heavy elements
// dE/dCN reaches hundreds of eV per unit CN, so an atom's CN force is a sum
// of terms of that size that cancel to a few eV/A, and float terms (relative
// error ~1e-6 here) leave errors of 1e-3 eV/A.  sigma (1 - sigma) is formed
// as e sigma^2 with e = exp(-s), which does not cancel where sigma -> 1.
// 1/sqrt(x) and 1/x for positive normal x in float range: the hardware
// float estimate (~1 ulp of float) and one Newton step in double, which
// squares the relative error to ~1e-14 (the CN term needs ~1e-10).
__device__ inline double rsq_d(double x) {
  const double y = (double)__builtin_amdgcn_rsqf((float)x);
  return y * (1.5 - 0.5 * x * y * y);
}
__device__ inline double rcp_d(double x) {
  const double y = (double)__builtin_amdgcn_rcpf((float)x);
  return y * (2.0 - x * y);
}

__device__ inline void cn_term(double r2, double rcs, double rcut, double* f, double* fpr) {
  if (!(r2 > 1e-12)) {
    *f = 0.0;
    *fpr = 0.0;
    return;
  }
  const double ir = rsq_d(r2);
  const double r = r2 * ir;
  const double r1 = 0.5 * rcs + 0.5 * rcut;
  const bool far = r > r1;
  const double rc = far ? r : r1;
  const double dr = rc - r1, dc = rcut - rc;
  const double iden = rcp_d(dc * dc + 1e-6);
  const double k = 16.0 + dr * dr * iden;
  const double ratio = (4.0 / 3.0) * rcs;
  const double arg = ratio * ir - 1.0;
  const double s = k * arg;
  if (s < -80.0) {  // sigma < 2e-35 and its derivative below 1e-22: zero (and exp(-s) stays finite)
    *f = 0.0;
    *fpr = 0.0;
    return;
  }
  const double e = exp(-s);
  const double sig = rcp_d(1.0 + e);
  *f = sig;
  const double dk = far ? 2.0 * dr * iden * (1.0 + dr * dc * iden) : 0.0;
  const double ds = dk * arg - k * ratio * ir * ir;
  *fpr = e * sig * sig * ds * ir;
}

// Take the cell and its inverse from the device cell when there is one.
__device__ inline void cell_refresh(CellArgs& c) {
  if (!c.dyn) return;
  for (int i = 0; i < 9; ++i) {
    c.cell[i] = c.dyn->cell[i];
    c.inv[i] = c.dyn->inv[i];
  }
}

__device__ inline int floordiv(int a, int b) {
  int q = a / b;
  if ((a % b) != 0 && ((a < 0) != (b < 0))) --q;
  return q;
}

// Visit every neighbour image (t, S) of slot s within the cutoff.
template <typename F>
__device__ inline void for_each_neighbour(const CellArgs& c, int64_t s, F&& fn) {
  const double4 ps = c.pos4[s];
  const int lin = c.sbin[s];
  const int bz = lin % c.nb[2], by = (lin / c.nb[2]) % c.nb[1], bx = lin / (c.nb[2] * c.nb[1]);
  for (int dx = -c.sr[0]; dx <= c.sr[0]; ++dx) {
    const int ux = bx + dx, wx = wrap(ux, c.nb[0]), Sx = floordiv(ux, c.nb[0]);
    for (int dy = -c.sr[1]; dy <= c.sr[1]; ++dy) {
      const int uy = by + dy, wy = wrap(uy, c.nb[1]), Sy = floordiv(uy, c.nb[1]);
      for (int dz = -c.sr[2]; dz <= c.sr[2]; ++dz) {
        const int uz = bz + dz, wz = wrap(uz, c.nb[2]), Sz = floordiv(uz, c.nb[2]);
        const double shx = Sx * c.cell[0] + Sy * c.cell[3] + Sz * c.cell[6];
        const double shy = Sx * c.cell[1] + Sy * c.cell[4] + Sz * c.cell[7];
        const double shz = Sx * c.cell[2] + Sy * c.cell[5] + Sz * c.cell[8];
        const bool zero = (Sx | Sy | Sz) == 0;
        const int wb = (wx * c.nb[1] + wy) * c.nb[2] + wz;
        const int t1 = c.bstart[wb + 1];
        for (int t = c.bstart[wb]; t < t1; ++t) {
          if (zero && t == s) continue;
          const double4 pt = c.pos4[t];
          const double vx = pt.x + shx - ps.x, vy = pt.y + shy - ps.y, vz = pt.z + shz - ps.z;
          const double r2 = vx * vx + vy * vy + vz * vz;
          if (r2 < c.rc2) fn(t, vx, vy, vz, r2, Sx, Sy, Sz);
        }
      }
    }
  }
}

// --------------------------------------------------------- build kernels ---
struct BinArgs {
  CellArgs c;
  const void* pos;
  int dtype;
  int64_t n;
  double4* posw;
  int4* flo;
  int32_t* abin;
  int32_t* bcount;
};

__device__ __forceinline__ void bin_body(const BlockInfo& bk, const BinArgs& a) {
  const int64_t i = (int64_t)bk.id.x * blockDim.x + threadIdx.x;
  if (i >= a.n) return;
  CellArgs c = a.c;
  cell_refresh(c);
  double r[3];
  load_pos(a.pos, a.dtype, i, r);
  double fw[3];
  int fl[3], b[3];
  for (int d = 0; d < 3; ++d) {
    const double f = r[0] * c.inv[0 * 3 + d] + r[1] * c.inv[1 * 3 + d] + r[2] * c.inv[2 * 3 + d];
    const double flf = floor(f);
    fl[d] = (int)flf;
    fw[d] = f - flf;
    if (fw[d] >= 1.0) {  // f just below an integer: f - floor(f) rounds up to 1
      fw[d] = 0.0;
      fl[d] += 1;
    }
    b[d] = min((int)(fw[d] * c.nb[d]), c.nb[d] - 1);
  }
  const double x = fw[0] * c.cell[0] + fw[1] * c.cell[3] + fw[2] * c.cell[6];
  const double y = fw[0] * c.cell[1] + fw[1] * c.cell[4] + fw[2] * c.cell[7];
  const double z = fw[0] * c.cell[2] + fw[1] * c.cell[5] + fw[2] * c.cell[8];
  a.posw[i] = make_double4(x, y, z, 0.0);
  a.flo[i] = make_int4(fl[0], fl[1], fl[2], 0);
  const int lin = (b[0] * c.nb[1] + b[1]) * c.nb[2] + b[2];
  a.abin[i] = lin;
  atomicAdd(a.bcount + lin, 1);
}

// Exclusive scan of n ints into out[0..n] (out[n] = total); one block of 1024.
// OutT = int64_t for the neighbour-pair offsets, whose total may pass 2^31.
template <typename OutT>
struct Scan1Args {
  const int32_t* in;
  int64_t n;
  OutT* out;
  int64_t* total64;
};

template <typename OutT>
__device__ __forceinline__ void scan1_body(const BlockInfo&, const Scan1Args<OutT>& a) {
  __shared__ int64_t part[1024];
  const int t = threadIdx.x;
  const int64_t n = a.n;
  const int64_t chunk = (n + 1023) / 1024;
  const int64_t lo = t * chunk, hi = min(n, lo + chunk);
  int64_t s = 0;
  for (int64_t i = lo; i < hi; ++i) s += a.in[i];
  part[t] = s;
  __syncthreads();
  for (int off = 1; off < 1024; off <<= 1) {
    int64_t v = t >= off ? part[t - off] : 0;
    __syncthreads();
    part[t] += v;
    __syncthreads();
  }
  int64_t run = t ? part[t - 1] : 0;
  for (int64_t i = lo; i < hi; ++i) {
    a.out[i] = (OutT)run;
    run += a.in[i];
  }
  if (t == 1023) {
    a.out[n] = (OutT)part[1023];
    if (a.total64) a.total64[0] = part[1023];
  }
}

// Device-wide exclusive scan in two steps: per-block scans of 1024
// elements, then the offset add (scan_add_body).
struct ScanArgs {
  const int32_t* in;
  int64_t n;
  int32_t* out;
  int32_t* bsum;
  int64_t* total64;
};

__device__ __forceinline__ void scan_blocks_body(const BlockInfo& bk, const ScanArgs& a) {
  __shared__ int32_t sh[256];
  const int64_t base = (int64_t)bk.id.x * 1024 + 4 * threadIdx.x;
  int32_t v[4];
  int32_t t = 0;
  for (int k = 0; k < 4; ++k) {
    v[k] = base + k < a.n ? a.in[base + k] : 0;
    t += v[k];
  }
  sh[threadIdx.x] = t;
  __syncthreads();
  for (int off = 1; off < 256; off <<= 1) {
    const int32_t x = threadIdx.x >= off ? sh[threadIdx.x - off] : 0;
    __syncthreads();
    sh[threadIdx.x] += x;
    __syncthreads();
  }
  int32_t run = sh[threadIdx.x] - t;
  for (int k = 0; k < 4; ++k) {
    if (base + k < a.n) a.out[base + k] = run;
    run += v[k];
  }
  if (threadIdx.x == 255) a.bsum[bk.id.x] = sh[255];
}

// Block b of 1024 elements adds the sum of the block totals before it (each
// block reduces that prefix itself, so no separate scan of the totals is
// launched); the last block also writes out[n] = total.
__device__ __forceinline__ void scan_add_body(const BlockInfo& bk, const ScanArgs& a) {
  __shared__ int64_t red[16];
  const int64_t b = bk.id.x;
  int64_t v = 0;
  for (int64_t i = threadIdx.x; i < b; i += 1024) v += a.bsum[i];
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
  __syncthreads();
  int64_t off = 0;
#pragma unroll
  for (int w = 0; w < 16; ++w) off += red[w];
  const int64_t i = b * 1024 + threadIdx.x;
  if (i < a.n) a.out[i] += (int32_t)off;
  if (b == (int64_t)bk.n.x - 1 && threadIdx.x == 0) {
    const int64_t tot = off + a.bsum[b];
    a.out[a.n] = (int32_t)tot;
    if (a.total64) a.total64[0] = tot;
  }
}

int scan_exclusive(const int32_t* in, int64_t n, int32_t* out, int32_t* tmp, int64_t* total, hipStream_t s) {
  const int64_t nb = std::max<int64_t>(1, cdiv(n, 1024));
  const ScanArgs a{in, n, out, tmp, total};
  int rc;
  if ((rc = launch<ScanArgs, scan_blocks_body, 256>("k_scan_blocks", dim3((unsigned)nb), 0, s, a))) return rc;
  return launch<ScanArgs, scan_add_body, 1024>("k_scan_add", dim3((unsigned)nb), 0, s, a);
}

size_t scan_tmp_elems(int64_t n) { return (size_t)std::max<int64_t>(1, cdiv(n, 1024)); }

struct ScatterArgs {
  const int32_t* abin;
  int64_t n;
  const int32_t* bstart;
  int32_t* bcur;
  int32_t* slot_atom;
};

__device__ __forceinline__ void scatter_body(const BlockInfo& bk, const ScatterArgs& a) {
  const int64_t i = (int64_t)bk.id.x * blockDim.x + threadIdx.x;
  if (i >= a.n) return;
  const int b = a.abin[i];
  const int slot = a.bstart[b] + atomicAdd(a.bcur + b, 1);
  a.slot_atom[slot] = (int32_t)i;
}

// Deterministic order inside each bin: ascending atom index.  One wave per
// bin: each atom's slot is lo + (number of atoms of the bin with a smaller
// index), counted against 64 atoms at a time through lane shuffles.  Reads the
// scattered (arbitrary) order from `in`, writes the sorted one to `out`.
struct SortBinsArgs {
  const int32_t* bstart;
  int nbins;
  const int32_t* in;
  int32_t* out;
};

__device__ __forceinline__ void sort_bins_body(const BlockInfo& bk, const SortBinsArgs& a) {
  const int lane = threadIdx.x & 63;
  const int b = bk.id.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
  if (b >= a.nbins) return;  // uniform per wave
  const int lo = a.bstart[b], cnt = a.bstart[b + 1] - lo;
  for (int i0 = 0; i0 < cnt; i0 += 64) {
    const bool mine = i0 + lane < cnt;
    const int vi = mine ? a.in[lo + i0 + lane] : INT_MAX;
    int rank = 0;
    for (int j0 = 0; j0 < cnt; j0 += 64) {
      const int vj = j0 + lane < cnt ? a.in[lo + j0 + lane] : INT_MAX;
      const int m = min(64, cnt - j0);
      for (int t = 0; t < m; ++t) rank += __shfl(vj, t) < vi ? 1 : 0;
    }
    if (mine) a.out[lo + rank] = vi;
  }
}

struct FillSortedArgs {
  const int32_t* slot_atom;
  int64_t n;
  const double4* posw;
  const int4* flo;
  const int32_t* abin;
  const int32_t* species;
  double4* pos4;
  int4* flo_s;
  int32_t* sbin;
  int32_t* spec_s;
};

__device__ __forceinline__ void fill_sorted_body(const BlockInfo& bk, const FillSortedArgs& a) {
  const int64_t s = (int64_t)bk.id.x * blockDim.x + threadIdx.x;
  if (s >= a.n) return;
  const int i = a.slot_atom[s];
  a.pos4[s] = a.posw[i];
  a.flo_s[s] = a.flo[i];
  a.sbin[s] = a.abin[i];
  if (a.species) a.spec_s[s] = a.species[i];
}

// ----------------------------------------------------------- CN kernels ---
// One block per home bin.  The images of the atoms in the neighbouring bins
// (candidates) are staged in LDS with their lattice shift applied, so every
// candidate is read from HBM once per home bin instead of once per home atom.
// Each wave then takes home atoms in turn: its 64 lanes test |r| < cutoff
// for up to 640 candidates (a few instructions each), compact the hits into a
// list in LDS with a ballot, and evaluate the CN term on that list with all
// lanes busy; the home atom's sums stay in registers until the wave moves to
// the next atom.  (Evaluating the term inside the distance loop would run it
// for almost every candidate: one of 64 lanes nearly always has a hit.)
// Forward:  CN_s = sum_t f(r_st)
// Backward: dE/dr_s -= sum_t (g_s + g_t) f'(r) v / r,
//           virial  += sum_t g_s f'(r) v v^T / r        (directed pairs s -> t)
// with g = dE/dCN (src/fourierd3/core.py:348-386 and its adjoint).
// Candidates staged per pass: CAP = 640 (27 neighbour bins of ~20 atoms at
// 1 g/cm^3 H/C/N/O fit one pass; LDS for five blocks per CU).  The
// candidates of a home bin form one list, neighbour bin after neighbour bin;
// a longer list (small boxes with many images per bin, dense regions) is
// staged in passes of CAP candidates.  Plans with more than kNbrMax
// neighbour bins stage one neighbour bin at a time instead.  (A second
// variant staging up to 1536 candidates at once ran two blocks per CU and
// was slower on every measured workload.)
constexpr int kCnCapSmall = 640;
constexpr int kNbrMax = 343;   // neighbour bins of a single-pass plan (reach <= 3)
constexpr int kCnHomes = 128;  // home atoms per pass
// ballot steps (of 64 candidates) per hit list: kCnCapSmall / 64, so the
// candidates of a single pass form one list (a second list of a few
// leftover candidates would cost a whole FP64 term step for a few lanes)
constexpr int kCnStep = kCnCapSmall / 64;

struct NbrBin {
  int wb;            // wrapped bin index
  int S[3];          // lattice shift in cells
  double sh[3];      // lattice shift vector (Bohr)
  bool zero;         // shift is zero
};

__device__ inline NbrBin nbr_bin(const CellArgs& c, int b, int k) {
  const int ny = 2 * c.sr[1] + 1, nz = 2 * c.sr[2] + 1;
  const int dz = k % nz - c.sr[2], dy = (k / nz) % ny - c.sr[1], dx = k / (nz * ny) - c.sr[0];
  const int bz = b % c.nb[2], by = (b / c.nb[2]) % c.nb[1], bx = b / (c.nb[2] * c.nb[1]);
  const int ux = bx + dx, uy = by + dy, uz = bz + dz;
  const int Sx = floordiv(ux, c.nb[0]), Sy = floordiv(uy, c.nb[1]), Sz = floordiv(uz, c.nb[2]);
  NbrBin r;
  r.wb = (wrap(ux, c.nb[0]) * c.nb[1] + wrap(uy, c.nb[1])) * c.nb[2] + wrap(uz, c.nb[2]);
  r.S[0] = Sx; r.S[1] = Sy; r.S[2] = Sz;
  for (int d = 0; d < 3; ++d) r.sh[d] = Sx * c.cell[d] + Sy * c.cell[3 + d] + Sz * c.cell[6 + d];
  r.zero = (Sx | Sy | Sz) == 0;
  return r;
}

// Lattice shift of a neighbour bin, 8 bits per axis (|S| <= reach <= 3 in a
// single-pass plan).
__device__ inline int pack_shift(const int* S) {
  return ((S[0] + 128) << 16) | ((S[1] + 128) << 8) | (S[2] + 128);
}

__device__ inline double wave_sum_d(double v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}

// Launched through the _w4 wrappers: at most 128 VGPRs in the backward form,
// four waves per SIMD, not three.
struct CnBinsArgs {
  CellArgs c;
  CnArgs a;
};

template <bool BWD, int kCnCap>
__device__ __forceinline__ void cn_bins_body(const BlockInfo& bk, const CnBinsArgs& args) {
  CellArgs c = args.c;
  const CnArgs& a = args.a;
  cell_refresh(c);
  constexpr int NW = 4;                   // waves per block
  // candidate image positions relative to the bin's first atom as float
  // pairs: hi (with rcov in w) drives the distance filter, hi + lo gives the
  // pair vector to double precision for the term itself
  __shared__ float4 cpos[kCnCap];
  __shared__ float4 clo[kCnCap];
  __shared__ double cg[BWD ? kCnCap : 1];         // dE/dCN of the candidate (backward)
  __shared__ double hacc[kCnHomes][BWD ? 3 : 1];  // per home atom: CN, or dE/dr
  __shared__ int bcnt[kNbrMax + 1];       // candidate offset of neighbour bin k
  __shared__ int nb0[kNbrMax];            // first slot of neighbour bin k
  __shared__ int nsh[kNbrMax];            // packed lattice shift of neighbour bin k
  __shared__ unsigned short wl[NW][kCnStep * 64];  // per wave: hit candidates
  __shared__ double red[16 * 6];
  const int b = bk.id.x;
  const int hs = c.bstart[b], he = c.bstart[b + 1];
  if (hs == he) {  // empty bin (uniform across the block): zero virial partials
    if (BWD && threadIdx.x < 8) a.partials[(size_t)b * 8 + threadIdx.x] = 0.0;
    return;
  }
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const int nbr = (2 * c.sr[0] + 1) * (2 * c.sr[1] + 1) * (2 * c.sr[2] + 1);
  // origin of the block's float coordinates: differences stay within a few
  // bin widths, so float keeps ~1e-7 relative precision on pair vectors
  const double4 org = c.pos4[hs];
  bool concat = false, single = false;  // one list; a list that fits one pass
  if (nbr <= kNbrMax) {
    for (int k = threadIdx.x; k < nbr; k += blockDim.x) {
      const NbrBin nb = nbr_bin(c, b, k);
      const int lo = c.bstart[nb.wb];
      nb0[k] = lo;
      nsh[k] = pack_shift(nb.S);
      bcnt[k + 1] = c.bstart[nb.wb + 1] - lo;
    }
    __syncthreads();
    if (threadIdx.x < 64) {  // wave 0: prefix sums of the counts, 64 bins at a time
      int carry = 0;
      for (int k0 = 0; k0 < nbr; k0 += 64) {
        const int k = k0 + lane;
        int v = k < nbr ? bcnt[k + 1] : 0;
#pragma unroll
        for (int o = 1; o < 64; o <<= 1) {
          const int u = __shfl_up(v, o, 64);
          if (lane >= o) v += u;
        }
        if (k < nbr) bcnt[k + 1] = carry + v;
        carry += __shfl(v, 63, 64);
      }
      if (lane == 0) bcnt[0] = 0;
    }
    __syncthreads();
    concat = true;
    single = bcnt[nbr] <= kCnCap;
  }
  const float rc2 = (float)c.rc2;
  double vir[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
  for (int h0 = hs; h0 < he; h0 += kCnHomes) {
    const int nh = min(kCnHomes, he - h0);
    for (int h = threadIdx.x; h < nh; h += blockDim.x)
      for (int d = 0; d < (BWD ? 3 : 1); ++d) hacc[h][d] = 0.0;
    // candidate passes of up to kCnCap: over the whole list, or over one
    // neighbour bin at a time
    const int npass_bins = concat ? 1 : nbr;
    for (int kb = 0; kb < npass_bins; ++kb) {
      int count;
      NbrBin fixed;
      if (concat) {
        count = bcnt[nbr];
      } else {
        fixed = nbr_bin(c, b, kb);
        count = c.bstart[fixed.wb + 1] - c.bstart[fixed.wb];
      }
      for (int off0 = 0; off0 < count; off0 += kCnCap) {
        const int nc = min(kCnCap, count - off0);
        __syncthreads();  // hacc cleared; the previous pass is done with cpos
        for (int e = threadIdx.x; e < nc; e += blockDim.x) {
          double sh[3];
          int t;
          if (concat) {
            const int eg = off0 + e;   // position in the list
            int lo = 0, hi = nbr - 1;  // the neighbour bin k with bcnt[k] <= eg < bcnt[k + 1]
            while (lo < hi) {
              const int mid = (lo + hi + 1) >> 1;
              if (bcnt[mid] <= eg) lo = mid; else hi = mid - 1;
            }
            const int ps = nsh[lo];
            const int Sx = (ps >> 16) - 128, Sy = ((ps >> 8) & 255) - 128, Sz = (ps & 255) - 128;
            for (int d = 0; d < 3; ++d) sh[d] = Sx * c.cell[d] + Sy * c.cell[3 + d] + Sz * c.cell[6 + d];
            t = nb0[lo] + (eg - bcnt[lo]);
          } else {
            for (int d = 0; d < 3; ++d) sh[d] = fixed.sh[d];
            t = c.bstart[fixed.wb] + off0 + e;
          }
          const double4 p = c.pos4[t];
          const double rx = p.x + sh[0] - org.x, ry = p.y + sh[1] - org.y, rz = p.z + sh[2] - org.z;
          const float hx = (float)rx, hy = (float)ry, hz = (float)rz;
          cpos[e] = make_float4(hx, hy, hz, a.rcov[c.spec_s[t]]);
          clo[e] = make_float4((float)(rx - (double)hx), (float)(ry - (double)hy), (float)(rz - (double)hz), 0.f);
          if (BWD) cg[e] = a.dEdcn[t];
        }
        __syncthreads();
        // single pass: the home atoms are the candidates of the centre
        // neighbour bin (zero shift), in home order from slot hc
        const int hc = single ? bcnt[(nbr - 1) >> 1] + (h0 - hs) : 0;
        for (int h = wv; h < nh; h += NW) {  // home atom h0 + h: wave-uniform
          const int sidx = h0 + h;
          float px, py, pz;
          double lx, ly, lz, ri, gs;
          if (single) {
            const float4 q = cpos[hc + h], ql = clo[hc + h];
            px = q.x; py = q.y; pz = q.z;
            lx = ql.x; ly = ql.y; lz = ql.z;
            ri = (double)q.w;
            gs = BWD ? cg[hc + h] : 0.0;
          } else {
            const double4 pd = c.pos4[sidx];
            const double hxd = pd.x - org.x, hyd = pd.y - org.y, hzd = pd.z - org.z;
            px = (float)hxd; py = (float)hyd; pz = (float)hzd;
            lx = (double)(float)(hxd - (double)px);
            ly = (double)(float)(hyd - (double)py);
            lz = (double)(float)(hzd - (double)pz);
            ri = (double)a.rcov[c.spec_s[sidx]];
            gs = BWD ? a.dEdcn[sidx] : 0.0;
          }
          double s0 = 0.0, s1 = 0.0, s2 = 0.0;
          for (int e0 = 0; e0 < nc; e0 += kCnStep * 64) {
            // in-range candidates of [e0, e0 + 640) -> wl[wv][0 .. cnt)
            int cnt = 0;
            const int nit = min(kCnStep, (nc - e0 + 63) >> 6);  // steps holding candidates
#pragma unroll
            for (int it = 0; it < kCnStep; ++it) {
              if (it >= nit) break;  // uniform
              // a lane past the end reads the last candidate and never hits
              const int e = e0 + it * 64 + lane;
              const float4 q = cpos[min(e, nc - 1)];
              const float vx = q.x - px, vy = q.y - py, vz = q.z - pz;
              const bool hit = (vx * vx + vy * vy + vz * vz < rc2) & (e < nc);  // includes the atom itself at r = 0
              const unsigned long long bal = __builtin_amdgcn_ballot_w64(hit);
              const unsigned below =
                  __builtin_amdgcn_mbcnt_hi((unsigned)(bal >> 32), __builtin_amdgcn_mbcnt_lo((unsigned)bal, 0u));
              if (hit) wl[wv][cnt + below] = (unsigned short)e;
              cnt += __popcll(bal);
            }
            __builtin_amdgcn_wave_barrier();  // the list is written and read by lanes of one wave
            for (int k = lane; k < cnt; k += 64) {
              const int e = wl[wv][k];
              const float4 q = cpos[e], ql = clo[e];
              const double vx = ((double)q.x - (double)px) + ((double)ql.x - lx);
              const double vy = ((double)q.y - (double)py) + ((double)ql.y - ly);
              const double vz = ((double)q.z - (double)pz) + ((double)ql.z - lz);
              double f, w;
              cn_term(vx * vx + vy * vy + vz * vz, ri + (double)q.w, a.rcut, &f, &w);  // 0, 0 at r = 0
              if (!BWD) {
                s0 += f;
              } else {
                const double wt = w * (gs + cg[e]);
                s0 -= wt * vx; s1 -= wt * vy; s2 -= wt * vz;
                const double wvv = w * gs;
                vir[0] += wvv * vx * vx; vir[1] += wvv * vy * vy; vir[2] += wvv * vz * vz;
                vir[3] += wvv * vx * vy; vir[4] += wvv * vx * vz; vir[5] += wvv * vy * vz;
              }
            }
            __builtin_amdgcn_wave_barrier();  // the next step overwrites the list
          }
          s0 = wave_sum_d(s0);
          if (BWD) { s1 = wave_sum_d(s1); s2 = wave_sum_d(s2); }
          if (lane == 0) {  // one wave per home atom and pass: plain updates
            hacc[h][0] += s0;
            if (BWD) { hacc[h][1] += s1; hacc[h][2] += s2; }
          }
        }
      }
    }
    __syncthreads();
    for (int h = threadIdx.x; h < nh; h += blockDim.x) {
      if (!BWD) {
        a.cn[h0 + h] = hacc[h][0];  // each atom is the home of exactly one block
      } else {
        // each atom is the home of exactly one block: a plain update, or
        // (gout) the final gradient straight into the output in atom order
        const int sl = h0 + h;
        if (a.gout) {
          const int64_t i = a.slot_atom ? a.slot_atom[sl] : sl;
          for (int d = 0; d < 3; ++d) a.gout[3 * i + d] = (a.gpos[3 * sl + d] + hacc[h][BWD ? d : 0]) * kA2B;
        } else {
          for (int d = 0; d < 3; ++d) a.gpos[3 * sl + d] += hacc[h][BWD ? d : 0];
        }
      }
    }
    __syncthreads();  // the next home pass clears hacc
  }
  if (BWD) block_sum_store<6>(vir, red, a.partials + (size_t)b * 8);
}

// Edge-list forms: pairs given by (src, dst, shift in Angstrom).
__device__ inline void edge_vec(const EdgeArgs& e, int64_t k, const double4* pos4, double* v) {
  const int64_t i = e.src[k], j = e.dst[k];
  double sh[3];
  load_pos(e.shifts, e.shifts_dtype, k, sh);  // Angstrom -> Bohr
  const double4 pi = pos4[i], pj = pos4[j];
  v[0] = pj.x - pi.x + sh[0];
  v[1] = pj.y - pi.y + sh[1];
  v[2] = pj.z - pi.z + sh[2];
}

__global__ void __launch_bounds__(256) k_cn_edges(EdgeArgs e, CnArgs a, const double4* __restrict__ pos4,
                                                  const int32_t* __restrict__ spec) {
  const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= e.n_edges) return;
  double v[3];
  edge_vec(e, k, pos4, v);
  const double r2 = v[0] * v[0] + v[1] * v[1] + v[2] * v[2];
  const int64_t i = e.src[k], j = e.dst[k];
  double f, fpr;
  cn_term(r2, (double)a.rcov[spec[i]] + (double)a.rcov[spec[j]], a.rcut, &f, &fpr);
  if (f != 0.0) atomicAdd(a.cn + i, f);
}

__global__ void __launch_bounds__(256) k_cn_bwd_edges(EdgeArgs e, CnArgs a, const double4* __restrict__ pos4,
                                                      const int32_t* __restrict__ spec, double* __restrict__ gshift) {
  __shared__ double red[16 * 6];
  const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  double vir[6] = {0, 0, 0, 0, 0, 0};
  if (k < e.n_edges) {
    double v[3];
    edge_vec(e, k, pos4, v);
    const double r2 = v[0] * v[0] + v[1] * v[1] + v[2] * v[2];
    const int64_t i = e.src[k], j = e.dst[k];
    double f, fpr;
    cn_term(r2, (double)a.rcov[spec[i]] + (double)a.rcov[spec[j]], a.rcut, &f, &fpr);
    const double w = a.dEdcn[i] * fpr;
    const double g0 = w * v[0], g1 = w * v[1], g2 = w * v[2];  // dE/dv, Hartree/Bohr
    if (w != 0.0) {
      atomicAdd(a.gpos + 3 * j + 0, g0);
      atomicAdd(a.gpos + 3 * j + 1, g1);
      atomicAdd(a.gpos + 3 * j + 2, g2);
      atomicAdd(a.gpos + 3 * i + 0, -g0);
      atomicAdd(a.gpos + 3 * i + 1, -g1);
      atomicAdd(a.gpos + 3 * i + 2, -g2);
    }
    if (gshift) {
      gshift[3 * k + 0] = g0 * kA2B;
      gshift[3 * k + 1] = g1 * kA2B;
      gshift[3 * k + 2] = g2 * kA2B;
    }
    vir[0] = g0 * v[0]; vir[1] = g1 * v[1]; vir[2] = g2 * v[2];
    vir[3] = g0 * v[1]; vir[4] = g0 * v[2]; vir[5] = g1 * v[2];
  }
  block_sum_store<6>(vir, red, a.partials + (size_t)blockIdx.x * 8);
}

// ----------------------------------------------- neighbour materialisation ---
__global__ void k_nb_count(CellArgs c, int32_t* __restrict__ count_atom) {
  const int64_t s = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= c.n) return;
  int cnt = 0;
  for_each_neighbour(c, s, [&](int, double, double, double, double, int, int, int) { ++cnt; });
  count_atom[c.slot_atom[s]] = cnt;
}

__global__ void k_nb_fill(CellArgs c, const int64_t* __restrict__ offset, int64_t capacity, int64_t* __restrict__ src,
                          int64_t* __restrict__ dst, int32_t* __restrict__ sh) {
  const int64_t s = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= c.n) return;
  const int i = c.slot_atom[s];
  int64_t o = offset[i];
  const int4 fi = c.flo_s[s];
  for_each_neighbour(c, s, [&](int t, double, double, double, double, int Sx, int Sy, int Sz) {
    if (o < capacity) {
      const int4 ft = c.flo_s[t];
      src[o] = i;
      dst[o] = c.slot_atom[t];
      sh[3 * o + 0] = Sx - ft.x + fi.x;
      sh[3 * o + 1] = Sy - ft.y + fi.y;
      sh[3 * o + 2] = Sz - ft.z + fi.z;
    }
    ++o;
  });
}

__device__ inline bool nb_less(int64_t j1, const int32_t* s1, int64_t j2, const int32_t* s2) {
  if (j1 != j2) return j1 < j2;
  for (int d = 0; d < 3; ++d)
    if (s1[d] != s2[d]) return s1[d] < s2[d];
  return false;
}

// Sort each atom's pairs by (j, S).
__global__ void k_nb_sort(const int64_t* __restrict__ offset, int64_t n, int64_t capacity, int64_t* __restrict__ dst,
                          int32_t* __restrict__ sh) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int64_t lo = offset[i], hi = min(offset[i + 1], capacity);
  for (int64_t a = lo + 1; a < hi; ++a) {
    const int64_t vj = dst[a];
    const int32_t vs[3] = {sh[3 * a], sh[3 * a + 1], sh[3 * a + 2]};
    int64_t b = a - 1;
    while (b >= lo && nb_less(vj, vs, dst[b], sh + 3 * b)) {
      dst[b + 1] = dst[b];
      for (int d = 0; d < 3; ++d) sh[3 * (b + 1) + d] = sh[3 * b + d];
      --b;
    }
    dst[b + 1] = vj;
    for (int d = 0; d < 3; ++d) sh[3 * (b + 1) + d] = vs[d];
  }
}

// ------------------------------------------------------------------ host ---
int cell_grid_make(const double cell_A[9], double r_cut_A, CellGrid* g, double reach) {
  double c[9];
  for (int i = 0; i < 9; ++i) c[i] = cell_A[i] * kA2B;
  const double det = c[0] * (c[4] * c[8] - c[5] * c[7]) - c[1] * (c[3] * c[8] - c[5] * c[6]) +
                     c[2] * (c[3] * c[7] - c[4] * c[6]);
  if (!(fabs(det) > 1e-12) || !std::isfinite(det)) {
    set_error("cell is singular or not finite (det = %g)", det);
    return FD3_ERR_ARG;
  }
  if (!(r_cut_A > 0.0)) {
    set_error("r_cut must be positive, got %g", r_cut_A);
    return FD3_ERR_ARG;
  }
  double inv[9];
  inv[0] = (c[4] * c[8] - c[5] * c[7]) / det;
  inv[1] = (c[2] * c[7] - c[1] * c[8]) / det;
  inv[2] = (c[1] * c[5] - c[2] * c[4]) / det;
  inv[3] = (c[5] * c[6] - c[3] * c[8]) / det;
  inv[4] = (c[0] * c[8] - c[2] * c[6]) / det;
  inv[5] = (c[2] * c[3] - c[0] * c[5]) / det;
  inv[6] = (c[3] * c[7] - c[4] * c[6]) / det;
  inv[7] = (c[1] * c[6] - c[0] * c[7]) / det;
  inv[8] = (c[0] * c[4] - c[1] * c[3]) / det;
  const double rc = r_cut_A * kA2B;
  if (!(reach >= 1.0) || !std::isfinite(reach)) {
    set_error("cell-list reach factor must be >= 1, got %g", reach);
    return FD3_ERR_ARG;
  }
  const double rs = rc * reach;  // bins and search ranges cover this radius
  long long total = 1;
  for (int a = 0; a < 3; ++a) {
    // plane spacing normal to lattice vector a: 1 / |column a of cell^-1|
    const double nrm = sqrt(inv[0 * 3 + a] * inv[0 * 3 + a] + inv[1 * 3 + a] * inv[1 * 3 + a] +
                            inv[2 * 3 + a] * inv[2 * 3 + a]);
    const double d = 1.0 / nrm;
    int nb = (int)std::max(1.0, floor(d / rs));
    // fewer than three bins of rs along this axis: bins of at least rs / 2
    // (search reach 2), so an atom scans the images within ~2.5 rs instead
    // of three whole-box images per axis (a 50-atom box: ~230 candidates
    // per atom instead of ~1350)
    if (nb < 3) nb = (int)std::max(1.0, floor(2.0 * d / rs));
    nb = std::min(nb, 256);
    g->nb[a] = nb;
    g->sr[a] = (int)ceil(rs * nb / d - 1e-12);
    if (g->sr[a] < 1) g->sr[a] = 1;
    if (g->sr[a] > 64) {
      set_error("box is too thin for r_cut along axis %d (%d images)", a, g->sr[a]);
      return FD3_ERR_ARG;
    }
    total *= nb;
  }
  for (int i = 0; i < 9; ++i) {
    g->cell[i] = c[i];
    g->inv[i] = inv[i];
  }
  g->nbins = (int)total;
  g->rc = rc;
  g->rc2 = rc * rc;
  g->dyn = nullptr;
  return FD3_OK;
}

void cell_carve(const CellGrid& g, int64_t n, Arena& ar, CellBuffers* b) {
  b->posw = ar.take<double4>(n);
  b->flo = ar.take<int4>(n);
  b->abin = ar.take<int32_t>(n);
  b->bcount = ar.take<int32_t>(2 * (size_t)g.nbins);  // bcount and bcur: cleared by one fill
  b->bcur = offset_or_null(b->bcount, (size_t)g.nbins);
  b->bstart = ar.take<int32_t>(g.nbins + 1);
  b->slot_atom = ar.take<int32_t>(n);
  b->pos4 = ar.take<double4>(n);
  b->flo_s = ar.take<int4>(n);
  b->sbin = ar.take<int32_t>(n);
  b->spec_s = ar.take<int32_t>(n);
}

CellArgs cell_args(const CellGrid& g, const CellBuffers& b, int64_t n) {
  CellArgs c;
  for (int i = 0; i < 9; ++i) {
    c.cell[i] = g.cell[i];
    c.inv[i] = g.inv[i];
  }
  for (int a = 0; a < 3; ++a) {
    c.nb[a] = g.nb[a];
    c.sr[a] = g.sr[a];
  }
  c.rc2 = g.rc2;
  c.dyn = g.dyn;
  c.n = n;
  c.pos4 = b.pos4;
  c.sbin = b.sbin;
  c.bstart = b.bstart;
  c.spec_s = b.spec_s;
  c.slot_atom = b.slot_atom;
  c.flo_s = b.flo_s;
  return c;
}

int cell_build(const CellGrid& g, const CellBuffers& b, int64_t n, const void* pos, int dtype, const int32_t* species,
               hipStream_t s) {
  if (n <= 0) return FD3_OK;
  int rc;
  if ((rc = launch_fill_i32(b.bcount, 2 * (size_t)g.nbins, s))) return rc;  // and bcur
  const CellArgs c = cell_args(g, b, n);
  ProfScope _ps("k_celllist_build", s);
  const BinArgs ba{c, pos, dtype, n, b.posw, b.flo, b.abin, b.bcount};
  if ((rc = launch<BinArgs, bin_body, 256>("k_bin", dim3(cdiv(n, 256)), 0, s, ba))) return rc;
  const Scan1Args<int32_t> sa{b.bcount, (int64_t)g.nbins, b.bstart, nullptr};
  if ((rc = launch<Scan1Args<int32_t>, scan1_body<int32_t>, 1024>("k_scan1", dim3(1), 0, s, sa))) return rc;
  // scatter into sbin (free until fill_sorted), then sort each bin into slot_atom
  const ScatterArgs sc{b.abin, n, b.bstart, b.bcur, b.sbin};
  if ((rc = launch<ScatterArgs, scatter_body, 256>("k_scatter", dim3(cdiv(n, 256)), 0, s, sc))) return rc;
  const SortBinsArgs so{b.bstart, g.nbins, b.sbin, b.slot_atom};
  if ((rc = launch<SortBinsArgs, sort_bins_body, 256>("k_sort_bins", dim3(cdiv(g.nbins, 4)), 0, s, so))) return rc;
  const FillSortedArgs fs{b.slot_atom, n, b.posw, b.flo, b.abin, species, b.pos4, b.flo_s, b.sbin, b.spec_s};
  return launch<FillSortedArgs, fill_sorted_body, 256>("k_fill_sorted", dim3(cdiv(n, 256)), 0, s, fs);
}

int cn_blocks(int64_t n) { return (int)std::max<int64_t>(1, cdiv(n, 256)); }

int launch_cn_cells(const CellArgs& c, const CnArgs& a, int nbins, hipStream_t s) {
  if (c.n <= 0) return FD3_OK;
  ProfScope _ps("k_cn_cells", s);
  const CnBinsArgs ca{c, a};
  return launch<CnBinsArgs, cn_bins_body<false, kCnCapSmall>, 256, 4>("k_cn_cells", dim3(nbins), 0, s, ca);
}

int launch_cn_bwd_cells(const CellArgs& c, const CnArgs& a, int nbins, hipStream_t s) {
  ProfScope _ps("k_cn_bwd_cells", s);
  const CnBinsArgs ca{c, a};
  return launch<CnBinsArgs, cn_bins_body<true, kCnCapSmall>, 256, 4>("k_cn_bwd_cells", dim3(nbins), 0, s, ca);
}

int launch_cn_edges(const EdgeArgs& e, const CnArgs& a, const double4* pos4, const int32_t* spec, hipStream_t s) {
  if (e.n_edges <= 0) return FD3_OK;
  ProfScope _ps("k_cn_edges", s);
  hipLaunchKernelGGL(k_cn_edges, dim3(cdiv(e.n_edges, 256)), dim3(256), 0, s, e, a, pos4, spec);
  FD3_LAUNCH_CHECK("k_cn_edges");
  return FD3_OK;
}

int launch_cn_bwd_edges(const EdgeArgs& e, const CnArgs& a, const double4* pos4, const int32_t* spec, double* gshift,
                        hipStream_t s) {
  ProfScope _ps("k_cn_bwd_edges", s);
  hipLaunchKernelGGL(k_cn_bwd_edges, dim3(cn_blocks(e.n_edges)), dim3(256), 0, s, e, a, pos4, spec, gshift);
  FD3_LAUNCH_CHECK("k_cn_bwd_edges");
  return FD3_OK;
}

int nb_count(const CellArgs& c, int32_t* count_atom, int64_t* offset, int64_t* total, hipStream_t s) {
  if (c.n <= 0) {
    return FD3_OK;
  }
  hipLaunchKernelGGL(k_nb_count, dim3(cdiv(c.n, 256)), dim3(256), 0, s, c, count_atom);
  FD3_LAUNCH_CHECK("k_nb_count");
  const Scan1Args<int64_t> sa{count_atom, c.n, offset, total};
  return launch<Scan1Args<int64_t>, scan1_body<int64_t>, 1024>("k_scan1", dim3(1), 0, s, sa);
}

int nb_fill(const CellArgs& c, const int64_t* offset, int64_t capacity, int64_t* src, int64_t* dst, int32_t* sh,
            hipStream_t s) {
  if (c.n <= 0) return FD3_OK;
  hipLaunchKernelGGL(k_nb_fill, dim3(cdiv(c.n, 256)), dim3(256), 0, s, c, offset, capacity, src, dst, sh);
  FD3_LAUNCH_CHECK("k_nb_fill");
  hipLaunchKernelGGL(k_nb_sort, dim3(cdiv(c.n, 256)), dim3(256), 0, s, offset, c.n, capacity, dst, sh);
  FD3_LAUNCH_CHECK("k_nb_sort");
  return FD3_OK;
}

}  // namespace fd3
