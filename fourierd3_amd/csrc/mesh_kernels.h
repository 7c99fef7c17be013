// Argument blocks and launchers of the per-atom mesh kernels (mesh.hip).
#pragma once

#include <algorithm>

#include "fd3_internal.h"

namespace fd3 {

struct MeshGeom {
  double inv[9];  // cell^-1 (Bohr^-1)
  int nx, ny, nz, order, interp;
};

inline MeshGeom mesh_geom(const Geom& g) {
  MeshGeom m;
  for (int i = 0; i < 9; ++i) m.inv[i] = g.inv[i];
  m.nx = g.nx; m.ny = g.ny; m.nz = g.nz; m.order = g.order; m.interp = g.interp;
  return m;
}

struct C6Args {
  int64_t n;
  int nS, r;
  const double* cn_raw;    // (n) slot order
  const int32_t* spec_s;   // (n)
  const int32_t* slot_atom;  // may be null (identity)
  const float* cnref;      // (nS, 7)
  const float* vq;         // (nS, 7, r)
  const float* eigs;       // (r)
  const float* qq;         // (nS, nS)
  const float* cncorr;     // (nS, 2) or null
  double s6, s8, a1, a2;
  float* charges;          // (n, r) out
  float* cn_out;           // (n) caller order, or null
  double* partials;        // (blocks, 8)
};

struct SpreadArgs {
  MeshGeom g;
  int64_t n;
  int r;
  const double4* pos4;   // Bohr, slot order; if null, raw_pos is read instead
  const void* raw_pos;   // (n, 3) Angstrom, dtype raw_dtype
  int raw_dtype;
  const int32_t* spec_s;
  const float* charges;
  float* rho;  // (nS*r, nx, ny, nz), zeroed
};

struct GatherArgs {
  MeshGeom g;
  int64_t n;
  int nS, r;
  const double4* pos4;   // as in SpreadArgs
  const void* raw_pos;
  int raw_dtype;
  const int32_t* spec_s;
  const float* mesh;   // (nS*r, nx, ny, nz)
};

int launch_fill_f32(float* p, size_t n, hipStream_t s);
int launch_fill_f64(double* p, size_t n, hipStream_t s);
int launch_fill_i32(int32_t* p, size_t n, hipStream_t s);
int launch_fill_i64(int64_t* p, size_t n, hipStream_t s);
int launch_prep_pos(const void* pos, int dtype, const int32_t* species, int64_t n, double4* pos4, int32_t* spec_s,
                    hipStream_t s);
int c6_blocks(int64_t n);
int gather_blocks(int64_t n, int order);
int launch_c6_fwd(const C6Args& a, hipStream_t s);
int launch_spread(const SpreadArgs& a, hipStream_t s);
// double -> float copy (n values)
int launch_to_f32(const double* src, float* dst, int64_t n, hipStream_t s);

// The softmax weights L_p = softmax_p(-4 (CN - ref_p)^2) over the valid
// references of species X (src/fourierd3/core.py:469-478), in double;
// returns 1 / (sum of the unnormalised weights).
__device__ inline double c6_softmax(const float* __restrict__ cnref, int X, double cn, double* L) {
  double mx = -INFINITY;
  for (int p = 0; p < kNRef; ++p) {
    const float ref = cnref[X * kNRef + p];
    const double d = cn - (double)ref;
    L[p] = ref == -1.0f ? -INFINITY : -4.0 * d * d;
    mx = fmax(mx, L[p]);
  }
  double sum = 0.0;
  for (int p = 0; p < kNRef; ++p) {
    L[p] = L[p] == -INFINITY ? 0.0 : exp(L[p] - mx);
    sum += L[p];
  }
  const double inv = 1.0 / sum;
  for (int p = 0; p < kNRef; ++p) L[p] *= inv;
  return inv;
}

int launch_gather_plain(const GatherArgs& a, float* out, hipStream_t s);
int launch_unpermute(const double* gpos, const int32_t* slot_atom, int64_t n, double* out, hipStream_t s);

}  // namespace fd3
