// Spreading fused into the forward z pass and gathering fused into the inverse
// z pass (fused.hip).
#pragma once

#include "fd3_internal.h"
#include "fft_plan.h"

namespace fd3 {

// Atoms sorted into mesh columns: key = (species, x node 0, y node 0).
struct ColBuffers {
  int32_t* key;       // (n) slot order
  int4* tmp_i;        // (n) slot order: node-0 indices (wrapped) and slot
  float4* tmp_f;      // (n) slot order: per-axis offsets x of the stencil
  int32_t* colcount;  // (nS * nx * ny)
  int32_t* colcur;    // (nS * nx * ny)
  int32_t* colstart;  // (nS * nx * ny + 1)
  int32_t* scantmp;   // scratch of the exclusive scan
  int4* rec_i;        // (n) column order
  float4* rec_f;      // (n) column order
  float* accA;        // (n, r) slot order: sum_nodes W dE/drho
  float* accG;        // (n, 3) slot order: sum_q c_q sum_nodes dW/du dE/drho
  float4* part;       // (order, 2, r, n), n in column order: the gather's per-block sums (accA_q, accG_q)
};

void col_carve(const FftPlan& p, int64_t n, int order, Arena& ar, ColBuffers* b);

struct FusedArgs {
  int64_t n;
  int nS, r;
  int nx, ny, nz, order, interp;
  double inv[9];
  const CellDyn* dyn = nullptr;  // when set, inv is read from it
  const double4* pos4;
  const int32_t* spec_s;
  const float* charges;  // (n, r) slot order
  ColBuffers cb;
  int det = 0;  // FD3_DETERMINISTIC: columns ordered by slot, spread in that order
};

// sort atoms into mesh columns (positions and species only: the z passes
// read each atom's charges through its slot index, so this may run before
// or beside the CN and C6 kernels)
int col_sort(const FusedArgs& a, hipStream_t s);
// spreading + forward z pass: A[c][x][kz][y]
int spread_r2c_z(const FusedArgs& a, const FftPlan& p, const FftTables& t, float2* A, hipStream_t s);
// inverse z pass + gathering into cb.accA / cb.accG
int c2r_z_gather(const FusedArgs& a, const FftPlan& p, const FftTables& t, const float2* A, hipStream_t s);

struct C6BwdArgs {
  int64_t n;
  int nS, r, nx, ny, nz;
  double inv[9];
  const CellDyn* dyn = nullptr;  // when set, inv is read from it
  const double4* pos4;
  const int32_t* spec_s;
  const float* charges;
  const float* accA;
  const float* accG;
  const float2* k0;  // (nS * r) or NULL: k = 0 part of dE/dc, added to accA
  const double* cn_raw;
  const float* cnref;
  const float* vq;
  const float* eigs;
  const float* qq;
  const float* cncorr;
  double s6, s8, a1, a2;
  double* gpos;      // (n, 3) out: mesh part of dE/dr, Hartree/Bohr
  double* dEdcn;     // (n) out
  double* partials;  // (c6_blocks(n), 16): sum_i f_i (x) dE/dr_i
  // When set, (n, r) dE/dc of the real-space sum: replaces accA, the self
  // term and k0; gpos and partials are then left alone (accG unused).
  const double* dEdq = nullptr;
};
int launch_c6_bwd(const C6BwdArgs& a, hipStream_t s);

}  // namespace fd3
