// Argument blocks and launchers of the cell-list / CN kernels (neigh.hip).
#pragma once

#include <algorithm>
#include <cmath>

#include "fd3_internal.h"

namespace fd3 {

struct CellGrid {
  double cell[9], inv[9];  // Bohr
  int nb[3], sr[3];        // bins and search reach per lattice direction
  int nbins;
  double rc, rc2;          // cutoff (Bohr)
  const CellDyn* dyn = nullptr;  // device cell or nullptr
};

struct CellBuffers {
  double4* posw;   // (n) caller order, wrapped positions (Bohr)
  int4* flo;       // (n) caller order, floor of fractional coordinates
  int32_t* abin;   // (n) caller order, bin of each atom
  int32_t* bcount; // (nbins)
  int32_t* bcur;   // (nbins)
  int32_t* bstart; // (nbins + 1)
  int32_t* slot_atom;  // (n) atom of each slot
  double4* pos4;   // (n) slot order, wrapped positions
  int4* flo_s;     // (n) slot order
  int32_t* sbin;   // (n) slot order
  int32_t* spec_s; // (n) slot order
};

struct CellArgs {
  double cell[9], inv[9];
  const CellDyn* dyn = nullptr;  // when set, cell and inv are read from it
  int nb[3], sr[3];
  double rc2;
  int64_t n;
  const double4* pos4;
  const int32_t* sbin;
  const int32_t* bstart;
  const int32_t* spec_s;
  const int32_t* slot_atom;
  const int4* flo_s;
};

struct CnArgs {
  const float* rcov;   // (nS) Bohr
  double rcut;         // Bohr
  double* cn;          // (n) out (fwd)
  const double* dEdcn; // (n) in (bwd)
  double* gpos;       // (n, 3) in/out (bwd), Hartree/Bohr
  double* partials;   // (blocks, 8) virial (bwd)
  double* gout;       // (n, 3) or NULL (cell kernels, bwd): the final dE/dr in atom order, Hartree/Angstrom
  const int32_t* slot_atom;  // slot -> atom for gout
};

struct EdgeArgs {
  int64_t n_edges;
  const int64_t* src;
  const int64_t* dst;
  const void* shifts;  // (E, 3) Angstrom
  int shifts_dtype;
};

// Bins of plane spacing >= reach * r_cut and search ranges that cover
// reach * r_cut: a cell whose plane spacings shrink by up to the factor
// reach still has every pair within r_cut in range.
int cell_grid_make(const double cell_A[9], double r_cut_A, CellGrid* g, double reach = 1.0);
void cell_carve(const CellGrid& g, int64_t n, Arena& ar, CellBuffers* b);
CellArgs cell_args(const CellGrid& g, const CellBuffers& b, int64_t n);
int cell_build(const CellGrid& g, const CellBuffers& b, int64_t n, const void* pos, int dtype, const int32_t* species,
               hipStream_t s);
int cn_blocks(int64_t n);
// one block per bin; the backward writes nbins rows of virial partials
int launch_cn_cells(const CellArgs& c, const CnArgs& a, int nbins, hipStream_t s);
int launch_cn_bwd_cells(const CellArgs& c, const CnArgs& a, int nbins, hipStream_t s);
int scan_exclusive(const int32_t* in, int64_t n, int32_t* out, int32_t* tmp, int64_t* total, hipStream_t s);
size_t scan_tmp_elems(int64_t n);
int launch_cn_edges(const EdgeArgs& e, const CnArgs& a, const double4* pos4, const int32_t* spec, hipStream_t s);
int launch_cn_bwd_edges(const EdgeArgs& e, const CnArgs& a, const double4* pos4, const int32_t* spec, double* gshift,
                        hipStream_t s);
// pair offsets are int64: the pair total of a large box may pass 2^31
int nb_count(const CellArgs& c, int32_t* count_atom, int64_t* offset, int64_t* total, hipStream_t s);
int nb_fill(const CellArgs& c, const int64_t* offset, int64_t capacity, int64_t* src, int64_t* dst, int32_t* sh,
            hipStream_t s);

// implemented in mesh.hip
int launch_fill_i32(int32_t* p, size_t n, hipStream_t s);
int launch_fill_f32(float* p, size_t n, hipStream_t s);

}  // namespace fd3
