// Real-space lattice sum of the BJ-damped pair energy (direct.hip).
#pragma once

#include "fd3_internal.h"

namespace fd3 {

// E = -1/2 sum_{i,j,n}' C6_ij [s6 / (r^6 + R0^6) + 3 s8 QQ_ij / (r^8 + R0^8)],
// r = |r_j + n.cell - r_i| < rmax, C6_ij = sum_q eigs_q c_iq c_jq, the prime
// dropping i == j at n = 0, plus the continuum tail beyond rmax (each atom
// sees the other species at the mean density 1/Omega).
struct DirectArgs {
  int64_t n;
  int nS, r;
  double cell[9];        // Bohr, rows = lattice vectors
  double volume;         // Bohr^3
  int nimg[3];           // lattice images -nimg..nimg along each axis
  double rmax;           // Bohr
  const double4* pos4;   // (n) slot order, Bohr
  const int32_t* spec_s; // (n)
  const float* charges;  // (n, r)
  const float* eigs;     // (r)
  const float* qq;       // (nS, nS)
  double s6, s8, a1, a2;
  double* tab;           // (nS * nS, 4): tail integrals I6, I8 and their R0 derivatives
  double* qsum;          // (nS, r): sum over the atoms of species X of c_iq
  double* rows;          // (n, 16): per atom E, dE/dstrain (xx yy zz xy xz yz), dE/d(s6 s8 a1 a2)
  double* gpos;          // (n, 3): dE/dr_i at fixed charges, Hartree/Bohr
  double* dEdq;          // (n, r): dE/dc_iq
  const double* cnv;     // (ncnv, 8): CN strain partials (xx yy zz xy xz yz), or null
  int ncnv;
  double* energy;        // (1) out
  double* virial;        // (3, 3) out, or null
  double* grad_params;   // (4) out, or null
};

// Image range along each axis that covers rmax for wrapped positions:
// ceil(rmax / d_a) + 1 with d_a the plane spacing of the cell.
void direct_images(const double cell_bohr[9], const double inv[9], double rmax, int nimg[3]);
int launch_direct(const DirectArgs& a, hipStream_t s);           // tables and pair sums
int launch_direct_finalize(const DirectArgs& a, hipStream_t s);  // energy, virial, parameter gradient

}  // namespace fd3
