/*
 * fd3.h - C ABI of libfd3hip.so, the MI355X (gfx950) implementation of
 * FourierD3's D3(BJ) energy / force / stress path.
 *
 * Conventions
 *  - Every pointer marked [device] is device memory; everything else is host
 *    memory read during the call only.
 *  - Nothing here allocates device memory or synchronises the host: the caller
 *    passes a workspace of the size the matching *_workspace function reports,
 *    and every kernel is enqueued on `stream` (a hipStream_t; NULL = default
 *    stream), so the calls can be captured in a HIP graph.
 *  - Units at the boundary follow the reference's FourierD3.forward
 *    (src/fourierd3/core.py:480-503): positions, shifts and cells in Angstrom,
 *    energies in Hartree, gradients in Hartree/Angstrom.  Internally the
 *    kernels work in Bohr (1 A = 1/0.52917726 Bohr, src/fourierd3/core.py:120).
 *  - Return value: FD3_OK (0) or a negative FD3_ERR_* code; fd3_last_error()
 *    returns a message for the calling thread.
 */
#ifndef FD3_H
#define FD3_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define FD3_OK 0
#define FD3_ERR_ARG (-1)         /* bad shape / value  -> ValueError          */
#define FD3_ERR_UNSUPPORTED (-2) /* option not built   -> NotImplementedError */
#define FD3_ERR_HIP (-3)         /* HIP runtime error  -> RuntimeError        */
#define FD3_ERR_WORKSPACE (-4)   /* workspace too small                       */

#define FD3_INTERP_P3M 0      /* method='spme': centred cardinal B-splines, sinc^2p division */
#define FD3_INTERP_LAGRANGE 1 /* method='pme' : Lagrange interpolation, no sinc division      */
#define FD3_INTERP_EWALD 2    /* method='ewald': structure-factor sum over the full k grid
                                 m_d in fftfreq(ns_d) * ns_d (src/fourierd3/core.py:225-242,
                                 563-577); ns 1..4096 per axis, rank <= 32, nS * rank <= 64 */

#define FD3_DTYPE_F32 0
#define FD3_DTYPE_F64 1

/* Flags of fd3_eval_* */
#define FD3_NEED_GRAD 1    /* compute forces / virial / cell and shift gradients */
#define FD3_NEED_PARAMS 2  /* compute dE/d(s6, s8, a1, a2)                        */
/* Deterministic order (no reference counterpart: the reference's index_add_
 * spread, src/fourierd3/mesh_interpolator_d3.py:128-133, is order-dependent
 * on the GPU).  The atoms of each mesh column are ordered by index and the
 * spread adds every atom's stencil in that order with plain read-modify-write
 * steps instead of concurrent compare-and-swap adds, so repeated calls on the
 * same inputs return bit-identical results.  Slower spread; a debugging and
 * reproducibility mode. */
#define FD3_DETERMINISTIC 4

/* Per-species model constants.  All arrays are float32 [device].
 * Mirrors the buffers FourierD3.__init__ registers (src/fourierd3/core.py:139-182)
 * and D3Potential's R0 / Q tables (src/fourierd3/pair_pot.py:49-89). */
typedef struct fd3_model {
  int32_t n_species; /* nS, 1..64                                         */
  int32_t n_rank;    /* r, 1..32                                          */
  const float* rcov;     /* (nS)      covalent radii, Bohr                */
  const float* cnref;    /* (nS, 7)   reference CNs, -1 = unused          */
  const float* vq;       /* (nS, 7, r) eigenvector rows V[Z, p, r]        */
  const float* eigs;     /* (r)       eigenvalues lambda_r                */
  const float* qq;       /* (nS, nS)  sqrt(Q_X) sqrt(Q_Y)                 */
  const float* cncorr;   /* (nS, 2)   [b, a] of CN' = a CN + b, or NULL   */
  double s6, s8, a1, a2; /* BJ parameters (src/fourierd3/core.py:165-170) */
  double r_cut;          /* CN cutoff, Angstrom                           */
} fd3_model;

/* Geometry of one periodic box. */
typedef struct fd3_geom {
  double cell[9];  /* Angstrom, row i = lattice vector a_i                   */
  int32_t ns[3];   /* mesh points per axis: powers of two, 2..1024 (Ewald: k grid) */
  int32_t order;   /* interpolation nodes p: 1..5 (P3M), 2..7 (Lagrange)      */
  int32_t interp;  /* FD3_INTERP_*                                            */
} fd3_geom;

/* Outputs of fd3_eval_*.  Any pointer may be NULL when not wanted. [device] */
typedef struct fd3_outputs {
  double* energy;      /* (1)    E in Hartree                                      */
  double* grad_pos;    /* (N,3)  dE/dpositions, Hartree/A                          */
  double* grad_shifts; /* (E,3)  dE/dshifts, Hartree/A (edge-list entry only)      */
  double* grad_cell;   /* (3,3)  dE/dcell at fixed positions and shifts, Hartree/A  */
  double* virial;      /* (3,3)  dE/dstrain for cell' = cell (1+eps)^T, r' = r (1+eps)^T, Hartree */
  double* grad_params; /* (4)    dE/d(s6, s8, a1, a2), Hartree                     */
  float* cn;           /* (N)    coordination numbers                              */
} fd3_outputs;

const char* fd3_last_error(void);
int fd3_version(void);

/* Per-kernel timing (no counterpart in the reference; used by the benchmark
 * to measure kernel durations on the launch stream).  fd3_profile_enable(1)
 * starts recording a HIP event pair around every kernel launch (do not enable
 * it during graph capture); fd3_profile_report synchronises on the recorded
 * events, writes up to max_entries kernel names (64 chars each into `names`),
 * total milliseconds and launch counts, and clears the records.
 * fd3_profile_only(name) limits the recording to launches of that kernel
 * (NULL or "" records every kernel again), so that a timed loop carries one
 * event pair per launch of the kernel of interest and nothing else. */
int fd3_profile_enable(int on);
int fd3_profile_only(const char* kernel);
int fd3_profile_report(int max_entries, char* names, double* total_ms, int64_t* counts, int* n_entries);

/* ---- whole-path entry points ------------------------------------------------
 * fd3_eval_edges: FourierD3.forward(positions, edge_index, shifts) with
 *   cnfunc='smooth_cut' and method='spme'/'pme' (src/fourierd3/core.py:480-583):
 *   energy and, with FD3_NEED_GRAD, its exact gradients with respect to
 *   positions, shifts and cell (the reference obtains them by autograd).
 *   edge_src/edge_dst are int64 [device] (rows of edge_index); shifts (E,3)
 *   are lattice translations in Angstrom [device], dtype FD3_DTYPE_*.
 * fd3_eval_celllist: FourierD3ASECalculator.calculate (src/fourierd3/calculator.py:125-192)
 *   without a materialised neighbour list: the directed pairs within r_cut
 *   (matscipy neighbour_list "ijS" semantics, src/fourierd3/calculator.py:94-123)
 *   are enumerated on the fly from a cell list.  grad_shifts and grad_cell are
 *   not produced; virial is.
 * species: int32 [device] compact species ids 0..nS-1 (ascending atomic number,
 *   src/fourierd3/core.py:104-114).  positions: (N,3) Angstrom [device].
 */
int fd3_eval_edges_workspace(const fd3_model* m, const fd3_geom* g, int64_t n_atoms, int64_t n_edges,
                             int flags, size_t* bytes);
int fd3_eval_edges(const fd3_model* m, const fd3_geom* g, int64_t n_atoms, const void* positions,
                   int pos_dtype, const int32_t* species, int64_t n_edges, const int64_t* edge_src,
                   const int64_t* edge_dst, const void* shifts, int shifts_dtype, int flags,
                   void* workspace, size_t workspace_bytes, const fd3_outputs* out, void* stream);

/* fd3_eval_edges_npt: fd3_eval_edges with the cell read from device memory
 *   (FourierD3._update_cell then forward, src/fourierd3/core.py:244-267,
 *   480-583; the MLFF-in-NPT caller shape): cell (3,3) Angstrom float64
 *   [device], g->cell the reference cell for the |k| tables (reach as in
 *   fd3_eval_celllist_npt), status [device] or NULL (bits 1 and 4).  The
 *   shifts stay Cartesian, as the caller computes them from its current
 *   cell; grad_cell is dE/dcell at the device cell.  Mesh methods only. */
int fd3_eval_edges_npt_workspace(const fd3_model* m, const fd3_geom* g, double reach, int64_t n_atoms,
                                 int64_t n_edges, int flags, size_t* bytes);
int fd3_eval_edges_npt(const fd3_model* m, const fd3_geom* g, double reach, const double* cell,
                       int32_t* status, int64_t n_atoms, const void* positions, int pos_dtype,
                       const int32_t* species, int64_t n_edges, const int64_t* edge_src,
                       const int64_t* edge_dst, const void* shifts, int shifts_dtype, int flags,
                       void* workspace, size_t workspace_bytes, const fd3_outputs* out, void* stream);

int fd3_eval_celllist_workspace(const fd3_model* m, const fd3_geom* g, int64_t n_atoms, int flags,
                                size_t* bytes);
int fd3_eval_celllist(const fd3_model* m, const fd3_geom* g, int64_t n_atoms, const void* positions,
                      int pos_dtype, const int32_t* species, int flags, void* workspace,
                      size_t workspace_bytes, const fd3_outputs* out, void* stream);

/* fd3_eval_celllist_npt: fd3_eval_celllist for NPT steps, after
 *   FourierD3._update_cell (src/fourierd3/core.py:244-267), with the cell read
 *   from device memory on the stream, so that no host round trip sits between
 *   a barostat's cell update and the evaluation and a captured graph replays
 *   with the current cell.  cell: (3,3) Angstrom, row i = a_i, float64
 *   [device].  As in the reference the mesh (g->ns) stays fixed; g->cell is a
 *   reference cell from which the cell list and the |k| tables are planned to
 *   cover cells whose plane spacings shrink, and largest |k| grows, by up to
 *   the factor reach (>= 1).  status: int32 [device] or NULL, written on the
 *   stream: 0, or 1 (cell singular or not finite) | 2 (plane spacing below
 *   the planned reach) | 4 (|k| beyond the tables); when nonzero the energy,
 *   virial, grad_params and grad_pos (and, for the edge-list form,
 *   grad_shifts) are NaN.  Mesh methods only (not Ewald). */
int fd3_eval_celllist_npt_workspace(const fd3_model* m, const fd3_geom* g, double reach, int64_t n_atoms,
                                    int flags, size_t* bytes);
int fd3_eval_celllist_npt(const fd3_model* m, const fd3_geom* g, double reach, const double* cell,
                          int32_t* status, int64_t n_atoms, const void* positions, int pos_dtype,
                          const int32_t* species, int flags, void* workspace, size_t workspace_bytes,
                          const fd3_outputs* out, void* stream);

/* fd3_eval_direct: real-space lattice sum of the same pair energy (no
 *   counterpart in the reference, whose energy is reciprocal-space plus the
 *   self term only, src/fourierd3/core.py:553-581).  For checking the mesh
 *   methods and for small boxes:
 *     E = -1/2 sum_{i,j,n}' C6_ij [s6/(r^6+R0^6) + 3 s8 QQ_ij/(r^8+R0^8)]
 *   over lattice images with r = |r_j + n.cell - r_i| < r_max (Angstrom),
 *   C6_ij = sum_q eigs_q c_iq c_jq from the cell-list CN (as
 *   fd3_eval_celllist), R0_ij = a1 sqrt(3 QQ_ij) + a2, plus the continuum
 *   tail beyond r_max at the mean density 1/Omega.  cell: (3,3) Angstrom,
 *   rows = lattice vectors.  Energy; with FD3_NEED_GRAD grad_pos and virial;
 *   with FD3_NEED_PARAMS grad_params; cn if set (grad_shifts and grad_cell
 *   are not produced).  The image cube, 2 ceil(r_max / d_a) + 3 per axis
 *   with d_a the plane spacings, is limited to 2^22 images.  Double
 *   precision throughout; the cost grows as N^2 times the images. */
int fd3_eval_direct_workspace(const fd3_model* m, const double cell[9], double r_max, int64_t n_atoms, int flags,
                              size_t* bytes);
int fd3_eval_direct(const fd3_model* m, const double cell[9], double r_max, int64_t n_atoms, const void* positions,
                    int pos_dtype, const int32_t* species, int flags, void* workspace, size_t workspace_bytes,
                    const fd3_outputs* out, void* stream);

/* fd3_eval_d4: FourierD3ASECalculator.calculate / FourierD3.forward(positions)
 *   with cnfunc='d4' and method='spme'/'pme' (src/fourierd3/core.py:389-447,
 *   480-583; src/fourierd3/calculator.py:169-171): the coordination numbers
 *   come from a k-space convolution of the species densities
 *   (src/fourierd3/kspace_filter_d3.py:101-158) with the transform of the D4
 *   pair function (src/fourierd3/pair_pot.py:235-402) on the same mesh, so no
 *   neighbour list is built.  Energy, and with FD3_NEED_GRAD grad_pos and
 *   virial (grad_shifts and grad_cell are not produced); cn receives the D4
 *   CN.  With method='ewald' (geom.interp = FD3_INTERP_EWALD) the CN comes
 *   from structure factors on the Ewald k grid (src/fourierd3/core.py:430-441).
 */
typedef struct {
  const float* r0;        /* (nS, nS) r_cov_X + r_cov_Y of the D4 CN radii, Bohr [device] */
  const float* delta;     /* (nS, nS) k1 exp(-(|EN_X - EN_Y| + k2)^2 / k3) [device] */
  const float* selfcont;  /* (nS) f_XX(0) = delta_XX / 2 erfc(-k0) [device] */
  double k0;              /* erfc steepness (7.5 in the reference) */
} fd3_cn_d4;

int fd3_eval_d4_workspace(const fd3_model* m, const fd3_cn_d4* d4, const fd3_geom* g, int64_t n_atoms,
                          int flags, size_t* bytes);
int fd3_eval_d4(const fd3_model* m, const fd3_cn_d4* d4, const fd3_geom* g, int64_t n_atoms,
                const void* positions, int pos_dtype, const int32_t* species, int flags, void* workspace,
                size_t workspace_bytes, const fd3_outputs* out, void* stream);

/* fd3_eval_d4_npt: fd3_eval_d4 with the cell read from device memory, as
 *   fd3_eval_celllist_npt (src/fourierd3/core.py:244-267, which also updates
 *   kspace_filter_cn): cell (3,3) Angstrom float64 [device], g->cell the
 *   reference cell from which the G and G^CN tables are planned for |k| up to
 *   reach times its largest; status [device] or NULL as there (bit 4: |k|
 *   beyond the tables).  Mesh methods only. */
int fd3_eval_d4_npt_workspace(const fd3_model* m, const fd3_cn_d4* d4, const fd3_geom* g, double reach,
                              int64_t n_atoms, int flags, size_t* bytes);
int fd3_eval_d4_npt(const fd3_model* m, const fd3_cn_d4* d4, const fd3_geom* g, double reach, const double* cell,
                    int32_t* status, int64_t n_atoms, const void* positions, int pos_dtype,
                    const int32_t* species, int flags, void* workspace, size_t workspace_bytes,
                    const fd3_outputs* out, void* stream);

/* fd3_eval_celllist_batch: the serial form of a batch, n_boxes independent
 *   periodic boxes that share one model (one species set), one
 *   fd3_eval_celllist call after another (fd3_batch_* merges them instead): box b owns atoms [atom_offsets[b], atom_offsets[b+1])
 *   of the packed positions (N_total, 3) and species (N_total) [device], has its
 *   own geometry geoms[b] (its own cell and mesh) and writes outs[b].  Same
 *   results as n_boxes fd3_eval_celllist calls; atom_offsets (n_boxes + 1),
 *   geoms and outs are host arrays.  The boxes run in order on `stream` and
 *   share one workspace of the size fd3_eval_celllist_batch_workspace reports
 *   (the largest single-box workspace).  No counterpart in the reference (it
 *   evaluates one box per call); this is the batched form of
 *   src/fourierd3/calculator.py:125-192 for many small cells.
 */
int fd3_eval_celllist_batch_workspace(const fd3_model* m, int32_t n_boxes, const fd3_geom* geoms,
                                      const int64_t* atom_offsets, int flags, size_t* bytes);
int fd3_eval_celllist_batch(const fd3_model* m, int32_t n_boxes, const fd3_geom* geoms, const int64_t* atom_offsets,
                            const void* positions, int pos_dtype, const int32_t* species, int flags, void* workspace,
                            size_t workspace_bytes, const fd3_outputs* outs, void* stream);

/* ---- batched evaluation (the batched form of src/fourierd3/calculator.py:125-192) ----
 * Many independent periodic boxes, of one or more species sets, evaluated by
 * one merged launch sequence: each stage of the cell-list pipeline is one
 * launch over the blocks of every box (instead of one launch per box and
 * stage), with each box's mesh, cell list and partial sums packed into one
 * workspace.  Box b uses models[model_of_box[b]] (so two species sets, each
 * with its own decomposition, share a batch), the geometry geoms[b] (its own
 * cell and mesh; mesh methods only) and atoms [atom_offsets[b],
 * atom_offsets[b+1]) of the packed positions (N_total, 3) and species
 * (N_total) [device]; species ids are compact within the box's model.
 * Results equal n_boxes fd3_eval_celllist calls.  No counterpart in the
 * reference, which evaluates one box per calculate() call.
 *
 * fd3_batch_create plans on the host (no device work; the model table
 *   pointers must stay valid for the plan's life).  The first run allocates
 *   the plan's pinned host copy of its launch tables (freed by destroy).
 * fd3_batch_workspace: bytes of device workspace a run needs.
 * fd3_batch_run: enqueues the copy of the launch tables into the workspace
 *   and the merged launches on `stream`; the tables are rebuilt on the host
 *   only when positions, species, workspace or outs change, so a run
 *   captured in a HIP graph replays as captured (keep one plan per graph).
 *   outs: n_boxes host structs of device pointers.
 * fd3_batch_info: merged launches per run and the launches n_boxes single-box
 *   calls would make.
 */
typedef struct fd3_batch fd3_batch;
int fd3_batch_create(int32_t n_models, const fd3_model* models, int32_t n_boxes, const int32_t* model_of_box,
                     const fd3_geom* geoms, const int64_t* atom_offsets, int flags, fd3_batch** plan);
int fd3_batch_workspace(const fd3_batch* plan, size_t* bytes);
int fd3_batch_run(fd3_batch* plan, const void* positions, int pos_dtype, const int32_t* species, void* workspace,
                  size_t workspace_bytes, const fd3_outputs* outs, void* stream);
int fd3_batch_info(const fd3_batch* plan, int32_t* merged_launches, int32_t* box_launches);
void fd3_batch_destroy(fd3_batch* plan);

/* ---- tile tuning (no counterpart in the reference) ----
 * Tile shapes of the mesh passes per (mesh ns, species, rank): tiles[0] = y
 * lines per z-pass block, tiles[1] = x lines per y-pass block, tiles[2] =
 * ranks per k-space tile, tiles[3] = k-space tiles per block; 0 keeps the
 * plan's own choice, and a value outside a kernel's limits is ignored.
 * fd3_tune_set records an override (the package loads the autotuner's
 * measured table at import), fd3_tune_clear drops them all, fd3_plan_tiles
 * reports the tiles a plan would use.  Results do not depend on the tiles. */
int fd3_tune_set(const int32_t ns[3], int32_t n_species, int32_t n_rank, const int32_t tiles[4]);
int fd3_tune_clear(void);
int fd3_plan_tiles(const int32_t ns[3], int32_t n_species, int32_t n_rank, int32_t tiles[4]);

/* fd3_eval_celllist_noalloc: fd3_eval_celllist with no workspace argument
 *   and no allocation of any kind.  No counterpart in the reference: it is for
 *   callers that cannot allocate device memory (code under HIP graph capture,
 *   tight serving loops, harnesses that forbid hipMalloc).  The workspace is a
 *   static device array of FD3_NOALLOC_SCRATCH_BYTES (256 MiB) that every GPU
 *   the library runs on holds until the process ends; a box whose
 *   fd3_eval_celllist_workspace exceeds it is rejected with
 *   FD3_ERR_WORKSPACE before any device work (e.g. the 100k-atom 256^3 box
 *   needs ~2.3 GB: use fd3_eval_celllist).  The current device must be the
 *   stream's device.  All calls on one GPU share its array: calls that
 *   overlap on one GPU (two streams, two threads) race and must be
 *   serialised by the caller. */
#define FD3_NOALLOC_SCRATCH_BYTES ((size_t)256 << 20)
int fd3_eval_celllist_noalloc(const fd3_model* m, const fd3_geom* g, int64_t n_atoms, const void* positions,
                              int pos_dtype, const int32_t* species, int flags, const fd3_outputs* out,
                              void* stream);

/* ---- component entry points (tests and callers that assemble their own path) ----
 * Neighbour list (matscipy "ijS", src/fourierd3/calculator.py:103-109):
 *   fd3_neighbours_count writes the number of directed pairs to *n_pairs [device int64];
 *   fd3_neighbours_fill writes up to `capacity` pairs, sorted by (i, j, S),
 *   as src/dst int64 and integer shifts S int32 (E,3).
 */
int fd3_neighbours_workspace(int64_t n_atoms, const double cell[9], double r_cut, size_t* bytes);
int fd3_neighbours_count(int64_t n_atoms, const void* positions, int pos_dtype, const double cell[9],
                         double r_cut, void* workspace, size_t workspace_bytes, int64_t* n_pairs,
                         void* stream);
int fd3_neighbours_fill(int64_t n_atoms, const void* positions, int pos_dtype, const double cell[9],
                        double r_cut, void* workspace, size_t workspace_bytes, int64_t capacity,
                        int64_t* src, int64_t* dst, int32_t* shifts, void* stream);

/* Spread (MeshInterpolatorD3.points_to_mesh, src/fourierd3/mesh_interpolator_d3.py:74-138):
 * rho[c, x, y, z] (c = species*r + rank, float32 [device], zeroed here) from
 * per-atom charges (N, r) float32 placed in the channel block of the atom's species. */
int fd3_spread(const fd3_geom* g, int32_t n_species, int32_t n_rank, int64_t n_atoms, const void* positions,
               int pos_dtype, const int32_t* species, const float* charges, float* rho, void* stream);

/* Gather (MeshInterpolatorD3.mesh_to_points, src/fourierd3/mesh_interpolator_d3.py:141-189):
 * out[i, q] = sum_nodes W * mesh[species_i * r + q, node]   (N, r) float32 [device]. */
int fd3_gather(const fd3_geom* g, int32_t n_species, int32_t n_rank, int64_t n_atoms, const void* positions,
               int pos_dtype, const int32_t* species, const float* mesh, float* out, void* stream);

/* 3-D real FFTs of C channels (torch.fft.rfftn / irfftn over the last three
 * dims, src/fourierd3/kspace_filter_d3.py:55), unnormalised in both
 * directions.  Spectrum layout: spec[c, kz, ky, kx] complex64 (kz = 0..nz/2),
 * i.e. kx fastest.  fd3_fft_workspace gives the scratch size. */
int fd3_fft_workspace(const int32_t ns[3], int32_t n_channels, size_t* bytes);
int fd3_fft_r2c(const int32_t ns[3], int32_t n_channels, const float* rho, float* spec, void* workspace,
                size_t workspace_bytes, void* stream);
int fd3_fft_c2r(const int32_t ns[3], int32_t n_channels, const float* spec, float* out, void* workspace,
                size_t workspace_bytes, void* stream);

/* Green's function G_XY(|k|) (D3Potential.lr_from_k_sq, src/fourierd3/pair_pot.py:142-178)
 * on n host values of |k| (Bohr^-1) for one species pair, host-side evaluation of
 * the same code the kernels run (float32); also dG/d|k|. */
int fd3_green_host(double s6, double s8, double r0, double qq, int64_t n, const double* k, float* g,
                   float* dg_dk);
/* The same G_XY(|k|) and dG/d|k| from the per-pair table over [0, kmax] that
 * the k-space kernel uses when no parameter gradients are asked for (host
 * evaluation of the kernel's code; |k| <= kmax). */
int fd3_green_table_host(double s6, double s8, double r0, double qq, double kmax, int64_t n, const double* k,
                         float* g, float* dg_dk);

/* 1-D interpolation weights (torch-pme MeshInterpolator._compute_1d_weights as
 * used at src/fourierd3/core.py:555), host-side evaluation of the kernel code:
 * for offsets x (n), w[j*n + i] and dw[j*n + i], j = 0..order-1. */
int fd3_interp_weights_host(int32_t order, int32_t interp, int64_t n, const double* x, float* w, float* dw);

#ifdef __cplusplus
}
#endif
#endif /* FD3_H */
