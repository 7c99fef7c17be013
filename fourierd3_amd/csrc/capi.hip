// C ABI of libfd3hip.so: validation, workspace layout, the two whole-path
// pipelines and the component entry points.  See include/fd3.h.

#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>
#include <utility>
#include <vector>

#include "fd3_internal.h"
#include "cn_d4.h"
#include "ewald.h"
#include "fft_plan.h"
#include "fused.h"
#include "direct.h"
#include "launch.h"
#include "mesh_kernels.h"
#include "neigh_kernels.h"

namespace fd3 {

static thread_local std::string g_err;

void set_error(const char* fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof(buf), fmt, ap);
  va_end(ap);
  g_err = buf;
}

int hip_fail(hipError_t e, const char* what) {
  set_error("%s: %s", what, hipGetErrorString(e));
  return FD3_ERR_HIP;
}

// ------------------------------------------------------------- profiling ---
struct ProfRec {
  const char* name;
  hipEvent_t start, stop;
};
static bool g_prof = false;
static std::string g_prof_only;  // record only this kernel name when non-empty
static std::vector<ProfRec> g_prof_recs;
static std::vector<std::pair<const char*, hipEvent_t>> g_prof_open;

// ---------------------------------------------------- launch recording ---
Recorder*& current_recorder() {
  static thread_local Recorder* r = nullptr;
  return r;
}

int lds_attr(const void* fn, size_t sh) {
  if (sh <= 64 * 1024) return FD3_OK;
  static std::mutex mu;
  static std::vector<std::pair<const void*, size_t>> done;
  std::lock_guard<std::mutex> lock(mu);
  for (auto& d : done)
    if (d.first == fn && d.second >= sh) return FD3_OK;
  FD3_HIP(hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)sh));
  done.push_back({fn, sh});
  return FD3_OK;
}

// no events while launches are only being recorded
bool prof_enabled() { return g_prof && !current_recorder(); }

void prof_record(const char* name, hipStream_t s, bool start) {
  if (!g_prof_only.empty() && g_prof_only != name) return;
  hipEvent_t ev;
  if (hipEventCreate(&ev) != hipSuccess) return;
  if (hipEventRecord(ev, s) != hipSuccess) {
    (void)hipEventDestroy(ev);
    return;
  }
  if (start) {
    g_prof_open.push_back({name, ev});
  } else if (!g_prof_open.empty()) {
    g_prof_recs.push_back({g_prof_open.back().first, g_prof_open.back().second, ev});
    g_prof_open.pop_back();
  } else {
    (void)hipEventDestroy(ev);
  }
}

static void prof_clear() {
  for (auto& r : g_prof_recs) {
    (void)hipEventDestroy(r.start);
    (void)hipEventDestroy(r.stop);
  }
  for (auto& o : g_prof_open) (void)hipEventDestroy(o.second);
  g_prof_recs.clear();
  g_prof_open.clear();
}

void prof_set(bool on) {
  prof_clear();
  g_prof = on;
}

void prof_only(const char* name) { g_prof_only = name ? name : ""; }

int prof_report(int max_entries, char* names, double* total_ms, int64_t* counts, int* n_entries) {
  std::vector<std::string> keys;
  std::vector<double> ms;
  std::vector<int64_t> cnt;
  for (auto& r : g_prof_recs) {
    FD3_HIP(hipEventSynchronize(r.stop));
    float t = 0.f;
    FD3_HIP(hipEventElapsedTime(&t, r.start, r.stop));
    size_t k = 0;
    while (k < keys.size() && keys[k] != r.name) ++k;
    if (k == keys.size()) {
      keys.push_back(r.name);
      ms.push_back(0.0);
      cnt.push_back(0);
    }
    ms[k] += t;
    cnt[k] += 1;
  }
  const int n = (int)std::min<size_t>(keys.size(), (size_t)max_entries);
  for (int i = 0; i < n; ++i) {
    std::strncpy(names + 64 * i, keys[i].c_str(), 63);
    names[64 * i + 63] = 0;
    total_ms[i] = ms[i];
    counts[i] = cnt[i];
  }
  *n_entries = n;
  prof_clear();
  return FD3_OK;
}

// Cell part of a geometry: Bohr cell, inverse, reciprocal rows, volume.
static int make_cell(const double cell_A[9], Geom* out) {
  if (!cell_A) {
    set_error("cell is NULL");
    return FD3_ERR_ARG;
  }
  double c[9];
  for (int i = 0; i < 9; ++i) {
    c[i] = cell_A[i] * kA2B;
    if (!std::isfinite(c[i])) {
      set_error("cell entry %d is not finite", i);
      return FD3_ERR_ARG;
    }
  }
  const double det = c[0] * (c[4] * c[8] - c[5] * c[7]) - c[1] * (c[3] * c[8] - c[5] * c[6]) +
                     c[2] * (c[3] * c[7] - c[4] * c[6]);
  if (!(fabs(det) > 1e-12)) {
    set_error("cell is singular (det = %g Bohr^3)", det);
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
  for (int i = 0; i < 9; ++i) {
    out->cell[i] = c[i];
    out->inv[i] = inv[i];
  }
  for (int i = 0; i < 3; ++i)
    for (int d = 0; d < 3; ++d) out->recip[i * 3 + d] = 2.0 * kPi * inv[d * 3 + i];
  out->volume = fabs(det);
  out->dyn = nullptr;
  return FD3_OK;
}

int make_geom(const fd3_geom* g, Geom* out) {
  if (!g) {
    set_error("geometry is NULL");
    return FD3_ERR_ARG;
  }
  int rc;
  if ((rc = make_cell(g->cell, out))) return rc;
  out->nx = g->ns[0];
  out->ny = g->ns[1];
  out->nz = g->ns[2];
  out->order = g->order;
  out->interp = g->interp;
  if (g->interp == FD3_INTERP_EWALD) return FD3_OK;  // k grid sizes: checked by ewald_plan_make
  if (g->interp != FD3_INTERP_P3M && g->interp != FD3_INTERP_LAGRANGE) {
    set_error("unknown interpolation %d", g->interp);
    return FD3_ERR_ARG;
  }
  if (g->interp == FD3_INTERP_P3M && (g->order < 1 || g->order > 5)) {
    set_error("P3M (spme) interpolation supports 1..5 nodes, got %d", g->order);
    return FD3_ERR_UNSUPPORTED;
  }
  if (g->interp == FD3_INTERP_LAGRANGE && (g->order < 2 || g->order > kMaxOrder)) {
    set_error("Lagrange (pme) interpolation supports 2..%d nodes, got %d", kMaxOrder, g->order);
    return FD3_ERR_UNSUPPORTED;
  }
  for (int i = 0; i < 3; ++i) {
    const int n = g->ns[i];
    if (n < 2 || n > 1024 || (n & (n - 1))) {
      set_error("mesh size %d on axis %d is not a power of two in [2, 1024]", n, i);
      return FD3_ERR_UNSUPPORTED;
    }
    if (n < g->order) {
      set_error("mesh size %d on axis %d is smaller than the %d interpolation nodes", n, i, g->order);
      return FD3_ERR_ARG;
    }
  }
  return FD3_OK;
}

static int check_model(const fd3_model* m) {
  if (!m) {
    set_error("model is NULL");
    return FD3_ERR_ARG;
  }
  if (m->n_species < 1 || m->n_rank < 1) {
    set_error("n_species and n_rank must be positive, got %d and %d", m->n_species, m->n_rank);
    return FD3_ERR_ARG;
  }
  if (m->n_species > 64 || m->n_rank > kMaxRank) {
    set_error("this build supports up to 64 species and rank 32, got %d and %d", m->n_species, m->n_rank);
    return FD3_ERR_UNSUPPORTED;
  }
  if (!m->rcov || !m->cnref || !m->vq || !m->eigs || !m->qq) {
    set_error("model table pointer is NULL");
    return FD3_ERR_ARG;
  }
  if (!(m->r_cut > 0.0)) {
    set_error("r_cut must be positive, got %g", m->r_cut);
    return FD3_ERR_ARG;
  }
  return FD3_OK;
}

// ------------------------------------------------------------ finalize ---
struct FinalArgs {
  const double* ks; int nks;     // (nks, 16)
  const double* c6; int nc6;     // (nc6, 8)
  const double* ga; int nga;     // (nga, 16) or null
  const double* cnv; int ncnv;   // (ncnv, 8) or null
  double volume;
  double inv[9];
  const CellDyn* dyn = nullptr;     // when set, volume and inv are read from it
  const int32_t* status = nullptr;  // device cell check: nonzero poisons the outputs with NaN
  double* energy;
  double* virial;
  double* grad_cell;
  double* grad_params;
};

__device__ inline double block_sum(double v, double* sh) {
  sh[threadIdx.x] = v;
  __syncthreads();
  for (int off = blockDim.x / 2; off > 0; off >>= 1) {
    if (threadIdx.x < off) sh[threadIdx.x] += sh[threadIdx.x + off];
    __syncthreads();
  }
  const double r = sh[0];
  __syncthreads();
  return r;
}

// First stage of the partial-sum reduction, for up to four (rows, stride)
// double arrays in one launch: job j owns blocks [b0[j], b0[j+1]); its block
// lb sums rows [lb rpb, (lb+1) rpb), columns 0..ncols-1, into
// dst[lb * stride + col].  256 threads = 16 row lanes x 16 columns; fixed
// order, so deterministic.
struct ReduceJobs {
  const double* src[4];
  double* dst[4];
  int rows[4], stride[4], ncols[4], rpb[4];
  int b0[5];
  int n;
};

__device__ __forceinline__ void reduce_rows_body(const BlockInfo& bk, const ReduceJobs& jb) {
  __shared__ double sh[16][17];
  int j = 0;
  while (j + 1 < jb.n && (int)bk.id.x >= jb.b0[j + 1]) ++j;
  const int lb = bk.id.x - jb.b0[j];
  const int stride = jb.stride[j], ncols = jb.ncols[j];
  const int col = threadIdx.x & 15, lane = threadIdx.x >> 4;
  const int r0 = lb * jb.rpb[j], r1 = min(jb.rows[j], r0 + jb.rpb[j]);
  double s = 0.0;
  if (col < ncols)
    for (int r = r0 + lane; r < r1; r += 16) s += jb.src[j][(size_t)r * stride + col];
  sh[lane][col] = s;
  __syncthreads();
  if (threadIdx.x < 16) {
    double t = 0.0;
    for (int l = 0; l < 16; ++l) t += sh[l][threadIdx.x];
    if (threadIdx.x < ncols) jb.dst[j][(size_t)lb * stride + threadIdx.x] = t;
  }
}

// Queue the reduction of (rows, stride) partials to at most 256 rows written
// to dst (room for 256 * stride doubles); *src and *rows are redirected to
// the result.  Partials with <= 256 rows are left as they are.
static void reduce_partials_add(ReduceJobs& jb, const double** src, int* rows, int stride, int ncols, double* dst) {
  if (*rows <= 256) return;
  const int rpb = (*rows + 255) / 256;
  const int nb = (*rows + rpb - 1) / rpb;
  const int j = jb.n++;
  jb.src[j] = *src; jb.dst[j] = dst;
  jb.rows[j] = *rows; jb.stride[j] = stride; jb.ncols[j] = ncols; jb.rpb[j] = rpb;
  jb.b0[j + 1] = jb.b0[j] + nb;
  *src = dst;
  *rows = nb;
}

static int reduce_partials_run(const ReduceJobs& jb, hipStream_t s) {
  if (jb.n == 0) return FD3_OK;
  ProfScope _ps("k_reduce_rows", s);
  return launch<ReduceJobs, reduce_rows_body, 256>("k_reduce_rows", dim3(jb.b0[jb.n]), 0, s, jb);
}

// One block of 1024 threads: wave w sums columns w and w + 16 of the 31
// partial columns (rows <= 256 after k_reduce_rows), thread 0 assembles.
__device__ __forceinline__ void finalize_body(const BlockInfo&, const FinalArgs& fa) {
  __shared__ double tot[32];
  FinalArgs a = fa;
  // ks 0..10 -> tot[0..10]; c6 0..4 -> tot[11..15]; ga 0..8 -> tot[16..24]; cnv 0..5 -> tot[25..30]
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  for (int k = w; k < 31; k += 16) {
    const double* src;
    int rows, stride, col;
    if (k < 11) { src = a.ks; rows = a.nks; stride = 16; col = k; }
    else if (k < 16) { src = a.c6; rows = a.nc6; stride = 8; col = k - 11; }
    else if (k < 25) { src = a.ga; rows = a.ga ? a.nga : 0; stride = 16; col = k - 16; }
    else { src = a.cnv; rows = a.cnv ? a.ncnv : 0; stride = 8; col = k - 25; }
    double v[4] = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      const int b = lane + 64 * i;
      if (b < rows) v[i] = src[(size_t)b * stride + col];
    }
    const double sum = warp_sum((v[0] + v[1]) + (v[2] + v[3]));
    if (lane == 0) tot[k] = sum;
  }
  __syncthreads();
  if (threadIdx.x != 0) return;
  if (a.dyn) {
    a.volume = a.dyn->volume;
    for (int i = 0; i < 9; ++i) a.inv[i] = a.dyn->inv[i];
  }
  if (a.status && *a.status != 0)  // the cell left the range the plan covers
    for (int k = 0; k < 31; ++k) tot[k] = __builtin_nan("");
  const double om = a.volume;
  const double ekp = -tot[0] / (2.0 * om);  // reciprocal-space energy
  if (a.energy) a.energy[0] = ekp + 0.5 * tot[11];
  // symmetric 3x3 from (xx, yy, zz, xy, xz, yz)
  auto sym = [](const double* v, double* m) {
    m[0] = v[0]; m[4] = v[1]; m[8] = v[2];
    m[1] = m[3] = v[3]; m[2] = m[6] = v[4]; m[5] = m[7] = v[5];
  };
  double kv[9], cv[9];
  sym(&tot[1], kv);
  sym(&tot[25], cv);
  if (a.virial) {
    for (int i = 0; i < 9; ++i) a.virial[i] = kv[i] / (2.0 * om) + cv[i];
    for (int i = 0; i < 3; ++i) a.virial[4 * i] -= ekp;
  }
  if (a.grad_cell) {
    // dE/dC_cd = -sum_i f_ic g_id + (1/2Om) (C^-T kv)_cd - E_k (C^-T)_cd   (Bohr), then per Angstrom
    for (int c = 0; c < 3; ++c)
      for (int d = 0; d < 3; ++d) {
        double v = -tot[16 + c * 3 + d];
        double t = 0.0;
        for (int q = 0; q < 3; ++q) t += a.inv[q * 3 + c] * kv[q * 3 + d];
        v += t / (2.0 * om) - ekp * a.inv[d * 3 + c];
        a.grad_cell[c * 3 + d] = v * kA2B;
      }
  }
  if (a.grad_params)
    for (int k = 0; k < 4; ++k) a.grad_params[k] = -tot[7 + k] / (2.0 * om) + 0.5 * tot[12 + k];
}

// ------------------------------------------------------------ pipeline ---
struct Common {
  FftPlan plan;
  Geom geom;
  FftTables tab;
  float* buf1;   // spectrum
  float2* buf2;  // z-pass output A
  ColBuffers cols;
  float2* k0;        // (C) k = 0 coefficients of dE/drho
  int32_t* ew_perm;  // Ewald: atoms by species
  int32_t* ew_seg;
  double* cn_raw;  // CN and dE/dCN in double: the CN -> C6 chain runs in double
  float* charges;
  double* dEdcn;
  double* gpos;
  double *p_ks, *p_c6, *p_ga, *p_cn, *p_red;
  int n_c6, n_ga, n_cn;
  const int32_t* status = nullptr;  // device cell check (NPT entry point)
  int det = 0;                      // FD3_DETERMINISTIC (column order and spread)
};

static void carve_common(Arena& ar, Common& c, int64_t n, int r, int cn_rows) {
  fft_tables_carve(c.plan, ar, &c.tab);
  // spectrum and z-pass output (mesh methods), or S and phi (Ewald)
  const size_t nbuf = c.geom.interp == FD3_INTERP_EWALD ? ewald_buffer_elems(c.plan) : c.plan.mesh_complex();
  c.buf1 = ar.take<float>(2 * nbuf);
  c.buf2 = ar.take<float2>(nbuf);
  c.k0 = ar.take<float2>(c.plan.C);
  c.ew_perm = ar.take<int32_t>(std::max<int64_t>(n, 1));
  c.ew_seg = ar.take<int32_t>(c.plan.nS + 1);
  col_carve(c.plan, n, c.geom.order, ar, &c.cols);
  c.cn_raw = ar.take<double>(std::max<int64_t>(n, 1));
  c.charges = ar.take<float>(std::max<int64_t>(n * r, 1));
  c.dEdcn = ar.take<double>(std::max<int64_t>(n, 1));
  c.gpos = ar.take<double>(std::max<int64_t>(3 * n, 3));
  c.n_c6 = c6_blocks(n);
  c.n_ga = c6_blocks(n);  // partials of k_c6_bwd
  c.n_cn = std::max(cn_rows, 1);
  c.p_ks = ar.take<double>((size_t)c.plan.ks_blocks * 16);
  c.p_c6 = ar.take<double>((size_t)c.n_c6 * 8);
  c.p_ga = ar.take<double>((size_t)c.n_ga * 16);
  c.p_cn = ar.take<double>((size_t)c.n_cn * 8);
  c.p_red = ar.take<double>(256 * 48);
}

static int setup_common(const fd3_model* m, const fd3_geom* g, Common& c) {
  int rc;
  if ((rc = check_model(m))) return rc;
  if ((rc = make_geom(g, &c.geom))) return rc;
  if (g->interp == FD3_INTERP_EWALD) return ewald_plan_make(g->ns, m->n_species, m->n_rank, &c.plan);
  if ((rc = fft_plan_make(g->ns, m->n_species, m->n_rank, &c.plan))) return rc;
  return FD3_OK;
}

static FusedArgs fused_args(const fd3_model* m, const Common& c, int64_t n, const double4* pos4,
                            const int32_t* spec_s) {
  FusedArgs fa;
  fa.n = n; fa.nS = m->n_species; fa.r = m->n_rank;
  fa.nx = c.geom.nx; fa.ny = c.geom.ny; fa.nz = c.geom.nz; fa.order = c.geom.order; fa.interp = c.geom.interp;
  for (int i = 0; i < 9; ++i) fa.inv[i] = c.geom.inv[i];
  fa.dyn = c.geom.dyn;
  fa.pos4 = pos4; fa.spec_s = spec_s; fa.charges = c.charges; fa.cb = c.cols;
  fa.det = c.det;
  return fa;
}

// The part of the mesh path that needs only the model, the cell, positions
// and species: FFT and Green's function tables, and the column sort of the
// atoms (mesh methods).
static int mesh_prep(const fd3_model* m, Common& c, int64_t n, const double4* pos4, const int32_t* spec_s, int flags,
                     hipStream_t s) {
  int rc;
  c.det = (flags & FD3_DETERMINISTIC) != 0;
  PairInit pi;
  pi.qq = m->qq; pi.nS = m->n_species;
  pi.s6 = (float)m->s6; pi.s8 = (float)m->s8; pi.a1 = (float)m->a1; pi.a2 = (float)m->a2;
  pi.gdk = c.geom.interp == FD3_INTERP_EWALD ? 0.0 : ks_gtab_dk(c.plan, c.geom);
  if ((rc = fft_tables_init(c.plan, c.tab, c.geom.order, c.geom.interp == FD3_INTERP_P3M, s, &pi))) return rc;
  if (c.geom.interp != FD3_INTERP_EWALD) {
    const FusedArgs fa = fused_args(m, c, n, pos4, spec_s);
    if ((rc = col_sort(fa, s))) return rc;
  }
  return FD3_OK;
}

// Everything after the CN: tables and column sort, C6 chain, spread, FFTs,
// k-space, adjoint.  pos4/spec_s in slot order; slot_atom null for the
// identity.
static int run_mesh_path(const fd3_model* m, Common& c, int64_t n, const double4* pos4, const int32_t* spec_s,
                         const int32_t* slot_atom, int flags, const fd3_outputs* out, hipStream_t s,
                         bool prepped = false) {
  int rc;
  const int r = m->n_rank, nS = m->n_species;
  const bool grad = flags & FD3_NEED_GRAD;
  c.det = (flags & FD3_DETERMINISTIC) != 0;
  if (!prepped && (rc = mesh_prep(m, c, n, pos4, spec_s, flags, s))) return rc;
  C6Args ca;
  ca.n = n; ca.nS = nS; ca.r = r; ca.cn_raw = c.cn_raw; ca.spec_s = spec_s; ca.slot_atom = slot_atom;
  ca.cnref = m->cnref; ca.vq = m->vq; ca.eigs = m->eigs; ca.qq = m->qq; ca.cncorr = m->cncorr;
  ca.s6 = m->s6; ca.s8 = m->s8; ca.a1 = m->a1; ca.a2 = m->a2;
  ca.charges = c.charges; ca.cn_out = out ? out->cn : nullptr; ca.partials = c.p_c6;
  if ((rc = launch_c6_fwd(ca, s))) return rc;
  if (c.geom.interp == FD3_INTERP_EWALD) {
    EwaldArgs ea;
    ea.n = n; ea.nS = nS; ea.r = r; ea.nx = c.geom.nx; ea.ny = c.geom.ny; ea.nz = c.geom.nz;
    for (int i = 0; i < 9; ++i) { ea.inv[i] = c.geom.inv[i]; ea.recip[i] = c.geom.recip[i]; }
    ea.volume = c.geom.volume;
    ea.pos4 = pos4; ea.spec_s = spec_s; ea.charges = c.charges; ea.eigs = m->eigs;
    ea.pc = c.tab.pc; ea.htab = c.tab.htab;
    ea.S = (float2*)c.buf1; ea.phi = c.buf2; ea.perm = c.ew_perm; ea.seg = c.ew_seg;
    ea.partials = c.p_ks; ea.accA = c.cols.accA; ea.accG = c.cols.accG;
    ea.need_params = (flags & FD3_NEED_PARAMS) ? 1 : 0;
    if ((rc = ewald_run(ea, grad, s))) return rc;
    if (grad && n > 0) {
      C6BwdArgs ba;
      ba.n = n; ba.nS = nS; ba.r = r; ba.nx = 1; ba.ny = 1; ba.nz = 1;  // accG holds dE/df
      for (int i = 0; i < 9; ++i) ba.inv[i] = c.geom.inv[i];
      ba.pos4 = pos4; ba.spec_s = spec_s; ba.charges = c.charges; ba.accA = c.cols.accA; ba.accG = c.cols.accG;
      ba.k0 = nullptr;  // the Ewald adjoint sums every k directly
      ba.cn_raw = c.cn_raw; ba.cnref = m->cnref; ba.vq = m->vq; ba.eigs = m->eigs; ba.qq = m->qq;
      ba.cncorr = m->cncorr; ba.s6 = m->s6; ba.s8 = m->s8; ba.a1 = m->a1; ba.a2 = m->a2;
      ba.gpos = c.gpos; ba.dEdcn = c.dEdcn; ba.partials = c.p_ga;
      if ((rc = launch_c6_bwd(ba, s))) return rc;
    }
    return FD3_OK;
  }
  // spreading fused into the forward z pass (atoms sorted into columns by mesh_prep)
  const FusedArgs fa = fused_args(m, c, n, pos4, spec_s);
  float2* spec = (float2*)c.buf1;
  if ((rc = spread_r2c_z(fa, c.plan, c.tab, c.buf2, s))) return rc;
  const bool xs = c.plan.xs != 0;  // split x FFT between the y passes and the k-space kernel
  if ((rc = fft_y(c.plan, c.tab, c.buf2, spec, -1, s, xs))) return rc;
  if ((rc = kspace_run(c.plan, c.tab, c.geom, spec, m->eigs, grad ? 1 : 0, (flags & FD3_NEED_PARAMS) ? 1 : 0, c.p_ks,
                       c.k0, s, xs)))
    return rc;
  if (grad) {
    if ((rc = fft_y(c.plan, c.tab, spec, c.buf2, 1, s, xs))) return rc;
    if ((rc = c2r_z_gather(fa, c.plan, c.tab, c.buf2, s))) return rc;
    C6BwdArgs ba;
    ba.n = n; ba.nS = nS; ba.r = r; ba.nx = c.geom.nx; ba.ny = c.geom.ny; ba.nz = c.geom.nz;
    for (int i = 0; i < 9; ++i) ba.inv[i] = c.geom.inv[i];
    ba.dyn = c.geom.dyn;
    ba.pos4 = pos4; ba.spec_s = spec_s; ba.charges = c.charges; ba.accA = c.cols.accA; ba.accG = c.cols.accG;
    ba.k0 = c.k0;
    ba.cn_raw = c.cn_raw; ba.cnref = m->cnref; ba.vq = m->vq; ba.eigs = m->eigs; ba.qq = m->qq;
    ba.cncorr = m->cncorr; ba.s6 = m->s6; ba.s8 = m->s8; ba.a1 = m->a1; ba.a2 = m->a2;
    ba.gpos = c.gpos; ba.dEdcn = c.dEdcn; ba.partials = c.p_ga;
    if (n > 0 && (rc = launch_c6_bwd(ba, s))) return rc;
  }
  return FD3_OK;
}

static int run_finalize(const Common& c, bool grad, bool params, bool have_cnv, const fd3_outputs* out,
                        hipStream_t s) {
  FinalArgs fa;
  fa.ks = c.p_ks; fa.nks = c.plan.ks_blocks;
  fa.c6 = c.p_c6; fa.nc6 = c.n_c6;
  fa.ga = grad ? c.p_ga : nullptr; fa.nga = c.n_ga;
  fa.cnv = (grad && have_cnv) ? c.p_cn : nullptr; fa.ncnv = c.n_cn;
  int rc;
  ReduceJobs jb{};
  reduce_partials_add(jb, &fa.ks, &fa.nks, 16, 11, c.p_red);
  reduce_partials_add(jb, &fa.c6, &fa.nc6, 8, 5, c.p_red + 256 * 16);
  if (fa.ga) reduce_partials_add(jb, &fa.ga, &fa.nga, 16, 9, c.p_red + 256 * 24);
  if (fa.cnv) reduce_partials_add(jb, &fa.cnv, &fa.ncnv, 8, 6, c.p_red + 256 * 40);
  if ((rc = reduce_partials_run(jb, s))) return rc;
  fa.volume = c.geom.volume;
  for (int i = 0; i < 9; ++i) fa.inv[i] = c.geom.inv[i];
  fa.dyn = c.geom.dyn;
  fa.status = c.status;
  fa.energy = out->energy;
  fa.virial = grad ? out->virial : nullptr;
  fa.grad_cell = grad ? out->grad_cell : nullptr;
  fa.grad_params = params ? out->grad_params : nullptr;
  ProfScope _ps("k_finalize", s);
  return launch<FinalArgs, finalize_body, 1024>("k_finalize", dim3(1), 0, s, fa);
}

// Device cell update of the NPT entry point: the cell (Angstrom, row i =
// a_i) is read from device memory, converted to Bohr, inverted, and checked
// against what the plan made from the reference cell covers: every cell-list
// axis still reaches r_cut (sr_a d_a / nb_a >= r_cut with d_a the plane
// spacing) and, for the per-pair G tables, the mesh's largest |k| stays in
// the tabulated range.  status: 0, or 1 (singular or not finite) | 2
// (cell-list reach) | 4 (|k| range).
struct CellCheck {
  int nb[3], sr[3];
  double rc;        // Bohr
  double kmax_tab;  // largest tabulated |k| (Bohr^-1), 0 = no table limit
  int nx, ny, nz;
};

__global__ void k_cell_update(const double* __restrict__ cell_A, CellCheck ck, CellDyn* __restrict__ out,
                              int32_t* __restrict__ status, int32_t* __restrict__ status_user) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  double c[9];
  for (int i = 0; i < 9; ++i) c[i] = cell_A[i] * kA2B;
  const double det = c[0] * (c[4] * c[8] - c[5] * c[7]) - c[1] * (c[3] * c[8] - c[5] * c[6]) +
                     c[2] * (c[3] * c[7] - c[4] * c[6]);
  int st = 0;
  double inv[9];
  if (!(fabs(det) > 1e-12) || !isfinite(det)) {
    st |= 1;
    for (int i = 0; i < 9; ++i) inv[i] = (i % 4 == 0) ? 1.0 : 0.0;  // keeps later kernels finite
  } else {
    inv[0] = (c[4] * c[8] - c[5] * c[7]) / det;
    inv[1] = (c[2] * c[7] - c[1] * c[8]) / det;
    inv[2] = (c[1] * c[5] - c[2] * c[4]) / det;
    inv[3] = (c[5] * c[6] - c[3] * c[8]) / det;
    inv[4] = (c[0] * c[8] - c[2] * c[6]) / det;
    inv[5] = (c[2] * c[3] - c[0] * c[5]) / det;
    inv[6] = (c[3] * c[7] - c[4] * c[6]) / det;
    inv[7] = (c[1] * c[6] - c[0] * c[7]) / det;
    inv[8] = (c[0] * c[4] - c[1] * c[3]) / det;
  }
  for (int a = 0; a < 3; ++a) {
    const double nrm = sqrt(inv[a] * inv[a] + inv[3 + a] * inv[3 + a] + inv[6 + a] * inv[6 + a]);
    if (!(ck.sr[a] / (nrm * ck.nb[a]) >= ck.rc)) st |= 2;
  }
  double recip[9];
  for (int i = 0; i < 3; ++i)
    for (int d = 0; d < 3; ++d) recip[i * 3 + d] = 2.0 * kPi * inv[d * 3 + i];
  if (ck.kmax_tab > 0.0) {
    for (int sx = -1; sx <= 1; sx += 2)
      for (int sy = -1; sy <= 1; sy += 2) {
        const double mm[3] = {0.5 * sx * ck.nx, 0.5 * sy * ck.ny, 0.5 * ck.nz};
        double k2 = 0.0;
        for (int d = 0; d < 3; ++d) {
          const double kc = mm[0] * recip[d] + mm[1] * recip[3 + d] + mm[2] * recip[6 + d];
          k2 += kc * kc;
        }
        if (!(sqrt(k2) <= ck.kmax_tab)) st |= 4;
      }
  }
  for (int i = 0; i < 9; ++i) {
    out->cell[i] = c[i];
    out->inv[i] = inv[i];
    out->recip[i] = recip[i];
    out->recip_f[i] = (float)recip[i];
  }
  out->volume = (st & 1) ? 1.0 : fabs(det);
  out->neg_inv_vol_f = (float)(-1.0 / out->volume);
  out->inv_volume = 1.0 / out->volume;
  *status = st;
  if (status_user) *status_user = st;
}

// A nonzero device-cell status means pairs beyond the cell-list reach were
// missed or G was read past its table: the per-atom gradients are as wrong as
// the energy, so they are poisoned with NaN like it (k_finalize does the rest).
__global__ void k_poison_grads(const int32_t* __restrict__ status, double* __restrict__ a, int64_t na,
                               double* __restrict__ b, int64_t nb) {
  if (*status == 0) return;
  const double nan = __builtin_nan("");
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < na + nb; i += (int64_t)gridDim.x * blockDim.x) {
    if (i < na) a[i] = nan;
    else b[i - na] = nan;
  }
}

static int launch_poison(const Common& c, double* gpos, int64_t n_pos, double* gsh, int64_t n_sh, hipStream_t s) {
  if (!c.status) return FD3_OK;
  const int64_t na = gpos ? 3 * n_pos : 0, nb = gsh ? 3 * n_sh : 0;
  if (na + nb == 0) return FD3_OK;
  ProfScope _ps("k_poison_grads", s);
  hipLaunchKernelGGL(k_poison_grads, dim3((unsigned)std::min<int64_t>(cdiv(na + nb, 256), 1024)), dim3(256), 0, s,
                     c.status, gpos, na, gsh, nb);
  FD3_LAUNCH_CHECK("k_poison_grads");
  return FD3_OK;
}

// Queue k_cell_update and point the pipeline at the device cell.
static int launch_cell_update(Common& c, CellCheck ck, const double* cell_dev, CellDyn* cd, int32_t* st,
                              int32_t* status_user, hipStream_t s) {
  ck.nx = c.geom.nx; ck.ny = c.geom.ny; ck.nz = c.geom.nz;
  ProfScope _ps("k_cell_update", s);
  hipLaunchKernelGGL(k_cell_update, dim3(1), dim3(64), 0, s, cell_dev, ck, cd, st, status_user);
  FD3_LAUNCH_CHECK("k_cell_update");
  c.geom.dyn = cd;
  c.status = st;
  return FD3_OK;
}

// ------------------------------------------------------------- D4 CN path ---
struct D4Bufs {
  FftPlan pcn;      // the mesh plan with one rank: channels = species
  CnKsPlan kp;
  float2* rho_hat;  // (nS, M1, ny, nx) x-transformed species density spectrum
  float* ones;      // (n) unit charges of the one-hot density
  float* g1;        // (n, 3) forward gather's sum_nodes dW/du phi
  float* wf;        // (n) dE/dCN as float: the charges of the backward density
  float4* tab;      // per-pair G^CN tables
};

static int d4_layout(const fd3_model* m, const fd3_cn_d4* d4, const fd3_geom* g, int64_t n, Common& c, Arena& ar,
                     double4** pos4, int32_t** spec_s, D4Bufs& db, double reach = 1.0, CellDyn** cd = nullptr,
                     int32_t** st = nullptr) {
  int rc;
  if ((rc = setup_common(m, g, c))) return rc;
  if (!(reach >= 1.0) || !std::isfinite(reach)) {
    set_error("reach factor must be >= 1, got %g", reach);
    return FD3_ERR_ARG;
  }
  c.geom.kreach = reach;
  if (!d4) {
    set_error("D4 CN parameters are NULL");
    return FD3_ERR_ARG;
  }
  if (n < 0) {
    set_error("negative atom count");
    return FD3_ERR_ARG;
  }
  const int nS = m->n_species;
  db.pcn = c.plan;
  db.pcn.r = 1;
  db.pcn.C = nS;
  size_t nrho;
  if (g->interp == FD3_INTERP_EWALD) {  // k-space rows: one per 256 k points
    if (nS > 64) {
      set_error("the Ewald D4 CN supports up to 64 species");
      return FD3_ERR_UNSUPPORTED;
    }
    db.kp = CnKsPlan{};
    db.kp.blocks = (int)cdiv((int64_t)c.plan.nx * c.plan.ny * c.plan.nz, 256);
    nrho = (size_t)c.plan.nx * c.plan.ny * c.plan.nz * nS;
  } else {
    if ((rc = cn_ks_plan(db.pcn, nS, &db.kp))) return rc;
    nrho = db.pcn.mesh_complex();
  }
  carve_common(ar, c, n, m->n_rank, db.kp.blocks + (int)cdiv(std::max<int64_t>(n, 1), 256));
  *pos4 = ar.take<double4>(std::max<int64_t>(n, 1));
  *spec_s = ar.take<int32_t>(std::max<int64_t>(n, 1));
  db.rho_hat = ar.take<float2>(nrho);
  db.ones = ar.take<float>(std::max<int64_t>(n, 1));
  db.g1 = ar.take<float>(std::max<int64_t>(3 * n, 3));
  db.wf = ar.take<float>(std::max<int64_t>(n, 1));
  db.tab = ar.take<float4>((size_t)nS * (nS + 1) * kGtabN);
  if (cd) {
    *cd = ar.take<CellDyn>(1);
    *st = ar.take<int32_t>(1);
  }
  return FD3_OK;
}

// The spread / y / k-space / y / gather passes of the D4 CN over the species
// channels: forward (charges 1, the CN and dW phi per atom) or backward
// (charges dE/dCN, the force and the strain derivative).
static CnKsArgs d4_ks_args(const fd3_model* m, const fd3_cn_d4* d4, const Common& c, const D4Bufs& db) {
  CnKsArgs ka;
  ka.spec = (float2*)c.buf1; ka.rho_hat = db.rho_hat; ka.twx = c.tab.twx;
  ka.scx = c.tab.scx; ka.scy = c.tab.scy; ka.scz = c.tab.scz;
  ka.r0 = d4->r0; ka.delta = d4->delta; ka.k0 = (float)d4->k0;
  ka.tab = db.tab; ka.tkinv = (float)(1.0 / ks_gtab_dk(c.plan, c.geom));
  ka.partials = c.p_cn; ka.inv_vol = 1.0 / c.geom.volume;
  for (int i = 0; i < 9; ++i) ka.recip[i] = (float)c.geom.recip[i];
  ka.nx = c.geom.nx; ka.ny = c.geom.ny; ka.nz = c.geom.nz; ka.M1 = c.plan.M1; ka.nS = m->n_species;
  ka.NK = db.kp.NK; ka.KT = db.kp.KT; ka.use_sinc = c.geom.interp == FD3_INTERP_P3M;
  ka.dyn = c.geom.dyn;
  ka.sink = c.tab.sink;
  return ka;
}

static CnAtomArgs d4_atom_args(const fd3_cn_d4* d4, const Common& c, const D4Bufs& db, int64_t n,
                               const int32_t* spec_s) {
  CnAtomArgs aa;
  aa.n = n;
  // Ewald: the per-atom sums are derivatives in fractional coordinates
  const bool ew = c.geom.interp == FD3_INTERP_EWALD;
  aa.nx = ew ? 1 : c.geom.nx; aa.ny = ew ? 1 : c.geom.ny; aa.nz = ew ? 1 : c.geom.nz;
  for (int i = 0; i < 9; ++i) aa.inv[i] = c.geom.inv[i];
  aa.inv_vol = 1.0 / c.geom.volume;
  aa.spec_s = spec_s; aa.selfcont = d4->selfcont;
  aa.accA = c.cols.accA; aa.accG = c.cols.accG;
  aa.cn = c.cn_raw; aa.g1 = db.g1; aa.dEdcn = c.dEdcn; aa.gpos = c.gpos;
  aa.partials = offset_or_null(c.p_cn, (size_t)db.kp.blocks * 8);
  aa.dyn = c.geom.dyn;
  return aa;
}

// The Ewald form: structure factors of the species weights on the full k
// grid, G^CN, per-atom sums.
static int d4_passes_ewald(const fd3_model* m, const fd3_cn_d4* d4, Common& c, D4Bufs& db, int64_t n,
                           const double4* pos4, const int32_t* spec_s, bool bwd, hipStream_t s) {
  int rc;
  EwCnArgs ea;
  ea.n = n; ea.nS = m->n_species; ea.nx = c.plan.nx; ea.ny = c.plan.ny; ea.nz = c.plan.nz;
  for (int i = 0; i < 9; ++i) { ea.inv[i] = c.geom.inv[i]; ea.recip[i] = c.geom.recip[i]; }
  ea.volume = c.geom.volume;
  ea.pos4 = pos4; ea.spec_s = spec_s; ea.w = bwd ? db.wf : db.ones;
  ea.perm = c.ew_perm; ea.seg = c.ew_seg;
  ea.tab = db.tab; ea.tkinv = (float)(1.0 / ks_gtab_dk(c.plan, c.geom));
  ea.S = (float2*)c.buf1; ea.Srho = db.rho_hat; ea.phi = c.buf2;
  ea.partials = c.p_cn; ea.accA = c.cols.accA; ea.accG = c.cols.accG;
  if (!bwd) {
    if ((rc = ewald_partition(spec_s, n, m->n_species, c.ew_perm, c.ew_seg, s))) return rc;
    CnKsArgs ka{};
    ka.r0 = d4->r0; ka.delta = d4->delta; ka.k0 = (float)d4->k0; ka.tab = db.tab; ka.nS = m->n_species;
    if ((rc = cn_tab_init(ka, ks_gtab_dk(c.plan, c.geom), s))) return rc;
  }
  if ((rc = ewcn_pass(ea, bwd, s))) return rc;
  const CnAtomArgs aa = d4_atom_args(d4, c, db, n, spec_s);
  return bwd ? cn_d4_force(aa, s) : cn_d4_fin(aa, s);
}

static int d4_passes(const fd3_model* m, const fd3_cn_d4* d4, Common& c, D4Bufs& db, int64_t n,
                     const double4* pos4, const int32_t* spec_s, bool bwd, hipStream_t s) {
  if (c.geom.interp == FD3_INTERP_EWALD) return d4_passes_ewald(m, d4, c, db, n, pos4, spec_s, bwd, s);
  int rc;
  FusedArgs fa = fused_args(m, c, n, pos4, spec_s);
  fa.r = 1;
  fa.charges = db.ones;
  FusedArgs fs = fa;
  if (bwd) fs.charges = db.wf;
  float2* spec = (float2*)c.buf1;
  if ((rc = spread_r2c_z(fs, db.pcn, c.tab, c.buf2, s))) return rc;
  if ((rc = fft_y(db.pcn, c.tab, c.buf2, spec, -1, s, db.kp.xs))) return rc;
  const CnKsArgs ka = d4_ks_args(m, d4, c, db);
  if (!bwd && (rc = cn_tab_init(ka, ks_gtab_dk(c.plan, c.geom), s))) return rc;
  if ((rc = cn_ks_run(ka, db.kp, bwd, s))) return rc;
  if ((rc = fft_y(db.pcn, c.tab, spec, c.buf2, 1, s, db.kp.xs))) return rc;
  if ((rc = c2r_z_gather(fa, db.pcn, c.tab, c.buf2, s))) return rc;  // and the per-atom sums
  const CnAtomArgs aa = d4_atom_args(d4, c, db, n, spec_s);
  return bwd ? cn_d4_force(aa, s) : cn_d4_fin(aa, s);
}

}  // namespace fd3

using namespace fd3;

extern "C" {

int fd3_eval_d4_workspace(const fd3_model* m, const fd3_cn_d4* d4, const fd3_geom* g, int64_t n_atoms, int flags,
                          size_t* bytes) {
  (void)flags;
  Common c;
  Arena ar;
  double4* p4;
  int32_t* sp;
  D4Bufs db;
  int rc = d4_layout(m, d4, g, n_atoms, c, ar, &p4, &sp, db);
  if (rc) return rc;
  *bytes = align_up(ar.off);
  return FD3_OK;
}

}  // extern "C"

// The D4 pipeline; with cell_dev the cell comes from device memory (tables
// planned from g->cell with the reach factor).
static int eval_d4(const fd3_model* m, const fd3_cn_d4* d4, const fd3_geom* g, double reach, const double* cell_dev,
                   int32_t* status_user, int64_t n, const void* positions, int pos_dtype, const int32_t* species,
                   int flags, void* workspace, size_t ws_bytes, const fd3_outputs* out, hipStream_t s) {
  Common c;
  Arena ar(workspace, ws_bytes);
  double4* pos4;
  int32_t* spec_s;
  D4Bufs db;
  CellDyn* cd = nullptr;
  int32_t* st = nullptr;
  int rc = d4_layout(m, d4, g, n, c, ar, &pos4, &spec_s, db, reach, cell_dev ? &cd : nullptr, &st);
  if (rc) return rc;
  if (!ar.ok() || !workspace) {
    set_error("workspace of %zu bytes is smaller than the %zu required", ws_bytes, align_up(ar.off));
    return FD3_ERR_WORKSPACE;
  }
  if (!out) {
    set_error("outputs are NULL");
    return FD3_ERR_ARG;
  }
  if (n > 0 && (!positions || !species)) {
    set_error("input pointer is NULL");
    return FD3_ERR_ARG;
  }
  if (!d4->r0 || !d4->delta || !d4->selfcont) {
    set_error("D4 CN tables are NULL");
    return FD3_ERR_ARG;
  }
  const bool grad = flags & FD3_NEED_GRAD;
  if (cell_dev) {
    CellCheck ck;  // no cell list; the CN tables always cover |k|
    for (int a = 0; a < 3; ++a) { ck.nb[a] = 1; ck.sr[a] = 1; }
    ck.rc = 0.0;
    ck.kmax_tab = ks_gtab_dk(c.plan, c.geom) * kGtabN;
    if ((rc = launch_cell_update(c, ck, cell_dev, cd, st, status_user, s))) return rc;
  }
  if ((rc = launch_prep_pos(positions, pos_dtype, species, n, pos4, spec_s, s))) return rc;
  if ((rc = mesh_prep(m, c, n, pos4, spec_s, flags, s))) return rc;
  if ((rc = fill_value(db.ones, n, 1.0f, s))) return rc;
  if ((rc = d4_passes(m, d4, c, db, n, pos4, spec_s, false, s))) return rc;
  if ((rc = run_mesh_path(m, c, n, pos4, spec_s, nullptr, flags, out, s, true))) return rc;
  if (grad && n > 0) {
    if ((rc = launch_to_f32(c.dEdcn, db.wf, n, s))) return rc;
    if ((rc = d4_passes(m, d4, c, db, n, pos4, spec_s, true, s))) return rc;
    if (out->grad_pos && (rc = launch_unpermute(c.gpos, nullptr, n, out->grad_pos, s))) return rc;
    if ((rc = launch_poison(c, out->grad_pos, n, nullptr, 0, s))) return rc;
  }
  fd3_outputs o2 = *out;
  o2.grad_cell = nullptr;  // the D4 path reports the strain derivative only
  return run_finalize(c, grad, (flags & FD3_NEED_PARAMS) != 0, n > 0, &o2, s);
}

extern "C" {

int fd3_eval_d4(const fd3_model* m, const fd3_cn_d4* d4, const fd3_geom* g, int64_t n, const void* positions,
                int pos_dtype, const int32_t* species, int flags, void* workspace, size_t ws_bytes,
                const fd3_outputs* out, void* stream) {
  return eval_d4(m, d4, g, 1.0, nullptr, nullptr, n, positions, pos_dtype, species, flags, workspace, ws_bytes, out,
                 (hipStream_t)stream);
}

int fd3_eval_d4_npt_workspace(const fd3_model* m, const fd3_cn_d4* d4, const fd3_geom* g, double reach,
                              int64_t n_atoms, int flags, size_t* bytes) {
  (void)flags;
  Common c;
  Arena ar;
  double4* p4;
  int32_t* sp;
  D4Bufs db;
  CellDyn* cd = nullptr;
  int32_t* st = nullptr;
  int rc = d4_layout(m, d4, g, n_atoms, c, ar, &p4, &sp, db, reach, &cd, &st);
  if (rc) return rc;
  *bytes = align_up(ar.off);
  return FD3_OK;
}

int fd3_eval_d4_npt(const fd3_model* m, const fd3_cn_d4* d4, const fd3_geom* g, double reach, const double* cell,
                    int32_t* status, int64_t n, const void* positions, int pos_dtype, const int32_t* species,
                    int flags, void* workspace, size_t ws_bytes, const fd3_outputs* out, void* stream) {
  if (!cell) {
    set_error("device cell pointer is NULL");
    return FD3_ERR_ARG;
  }
  if (g && g->interp == FD3_INTERP_EWALD) {
    set_error("the device-cell entry point supports the mesh methods (spme, pme)");
    return FD3_ERR_UNSUPPORTED;
  }
  return eval_d4(m, d4, g, reach, cell, status, n, positions, pos_dtype, species, flags, workspace, ws_bytes, out,
                 (hipStream_t)stream);
}

const char* fd3_last_error(void) { return g_err.c_str(); }
int fd3_version(void) { return 1; }

int fd3_profile_enable(int on) {
  fd3::prof_set(on != 0);
  return FD3_OK;
}

int fd3_profile_only(const char* kernel) {
  fd3::prof_only(kernel);
  return FD3_OK;
}

int fd3_profile_report(int max_entries, char* names, double* total_ms, int64_t* counts, int* n_entries) {
  return fd3::prof_report(max_entries, names, total_ms, counts, n_entries);
}

// ------------------------------------------------------------ edges path ---
static int edges_layout(const fd3_model* m, const fd3_geom* g, int64_t n, int64_t ne, Common& c, Arena& ar,
                        double4** pos4, int32_t** spec_s, double reach = 1.0, CellDyn** cd = nullptr,
                        int32_t** st = nullptr) {
  int rc;
  if ((rc = setup_common(m, g, c))) return rc;
  if (n < 0 || ne < 0) {
    set_error("negative atom or edge count");
    return FD3_ERR_ARG;
  }
  if (!(reach >= 1.0) || !std::isfinite(reach)) {
    set_error("reach factor must be >= 1, got %g", reach);
    return FD3_ERR_ARG;
  }
  c.geom.kreach = reach;
  carve_common(ar, c, n, m->n_rank, cn_blocks(ne));
  *pos4 = ar.take<double4>(std::max<int64_t>(n, 1));
  *spec_s = ar.take<int32_t>(std::max<int64_t>(n, 1));
  if (cd) {
    *cd = ar.take<CellDyn>(1);
    *st = ar.take<int32_t>(1);
  }
  return FD3_OK;
}

int fd3_eval_edges_workspace(const fd3_model* m, const fd3_geom* g, int64_t n_atoms, int64_t n_edges, int flags,
                             size_t* bytes) {
  (void)flags;
  Common c;
  Arena ar;
  double4* p4;
  int32_t* sp;
  int rc = edges_layout(m, g, n_atoms, n_edges, c, ar, &p4, &sp);
  if (rc) return rc;
  *bytes = align_up(ar.off);
  return FD3_OK;
}

}  // extern "C"

// The edge-list pipeline; with cell_dev the cell comes from device memory
// (shifts are Cartesian, so the CN kernels need no cell; the mesh kernels
// and the cell gradient read the device cell).
static int eval_edges(const fd3_model* m, const fd3_geom* g, double reach, const double* cell_dev,
                      int32_t* status_user, int64_t n, const void* positions, int pos_dtype,
                      const int32_t* species, int64_t ne, const int64_t* edge_src, const int64_t* edge_dst,
                      const void* shifts, int shifts_dtype, int flags, void* workspace, size_t ws_bytes,
                      const fd3_outputs* out, hipStream_t s) {
  Common c;
  Arena ar(workspace, ws_bytes);
  double4* pos4;
  int32_t* spec_s;
  CellDyn* cd = nullptr;
  int32_t* st = nullptr;
  int rc = edges_layout(m, g, n, ne, c, ar, &pos4, &spec_s, reach, cell_dev ? &cd : nullptr, &st);
  if (rc) return rc;
  if (!ar.ok() || !workspace) {
    set_error("workspace of %zu bytes is smaller than the %zu required", ws_bytes, align_up(ar.off));
    return FD3_ERR_WORKSPACE;
  }
  if (!out) {
    set_error("outputs are NULL");
    return FD3_ERR_ARG;
  }
  if ((n > 0 && (!positions || !species)) || (ne > 0 && (!edge_src || !edge_dst || !shifts))) {
    set_error("input pointer is NULL");
    return FD3_ERR_ARG;
  }
  const bool grad = flags & FD3_NEED_GRAD;
  if (cell_dev) {
    CellCheck ck;  // no cell list
    for (int a = 0; a < 3; ++a) { ck.nb[a] = 1; ck.sr[a] = 1; }
    ck.rc = 0.0;
    ck.kmax_tab = c.plan.ks_reg ? ks_gtab_dk(c.plan, c.geom) * kGtabN : 0.0;
    if ((rc = launch_cell_update(c, ck, cell_dev, cd, st, status_user, s))) return rc;
  }
  if ((rc = launch_prep_pos(positions, pos_dtype, species, n, pos4, spec_s, s))) return rc;
  if ((rc = launch_fill_f64(c.cn_raw, std::max<int64_t>(n, 1), s))) return rc;
  CnArgs cna{};
  cna.rcov = m->rcov; cna.rcut = m->r_cut * kA2B; cna.cn = c.cn_raw; cna.dEdcn = c.dEdcn;
  cna.gpos = c.gpos; cna.partials = c.p_cn;
  EdgeArgs ea;
  ea.n_edges = ne; ea.src = edge_src; ea.dst = edge_dst; ea.shifts = shifts; ea.shifts_dtype = shifts_dtype;
  if ((rc = launch_cn_edges(ea, cna, pos4, spec_s, s))) return rc;
  if ((rc = run_mesh_path(m, c, n, pos4, spec_s, nullptr, flags, out, s))) return rc;
  if (grad) {
    if ((rc = launch_cn_bwd_edges(ea, cna, pos4, spec_s, out->grad_shifts, s))) return rc;
    if (out->grad_pos && (rc = launch_unpermute(c.gpos, nullptr, n, out->grad_pos, s))) return rc;
    if ((rc = launch_poison(c, out->grad_pos, n, out->grad_shifts, ne, s))) return rc;
  }
  return run_finalize(c, grad, (flags & FD3_NEED_PARAMS) != 0, true, out, s);
}

extern "C" {

int fd3_eval_edges(const fd3_model* m, const fd3_geom* g, int64_t n, const void* positions, int pos_dtype,
                   const int32_t* species, int64_t ne, const int64_t* edge_src, const int64_t* edge_dst,
                   const void* shifts, int shifts_dtype, int flags, void* workspace, size_t ws_bytes,
                   const fd3_outputs* out, void* stream) {
  return eval_edges(m, g, 1.0, nullptr, nullptr, n, positions, pos_dtype, species, ne, edge_src, edge_dst, shifts,
                    shifts_dtype, flags, workspace, ws_bytes, out, (hipStream_t)stream);
}

int fd3_eval_edges_npt_workspace(const fd3_model* m, const fd3_geom* g, double reach, int64_t n_atoms,
                                 int64_t n_edges, int flags, size_t* bytes) {
  (void)flags;
  Common c;
  Arena ar;
  double4* p4;
  int32_t* sp;
  CellDyn* cd = nullptr;
  int32_t* st = nullptr;
  int rc = edges_layout(m, g, n_atoms, n_edges, c, ar, &p4, &sp, reach, &cd, &st);
  if (rc) return rc;
  *bytes = align_up(ar.off);
  return FD3_OK;
}

int fd3_eval_edges_npt(const fd3_model* m, const fd3_geom* g, double reach, const double* cell, int32_t* status,
                       int64_t n, const void* positions, int pos_dtype, const int32_t* species, int64_t ne,
                       const int64_t* edge_src, const int64_t* edge_dst, const void* shifts, int shifts_dtype,
                       int flags, void* workspace, size_t ws_bytes, const fd3_outputs* out, void* stream) {
  if (!cell) {
    set_error("device cell pointer is NULL");
    return FD3_ERR_ARG;
  }
  if (g && g->interp == FD3_INTERP_EWALD) {
    set_error("the device-cell entry point supports the mesh methods (spme, pme)");
    return FD3_ERR_UNSUPPORTED;
  }
  return eval_edges(m, g, reach, cell, status, n, positions, pos_dtype, species, ne, edge_src, edge_dst, shifts,
                    shifts_dtype, flags, workspace, ws_bytes, out, (hipStream_t)stream);
}

// --------------------------------------------------------- cell-list path ---
static int cells_layout(const fd3_model* m, const fd3_geom* g, int64_t n, double reach, bool dyn, Common& c,
                        CellGrid& cg, CellBuffers& cb, Arena& ar, CellDyn** cd, int32_t** st) {
  int rc;
  if ((rc = setup_common(m, g, c))) return rc;
  if (n < 0) {
    set_error("negative atom count");
    return FD3_ERR_ARG;
  }
  if ((rc = cell_grid_make(g->cell, m->r_cut, &cg, reach))) return rc;
  c.geom.kreach = reach;
  carve_common(ar, c, n, m->n_rank, cg.nbins);
  cell_carve(cg, std::max<int64_t>(n, 1), ar, &cb);
  if (dyn) {
    *cd = ar.take<CellDyn>(1);
    *st = ar.take<int32_t>(1);
  }
  return FD3_OK;
}

int fd3_eval_celllist_workspace(const fd3_model* m, const fd3_geom* g, int64_t n_atoms, int flags, size_t* bytes) {
  (void)flags;
  Common c;
  CellGrid cg;
  CellBuffers cb;
  Arena ar;
  CellDyn* cd = nullptr;
  int32_t* st = nullptr;
  int rc = cells_layout(m, g, n_atoms, 1.0, false, c, cg, cb, ar, &cd, &st);
  if (rc) return rc;
  *bytes = align_up(ar.off);
  return FD3_OK;
}

int fd3_eval_celllist_npt_workspace(const fd3_model* m, const fd3_geom* g, double reach, int64_t n_atoms, int flags,
                                    size_t* bytes) {
  (void)flags;
  Common c;
  CellGrid cg;
  CellBuffers cb;
  Arena ar;
  CellDyn* cd = nullptr;
  int32_t* st = nullptr;
  int rc = cells_layout(m, g, n_atoms, reach, true, c, cg, cb, ar, &cd, &st);
  if (rc) return rc;
  *bytes = align_up(ar.off);
  return FD3_OK;
}

// The cell-list pipeline; with cell_dev the cell comes from device memory
// (bins, search ranges and tables planned from g->cell with the reach factor).
static int eval_celllist(const fd3_model* m, const fd3_geom* g, double reach, const double* cell_dev,
                         int32_t* status_user, int64_t n, const void* positions, int pos_dtype,
                         const int32_t* species, int flags, void* workspace, size_t ws_bytes,
                         const fd3_outputs* out, hipStream_t s) {
  Common c;
  CellGrid cg;
  CellBuffers cb;
  Arena ar(workspace, ws_bytes);
  CellDyn* cd = nullptr;
  int32_t* st = nullptr;
  int rc = cells_layout(m, g, n, reach, cell_dev != nullptr, c, cg, cb, ar, &cd, &st);
  if (rc) return rc;
  if (!ar.ok() || !workspace) {
    set_error("workspace of %zu bytes is smaller than the %zu required", ws_bytes, align_up(ar.off));
    return FD3_ERR_WORKSPACE;
  }
  if (!out) {
    set_error("outputs are NULL");
    return FD3_ERR_ARG;
  }
  if (n > 0 && (!positions || !species)) {
    set_error("input pointer is NULL");
    return FD3_ERR_ARG;
  }
  const bool grad = flags & FD3_NEED_GRAD;
  if (cell_dev) {
    CellCheck ck;
    for (int a = 0; a < 3; ++a) { ck.nb[a] = cg.nb[a]; ck.sr[a] = cg.sr[a]; }
    ck.rc = cg.rc;
    ck.kmax_tab = c.plan.ks_reg ? ks_gtab_dk(c.plan, c.geom) * kGtabN : 0.0;
    if ((rc = launch_cell_update(c, ck, cell_dev, cd, st, status_user, s))) return rc;
    cg.dyn = cd;
  }
  if ((rc = cell_build(cg, cb, n, positions, pos_dtype, species, s))) return rc;
  const CellArgs ca = cell_args(cg, cb, n);
  CnArgs cna{};
  cna.rcov = m->rcov; cna.rcut = m->r_cut * kA2B; cna.cn = c.cn_raw; cna.dEdcn = c.dEdcn;
  cna.gpos = c.gpos; cna.partials = c.p_cn;
  if ((rc = launch_cn_cells(ca, cna, cg.nbins, s))) return rc;
  if ((rc = run_mesh_path(m, c, n, cb.pos4, cb.spec_s, cb.slot_atom, flags, out, s))) return rc;
  if (grad && n > 0) {
    // the CN backward adds the last term of dE/dr and writes grad_pos (atom order, per Angstrom)
    cna.gout = out->grad_pos; cna.slot_atom = cb.slot_atom;
    if ((rc = launch_cn_bwd_cells(ca, cna, cg.nbins, s))) return rc;
    if ((rc = launch_poison(c, out->grad_pos, n, nullptr, 0, s))) return rc;
  }
  fd3_outputs o2 = *out;
  o2.grad_cell = nullptr;  // the cell-list path reports the strain derivative only
  return run_finalize(c, grad, (flags & FD3_NEED_PARAMS) != 0, n > 0, &o2, s);
}

int fd3_eval_celllist(const fd3_model* m, const fd3_geom* g, int64_t n, const void* positions, int pos_dtype,
                      const int32_t* species, int flags, void* workspace, size_t ws_bytes, const fd3_outputs* out,
                      void* stream) {
  return eval_celllist(m, g, 1.0, nullptr, nullptr, n, positions, pos_dtype, species, flags, workspace, ws_bytes,
                       out, (hipStream_t)stream);
}

int fd3_eval_celllist_npt(const fd3_model* m, const fd3_geom* g, double reach, const double* cell, int32_t* status,
                          int64_t n, const void* positions, int pos_dtype, const int32_t* species, int flags,
                          void* workspace, size_t ws_bytes, const fd3_outputs* out, void* stream) {
  if (!cell) {
    set_error("device cell pointer is NULL");
    return FD3_ERR_ARG;
  }
  if (g && g->interp == FD3_INTERP_EWALD) {
    set_error("the device-cell entry point supports the mesh methods (spme, pme)");
    return FD3_ERR_UNSUPPORTED;
  }
  return eval_celllist(m, g, reach, cell, status, n, positions, pos_dtype, species, flags, workspace, ws_bytes, out,
                       (hipStream_t)stream);
}

// ------------------------------------------------------ real-space sum ---
// The pair energy summed over lattice images within r_max (direct.hip), with
// the cell-list CN and the same CN -> C6 chain and adjoint as the mesh path.
constexpr int64_t kDirectMaxImages = (int64_t)1 << 22;

static int direct_layout(const fd3_model* m, const double cell_A[9], double r_max_A, int64_t n, Common& c,
                         CellGrid& cg, CellBuffers& cb, Arena& ar, DirectArgs& da) {
  int rc;
  if ((rc = check_model(m))) return rc;
  if ((rc = make_cell(cell_A, &c.geom))) return rc;
  if (n < 0) {
    set_error("negative atom count");
    return FD3_ERR_ARG;
  }
  if (!(r_max_A > 0.0) || !std::isfinite(r_max_A)) {
    set_error("r_max must be positive and finite, got %g", r_max_A);
    return FD3_ERR_ARG;
  }
  da = DirectArgs{};
  da.rmax = r_max_A * kA2B;
  direct_images(c.geom.cell, c.geom.inv, da.rmax, da.nimg);
  const int64_t ncube = (int64_t)(2 * da.nimg[0] + 1) * (2 * da.nimg[1] + 1) * (2 * da.nimg[2] + 1);
  if (ncube > kDirectMaxImages) {
    set_error("r_max %g A needs %lld lattice images for this cell (at most %lld)", r_max_A, (long long)ncube,
              (long long)kDirectMaxImages);
    return FD3_ERR_UNSUPPORTED;
  }
  if ((rc = cell_grid_make(cell_A, m->r_cut, &cg, 1.0))) return rc;
  const int64_t n1 = std::max<int64_t>(n, 1);
  const int r = m->n_rank, nS = m->n_species;
  c.cn_raw = ar.take<double>(n1);
  c.charges = ar.take<float>(n1 * r);
  c.dEdcn = ar.take<double>(n1);
  c.gpos = ar.take<double>(3 * n1);
  c.n_c6 = c6_blocks(n);
  c.p_c6 = ar.take<double>((size_t)std::max(c.n_c6, 1) * 8);
  c.n_cn = std::max(cg.nbins, 1);
  c.p_cn = ar.take<double>((size_t)c.n_cn * 8);
  da.tab = ar.take<double>((size_t)4 * nS * nS);
  da.qsum = ar.take<double>((size_t)nS * r);
  da.rows = ar.take<double>((size_t)16 * n1);
  da.dEdq = ar.take<double>((size_t)r * n1);
  cell_carve(cg, n1, ar, &cb);
  return FD3_OK;
}

int fd3_eval_direct_workspace(const fd3_model* m, const double cell[9], double r_max, int64_t n_atoms, int flags,
                              size_t* bytes) {
  (void)flags;
  if (!bytes) {
    set_error("bytes is NULL");
    return FD3_ERR_ARG;
  }
  Common c;
  CellGrid cg;
  CellBuffers cb;
  Arena ar;
  DirectArgs da;
  const int rc = direct_layout(m, cell, r_max, n_atoms, c, cg, cb, ar, da);
  if (rc) return rc;
  *bytes = align_up(ar.off);
  return FD3_OK;
}

int fd3_eval_direct(const fd3_model* m, const double cell[9], double r_max, int64_t n, const void* positions,
                    int pos_dtype, const int32_t* species, int flags, void* workspace, size_t ws_bytes,
                    const fd3_outputs* out, void* stream) {
  const hipStream_t s = (hipStream_t)stream;
  Common c;
  CellGrid cg;
  CellBuffers cb;
  Arena ar(workspace, ws_bytes);
  DirectArgs da;
  int rc = direct_layout(m, cell, r_max, n, c, cg, cb, ar, da);
  if (rc) return rc;
  if (!ar.ok() || !workspace) {
    set_error("workspace of %zu bytes is smaller than the %zu required", ws_bytes, align_up(ar.off));
    return FD3_ERR_WORKSPACE;
  }
  if (!out || !out->energy) {
    set_error("outputs or the energy pointer are NULL");
    return FD3_ERR_ARG;
  }
  if (n > 0 && (!positions || !species)) {
    set_error("input pointer is NULL");
    return FD3_ERR_ARG;
  }
  const bool grad = flags & FD3_NEED_GRAD;
  if ((rc = cell_build(cg, cb, n, positions, pos_dtype, species, s))) return rc;
  const CellArgs ca = cell_args(cg, cb, n);
  CnArgs cna{};
  cna.rcov = m->rcov; cna.rcut = m->r_cut * kA2B; cna.cn = c.cn_raw; cna.dEdcn = c.dEdcn;
  cna.gpos = c.gpos; cna.partials = c.p_cn;
  if ((rc = launch_cn_cells(ca, cna, cg.nbins, s))) return rc;
  const int r = m->n_rank, nS = m->n_species;
  if (n > 0) {
    C6Args c6a;
    c6a.n = n; c6a.nS = nS; c6a.r = r; c6a.cn_raw = c.cn_raw; c6a.spec_s = cb.spec_s; c6a.slot_atom = cb.slot_atom;
    c6a.cnref = m->cnref; c6a.vq = m->vq; c6a.eigs = m->eigs; c6a.qq = m->qq; c6a.cncorr = m->cncorr;
    c6a.s6 = m->s6; c6a.s8 = m->s8; c6a.a1 = m->a1; c6a.a2 = m->a2;
    c6a.charges = c.charges; c6a.cn_out = out->cn; c6a.partials = c.p_c6;
    if ((rc = launch_c6_fwd(c6a, s))) return rc;
  }
  da.n = n; da.nS = nS; da.r = r;
  for (int i = 0; i < 9; ++i) da.cell[i] = c.geom.cell[i];
  da.volume = c.geom.volume;
  da.pos4 = cb.pos4; da.spec_s = cb.spec_s; da.charges = c.charges; da.eigs = m->eigs; da.qq = m->qq;
  da.s6 = m->s6; da.s8 = m->s8; da.a1 = m->a1; da.a2 = m->a2;
  da.gpos = c.gpos;
  if ((rc = launch_direct(da, s))) return rc;
  if (grad && n > 0) {
    C6BwdArgs ba{};
    ba.n = n; ba.nS = nS; ba.r = r; ba.nx = 1; ba.ny = 1; ba.nz = 1;
    for (int i = 0; i < 9; ++i) ba.inv[i] = c.geom.inv[i];
    ba.pos4 = cb.pos4; ba.spec_s = cb.spec_s; ba.charges = c.charges;
    ba.cn_raw = c.cn_raw; ba.cnref = m->cnref; ba.vq = m->vq; ba.eigs = m->eigs; ba.qq = m->qq;
    ba.cncorr = m->cncorr; ba.s6 = m->s6; ba.s8 = m->s8; ba.a1 = m->a1; ba.a2 = m->a2;
    ba.gpos = c.gpos; ba.dEdcn = c.dEdcn; ba.partials = nullptr; ba.dEdq = da.dEdq;
    if ((rc = launch_c6_bwd(ba, s))) return rc;
    // the CN backward adds the CN term of dE/dr and writes grad_pos (atom order, per Angstrom)
    cna.gout = out->grad_pos; cna.slot_atom = cb.slot_atom;
    if ((rc = launch_cn_bwd_cells(ca, cna, cg.nbins, s))) return rc;
  }
  da.cnv = (grad && n > 0) ? c.p_cn : nullptr;
  da.ncnv = c.n_cn;
  da.energy = out->energy;
  da.virial = grad ? out->virial : nullptr;
  da.grad_params = (flags & FD3_NEED_PARAMS) ? out->grad_params : nullptr;
  return launch_direct_finalize(da, s);
}

// ---------------------------------------------------------- batched path ---
static int batch_check(int32_t nb, const fd3_geom* geoms, const int64_t* offs) {
  if (nb < 0 || (nb > 0 && (!geoms || !offs))) {
    set_error("bad batch description (n_boxes %d)", nb);
    return FD3_ERR_ARG;
  }
  if (nb > 0 && offs[0] != 0) {
    set_error("atom_offsets[0] must be 0");
    return FD3_ERR_ARG;
  }
  for (int32_t b = 0; b < nb; ++b)
    if (offs[b + 1] < offs[b]) {
      set_error("atom_offsets must be non-decreasing (box %d)", b);
      return FD3_ERR_ARG;
    }
  return FD3_OK;
}

int fd3_eval_celllist_batch_workspace(const fd3_model* m, int32_t n_boxes, const fd3_geom* geoms,
                                      const int64_t* atom_offsets, int flags, size_t* bytes) {
  int rc;
  if ((rc = batch_check(n_boxes, geoms, atom_offsets))) return rc;
  size_t mx = 256;
  for (int32_t b = 0; b < n_boxes; ++b) {
    size_t sz = 0;
    if ((rc = fd3_eval_celllist_workspace(m, &geoms[b], atom_offsets[b + 1] - atom_offsets[b], flags, &sz))) return rc;
    mx = std::max(mx, sz);
  }
  *bytes = mx;
  return FD3_OK;
}

int fd3_eval_celllist_batch(const fd3_model* m, int32_t n_boxes, const fd3_geom* geoms, const int64_t* atom_offsets,
                            const void* positions, int pos_dtype, const int32_t* species, int flags, void* workspace,
                            size_t workspace_bytes, const fd3_outputs* outs, void* stream) {
  int rc;
  if ((rc = batch_check(n_boxes, geoms, atom_offsets))) return rc;
  if (n_boxes > 0 && !outs) {
    set_error("outputs are NULL");
    return FD3_ERR_ARG;
  }
  const size_t esz = pos_dtype == FD3_DTYPE_F64 ? sizeof(double) : sizeof(float);
  for (int32_t b = 0; b < n_boxes; ++b) {
    const int64_t a0 = atom_offsets[b], n = atom_offsets[b + 1] - a0;
    const void* pb = positions ? (const char*)positions + (size_t)a0 * 3 * esz : nullptr;
    const int32_t* sb = species ? species + a0 : nullptr;
    // boxes run one after another on the stream, so they share the workspace
    if ((rc = fd3_eval_celllist(m, &geoms[b], n, pb, pos_dtype, sb, flags, workspace, workspace_bytes, &outs[b],
                                stream)))
      return rc;
  }
  return FD3_OK;
}

// ------------------------------------------------------------ neighbours ---
static int nb_layout(int64_t n, const double cell[9], double r_cut, CellGrid& cg, CellBuffers& cb, Arena& ar,
                     int32_t** count_atom, int64_t** offset) {
  if (n < 0) {
    set_error("negative atom count");
    return FD3_ERR_ARG;
  }
  int rc;
  if ((rc = cell_grid_make(cell, r_cut, &cg))) return rc;
  cell_carve(cg, std::max<int64_t>(n, 1), ar, &cb);
  *count_atom = ar.take<int32_t>(std::max<int64_t>(n, 1));
  *offset = ar.take<int64_t>(n + 1);
  return FD3_OK;
}

int fd3_neighbours_workspace(int64_t n, const double cell[9], double r_cut, size_t* bytes) {
  CellGrid cg;
  CellBuffers cb;
  Arena ar;
  int32_t* ca;
  int64_t* of;
  int rc = nb_layout(n, cell, r_cut, cg, cb, ar, &ca, &of);
  if (rc) return rc;
  *bytes = align_up(ar.off);
  return FD3_OK;
}

static int nb_prepare(int64_t n, const void* pos, int dtype, const double cell[9], double r_cut, void* ws,
                      size_t ws_bytes, int64_t* total, CellGrid& cg, CellBuffers& cb, int64_t** offset,
                      hipStream_t s) {
  Arena ar(ws, ws_bytes);
  int32_t* count_atom;
  int rc = nb_layout(n, cell, r_cut, cg, cb, ar, &count_atom, offset);
  if (rc) return rc;
  if (!ar.ok() || !ws) {
    set_error("workspace of %zu bytes is smaller than the %zu required", ws_bytes, align_up(ar.off));
    return FD3_ERR_WORKSPACE;
  }
  if (n == 0) {
    if (total) return launch_fill_i64(total, 1, s);
    return FD3_OK;
  }
  if ((rc = cell_build(cg, cb, n, pos, dtype, nullptr, s))) return rc;
  const CellArgs ca = cell_args(cg, cb, n);
  return nb_count(ca, count_atom, *offset, total, s);
}

int fd3_neighbours_count(int64_t n, const void* pos, int dtype, const double cell[9], double r_cut, void* ws,
                         size_t ws_bytes, int64_t* n_pairs, void* stream) {
  CellGrid cg;
  CellBuffers cb;
  int64_t* offset;
  return nb_prepare(n, pos, dtype, cell, r_cut, ws, ws_bytes, n_pairs, cg, cb, &offset, (hipStream_t)stream);
}

int fd3_neighbours_fill(int64_t n, const void* pos, int dtype, const double cell[9], double r_cut, void* ws,
                        size_t ws_bytes, int64_t capacity, int64_t* src, int64_t* dst, int32_t* shifts,
                        void* stream) {
  CellGrid cg;
  CellBuffers cb;
  int64_t* offset;
  hipStream_t s = (hipStream_t)stream;
  int rc = nb_prepare(n, pos, dtype, cell, r_cut, ws, ws_bytes, nullptr, cg, cb, &offset, s);
  if (rc || n == 0) return rc;
  return nb_fill(cell_args(cg, cb, n), offset, capacity, src, dst, shifts, s);
}

// ------------------------------------------------------ spread / gather ---
int fd3_spread(const fd3_geom* g, int32_t nS, int32_t r, int64_t n, const void* positions, int pos_dtype,
               const int32_t* species, const float* charges, float* rho, void* stream) {
  Geom geom;
  int rc;
  if ((rc = make_geom(g, &geom))) return rc;
  if (nS < 1 || r < 1) {
    set_error("n_species and n_rank must be positive");
    return FD3_ERR_ARG;
  }
  hipStream_t s = (hipStream_t)stream;
  const size_t M = (size_t)geom.nx * geom.ny * geom.nz;
  if ((rc = launch_fill_f32(rho, M * nS * r, s))) return rc;
  SpreadArgs sa;
  sa.g = mesh_geom(geom); sa.n = n; sa.r = r; sa.pos4 = nullptr; sa.raw_pos = positions; sa.raw_dtype = pos_dtype;
  sa.spec_s = species; sa.charges = charges; sa.rho = rho;
  return launch_spread(sa, s);
}

int fd3_gather(const fd3_geom* g, int32_t nS, int32_t r, int64_t n, const void* positions, int pos_dtype,
               const int32_t* species, const float* mesh, float* out, void* stream) {
  Geom geom;
  int rc;
  if ((rc = make_geom(g, &geom))) return rc;
  if (nS < 1 || r < 1) {
    set_error("n_species and n_rank must be positive");
    return FD3_ERR_ARG;
  }
  GatherArgs ga;
  std::memset(&ga, 0, sizeof(ga));
  ga.g = mesh_geom(geom); ga.n = n; ga.nS = nS; ga.r = r; ga.pos4 = nullptr; ga.raw_pos = positions;
  ga.raw_dtype = pos_dtype; ga.spec_s = species; ga.mesh = mesh;
  return launch_gather_plain(ga, out, (hipStream_t)stream);
}

// ------------------------------------------------------------------ FFT ---
static int fft_layout(const int32_t ns[3], int32_t C, FftPlan& p, FftTables& t, Arena& ar, float2** A,
                      float2** copy) {
  if (C < 1) {
    set_error("n_channels must be positive");
    return FD3_ERR_ARG;
  }
  int rc;
  if ((rc = fft_plan_make(ns, 1, C, &p))) return rc;
  fft_tables_carve(p, ar, &t);
  *A = ar.take<float2>(p.mesh_complex());
  *copy = ar.take<float2>(p.mesh_complex());
  return FD3_OK;
}

int fd3_fft_workspace(const int32_t ns[3], int32_t C, size_t* bytes) {
  FftPlan p;
  FftTables t;
  Arena ar;
  float2 *A, *cp;
  int rc = fft_layout(ns, C, p, t, ar, &A, &cp);
  if (rc) return rc;
  *bytes = align_up(ar.off);
  return FD3_OK;
}

int fd3_fft_r2c(const int32_t ns[3], int32_t C, const float* rho, float* spec, void* ws, size_t ws_bytes,
                void* stream) {
  FftPlan p;
  FftTables t;
  Arena ar(ws, ws_bytes);
  float2 *A, *cp;
  int rc = fft_layout(ns, C, p, t, ar, &A, &cp);
  if (rc) return rc;
  if (!ar.ok() || !ws) {
    set_error("FFT workspace too small");
    return FD3_ERR_WORKSPACE;
  }
  hipStream_t s = (hipStream_t)stream;
  if ((rc = fft_tables_init(p, t, 1, 0, s))) return rc;
  if ((rc = fft_forward_zy(p, t, rho, A, (float2*)spec, s))) return rc;
  return fft_x(p, t, (float2*)spec, -1, s);
}

int fd3_fft_c2r(const int32_t ns[3], int32_t C, const float* spec, float* out, void* ws, size_t ws_bytes,
                void* stream) {
  FftPlan p;
  FftTables t;
  Arena ar(ws, ws_bytes);
  float2 *A, *cp;
  int rc = fft_layout(ns, C, p, t, ar, &A, &cp);
  if (rc) return rc;
  if (!ar.ok() || !ws) {
    set_error("FFT workspace too small");
    return FD3_ERR_WORKSPACE;
  }
  hipStream_t s = (hipStream_t)stream;
  if ((rc = fft_tables_init(p, t, 1, 0, s))) return rc;
  FD3_HIP(hipMemcpyAsync(cp, spec, p.mesh_complex() * sizeof(float2), hipMemcpyDeviceToDevice, s));
  if ((rc = fft_x(p, t, cp, 1, s))) return rc;
  return fft_inverse_yz(p, t, cp, A, out, s);
}

// ------------------------------------------------------------ host checks ---
int fd3_green_host(double s6, double s8, double r0, double qq, int64_t n, const double* k, float* g, float* dg) {
  // R0 = a1 sqrt(3 qq) + a2 with a1 = 0, a2 = r0: the kernel's pair constants for this R0
  const KsPair pk = ks_pair((float)qq, (float)s6, (float)s8, 0.0f, (float)r0);
  float4 tab[3 * (kHtabN + 1)];  // the table the k-space kernel reads, built by the same code
  for (int i = 0; i <= kHtabN; ++i) htab_interval(i, tab + 3 * i);
  for (int64_t i = 0; i < n; ++i) {
    float h6, d6, h8, d8;
    green_eval_tab(pk, tab, (float)k[i], g + i, dg + i, &h6, &d6, &h8, &d8);
  }
  return FD3_OK;
}

int fd3_green_table_host(double s6, double s8, double r0, double qq, double kmax, int64_t n, const double* k,
                         float* g, float* dg) {
  if (!(kmax > 0.0)) {
    set_error("kmax must be positive");
    return FD3_ERR_ARG;
  }
  const KsPair pk = ks_pair((float)qq, (float)s6, (float)s8, 0.0f, (float)r0);
  const double dk = kmax * (1.0 + 1e-6) / kGtabN;
  const float kinv = (float)(1.0 / dk);
  float4 tab[2 * kGtabN];  // the table of this pair, built by the kernels' code
  for (int i = 0; i < kGtabN; ++i) gtab_interval(dk, i, (double)pk.Rs * kHtabDx, pk.p6, pk.p8, tab + 2 * i);
  for (int64_t j = 0; j < n; ++j) {
    const float fk = (float)k[j] * kinv;
    const int i = std::min((int)fk, kGtabN - 1);
    float dgdw;
    gtab_eval(tab[2 * i], tab[2 * i + 1], fk - (float)i, g + j, &dgdw);
    dg[j] = dgdw * kinv;
  }
  return FD3_OK;
}

int fd3_interp_weights_host(int32_t order, int32_t interp, int64_t n, const double* x, float* w, float* dw) {
  if (order < 1 || order > kMaxOrder || (interp == FD3_INTERP_P3M && order > 5)) {
    set_error("unsupported interpolation order %d", order);
    return FD3_ERR_UNSUPPORTED;
  }
  float ww[kMaxOrder], dd[kMaxOrder];
  for (int64_t i = 0; i < n; ++i) {
    interp_weights<float>(order, interp, (float)x[i], ww, dd);
    for (int j = 0; j < order; ++j) {
      w[j * n + i] = ww[j];
      dw[j * n + i] = dd[j];
    }
  }
  return FD3_OK;
}

}  // extern "C"
