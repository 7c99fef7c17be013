// Host-side checks of libfd3hip built with AddressSanitizer on the host code
// (test infrastructure: tests/test_asan_host.py builds and runs it; no GPU).
// Exercises every entry point's host logic that runs before any device work:
// plans and tile queries (with tuning overrides), workspace queries of all
// paths, the host builds of the Green's-function and interpolation code, the
// batch planner (dry recording and launch merging), and the argument checks.
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "fd3.h"

static int g_fail = 0;
#define CHECK(cond)                                                        \
  do {                                                                     \
    if (!(cond)) {                                                         \
      std::printf("FAIL %s:%d %s (%s)\n", __FILE__, __LINE__, #cond, fd3_last_error()); \
      ++g_fail;                                                            \
    }                                                                      \
  } while (0)

// Never dereferenced by the host code: model arrays live on the device.
static const float* fake_dev(size_t k) { return (const float*)(uintptr_t)(4096 * (k + 1)); }

static fd3_model model(int nS, int r) {
  fd3_model m;
  std::memset(&m, 0, sizeof(m));
  m.n_species = nS;
  m.n_rank = r;
  m.rcov = fake_dev(1); m.cnref = fake_dev(2); m.vq = fake_dev(3); m.eigs = fake_dev(4); m.qq = fake_dev(5);
  m.s6 = 1.0; m.s8 = 0.7875; m.a1 = 0.4289; m.a2 = 4.4407;
  m.r_cut = 6.0;
  return m;
}

static fd3_geom geom(double edge, int ns, int order, int interp) {
  fd3_geom g;
  std::memset(&g, 0, sizeof(g));
  g.cell[0] = edge; g.cell[4] = edge * 1.1; g.cell[8] = edge * 0.9; g.cell[3] = 0.2 * edge;
  g.ns[0] = g.ns[1] = g.ns[2] = ns;
  g.order = order;
  g.interp = interp;
  return g;
}

int main() {
  CHECK(fd3_version() >= 1);
  size_t b = 0;
  // plans and tiles, with and without tuning overrides
  const int meshes[][3] = {{16, 16, 16}, {64, 64, 64}, {256, 256, 256}, {256, 256, 32}, {512, 512, 512},
                           {128, 256, 1024}, {2, 4, 8}, {1024, 8, 2}};
  for (const auto& ns : meshes)
    for (int nS = 1; nS <= 12; nS += 3)
      for (int r = 1; r <= 8; r *= 2) {
        int32_t t[4] = {0, 0, 0, 0};
        const int rc = fd3_plan_tiles(ns, nS, r, t);
        CHECK(rc == FD3_OK || rc == FD3_ERR_UNSUPPORTED);
        if (rc == FD3_OK) CHECK(t[0] >= 1 && t[1] >= 1 && t[2] >= 1 && t[3] >= 1);
        CHECK(fd3_fft_workspace(ns, nS * r, &b) == FD3_OK || fd3_fft_workspace(ns, nS * r, &b) == FD3_ERR_UNSUPPORTED);
      }
  const int32_t ns256[3] = {256, 256, 256}, tiles[4] = {8, 16, 1, 2};
  CHECK(fd3_tune_set(ns256, 4, 2, tiles) == FD3_OK);
  int32_t t[4];
  CHECK(fd3_plan_tiles(ns256, 4, 2, t) == FD3_OK);
  CHECK(fd3_tune_clear() == FD3_OK);
  CHECK(fd3_tune_set(nullptr, 4, 2, tiles) == FD3_ERR_ARG);
  const int32_t bad[3] = {12, 16, 16};
  CHECK(fd3_plan_tiles(bad, 1, 1, t) == FD3_ERR_UNSUPPORTED);
  CHECK(std::strlen(fd3_last_error()) > 0);

  // workspace queries of every path
  const int interps[3] = {FD3_INTERP_P3M, FD3_INTERP_LAGRANGE, FD3_INTERP_EWALD};
  for (int nS : {1, 4, 7})
    for (int ii = 0; ii < 3; ++ii) {
      fd3_model m = model(nS, 3);
      const fd3_geom g = geom(30.0, interps[ii] == FD3_INTERP_EWALD ? 12 : 32, interps[ii] == FD3_INTERP_LAGRANGE ? 6 : 5,
                              interps[ii]);
      for (int flags : {0, FD3_NEED_GRAD, FD3_NEED_GRAD | FD3_NEED_PARAMS | FD3_DETERMINISTIC}) {
        CHECK(fd3_eval_celllist_workspace(&m, &g, 1000, flags, &b) == FD3_OK && b > 0);
        CHECK(fd3_eval_edges_workspace(&m, &g, 1000, 50000, flags, &b) == FD3_OK && b > 0);
        const int rn = fd3_eval_celllist_npt_workspace(&m, &g, 1.1, 1000, flags, &b);
        CHECK(interps[ii] == FD3_INTERP_EWALD || rn == FD3_OK);  // Ewald is refused by the evaluation
        fd3_cn_d4 d4;
        std::memset(&d4, 0, sizeof(d4));
        d4.r0 = fake_dev(6); d4.delta = fake_dev(7); d4.selfcont = fake_dev(8); d4.k0 = 7.5;
        CHECK(fd3_eval_d4_workspace(&m, &d4, &g, 1000, flags, &b) == FD3_OK);
        CHECK(fd3_eval_direct_workspace(&m, g.cell, 20.0, 100, flags, &b) == FD3_OK);
      }
    }
  double cell[9] = {20, 0, 0, 0, 21, 0, 1, 0, 19};
  CHECK(fd3_neighbours_workspace(5000, cell, 6.0, &b) == FD3_OK);
  CHECK(fd3_neighbours_workspace(5000, cell, -1.0, &b) == FD3_ERR_ARG);

  // host builds of the device math
  std::vector<double> k(257);
  for (size_t i = 0; i < k.size(); ++i) k[i] = 0.02 * (double)i;
  std::vector<float> g(k.size()), dg(k.size());
  CHECK(fd3_green_host(1.0, 0.79, 5.3, 2.1, (int64_t)k.size(), k.data(), g.data(), dg.data()) == FD3_OK);
  CHECK(fd3_green_table_host(1.0, 0.79, 5.3, 2.1, 6.0, (int64_t)k.size(), k.data(), g.data(), dg.data()) == FD3_OK);
  for (size_t i = 0; i < k.size(); ++i) CHECK(std::isfinite(g[i]) && std::isfinite(dg[i]));
  std::vector<double> x(64);
  for (size_t i = 0; i < x.size(); ++i) x[i] = (double)i / 64.0;
  for (int order = 1; order <= 7; ++order)
    for (int interp : {FD3_INTERP_P3M, FD3_INTERP_LAGRANGE}) {
      std::vector<float> w(x.size() * order), dw(x.size() * order);
      const int rc = fd3_interp_weights_host(order, interp, (int64_t)x.size(), x.data(), w.data(), dw.data());
      CHECK(rc == FD3_OK || rc == FD3_ERR_UNSUPPORTED);
    }

  // batch planner: dry recording of many boxes of two models, launch merging
  {
    std::vector<fd3_model> models = {model(2, 2), model(4, 3)};
    const int nb = 40;
    std::vector<int32_t> mob(nb);
    std::vector<fd3_geom> geoms(nb);
    std::vector<int64_t> offs(nb + 1, 0);
    for (int i = 0; i < nb; ++i) {
      mob[i] = i % 2;
      const int ns = 16 << (i % 3);
      geoms[i] = geom(10.0 + i, ns, 5, i % 4 == 3 ? FD3_INTERP_LAGRANGE : FD3_INTERP_P3M);
      offs[i + 1] = offs[i] + 50 + 37 * i;
    }
    fd3_batch* plan = nullptr;
    CHECK(fd3_batch_create(2, models.data(), nb, mob.data(), geoms.data(), offs.data(), FD3_NEED_GRAD, &plan) ==
          FD3_OK);
    if (plan) {
      int32_t merged = 0, single = 0;
      CHECK(fd3_batch_info(plan, &merged, &single) == FD3_OK && merged > 0 && single > merged);
      CHECK(fd3_batch_workspace(plan, &b) == FD3_OK && b > 0);
      fd3_batch_destroy(plan);
    }
    mob[3] = 5;  // no such model
    CHECK(fd3_batch_create(2, models.data(), nb, mob.data(), geoms.data(), offs.data(), 0, &plan) == FD3_ERR_ARG);
    CHECK(plan == nullptr);
    CHECK(fd3_eval_celllist_batch_workspace(&models[0], 3, geoms.data(), offs.data(), 0, &b) == FD3_OK);
  }

  // argument checks before any device work
  fd3_model m = model(4, 2);
  fd3_geom g4 = geom(20.0, 16, 5, FD3_INTERP_P3M);
  fd3_outputs o;
  std::memset(&o, 0, sizeof(o));
  CHECK(fd3_eval_celllist(&m, &g4, 10, nullptr, FD3_DTYPE_F64, nullptr, 0, nullptr, 0, &o, nullptr) != FD3_OK);
  CHECK(fd3_eval_celllist(nullptr, &g4, 10, nullptr, FD3_DTYPE_F64, nullptr, 0, nullptr, 0, &o, nullptr) ==
        FD3_ERR_ARG);
  fd3_model m0 = model(0, 1);
  CHECK(fd3_eval_celllist_workspace(&m0, &g4, 10, 0, &b) == FD3_ERR_ARG);
  fd3_model m65 = model(65, 1);
  CHECK(fd3_eval_celllist_workspace(&m65, &g4, 10, 0, &b) != FD3_OK);
  CHECK(fd3_eval_celllist_npt_workspace(&m, &g4, 0.5, 10, 0, &b) == FD3_ERR_ARG);
  CHECK(fd3_eval_celllist_noalloc(&m, &g4, 10, nullptr, FD3_DTYPE_F64, nullptr, 0, &o, nullptr) != FD3_OK);

  std::printf("host checks: %d failure(s)\n", g_fail);
  return g_fail == 0 ? 0 : 1;
}
