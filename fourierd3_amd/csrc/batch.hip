// Batched evaluation: many periodic boxes (of one or more species sets) in
// one launch sequence.
//
// A plan records each box's cell-list pipeline (the launches fd3_eval_celllist
// would make, with that box's workspace carved from a shared buffer) and
// merges the boxes' launches stage by stage: launches of the same kernel
// variant from different boxes become one k_many launch whose grid is the
// boxes' blocks one after another, so a stage costs one launch however many
// boxes there are.  Each box keeps its own mesh, cell list, tables and
// partial sums, packed one after another in the workspace; within a box the
// launches keep their order, and boxes never read each other's buffers, so
// any merge that keeps each box's order is exact.
//
// The per-box argument structs, block offsets and grids of every merged
// launch form one table blob, built on the host (pinned memory owned by the
// plan) and copied into the workspace on the stream at the start of each
// run; the blob is rebuilt only when the positions, species, workspace or
// output pointers change, so a run captured in a HIP graph replays with the
// table it was captured with.

#include <algorithm>
#include <cstdio>
#include <cstring>
#include <map>
#include <string>
#include <vector>

#include "fd3_internal.h"
#include "launch.h"

namespace fd3 {

struct MergedLaunch {
  const char* tag;
  ManyLauncher many;
  int nt;
  size_t sh;
  std::vector<std::pair<int, int>> members;  // (box, record index)
  size_t off_args, off_bstart, off_grids;    // into the blob
  unsigned nblocks;
  int stride;
};

}  // namespace fd3

struct fd3_batch {
  std::vector<fd3_model> models;
  std::vector<int32_t> model_of_box;
  std::vector<fd3_geom> geoms;
  std::vector<int64_t> offs;
  int flags = 0;
  std::vector<size_t> box_off, box_bytes;  // per-box workspace regions
  size_t blob_off = 0, blob_cap = 0, ws_bytes = 0;
  char* pinned = nullptr;                  // host copy of the table blob
  // what the current blob was built for
  const void* k_pos = nullptr;
  const void* k_spec = nullptr;
  const void* k_ws = nullptr;
  int k_dtype = -1;
  std::vector<fd3_outputs> k_outs;
  std::vector<fd3::MergedLaunch> launches;
  size_t blob_size = 0;
  int n_records = 0;
};

namespace fd3 {

namespace {

// Records every box's pipeline; on success recs[b] holds box b's launches.
int record_boxes(const fd3_batch& p, const void* pos, int dtype, const int32_t* spec, char* ws,
                 const fd3_outputs* outs, std::vector<Recorder>& recs) {
  const int nb = (int)p.geoms.size();
  recs.assign(nb, Recorder{});
  const size_t esz = dtype == FD3_DTYPE_F64 ? sizeof(double) : sizeof(float);
  for (int b = 0; b < nb; ++b) {
    const int64_t a0 = p.offs[b], n = p.offs[b + 1] - a0;
    const void* pb = pos ? (const char*)pos + (size_t)a0 * 3 * esz : nullptr;
    const int32_t* sb = spec ? spec + a0 : nullptr;
    current_recorder() = &recs[b];
    const int rc = fd3_eval_celllist(&p.models[p.model_of_box[b]], &p.geoms[b], n, pb, dtype, sb, p.flags,
                                     ws + p.box_off[b], p.box_bytes[b], &outs[b], nullptr);
    current_recorder() = nullptr;
    if (rc) {
      char msg[600];
      std::snprintf(msg, sizeof(msg), "box %d: %s", b, fd3_last_error());
      set_error("%s", msg);
      return rc;
    }
  }
  return FD3_OK;
}

// Merge the boxes' launch sequences: repeatedly take, among the boxes' next
// launches, the one whose stage comes first in a reference order (the
// longest box's sequence; a stage is a launch site's tag and its
// occurrence), and launch it together with every box whose next launch is
// the same kernel (and occurrence of it).  Boxes that run another kernel
// variant at that stage (another candidate capacity of the CN kernel, say)
// rank the same and follow right after, so the stages after it still merge.
// Each box's own order is kept whatever the reference order is.
std::vector<MergedLaunch> merge(const std::vector<Recorder>& recs) {
  const int nb = (int)recs.size();
  using Ident = std::pair<const void*, int>;
  using Stage = std::pair<std::string, int>;
  std::vector<std::vector<Ident>> ids(nb);
  std::vector<std::vector<Stage>> stages(nb);
  for (int b = 0; b < nb; ++b) {
    std::map<const void*, int> occ;
    std::map<std::string, int> tocc;
    for (const LaunchRec& r : recs[b].recs) {
      ids[b].push_back({r.key, occ[r.key]++});
      const std::string t(r.tag);
      stages[b].push_back({t, tocc[t]++});
    }
  }
  std::vector<int> order(nb);
  for (int b = 0; b < nb; ++b) order[b] = b;
  std::stable_sort(order.begin(), order.end(),
                   [&](int x, int y) { return recs[x].recs.size() > recs[y].recs.size(); });
  std::map<Stage, int> rank;
  for (int b : order)
    for (const Stage& st : stages[b])
      if (!rank.count(st)) {
        const int r = (int)rank.size();
        rank[st] = r;
      }
  std::vector<size_t> cur(nb, 0);
  std::vector<MergedLaunch> out;
  for (;;) {
    int best = -1, best_rank = 0;
    for (int b = 0; b < nb; ++b) {
      if (cur[b] >= ids[b].size()) continue;
      const int r = rank[stages[b][cur[b]]];
      if (best < 0 || r < best_rank) { best = b; best_rank = r; }
    }
    if (best < 0) break;
    const Ident id = ids[best][cur[best]];
    const LaunchRec& r0 = recs[best].recs[cur[best]];
    MergedLaunch m;
    m.tag = r0.tag; m.many = r0.many; m.nt = r0.nt; m.sh = 0; m.stride = (int)r0.arg_size;
    m.off_args = m.off_bstart = m.off_grids = 0; m.nblocks = 0;
    for (int b = 0; b < nb; ++b) {
      if (cur[b] >= ids[b].size() || ids[b][cur[b]] != id) continue;
      const LaunchRec& r = recs[b].recs[cur[b]];
      m.members.push_back({b, (int)cur[b]});
      m.sh = std::max(m.sh, r.sh);
      ++cur[b];
    }
    out.push_back(std::move(m));
  }
  return out;
}

// Lay out the blob (per merged launch: argument structs, block offsets,
// grids) and fill it into `blob` (nullptr: sizes only).  Returns the size.
size_t build_blob(std::vector<MergedLaunch>& ls, const std::vector<Recorder>& recs, char* blob) {
  size_t off = 0;
  for (MergedLaunch& m : ls) {
    const int nbx = (int)m.members.size();
    m.off_args = align_up(off, 16);
    off = m.off_args + (size_t)nbx * m.stride;
    m.off_bstart = align_up(off, 16);
    off = m.off_bstart + (size_t)(nbx + 1) * sizeof(int);
    m.off_grids = align_up(off, 16);
    off = m.off_grids + (size_t)nbx * sizeof(int4);
    long long total = 0;
    for (int i = 0; i < nbx; ++i) {
      const LaunchRec& r = recs[m.members[i].first].recs[m.members[i].second];
      const long long nblk = (long long)r.grid.x * r.grid.y * r.grid.z;
      if (blob) {
        std::memcpy(blob + m.off_args + (size_t)i * m.stride, recs[m.members[i].first].pool.data() + r.arg_off,
                    r.arg_size);
        ((int*)(blob + m.off_bstart))[i] = (int)total;
        ((int4*)(blob + m.off_grids))[i] = make_int4((int)r.grid.x, (int)r.grid.y, (int)r.grid.z, 0);
      }
      total += nblk;
    }
    if (blob) ((int*)(blob + m.off_bstart))[nbx] = (int)total;
    m.nblocks = (unsigned)total;
  }
  return align_up(off, 256);
}

int check_inputs(int32_t n_models, const fd3_model* models, int32_t nb, const int32_t* mob, const fd3_geom* geoms,
                 const int64_t* offs) {
  if (n_models < 1 || !models) {
    set_error("a batch needs at least one model");
    return FD3_ERR_ARG;
  }
  if (nb < 1 || !mob || !geoms || !offs) {
    set_error("bad batch description (n_boxes %d)", nb);
    return FD3_ERR_ARG;
  }
  if (offs[0] != 0) {
    set_error("atom_offsets[0] must be 0");
    return FD3_ERR_ARG;
  }
  for (int32_t b = 0; b < nb; ++b) {
    if (offs[b + 1] < offs[b]) {
      set_error("atom_offsets must be non-decreasing (box %d)", b);
      return FD3_ERR_ARG;
    }
    if (mob[b] < 0 || mob[b] >= n_models) {
      set_error("box %d uses model %d of %d", b, mob[b], n_models);
      return FD3_ERR_ARG;
    }
    if (geoms[b].interp == FD3_INTERP_EWALD) {
      set_error("the batched form supports the mesh methods (spme, pme); box %d asks for Ewald", b);
      return FD3_ERR_UNSUPPORTED;
    }
  }
  return FD3_OK;
}

}  // namespace
}  // namespace fd3

using namespace fd3;

extern "C" {

int fd3_batch_create(int32_t n_models, const fd3_model* models, int32_t n_boxes, const int32_t* model_of_box,
                     const fd3_geom* geoms, const int64_t* atom_offsets, int flags, fd3_batch** plan) {
  if (!plan) {
    set_error("plan pointer is NULL");
    return FD3_ERR_ARG;
  }
  *plan = nullptr;
  int rc;
  if ((rc = check_inputs(n_models, models, n_boxes, model_of_box, geoms, atom_offsets))) return rc;
  fd3_batch* p = new fd3_batch();
  p->models.assign(models, models + n_models);
  p->model_of_box.assign(model_of_box, model_of_box + n_boxes);
  p->geoms.assign(geoms, geoms + n_boxes);
  p->offs.assign(atom_offsets, atom_offsets + n_boxes + 1);
  p->flags = flags;
  size_t off = 0;
  for (int32_t b = 0; b < n_boxes; ++b) {
    size_t sz = 0;
    rc = fd3_eval_celllist_workspace(&p->models[model_of_box[b]], &geoms[b], atom_offsets[b + 1] - atom_offsets[b],
                                     flags, &sz);
    if (rc) {
      char msg[600];
      std::snprintf(msg, sizeof(msg), "box %d: %s", b, fd3_last_error());
      set_error("%s", msg);
      delete p;
      return rc;
    }
    p->box_off.push_back(off);
    p->box_bytes.push_back(sz);
    off = align_up(off + sz);
  }
  p->blob_off = off;
  // a dry recording (placeholder addresses) fixes the merged launches and the blob size
  std::vector<Recorder> recs;
  std::vector<fd3_outputs> outs(n_boxes);
  for (auto& o : outs) std::memset(&o, 0, sizeof(o));
  char* fake = (char*)(uintptr_t)4096;
  const int64_t ntot = atom_offsets[n_boxes];
  const void* fpos = ntot > 0 ? (const void*)fake : nullptr;
  const int32_t* fspec = ntot > 0 ? (const int32_t*)fake : nullptr;
  if ((rc = record_boxes(*p, fpos, FD3_DTYPE_F64, fspec, fake, outs.data(), recs))) {
    delete p;
    return rc;
  }
  p->launches = merge(recs);
  p->blob_cap = build_blob(p->launches, recs, nullptr);
  p->n_records = 0;
  for (auto& r : recs) p->n_records += (int)r.recs.size();
  p->ws_bytes = p->blob_off + p->blob_cap;
  *plan = p;  // the pinned table copy is allocated by the first run: planning touches no device
  return FD3_OK;
}

int fd3_batch_workspace(const fd3_batch* plan, size_t* bytes) {
  if (!plan || !bytes) {
    set_error("plan or output pointer is NULL");
    return FD3_ERR_ARG;
  }
  *bytes = plan->ws_bytes;
  return FD3_OK;
}

int fd3_batch_info(const fd3_batch* plan, int32_t* merged_launches, int32_t* box_launches) {
  if (!plan) {
    set_error("plan is NULL");
    return FD3_ERR_ARG;
  }
  if (merged_launches) *merged_launches = (int32_t)plan->launches.size();
  if (box_launches) *box_launches = (int32_t)plan->n_records;
  return FD3_OK;
}

int fd3_batch_run(fd3_batch* p, const void* positions, int pos_dtype, const int32_t* species, void* workspace,
                  size_t workspace_bytes, const fd3_outputs* outs, void* stream) {
  if (!p) {
    set_error("plan is NULL");
    return FD3_ERR_ARG;
  }
  if (!workspace || workspace_bytes < p->ws_bytes) {
    set_error("workspace of %zu bytes is smaller than the %zu required", workspace_bytes, p->ws_bytes);
    return FD3_ERR_WORKSPACE;
  }
  if (!outs) {
    set_error("outputs are NULL");
    return FD3_ERR_ARG;
  }
  if (pos_dtype != FD3_DTYPE_F32 && pos_dtype != FD3_DTYPE_F64) {
    set_error("unknown position dtype %d", pos_dtype);
    return FD3_ERR_ARG;
  }
  hipStream_t s = (hipStream_t)stream;
  const int nb = (int)p->geoms.size();
  bool same = p->k_pos == positions && p->k_spec == species && p->k_ws == workspace && p->k_dtype == pos_dtype &&
              (int)p->k_outs.size() == nb;
  for (int b = 0; same && b < nb; ++b) same = std::memcmp(&p->k_outs[b], &outs[b], sizeof(fd3_outputs)) == 0;
  if (!p->pinned) {
    const hipError_t e = hipHostMalloc((void**)&p->pinned, p->blob_cap);
    if (e != hipSuccess) {
      p->pinned = nullptr;
      return hip_fail(e, "hipHostMalloc (batch launch tables)");
    }
  }
  if (!same) {
    std::vector<Recorder> recs;
    int rc = record_boxes(*p, positions, pos_dtype, species, (char*)workspace, outs, recs);
    if (rc) return rc;
    std::vector<MergedLaunch> ls = merge(recs);
    const size_t sz = build_blob(ls, recs, nullptr);
    if (ls.size() != p->launches.size() || sz > p->blob_cap) {
      set_error("the batch recorded a different launch sequence than when it was planned");
      return FD3_ERR_ARG;
    }
    build_blob(ls, recs, p->pinned);
    p->launches = std::move(ls);
    p->blob_size = sz;
    p->k_pos = positions; p->k_spec = species; p->k_ws = workspace; p->k_dtype = pos_dtype;
    p->k_outs.assign(outs, outs + nb);
  }
  char* dev = (char*)workspace + p->blob_off;
  FD3_HIP(hipMemcpyAsync(dev, p->pinned, p->blob_size, hipMemcpyHostToDevice, s));
  for (const MergedLaunch& m : p->launches) {
    if (m.nblocks == 0) continue;
    BatchTable t;
    t.args = dev + m.off_args;
    t.bstart = (const int*)(dev + m.off_bstart);
    t.grids = (const int4*)(dev + m.off_grids);
    t.nbox = (int)m.members.size();
    t.stride = m.stride;
    ProfScope _ps(m.tag, s);
    int rc;
    if ((rc = m.many(t, m.nblocks, m.sh, s))) return rc;
  }
  return FD3_OK;
}

void fd3_batch_destroy(fd3_batch* plan) {
  if (!plan) return;
  if (plan->pinned) (void)hipHostFree(plan->pinned);
  delete plan;
}

}  // extern "C"
