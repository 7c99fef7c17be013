// Kernel launches that serve one box or many.
//
// Every kernel of the cell-list pipeline is written as a __device__ body
//   void BODY(const BlockInfo& bk, const P& a)
// that reads its block index and grid size from bk instead of blockIdx and
// gridDim, with all its arguments in one struct P.  Two __global__ wrappers
// run it:
//   k_one<P, BODY, NT>   one box: bk = (blockIdx, gridDim), a = the kernel argument;
//   k_many<P, BODY, NT>  many boxes in one grid: a 1-D grid of the boxes'
//                        blocks one after another; each block finds its box
//                        in a device table of block offsets (binary search),
//                        rebuilds its (x, y, z) index in that box's grid and
//                        reads that box's P from the table.
// fd3::launch() launches k_one, or, while a Recorder is installed on the
// calling thread, appends the launch (its grid, LDS size and argument bytes)
// to the recorder instead.  A batch records each box's pipeline, then merges
// the boxes' launches of the same kernel into one k_many launch per stage.
#pragma once

#include <cstring>
#include <vector>

#include "fd3_internal.h"

namespace fd3 {

struct BlockInfo {
  uint3 id;  // block index within this box's grid
  uint3 n;   // this box's grid size
};

// Per-box tables of one batched launch (device memory).
struct BatchTable {
  const char* args;    // nbox * stride bytes: each box's argument struct
  const int* bstart;   // nbox + 1: first block of each box in the 1-D grid
  const int4* grids;   // nbox: each box's grid (x, y, z)
  int nbox, stride;
};

template <class P>
__device__ inline const P* batch_locate(const BatchTable& t, BlockInfo* bk) {
  const int b = (int)blockIdx.x;
  int lo = 0, hi = t.nbox - 1;
  while (lo < hi) {  // the last box whose first block is <= b (uniform per block)
    const int mid = (lo + hi + 1) >> 1;
    if (t.bstart[mid] <= b) lo = mid;
    else hi = mid - 1;
  }
  const int local = b - t.bstart[lo];
  const int4 g = t.grids[lo];
  bk->n = make_uint3((unsigned)g.x, (unsigned)g.y, (unsigned)g.z);
  bk->id = make_uint3((unsigned)(local % g.x), (unsigned)((local / g.x) % g.y), (unsigned)(local / (g.x * g.y)));
  return (const P*)(t.args + (size_t)lo * t.stride);
}

template <class P, void (*BODY)(const BlockInfo&, const P&), int NT>
__global__ void __launch_bounds__(NT) k_one(P a) {
  const BlockInfo bk{make_uint3(blockIdx.x, blockIdx.y, blockIdx.z), make_uint3(gridDim.x, gridDim.y, gridDim.z)};
  BODY(bk, a);
}

template <class P, void (*BODY)(const BlockInfo&, const P&), int NT>
__global__ void __launch_bounds__(NT) k_many(BatchTable t) {
  BlockInfo bk;
  const P a = *batch_locate<P>(t, &bk);  // one copy, read before any store
  BODY(bk, a);
}

// The same wrappers for bodies held to at least WPE waves per SIMD (at most
// 512 / WPE VGPRs): WPE = 4 (128 VGPRs) or 8 (64 VGPRs).
template <class P, void (*BODY)(const BlockInfo&, const P&), int NT, int WPE>
__global__ void __launch_bounds__(NT) __attribute__((amdgpu_waves_per_eu(WPE, 8))) k_one_w(P a) {
  const BlockInfo bk{make_uint3(blockIdx.x, blockIdx.y, blockIdx.z), make_uint3(gridDim.x, gridDim.y, gridDim.z)};
  BODY(bk, a);
}

template <class P, void (*BODY)(const BlockInfo&, const P&), int NT, int WPE>
__global__ void __launch_bounds__(NT) __attribute__((amdgpu_waves_per_eu(WPE, 8))) k_many_w(BatchTable t) {
  BlockInfo bk;
  const P a = *batch_locate<P>(t, &bk);
  BODY(bk, a);
}

// ------------------------------------------------------------ recording ---
using ManyLauncher = int (*)(const BatchTable& t, unsigned nblocks, size_t sh, hipStream_t s);

struct LaunchRec {
  const char* tag;      // launch site (the profiling name)
  const void* key;      // the single-box kernel: launches with equal keys merge
  ManyLauncher many;
  dim3 grid;
  int nt;
  size_t sh;
  size_t arg_off, arg_size;  // into Recorder::pool
};

struct Recorder {
  std::vector<LaunchRec> recs;
  std::vector<char> pool;
  void add(const char* tag, const void* key, ManyLauncher many, dim3 grid, int nt, size_t sh, const void* args,
           size_t size) {
    const size_t off = pool.size();
    pool.resize(off + size);
    std::memcpy(pool.data() + off, args, size);
    recs.push_back(LaunchRec{tag, key, many, grid, nt, sh, off, size});
  }
};

// The recorder installed on the calling thread (nullptr: launch for real).
Recorder*& current_recorder();

// Raise a kernel's dynamic-LDS limit once per kernel and size (above 64 KB).
int lds_attr(const void* fn, size_t sh);

template <class P, void (*BODY)(const BlockInfo&, const P&), int NT, int WPE>
int launch_many_impl(const BatchTable& t, unsigned nblocks, size_t sh, hipStream_t s) {
  int rc;
  if constexpr (WPE > 0) {
    if ((rc = lds_attr((const void*)k_many_w<P, BODY, NT, WPE>, sh))) return rc;
    hipLaunchKernelGGL((k_many_w<P, BODY, NT, WPE>), dim3(nblocks), dim3(NT), sh, s, t);
  } else {
    if ((rc = lds_attr((const void*)k_many<P, BODY, NT>, sh))) return rc;
    hipLaunchKernelGGL((k_many<P, BODY, NT>), dim3(nblocks), dim3(NT), sh, s, t);
  }
  FD3_LAUNCH_CHECK("batched launch");
  return FD3_OK;
}

// Launch BODY over `grid` blocks of NT threads (or record it); WPE > 0 asks
// the compiler for at least that many waves per SIMD.
template <class P, void (*BODY)(const BlockInfo&, const P&), int NT, int WPE = 0>
int launch(const char* tag, dim3 grid, size_t sh, hipStream_t s, const P& a) {
  static_assert(WPE == 0 || WPE == 4 || WPE == 8, "waves per SIMD: 0, 4 or 8");
  const void* one;
  if constexpr (WPE > 0) one = (const void*)k_one_w<P, BODY, NT, WPE>;
  else one = (const void*)k_one<P, BODY, NT>;
  if (Recorder* r = current_recorder()) {
    r->add(tag, one, &launch_many_impl<P, BODY, NT, WPE>, grid, NT, sh, &a, sizeof(P));
    return FD3_OK;
  }
  int rc;
  if ((rc = lds_attr(one, sh))) return rc;
  if constexpr (WPE > 0) hipLaunchKernelGGL((k_one_w<P, BODY, NT, WPE>), grid, dim3(NT), sh, s, a);
  else hipLaunchKernelGGL((k_one<P, BODY, NT>), grid, dim3(NT), sh, s, a);
  FD3_LAUNCH_CHECK(tag);
  return FD3_OK;
}

}  // namespace fd3
