// Allocation-free form of fd3_eval_celllist.
//
// The workspace of the cell-list pipeline lives in a file-scope device array
// (`scratch`, kScratchBytes) instead of a caller buffer, for callers that
// cannot allocate device memory at all.  Each GPU the library runs on holds
// its own copy of the array until the process ends.  Host code never passes
// the array's name to a kernel (on the host it is only a stand-in symbol):
// the device address is looked up once per device with hipGetSymbolAddress
// and kept in g_scratch_ptr.

#include <mutex>

#include "fd3_internal.h"

namespace fd3 {

constexpr size_t kScratchBytes = FD3_NOALLOC_SCRATCH_BYTES;  // 256 MiB per GPU

static __device__ __align__(256) unsigned char scratch[kScratchBytes];

namespace {

constexpr int kMaxDevices = 64;
void* g_scratch_ptr[kMaxDevices] = {};
std::mutex g_scratch_mu;

// Device address of `scratch` on the current device, looked up on the first
// call made with that device current.
int scratch_on_current_device(void** out) {
  int dev = -1;
  FD3_HIP(hipGetDevice(&dev));
  if (dev < 0 || dev >= kMaxDevices) {
    set_error("device %d is beyond the %d devices the static scratch table covers", dev, kMaxDevices);
    return FD3_ERR_UNSUPPORTED;
  }
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess) return hip_fail(e, "hipGetSymbolAddress(scratch)");
    if (!p) {
      set_error("hipGetSymbolAddress(scratch) returned NULL on device %d", dev);
      return FD3_ERR_HIP;
    }
    g_scratch_ptr[dev] = p;
  }
  *out = g_scratch_ptr[dev];
  return FD3_OK;
}

}  // namespace
}  // namespace fd3

using namespace fd3;

extern "C" {

int fd3_eval_celllist_noalloc(const fd3_model* m, const fd3_geom* g, int64_t n_atoms, const void* positions,
                              int pos_dtype, const int32_t* species, int flags, const fd3_outputs* out,
                              void* stream) {
  size_t need = 0;
  int rc = fd3_eval_celllist_workspace(m, g, n_atoms, flags, &need);
  if (rc) return rc;
  if (need > kScratchBytes) {  // checked before any device call
    set_error("this box needs %zu bytes of workspace; the static scratch of the allocation-free entry point "
              "holds %zu (256 MiB): use fd3_eval_celllist with a caller workspace",
              need, kScratchBytes);
    return FD3_ERR_WORKSPACE;
  }
  void* ws = nullptr;
  if ((rc = scratch_on_current_device(&ws))) return rc;
  // every buffer the pipeline needs cleared (bin and column counts) is
  // cleared on the stream by the pipeline itself, so the previous call's
  // contents do not matter
  return fd3_eval_celllist(m, g, n_atoms, positions, pos_dtype, species, flags, ws, kScratchBytes, out, stream);
}

}  // extern "C"
