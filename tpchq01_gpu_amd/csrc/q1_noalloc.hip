// Scratch memory of the entry point that allocates nothing (q1_enqueue_noalloc): a static
// per-device array, so that a caller that cannot allocate device memory (code under HIP
// graph capture, serving loops that forbid hipMalloc) can still run the scan kernels that
// need scratch -- the spread-merge tables of grids of kSpreadMinBlocks blocks or more, and
// the global placement's per-block tables.
//
// Every GPU the library runs on holds its own copy of `scratch` until the process ends.
// In host code `scratch` names a host-side stand-in, not a device address: q1_init looks up
// each device's address once (hipGetSymbolAddress, with that device current) and every use
// goes through that address.  All calls on one GPU share the array, which keeps the last
// call's contents: the spread tables are cleared at set-up and left clear by every launch
// that uses them (a failed launch marks them for clearing), and the global placement's
// per-block tables are cleared by the kernel itself.  Calls that overlap on one GPU race.
#include <hip/hip_runtime.h>

#include <mutex>

#include "q1_internal.h"

namespace {

constexpr uint64_t kSpreadBytes = q1::kSpreadTables * sizeof(q1_device_aggregates);
constexpr uint64_t kScratchBytes = kSpreadBytes + q1::kNoallocWorkspaceBytes;
constexpr int kMaxDevices = 64;

static __device__ __align__(256) uint64_t scratch[kScratchBytes / 8];

std::mutex g_lock;
void* g_scratch[kMaxDevices];   // device address of `scratch` on each device
bool g_dirty[kMaxDevices];      // the spread tables may hold a failed launch's partial sums

}  // namespace

namespace q1 {

q1_status noalloc_setup(int device, void* stream) {
  if (device < 0 || device >= kMaxDevices) return Q1_EINVAL;
  std::lock_guard<std::mutex> g(g_lock);
  if (g_scratch[device]) return Q1_OK;
  void* p = nullptr;
  if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) != hipSuccess || !p) return Q1_EDEVICE;
  // the spread tables start at zero (a fill kernel, no memset node)
  q1_status s = launch_zero(reinterpret_cast<q1_device_aggregates*>(p), kSpreadTables, stream);
  if (s == Q1_OK && hipStreamSynchronize((hipStream_t)stream) != hipSuccess) s = Q1_EDEVICE;
  if (s != Q1_OK) return s;
  g_scratch[device] = p;
  return Q1_OK;
}

q1_status noalloc_launch_scan(int device, q1_placement placement, q1_layout layout, bool fast, const ScanArgs& a,
                              const LaunchShape& shape, void* stream, const char** why) {
  if (device < 0 || device >= kMaxDevices) {
    *why = "device out of range";
    return Q1_EINVAL;
  }
  void* base = nullptr;
  bool dirty = false;
  {
    std::lock_guard<std::mutex> g(g_lock);
    base = g_scratch[device];
    dirty = g_dirty[device];
  }
  if (!base) {
    *why = "the scratch array of this device is not set up (q1_init does it)";
    return Q1_EDEVICE;
  }
  if (scan_workspace_bytes(placement, shape) > kNoallocWorkspaceBytes) {
    *why = "the launch needs more scratch than the static array holds (the global placement: at most "
           "34952 blocks)";
    return Q1_EINVAL;
  }
  Spread spread{reinterpret_cast<q1_device_aggregates*>(base)};
  void* workspace = static_cast<char*>(base) + kSpreadBytes;
  if (dirty) {
    const q1_status z = launch_zero(spread.tables, kSpreadTables, stream);
    if (z != Q1_OK) {
      *why = "clearing the spread tables failed";
      return z;
    }
    std::lock_guard<std::mutex> g(g_lock);
    g_dirty[device] = false;
  }
  const q1_status s = launch_scan(placement, layout, fast, a, shape, workspace, stream, &spread, false);
  if (s != Q1_OK) {
    std::lock_guard<std::mutex> g(g_lock);
    g_dirty[device] = true;
    *why = "scan kernel launch failed";
  }
  return s;
}

}  // namespace q1
