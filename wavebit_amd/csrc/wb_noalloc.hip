// The allocation-free batch solve: wb_solve_noalloc runs wb_solve's launch sequence on a
// file-scope device array instead of a caller workspace, for callers that cannot allocate
// device memory (code under HIP graph capture, tight serving loops, harnesses that forbid
// hipMalloc). The array is WB_SOLVE_NOALLOC_BYTES on every device the library runs on, held
// until the process ends. Its device address is taken once per device with
// hipGetSymbolAddress at context creation (wb_ctx_create, with that device current) and kept
// in g_scratch; host code uses only that address, never the array's own name.
//
// Every call on one device shares that device's array, which keeps the last call's contents:
// the solve clears what must start at zero (the flags) with a fill kernel, and reads nothing
// of the array that it has not written in the same call. Calls that overlap on one device (two
// streams, two host threads) race on the array: serialize them.
#include <mutex>

#include "wb_rcwa.h"

static __device__ __align__(256) unsigned char scratch[WB_SOLVE_NOALLOC_BYTES];

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_lock;
void* g_scratch[kMaxDevices] = {};
}  // namespace

// Called by wb_ctx_create with `device` current. Returns hipSuccess once the device's address
// is known (taken on the first call for that device).
hipError_t wb_noalloc_attach(int device) {
  if (device < 0 || device >= kMaxDevices) return hipErrorInvalidDevice;
  std::lock_guard<std::mutex> guard(g_lock);
  if (g_scratch[device]) return hipSuccess;
  void* ptr = nullptr;
  hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
  if (e != hipSuccess) return e;
  if (!ptr) return hipErrorInvalidSymbol;
  g_scratch[device] = ptr;
  return hipSuccess;
}

// The device address of `device`'s array, or nullptr before wb_noalloc_attach(device).
void* wb_noalloc_scratch(int device) {
  if (device < 0 || device >= kMaxDevices) return nullptr;
  std::lock_guard<std::mutex> guard(g_lock);
  return g_scratch[device];
}
