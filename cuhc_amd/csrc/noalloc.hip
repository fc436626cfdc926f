// cuhc_linkage_device_noalloc: single linkage (Boruvka MST) on points already
// in device memory, with every device buffer of the call carved out of one
// fixed array of this library (`scratch`) instead of allocated. For callers
// that cannot allocate device memory while they run: code being captured
// into a HIP graph, tight serving loops, harnesses that forbid hipMalloc.
//
// The array is a file-scope __device__ variable: every GPU the library runs
// on holds its own copy, at its own address, until the process ends. Host
// code never passes the array's name to a kernel or a copy (on the host it
// is a stand-in, not a device address); it looks up the current device's
// address once with hipGetSymbolAddress and hands out slices of that.
// Every call on a device reuses the same array: the passes clear what they
// need cleared, as they do for reused cached blocks. Calls are serialised by
// a mutex within the process; calls that overlap on one device through other
// means race on the array.
#include "../../include/cuhc.h"
#include "common.hpp"
#include "linkage.hpp"
#include "runtime.hpp"

#include <mutex>

using namespace cuhc;

namespace {

constexpr size_t kScratchBytes = (size_t)256 << 20;
constexpr int kMaxPoints = 65536;
constexpr int kMaxFeatures = 64;
constexpr int kMaxDevices = 64;

static __device__ __align__(256) unsigned char scratch[kScratchBytes];

std::mutex gNoAllocMutex;
void *gScratchOf[kMaxDevices]; // the array's address on each device, once looked up

// The array's address on the current device (looked up on that device's
// first call, with the device current).
int scratchAddress(void **out) {
  int device = 0;
  hipError_t e = hipGetDevice(&device);
  if (e != hipSuccess)
    return (int)e;
  if (device < 0 || device >= kMaxDevices)
    return (int)hipErrorInvalidDevice;
  if (!gScratchOf[device]) {
    void *ptr = nullptr;
    e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (e != hipSuccess)
      return (int)e;
    gScratchOf[device] = ptr;
  }
  *out = gScratchOf[device];
  return 0;
}

} // namespace

extern "C" int cuhc_linkage_device_noalloc(const float *deviceX, int n, int d, int metricId,
                                           float metricParam, int linkageId, int algorithmId,
                                           double *tree) {
  if (n <= 0 || d <= 0)
    return kErrInvalidShape;
  if (linkageId != kSingle)
    return kErrInvalidLinkage;
  if (algorithmId != kAlgMst)
    return kErrInvalidAlgorithm;
  if (n > kMaxPoints || d > kMaxFeatures)
    return kErrWorkspace;
  std::lock_guard<std::mutex> lock(gNoAllocMutex);
  void *base = nullptr;
  if (const int s = scratchAddress(&base))
    return s;
  // The library stream exists before the workspace is in use (its creation
  // is set-up, not part of the call's work).
  if (!stream())
    return (int)hipErrorOutOfMemory;
  WorkspaceScope scope(base, kScratchBytes);
  return linkageFromDevicePoints(deviceX, n, d, metricId, metricParam, linkageId, algorithmId,
                                 tree);
}
