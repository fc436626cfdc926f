#include "runtime.hpp"

#include "common.hpp"

#include <algorithm>
#include <atomic>
#include <cstdlib>
#include <mutex>
#include <unordered_map>

namespace cuhc {

namespace {
std::atomic<hipStream_t> gStream{nullptr};
std::mutex gStreamMutex;
bool gExitHandler = false;
int gNumCUs = 0;

struct PendingEvent {
  int phase;
  hipEvent_t start, stop;
};
bool gTiming = false;
std::vector<PendingEvent> gPending;
double gMs[kNumPhases];
long long gLaunches[kNumPhases];
long long gUnits[kNumPhases];
std::mutex gTimingMutex;
} // namespace

hipStream_t stream() {
  hipStream_t s = gStream.load(std::memory_order_acquire);
  if (s)
    return s;
  std::lock_guard<std::mutex> lock(gStreamMutex);
  s = gStream.load(std::memory_order_relaxed);
  if (!s) {
    if (hipStreamCreateWithFlags(&s, hipStreamNonBlocking) != hipSuccess)
      return nullptr;
    gStream.store(s, std::memory_order_release);
    // The HIP runtime is initialised by now, so this handler runs before its
    // teardown (exit handlers run in reverse order of registration).
    if (!gExitHandler) {
      gExitHandler = true;
      std::atexit(releaseLibraryState);
    }
  }
  return s;
}

int numCUs() {
  if (gNumCUs == 0) {
    int cus = 0;
    if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, 0) != hipSuccess ||
        cus <= 0)
      cus = 256;
    gNumCUs = cus;
  }
  return gNumCUs;
}

// Device memory cache. hipMalloc and hipFree of large blocks take
// milliseconds each (page mapping; hipFree also waits for the device), and
// every library call allocates its working buffers afresh, so freed blocks
// are kept per size class (the request rounded up to 2 MiB above 1 MiB, to
// 256 B below) and handed to the next request of the same class. All work
// runs on the one library stream, so a reused block is only touched by work
// ordered after every earlier use of it. The cache holds at most a quarter of
// device memory; when hipMalloc fails, the cache is released and the request
// retried. releaseLibraryState() frees it (at exit, before the runtime's own
// teardown).
namespace {
struct DeviceCache {
  std::mutex mutex;
  std::unordered_map<size_t, std::vector<void *>> blocks;
  size_t bytes = 0, limit = 0;
};

DeviceCache &deviceCache() {
  static DeviceCache *cache = new DeviceCache;
  return *cache;
}

size_t sizeClass(size_t bytes) {
  constexpr size_t kLarge = (size_t)2 << 20, kSmall = 256;
  return bytes > ((size_t)1 << 20) ? (bytes + kLarge - 1) / kLarge * kLarge
                                    : (bytes + kSmall - 1) / kSmall * kSmall;
}

void releaseCache(DeviceCache &cache) {
  if (hipStream_t st = gStream.load())
    (void)hipStreamSynchronize(st);
  for (auto &entry : cache.blocks)
    for (void *p : entry.second)
      (void)hipFree(p);
  cache.blocks.clear();
  cache.bytes = 0;
}
} // namespace

namespace {
struct Workspace {
  char *base = nullptr;
  size_t bytes = 0, used = 0;
};
thread_local Workspace *tWorkspace = nullptr;
constexpr size_t kSliceAlign = 256;
size_t sliceBytes(size_t bytes) { return (bytes + kSliceAlign - 1) / kSliceAlign * kSliceAlign; }
} // namespace

WorkspaceScope::WorkspaceScope(void *base, size_t bytes) {
  Workspace *w = new Workspace;
  w->base = static_cast<char *>(base);
  w->bytes = bytes;
  tWorkspace = w;
}

WorkspaceScope::~WorkspaceScope() {
  delete tWorkspace;
  tWorkspace = nullptr;
}

bool inWorkspace() { return tWorkspace != nullptr; }

int deviceAlloc(void **p, size_t bytes) {
  if (Workspace *w = tWorkspace) {
    const size_t want = sliceBytes(bytes);
    if (want > w->bytes - w->used) {
      *p = nullptr;
      return kErrWorkspace;
    }
    *p = w->base + w->used;
    w->used += want;
    return 0;
  }
  DeviceCache &cache = deviceCache();
  const size_t cls = sizeClass(bytes);
  std::lock_guard<std::mutex> lock(cache.mutex);
  auto it = cache.blocks.find(cls);
  if (it != cache.blocks.end() && !it->second.empty()) {
    *p = it->second.back();
    it->second.pop_back();
    cache.bytes -= cls;
    return 0;
  }
  hipError_t e = hipMalloc(p, cls);
  if (e == hipErrorOutOfMemory && cache.bytes > 0) {
    (void)hipGetLastError();
    releaseCache(cache);
    e = hipMalloc(p, cls);
  }
  return (int)e;
}

void deviceFree(void *p, size_t bytes) {
  if (Workspace *w = tWorkspace) {
    char *q = static_cast<char *>(p);
    if (q >= w->base && q < w->base + w->bytes) {
      if (q + sliceBytes(bytes) == w->base + w->used)
        w->used = (size_t)(q - w->base);
      return;
    }
  }
  DeviceCache &cache = deviceCache();
  const size_t cls = sizeClass(bytes);
  std::lock_guard<std::mutex> lock(cache.mutex);
  if (cache.limit == 0) {
    size_t freeBytes = 0, totalBytes = 0;
    if (hipMemGetInfo(&freeBytes, &totalBytes) == hipSuccess && totalBytes > 0)
      cache.limit = totalBytes / 4;
    else
      cache.limit = (size_t)16 << 30;
  }
  if (cache.bytes + cls > cache.limit) {
    (void)hipFree(p);
    return;
  }
  cache.blocks[cls].push_back(p);
  cache.bytes += cls;
}

size_t cachedDeviceBytes() {
  DeviceCache &cache = deviceCache();
  std::lock_guard<std::mutex> lock(cache.mutex);
  return cache.bytes;
}

int upload(void *dst, const void *src, size_t bytes) {
  if (bytes == 0)
    return 0;
  int s = (int)hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, stream());
  return s ? s : (int)hipStreamSynchronize(stream());
}

int download(void *dst, const void *src, size_t bytes) {
  if (bytes == 0)
    return 0;
  int s = (int)hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, stream());
  return s ? s : (int)hipStreamSynchronize(stream());
}

namespace {

struct PinnedStage {
  std::mutex mutex;
  void *p = nullptr;
  size_t bytes = 0;
};

PinnedStage &pinnedStage() {
  static PinnedStage *stage = new PinnedStage; // the buffer is freed by releaseLibraryState
  return *stage;
}

constexpr size_t kMaxPinnedStage = (size_t)256 << 20;

} // namespace

int downloadStaged(const void *src, size_t bytes,
                   const std::function<void(const void *)> &consume) {
  if (bytes == 0) {
    consume(nullptr);
    return 0;
  }
  if (inWorkspace()) {
    std::vector<char> host(bytes);
    const int s = download(host.data(), src, bytes);
    if (s == 0)
      consume(host.data());
    return s;
  }
  PinnedStage &st = pinnedStage();
  std::lock_guard<std::mutex> lock(st.mutex);
  if (bytes <= kMaxPinnedStage && st.bytes < bytes) {
    if (st.p)
      (void)hipHostFree(st.p);
    st.p = nullptr;
    st.bytes = 0;
    const size_t want = std::max(bytes, (size_t)1 << 20);
    if (hipHostMalloc(&st.p, want, hipHostMallocDefault) == hipSuccess)
      st.bytes = want;
    else
      st.p = nullptr;
  }
  if (st.bytes < bytes) {
    std::vector<char> host(bytes);
    const int s = download(host.data(), src, bytes);
    if (s == 0)
      consume(host.data());
    return s;
  }
  const int s = download(st.p, src, bytes);
  if (s == 0)
    consume(st.p);
  return s;
}

namespace {

// flag = 1 when some value has all exponent bits set (infinity or NaN). The
// test is on the bits, so it holds under any floating-point compile flags.
__global__ void __launch_bounds__(256)
    nonFiniteKernel(const unsigned int *__restrict__ bits, size_t count, int *flag) {
  const size_t stride = (size_t)gridDim.x * blockDim.x;
  bool bad = false;
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < count; i += stride)
    bad |= (bits[i] & 0x7f800000u) == 0x7f800000u;
  if (bad)
    *flag = 1;
}

} // namespace

int checkFinite(const float *X, size_t count) {
  if (count == 0)
    return 0;
  DevArray<int> flag;
  Status s;
  s(flag.alloc(1));
  s(fill(flag, 0, sizeof(int)));
  if (!s.ok())
    return s.code;
  const size_t blocks = std::min<size_t>((count + 255) / 256, (size_t)numCUs() * 16);
  hipLaunchKernelGGL(nonFiniteKernel, dim3((unsigned)blocks), dim3(256), 0, stream(),
                     reinterpret_cast<const unsigned int *>(X), count, flag.get());
  s((int)hipGetLastError());
  int bad = 0;
  s(download(&bad, flag, sizeof(int)));
  if (!s.ok())
    return s.code;
  return bad ? kErrNonFinite : 0;
}

namespace {

// Order-preserving 32-bit image of a float (any sign) for integer atomics.
HD unsigned int rangeBits(float f) {
  const unsigned int b = bitsOf(f);
  return (b & 0x80000000u) ? ~b : (b | 0x80000000u);
}
HD float rangeFloat(unsigned int o) {
  return floatOf((o & 0x80000000u) ? (o & 0x7fffffffu) : ~o);
}

constexpr int kRangeRows = 1024;

// Workgroup g covers rows [g * 1024, ...); thread t features t, t + 256, ...
// (consecutive threads read consecutive floats of a row).
__global__ void __launch_bounds__(256)
    featureRangesKernel(const float *__restrict__ X, int n, int d, unsigned int *lowHigh) {
  const int r0 = blockIdx.x * kRangeRows, r1 = min(n, r0 + kRangeRows);
  for (int f = threadIdx.x; f < d; f += blockDim.x) {
    float lo = INFINITY, hi = -INFINITY;
    for (int r = r0; r < r1; ++r) {
      const float v = X[(size_t)r * d + f];
      lo = fminf(lo, v);
      hi = fmaxf(hi, v);
    }
    atomicMin(&lowHigh[f], rangeBits(lo));
    atomicMax(&lowHigh[d + f], rangeBits(hi));
  }
}

__global__ void rangeInitKernel(unsigned int *lowHigh, int d) {
  for (int f = threadIdx.x + blockIdx.x * blockDim.x; f < d; f += gridDim.x * blockDim.x) {
    lowHigh[f] = 0xffffffffu;
    lowHigh[d + f] = 0u;
  }
}

} // namespace

int featureRanges(const float *X, int n, int d, float *low, float *high) {
  DevArray<unsigned int> lowHigh;
  Status s;
  s(lowHigh.alloc((size_t)2 * d));
  if (!s.ok())
    return s.code;
  hipLaunchKernelGGL(rangeInitKernel, dim3((d + 255) / 256), dim3(256), 0, stream(),
                     lowHigh.get(), d);
  hipLaunchKernelGGL(featureRangesKernel, dim3((n + kRangeRows - 1) / kRangeRows), dim3(256), 0,
                     stream(), X, n, d, lowHigh.get());
  s((int)hipGetLastError());
  std::vector<unsigned int> h((size_t)2 * d);
  s(download(h.data(), lowHigh, h.size() * sizeof(unsigned int)));
  if (!s.ok())
    return s.code;
  for (int f = 0; f < d; ++f) {
    low[f] = rangeFloat(h[f]);
    high[f] = rangeFloat(h[d + f]);
  }
  return 0;
}

void releaseCachedBlocks() {
  DeviceCache &cache = deviceCache();
  std::lock_guard<std::mutex> lock(cache.mutex);
  releaseCache(cache);
}

void releaseLibraryState() {
  releaseCachedBlocks();
  {
    PinnedStage &st = pinnedStage();
    std::lock_guard<std::mutex> lock(st.mutex);
    if (st.p)
      (void)hipHostFree(st.p);
    st.p = nullptr;
    st.bytes = 0;
  }
  std::lock_guard<std::mutex> lock(gStreamMutex);
  if (hipStream_t s = gStream.exchange(nullptr)) {
    (void)hipStreamSynchronize(s);
    (void)hipStreamDestroy(s);
  }
}

namespace {

// Byte fill: 16-byte stores over the aligned middle, single bytes at the ends.
__global__ void __launch_bounds__(256)
    fillKernel(unsigned char *__restrict__ dst, size_t bytes, unsigned int byte) {
  const size_t head = std::min<size_t>(bytes, (16 - ((uintptr_t)dst & 15)) & 15);
  const size_t words = (bytes - head) / 16;
  const size_t tail = head + words * 16;
  const size_t stride = (size_t)gridDim.x * blockDim.x;
  const size_t tid = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  const unsigned int v = byte * 0x01010101u;
  uint4 *body = reinterpret_cast<uint4 *>(dst + head);
  for (size_t i = tid; i < words; i += stride)
    body[i] = make_uint4(v, v, v, v);
  if (tid < head)
    dst[tid] = (unsigned char)byte;
  if (tid < bytes - tail)
    dst[tail + tid] = (unsigned char)byte;
}

} // namespace

// A kernel rather than hipMemsetAsync: the same stream-ordered semantics, and
// a plain kernel node wherever a caller captures the library's work.
int fill(void *dst, int byte, size_t bytes) {
  if (bytes == 0)
    return 0;
  const size_t words = bytes / 16 + 1;
  const size_t blocks = std::min<size_t>((words + 255) / 256, (size_t)numCUs() * 16);
  hipLaunchKernelGGL(fillKernel, dim3((unsigned)std::max<size_t>(blocks, 1)), dim3(256), 0,
                     stream(), static_cast<unsigned char *>(dst), bytes,
                     (unsigned int)(byte & 0xff));
  return (int)hipGetLastError();
}

int sync() { return (int)hipStreamSynchronize(stream()); }

bool timingEnabled() { return gTiming; }

void setTiming(bool on) {
  std::lock_guard<std::mutex> lock(gTimingMutex);
  gTiming = on;
}

thread_local int tDepth[kNumPhases];

PhaseTimer::PhaseTimer(int p)
    : phase(p), active(gTiming && !inWorkspace()), counted(gTiming && !inWorkspace()) {
  if (!active)
    return;
  if (tDepth[p]++ > 0) {
    active = false;
    return;
  }
  PendingEvent e{p, nullptr, nullptr};
  if (hipEventCreate(&e.start) != hipSuccess || hipEventCreate(&e.stop) != hipSuccess) {
    active = false;
    return;
  }
  (void)hipEventRecord(e.start, stream());
  std::lock_guard<std::mutex> lock(gTimingMutex);
  gPending.push_back(e);
}

PhaseTimer::~PhaseTimer() {
  if (counted)
    --tDepth[phase];
  if (!active)
    return;
  std::lock_guard<std::mutex> lock(gTimingMutex);
  for (auto it = gPending.rbegin(); it != gPending.rend(); ++it)
    if (it->phase == phase) {
      (void)hipEventRecord(it->stop, stream());
      break;
    }
}

void collectTimings() {
  (void)hipStreamSynchronize(stream());
  std::lock_guard<std::mutex> lock(gTimingMutex);
  for (PendingEvent &e : gPending) {
    float ms = 0.0f;
    if (hipEventElapsedTime(&ms, e.start, e.stop) == hipSuccess) {
      gMs[e.phase] += ms;
      gLaunches[e.phase] += 1;
    }
    (void)hipEventDestroy(e.start);
    (void)hipEventDestroy(e.stop);
  }
  gPending.clear();
}

void readTimings(double *ms, long long *launches, long long *units, int count) {
  collectTimings();
  std::lock_guard<std::mutex> lock(gTimingMutex);
  for (int i = 0; i < count && i < kNumPhases; ++i) {
    if (ms)
      ms[i] = gMs[i];
    if (launches)
      launches[i] = gLaunches[i];
    if (units)
      units[i] = gUnits[i];
  }
}

void addUnits(int phase, long long units) {
  if (!gTiming)
    return;
  std::lock_guard<std::mutex> lock(gTimingMutex);
  gUnits[phase] += units;
}

void addMs(int phase, double ms, long long launches) {
  if (!gTiming)
    return;
  std::lock_guard<std::mutex> lock(gTimingMutex);
  gMs[phase] += ms;
  gLaunches[phase] += launches;
}

double wallClockKHz() {
  int khz = 0;
  if (hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, 0) != hipSuccess || khz <= 0)
    khz = 100000;
  return (double)khz;
}

void resetTimings() {
  collectTimings();
  std::lock_guard<std::mutex> lock(gTimingMutex);
  for (int i = 0; i < kNumPhases; ++i) {
    gMs[i] = 0.0;
    gLaunches[i] = 0;
    gUnits[i] = 0;
  }
}

} // namespace cuhc
