// Host-side runtime shared by the entry points: the library's HIP stream,
// device buffers, first-error status, and per-phase kernel timing.
#pragma once

#include <hip/hip_runtime.h>
#include <stddef.h>
#include <stdint.h>
#include <functional>
#include <vector>

namespace cuhc {

// The one stream every kernel and copy of the library runs on (created on
// first use, default priority, non-blocking with respect to the null stream).
hipStream_t stream();

// Number of compute units of device 0 (cached).
int numCUs();

// Keeps the first non-zero status of a sequence of calls.
struct Status {
  int code = 0;
  void operator()(int result) {
    if (code == 0)
      code = result;
  }
  bool ok() const { return code == 0; }
};

// Fixed device workspace. While a WorkspaceScope lives on a thread,
// deviceAlloc on that thread hands out 256-byte aligned slices of
// [base, base + bytes) in order and allocates nothing; a request the rest of
// the workspace cannot hold fails with kErrWorkspace. deviceFree of the most
// recent slice gives its space back (a slice freed out of order is reclaimed
// when the scope ends). Inside a scope, downloads do not use the pinned
// staging buffer and phase timing records no events.
struct WorkspaceScope {
  WorkspaceScope(void *base, size_t bytes);
  ~WorkspaceScope();
  WorkspaceScope(const WorkspaceScope &) = delete;
  WorkspaceScope &operator=(const WorkspaceScope &) = delete;
};
bool inWorkspace();

// Device memory through the library's cache of freed blocks (runtime.hip).
// deviceFree takes the size given to deviceAlloc. cachedDeviceBytes: memory
// held by the cache (free for the library's own use).
int deviceAlloc(void **p, size_t bytes);
void deviceFree(void *p, size_t bytes);
size_t cachedDeviceBytes();

// Owning device allocation.
template <class T> struct DevArray {
  T *p = nullptr;
  size_t count = 0;
  DevArray() = default;
  DevArray(const DevArray &) = delete;
  DevArray &operator=(const DevArray &) = delete;
  ~DevArray() { release(); }
  int alloc(size_t n) {
    release();
    if (n == 0)
      return 0;
    void *q = nullptr;
    const int s = deviceAlloc(&q, n * sizeof(T));
    if (s == 0) {
      p = static_cast<T *>(q);
      count = n;
    }
    return s;
  }
  void release() {
    if (p)
      deviceFree(p, count * sizeof(T));
    p = nullptr;
    count = 0;
  }
  T *get() const { return p; }
  operator T *() const { return p; }
  size_t bytes() const { return count * sizeof(T); }
};

// Synchronous helpers on the library stream.
int upload(void *dst, const void *src, size_t bytes);
int download(void *dst, const void *src, size_t bytes);
// Downloads bytes into a cached pinned host buffer (one DMA instead of the
// staged chunks of a pageable copy) and hands it to consume while holding the
// buffer; above 256 MiB, or if pinned memory is refused, a pageable copy.
int downloadStaged(const void *src, size_t bytes, const std::function<void(const void *)> &consume);
int fill(void *dst, int byte, size_t bytes);
int sync();

// Per-feature minimum and maximum of the n x d points at X (device) into the
// host arrays low[d], high[d] (NaN values are ignored).
int featureRanges(const float *X, int n, int d, float *low, float *high);

// 0 when all count floats at X (device) are finite, kErrNonFinite (-6) when
// one is NaN or infinite, or a HIP error. One read of the array.
int checkFinite(const float *X, size_t count);

// Releases what the library holds between calls: the cached device blocks,
// the pinned staging buffer and the library stream (recreated on next use).
// Runs from an exit handler registered when the stream is first created, so
// that it precedes the HIP runtime's own teardown.
void releaseLibraryState();
// Frees the cached device blocks only.
void releaseCachedBlocks();

// Per-phase GPU timing with HIP events on the library stream. Off by default;
// cuhc_set_timing(1) turns it on. A phase is bracketed by begin()/end(); the
// elapsed time is collected at the next collect() (after a stream sync).
enum Phase : int {
  kPhaseDistanceTiles = 0,  // FP32 direct-form all-pairs passes (Boruvka, Ward)
  kPhaseMfmaCandidates = 1, // FP16 MFMA candidate pass
  kPhaseRecheck = 2,        // FP32 recheck of certified candidates
  kPhaseCore = 3,           // core distances
  kPhaseContract = 4,       // component bookkeeping between rounds
  kPhaseDense = 5,          // dense Lance-Williams loop
  kPhaseRnn = 6,            // RNN merge/compaction kernels
  kPhaseMlp = 7,            // learned-radius network
  kPhaseDenseSelect = 8,    // dense loop, device clock: pair selection
  kPhaseDenseUpdate = 9,    // dense loop, device clock: Lance-Williams update
  kPhaseDenseRescan = 10,   // dense loop, device clock: row rescans and selection records
  kPhaseDenseBarrier = 11,  // dense loop, device clock: grid barriers
  kPhaseHostTree = 12,      // host clock: tree assembly from the downloaded edges
  kNumPhases = 13,
};

// Times the enclosed launches of one phase; a timer nested inside another of
// the same phase (on the same host thread) records nothing.
struct PhaseTimer {
  int phase;
  bool active;
  bool counted;
  explicit PhaseTimer(int phase);
  ~PhaseTimer();
};

bool timingEnabled();
void setTiming(bool on);
void collectTimings(); // syncs the stream and folds finished events into totals
void readTimings(double *ms, long long *launches, long long *units, int count);
void addUnits(int phase, long long units); // algorithmic work units (pairs) per phase
void addMs(int phase, double ms, long long launches); // time measured on the device clock
double wallClockKHz();                              // frequency of wall_clock64()
void resetTimings();

} // namespace cuhc
