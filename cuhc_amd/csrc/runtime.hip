#include "runtime.hpp"

#include <mutex>

namespace cuhc {

namespace {
hipStream_t gStream = nullptr;
std::once_flag gStreamOnce;
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
  std::call_once(gStreamOnce, [] {
    if (hipStreamCreateWithFlags(&gStream, hipStreamNonBlocking) != hipSuccess)
      gStream = nullptr;
  });
  return gStream;
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

int fill(void *dst, int byte, size_t bytes) {
  if (bytes == 0)
    return 0;
  return (int)hipMemsetAsync(dst, byte, bytes, stream());
}

int sync() { return (int)hipStreamSynchronize(stream()); }

bool timingEnabled() { return gTiming; }

void setTiming(bool on) {
  std::lock_guard<std::mutex> lock(gTimingMutex);
  gTiming = on;
}

thread_local int tDepth[kNumPhases];

PhaseTimer::PhaseTimer(int p) : phase(p), active(gTiming), counted(gTiming) {
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
