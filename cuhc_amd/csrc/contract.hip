// Boruvka's contraction between rounds, on the GPU (the reference does this
// step on the host, src/mst.cuh:130-181, with the same result).
//
// From nearest[p] (the key of p's nearest point in another component: weight
// bits << 32 | original index) every component takes its smallest outgoing
// edge under the global order (weight, smaller endpoint, larger endpoint) in
// two atomicMin passes: (weight, smaller endpoint) as one u64, then the
// larger endpoint among the points that reached that minimum. Each component
// hooks onto the component at the edge's other end. With a strict total order
// the only cycles are pairs that chose the same edge; the lower component of
// such a pair stays a root and emits the edge, every other component emits
// its own edge. Pointer jumping turns the hooks into roots, an exclusive scan
// numbers the roots, and every point is relabelled. The host receives the new
// component count and the number of edges added; the edges stay in device
// memory until the tree is assembled.
#include "mst.hpp"

#include "primitives.hpp"

namespace cuhc {

namespace {

__global__ void iotaIntKernel(int *out, int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n)
    out[i] = i;
}

__global__ void resetComponentsKernel(unsigned long long *best1, unsigned int *best2,
                                      int numComp) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c < numComp) {
    best1[c] = kEmptyKey;
    best2[c] = 0xffffffffu;
  }
}

DEV unsigned long long lowKey(unsigned long long key, int a) {
  const int b = (int)keyIndex(key);
  return (key & 0xffffffff00000000ull) | (unsigned int)min(a, b);
}

__global__ void componentMinKernel(const unsigned long long *__restrict__ nearest,
                                   const int *__restrict__ comp, const int *__restrict__ orig,
                                   int n, unsigned long long *best1) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= n)
    return;
  const unsigned long long key = nearest[p];
  if (key == kEmptyKey)
    return;
  atomicMin(&best1[comp[p]], lowKey(key, orig ? orig[p] : p));
}

// The same with few components: same-address atomics from many waves
// serialise (one component holding every point would take one atomic per
// point), so each workgroup first reduces into LDS and then flushes one atomic
// per component it saw.
constexpr int kLdsComponents = 4096;

__global__ void __launch_bounds__(256)
    componentMinLdsKernel(const unsigned long long *__restrict__ nearest,
                          const int *__restrict__ comp, const int *__restrict__ orig, int n,
                          int numComp, unsigned long long *best1) {
  __shared__ unsigned long long local[kLdsComponents];
  for (int c = threadIdx.x; c < numComp; c += blockDim.x)
    local[c] = kEmptyKey;
  __syncthreads();
  for (int p = blockIdx.x * blockDim.x + threadIdx.x; p < n; p += gridDim.x * blockDim.x) {
    const unsigned long long key = nearest[p];
    if (key != kEmptyKey)
      atomicMin(&local[comp[p]], lowKey(key, orig ? orig[p] : p));
  }
  __syncthreads();
  for (int c = threadIdx.x; c < numComp; c += blockDim.x)
    if (local[c] != kEmptyKey)
      atomicMin(&best1[c], local[c]);
}

__global__ void componentMin2Kernel(const unsigned long long *__restrict__ nearest,
                                    const int *__restrict__ comp, const int *__restrict__ orig,
                                    int n, const unsigned long long *__restrict__ best1,
                                    unsigned int *best2) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= n)
    return;
  const unsigned long long key = nearest[p];
  if (key == kEmptyKey)
    return;
  const int a = orig ? orig[p] : p;
  const int c = comp[p];
  if (lowKey(key, a) == best1[c])
    atomicMin(&best2[c], (unsigned int)max(a, (int)keyIndex(key)));
}

__global__ void hookKernel(const unsigned long long *__restrict__ best1,
                           const unsigned int *__restrict__ best2, const int *__restrict__ comp,
                           const int *__restrict__ posOf, int n, int numComp, int *parent) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= numComp)
    return;
  const unsigned long long k1 = best1[c];
  const unsigned int lo32 = (unsigned int)(k1 & 0xffffffffu), hi32 = best2[c];
  if (k1 == kEmptyKey || lo32 >= (unsigned)n || hi32 >= (unsigned)n) {
    parent[c] = c; // no edge leaves this component
    return;
  }
  const int lo = (int)lo32, hi = (int)hi32;
  const int cl = comp[posOf ? posOf[lo] : lo], ch = comp[posOf ? posOf[hi] : hi];
  parent[c] = cl == c ? ch : cl;
}

__global__ void emitKernel(const unsigned long long *__restrict__ best1,
                           const unsigned int *__restrict__ best2,
                           const int *__restrict__ parent, int numComp, int *root, int *eu,
                           int *ev, float *ew, int *edgeCount, int capacity) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= numComp)
    return;
  const int pc = parent[c];
  bool emit;
  if (pc == c) {
    root[c] = c;
    emit = false;
  } else if (parent[pc] == c) { // both chose the same edge
    root[c] = c < pc ? c : pc;
    emit = c < pc;
  } else {
    root[c] = pc;
    emit = true;
  }
  if (emit) {
    const unsigned long long k1 = best1[c];
    const int at = atomicAdd(edgeCount, 1);
    if (at < capacity) {
      eu[at] = (int)(k1 & 0xffffffffu);
      ev[at] = (int)best2[c];
      ew[at] = floatOf((unsigned int)(k1 >> 32));
    }
  }
}

// One doubling step: root[c] = root[root[c]]. Reads may see values other
// threads already advanced; every value read is an ancestor, so the loop still
// ends with every entry at its root.
__global__ void jumpKernel(int *root, int numComp, int *changed) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= numComp)
    return;
  const int r = root[c];
  const int rr = root[r];
  if (rr != r) {
    root[c] = rr;
    *changed = 1;
  }
}

__global__ void rootFlagsKernel(const int *__restrict__ root, int numComp, int *flags) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c <= numComp)
    flags[c] = c < numComp && root[c] == c ? 1 : 0;
}

__global__ void relabelKernel(int *comp, int n, const int *__restrict__ root,
                              const int *__restrict__ newId) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p < n)
    comp[p] = newId[root[comp[p]]];
}

inline dim3 gridFor(long long count) { return dim3((unsigned)((count + 255) / 256)); }

} // namespace

int Contraction::alloc(int points) {
  n = points;
  Status s;
  s(best1.alloc(n));
  s(best2.alloc(n));
  s(parent.alloc(n));
  s(root.alloc(n));
  s(flags.alloc(n + 1));
  s(newId.alloc(n + 1));
  s(changed.alloc(1));
  s(counters.alloc(1));
  s(edges.u.alloc(std::max(1, n - 1)));
  s(edges.v.alloc(std::max(1, n - 1)));
  s(edges.w.alloc(std::max(1, n - 1)));
  const size_t bytes = scanTempBytes(n + 1);
  s(scanTemp.alloc(bytes));
  scanBytes = bytes;
  s(fill(counters, 0, sizeof(int)));
  edges.count = 0;
  return s.code;
}

int resetComponents(int *comp, int n) {
  hipLaunchKernelGGL(iotaIntKernel, gridFor(n), dim3(256), 0, stream(), comp, n);
  return (int)hipGetLastError();
}

int contractRound(Contraction &ct, const unsigned long long *nearest, int *comp, const int *orig,
                  const int *posOf, int numComp, int *newNumComp, int *added) {
  const int n = ct.n;
  Status s;
  hipLaunchKernelGGL(resetComponentsKernel, gridFor(numComp), dim3(256), 0, stream(),
                     ct.best1.get(), ct.best2.get(), numComp);
  if (numComp <= kLdsComponents)
    hipLaunchKernelGGL(componentMinLdsKernel, dim3(std::max(1, std::min(1024, (n + 255) / 256))),
                       dim3(256), 0, stream(), nearest, comp, orig, n, numComp, ct.best1.get());
  else
    hipLaunchKernelGGL(componentMinKernel, gridFor(n), dim3(256), 0, stream(), nearest, comp,
                       orig, n, ct.best1.get());
  hipLaunchKernelGGL(componentMin2Kernel, gridFor(n), dim3(256), 0, stream(), nearest, comp, orig,
                     n, ct.best1.get(), ct.best2.get());
  hipLaunchKernelGGL(hookKernel, gridFor(numComp), dim3(256), 0, stream(), ct.best1.get(),
                     ct.best2.get(), comp, posOf, n, numComp, ct.parent.get());
  hipLaunchKernelGGL(emitKernel, gridFor(numComp), dim3(256), 0, stream(), ct.best1.get(),
                     ct.best2.get(), ct.parent.get(), numComp, ct.root.get(), ct.edges.u.get(),
                     ct.edges.v.get(), ct.edges.w.get(), ct.counters.get(), std::max(1, n - 1));
  s((int)hipGetLastError());
  int count = 0;
  s(download(&count, ct.counters, sizeof(int)));
  if (!s.ok())
    return s.code;
  if (count > n - 1)
    return kErrNonFinite; // more edges than a spanning forest holds: inconsistent keys
  *added = count - ct.edges.count;
  ct.edges.count = count;
  // Pointer jumping until every component points at its root (log depth;
  // 40 doublings cover any depth below 2^31). The jumps run in batches of
  // four with one read of the flag per batch: it is cleared before each jump,
  // so it tells whether the batch's last jump still moved a pointer.
  constexpr int kJumpsPerRead = 4;
  bool converged = false;
  for (int iter = 0; iter < 40 && !converged; iter += kJumpsPerRead) {
    for (int j = 0; j < kJumpsPerRead; ++j) {
      s(fill(ct.changed, 0, sizeof(int)));
      hipLaunchKernelGGL(jumpKernel, gridFor(numComp), dim3(256), 0, stream(), ct.root.get(),
                         numComp, ct.changed.get());
    }
    s((int)hipGetLastError());
    int changed = 0;
    s(download(&changed, ct.changed, sizeof(int)));
    if (!s.ok())
      return s.code;
    converged = changed == 0;
  }
  if (!converged)
    return kErrNonFinite; // hooks with a cycle: inconsistent keys
  hipLaunchKernelGGL(rootFlagsKernel, gridFor(numComp + 1), dim3(256), 0, stream(),
                     ct.root.get(), numComp, ct.flags.get());
  s(exclusiveScan(ct.flags.get(), ct.newId.get(), numComp + 1, ct.scanTemp.get()));
  hipLaunchKernelGGL(relabelKernel, gridFor(n), dim3(256), 0, stream(), comp, n, ct.root.get(),
                     ct.newId.get());
  s((int)hipGetLastError());
  s(download(newNumComp, ct.newId.get() + numComp, sizeof(int)));
  return s.code;
}

} // namespace cuhc
