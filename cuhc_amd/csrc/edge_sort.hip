// MST edges into the global edge order (weight, smaller endpoint, larger
// endpoint) on the GPU: the single-linkage tree needs the n - 1 edges sorted,
// which is a radix sort of (weight bits, smaller endpoint) keys here instead
// of a comparison sort on the host. Equal keys (two edges of one point with
// the same weight) are ordered by the larger endpoint on the host afterwards.
#include "mst.hpp"

#include "primitives.hpp"

#include <algorithm>
#include <chrono>

namespace cuhc {

namespace {

using PackedEdge = Edge32;

__global__ void edgeKeysKernel(const int *__restrict__ u, const float *__restrict__ w, int m,
                               int uBits, unsigned long long *__restrict__ keys,
                               unsigned int *__restrict__ idx) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= m)
    return;
  // Weights are non-negative floats (packed keys clamp): their bits order
  // like the values; +0 stands for -0.
  const float wi = w[i] == 0.0f ? 0.0f : w[i];
  keys[i] = (unsigned long long)__float_as_uint(wi) << uBits | (unsigned int)u[i];
  idx[i] = (unsigned int)i;
}

__global__ void edgeGatherKernel(const int *__restrict__ u, const int *__restrict__ v,
                                 const float *__restrict__ w, const unsigned int *__restrict__ idx,
                                 int m, PackedEdge *__restrict__ out) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < m) {
    const unsigned int j = idx[i];
    out[i] = PackedEdge{u[j], v[j], w[j]};
  }
}

} // namespace

int treeFromDeviceEdges(const DeviceEdges &edges, int n, double *tree) {
  const int m = edges.count;
  if (m <= 0)
    return 0;
  Status s;
  if (m < 65536) {
    std::vector<Edge> host(m);
    std::vector<int> hu(m), hv(m);
    std::vector<float> hw(m);
    s(download(hu.data(), edges.u, m * sizeof(int)));
    s(download(hv.data(), edges.v, m * sizeof(int)));
    s(download(hw.data(), edges.w, m * sizeof(float)));
    if (!s.ok())
      return s.code;
    for (int i = 0; i < m; ++i)
      host[i] = Edge{hu[i], hv[i], (double)hw[i]};
    sortEdges(host);
    treeFromOrderedMerges(host, n, tree);
    return 0;
  }
  int uBits = 1;
  while (uBits < 31 && (n - 1) >> uBits)
    ++uBits;
  DevArray<unsigned long long> keys, keysOut;
  DevArray<unsigned int> idx, idxOut;
  DevArray<PackedEdge> packed;
  DevArray<unsigned char> temp;
  s(keys.alloc(m));
  s(keysOut.alloc(m));
  s(idx.alloc(m));
  s(idxOut.alloc(m));
  s(packed.alloc(m));
  s(temp.alloc(sortTempBytes64(m)));
  if (!s.ok())
    return s.code;
  const dim3 grid((m + 255) / 256);
  hipLaunchKernelGGL(edgeKeysKernel, grid, dim3(256), 0, stream(), edges.u.get(),
                     edges.w.get(), m, uBits, keys.get(), idx.get());
  s((int)hipGetLastError());
  s(radixSortPairs(keys.get(), keysOut.get(), idx.get(), idxOut.get(), m, 32 + uBits,
                   temp.get()));
  hipLaunchKernelGGL(edgeGatherKernel, grid, dim3(256), 0, stream(), edges.u.get(),
                     edges.v.get(), edges.w.get(), idxOut.get(), m, packed.get());
  s((int)hipGetLastError());
  // The tree is written straight from the staged edges; runs of equal
  // (weight, u) are ordered by v in place as the assembly reaches them.
  s(downloadStaged(packed, m * sizeof(PackedEdge), [&](const void *p) {
    const auto t0 = std::chrono::steady_clock::now();
    PackedEdge *hp = static_cast<PackedEdge *>(const_cast<void *>(p));
    treeFromOrderedEdges(hp, (size_t)m, n, tree);
    if (timingEnabled())
      addMs(kPhaseHostTree,
            std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0)
                .count(),
            1);
  }));
  return s.code;
}

} // namespace cuhc
