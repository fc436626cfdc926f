// MST edges into the global edge order (weight, smaller endpoint, larger
// endpoint) on the GPU: the single-linkage tree needs the n - 1 edges sorted,
// which is a radix sort of (weight bits, smaller endpoint) keys here instead
// of a comparison sort on the host. Equal keys (two edges of one point with
// the same weight) are ordered by the larger endpoint on the host afterwards.
#include "host_tree.hpp"
#include "runtime.hpp"

#include <hipcub/hipcub.hpp>

#include <algorithm>
#include <cstring>

namespace cuhc {

namespace {

struct PackedEdge {
  int u, v;
  float w;
};

__global__ void edgeKeysKernel(const PackedEdge *__restrict__ edges, long long m, int uBits,
                               unsigned long long *__restrict__ keys,
                               unsigned int *__restrict__ idx) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= m)
    return;
  // Non-negative float bits order like the floats; +0 for -0.
  const float w = edges[i].w == 0.0f ? 0.0f : edges[i].w;
  keys[i] = (unsigned long long)__float_as_uint(w) << uBits | (unsigned int)edges[i].u;
  idx[i] = (unsigned int)i;
}

__global__ void edgeGatherKernel(const PackedEdge *__restrict__ edges,
                                 const unsigned int *__restrict__ idx, long long m,
                                 PackedEdge *__restrict__ out) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < m)
    out[i] = edges[idx[i]];
}

} // namespace

int sortEdgesGpu(std::vector<Edge> &edges, bool &done) {
  done = false;
  const long long m = (long long)edges.size();
  if (m < 2 || m > 0x7fffffffLL)
    return 0;
  std::vector<PackedEdge> host(m);
  int uBits = 1;
  for (long long i = 0; i < m; ++i) {
    const Edge &e = edges[i];
    const float f = (float)e.w;
    if (!(f >= 0.0f) || (double)f != e.w)
      return 0; // not float32 weights: the caller sorts on the host
    host[i] = PackedEdge{e.u, e.v, f};
    while ((unsigned)e.u >> uBits)
      ++uBits;
  }
  DevArray<PackedEdge> dIn, dOut;
  DevArray<unsigned long long> keys, keysOut;
  DevArray<unsigned int> idx, idxOut;
  DevArray<unsigned char> temp;
  Status s;
  s(dIn.alloc(m));
  s(dOut.alloc(m));
  s(keys.alloc(m));
  s(keysOut.alloc(m));
  s(idx.alloc(m));
  s(idxOut.alloc(m));
  size_t tempBytes = 0;
  s((int)hipcub::DeviceRadixSort::SortPairs(nullptr, tempBytes, (unsigned long long *)nullptr,
                                            (unsigned long long *)nullptr,
                                            (unsigned int *)nullptr, (unsigned int *)nullptr,
                                            (int)m, 0, 32 + uBits, stream()));
  s(temp.alloc(std::max<size_t>(tempBytes, 1)));
  s(upload(dIn, host.data(), m * sizeof(PackedEdge)));
  if (!s.ok())
    return s.code;
  const int blocks = (int)((m + 255) / 256);
  hipLaunchKernelGGL(edgeKeysKernel, dim3(blocks), dim3(256), 0, stream(), dIn.get(), m, uBits,
                     keys.get(), idx.get());
  s((int)hipGetLastError());
  s((int)hipcub::DeviceRadixSort::SortPairs(temp.get(), tempBytes, keys.get(), keysOut.get(),
                                            idx.get(), idxOut.get(), (int)m, 0, 32 + uBits,
                                            stream()));
  hipLaunchKernelGGL(edgeGatherKernel, dim3(blocks), dim3(256), 0, stream(), dIn.get(),
                     idxOut.get(), m, dOut.get());
  s((int)hipGetLastError());
  s(download(host.data(), dOut, m * sizeof(PackedEdge)));
  if (!s.ok())
    return s.code;
  for (long long i = 0; i < m; ++i)
    edges[i] = Edge{host[i].u, host[i].v, (double)host[i].w};
  // Runs of equal (weight, u): order by v.
  for (long long i = 0; i < m;) {
    long long j = i + 1;
    while (j < m && edges[j].w == edges[i].w && edges[j].u == edges[i].u)
      ++j;
    if (j - i > 1)
      std::sort(edges.begin() + i, edges.begin() + j, edgeLess);
    i = j;
  }
  done = true;
  return 0;
}

int treeFromMst(std::vector<Edge> edges, int n, double *tree) {
  bool sorted = false;
  if (edges.size() >= 65536) {
    const int status = sortEdgesGpu(edges, sorted);
    if (status != 0)
      return status;
  }
  if (!sorted)
    sortEdges(edges);
  treeFromOrderedMerges(edges, n, tree);
  return 0;
}

} // namespace cuhc
