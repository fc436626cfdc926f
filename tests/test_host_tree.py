"""The host assembly of merge trees (cuhc_amd/csrc/host_tree.cpp), checked on
the CPU against a plain union-find.

treeFromOrderedEdges writes scipy linkage rows from edges sorted by
(weight, smaller endpoint): it orders each run of equal (weight, u) by v as it
reaches the run, keeps size and id records only for merged clusters (a point
not merged yet is its own root of size 1 and id itself), halves find paths and
prefetches ahead. The reference here sorts by the full (weight, u, v) order
first and runs a union-find with a record for every point and no path
shortcuts. Integer weights make most edges tie. The library source is compiled
with g++ into a small driver (skipped when g++ is absent).
"""

import shutil
import subprocess
from pathlib import Path

import numpy as np
import pytest

CSRC = Path(__file__).resolve().parent.parent / "cuhc_amd" / "csrc"

DRIVER = r"""
#include "host_tree.hpp"
#include <cstdio>
#include <vector>
// stdin: n, m, then m lines "u v w"; stdout: m rows "a b h size".
int main() {
  int n = 0;
  size_t m = 0;
  if (std::scanf("%d %zu", &n, &m) != 2)
    return 1;
  std::vector<cuhc::Edge32> e(m);
  for (size_t i = 0; i < m; ++i)
    if (std::scanf("%d %d %f", &e[i].u, &e[i].v, &e[i].w) != 3)
      return 1;
  std::vector<double> tree(m * 4);
  for (int rep = 0; rep < 2; ++rep) { // the per-thread buffers are reused
    std::vector<cuhc::Edge32> work(e);
    cuhc::treeFromOrderedEdges(work.data(), m, n, tree.data());
  }
  for (size_t i = 0; i < m; ++i)
    std::printf("%.17g %.17g %.17g %.17g\n", tree[4 * i], tree[4 * i + 1], tree[4 * i + 2],
                tree[4 * i + 3]);
  return 0;
}
"""


@pytest.fixture(scope="module")
def driver(tmp_path_factory):
    gxx = shutil.which("g++")
    if gxx is None:
        pytest.skip("g++ not available")
    d = tmp_path_factory.mktemp("host_tree")
    (d / "driver.cpp").write_text(DRIVER)
    exe = d / "driver"
    subprocess.run([gxx, "-O2", "-std=c++17", "-I", str(CSRC), str(d / "driver.cpp"),
                    str(CSRC / "host_tree.cpp"), str(CSRC / "extract.cpp"), "-o", str(exe)],
                   check=True, capture_output=True, text=True)
    return exe


def run_driver(exe, n, edges):
    text = f"{n} {len(edges)}\n" + "".join(f"{u} {v} {w!r}\n" for u, v, w in edges)
    out = subprocess.run([str(exe)], input=text, capture_output=True, text=True, check=True)
    return np.array([[float(x) for x in line.split()] for line in out.stdout.splitlines()])


def reference_tree(n, edges):
    """Rows {min id, max id, weight, size} from a plain union-find over the
    edges in (weight, u, v) order; merge s makes cluster n + s."""
    parent = list(range(n))
    size = [1] * n
    ident = list(range(n))

    def find(x):
        while parent[x] != x:
            x = parent[x]
        return x

    rows = []
    for s, (u, v, w) in enumerate(sorted(edges, key=lambda e: (e[2], e[0], e[1]))):
        a, b = find(u), find(v)
        rows.append([min(ident[a], ident[b]), max(ident[a], ident[b]), w, size[a] + size[b]])
        if size[a] < size[b]:
            a, b = b, a
        parent[b] = a
        size[a] += size[b]
        ident[a] = n + s
    return np.array(rows, dtype=np.float64)


def random_spanning_tree(rng, n, weights):
    """A random tree on n points (each point joins an earlier one, in shuffled
    order) with weights drawn from `weights`, as (u < v, w) edges."""
    order = rng.permutation(n)
    edges = []
    for i in range(1, n):
        a, b = int(order[i]), int(order[rng.integers(0, i)])
        edges.append((min(a, b), max(a, b), float(np.float32(rng.choice(weights)))))
    return edges


@pytest.mark.parametrize("n, levels, seed", [(2, 1, 0), (50, 3, 1), (2000, 5, 2), (3000, 400, 3)])
def test_tree_from_ordered_edges_matches_plain_union_find(driver, n, levels, seed):
    rng = np.random.default_rng(seed)
    edges = random_spanning_tree(rng, n, np.arange(levels, dtype=np.float32) * 0.5)
    # The input order the device sort produces: (weight, u) only, v unordered.
    by_weight_u = sorted(edges, key=lambda e: (e[2], e[0], rng.random()))
    got = run_driver(driver, n, by_weight_u)
    np.testing.assert_array_equal(got, reference_tree(n, edges))


def test_chain_into_one_giant_cluster(driver):
    # Single-linkage shape of high-dimensional data: one cluster absorbs the
    # points one at a time (long find paths without compression), with ties.
    n = 4000
    rng = np.random.default_rng(9)
    order = rng.permutation(n)
    edges = []
    for i in range(1, n):
        a, b = int(order[i]), int(order[rng.integers(max(0, i - 3), i)])
        edges.append((min(a, b), max(a, b), float(i // 7)))
    got = run_driver(driver, n, sorted(edges, key=lambda e: (e[2], e[0])))
    np.testing.assert_array_equal(got, reference_tree(n, edges))
