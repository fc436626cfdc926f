"""Benchmark: seconds per full single-linkage tree of 1,000,000 points in 32-d.

Workload (the 1M-point single-linkage configuration of the headline metric,
the GEMM-shaped one): n = 1,000,000 points, d = 32, float32 i.i.d. N(0,1) from
numpy's default_rng(3); exact single linkage through Boruvka's MST.
A step is one full tree (scipy-format Z back in host memory) from points
already resident in HBM (cuhc_linkage_device). The timed form is the mixed
one: FP16 MFMA candidate tiles + certified FP32 recheck; the plain FP32 form is
timed once beside it and the trees must be identical.

Prints one JSON line (rank 0). --gpus > 1 is not supported (one GPU).

The same line carries the host-array time of the drop-in entry point
(cuhc_linkage, PCIe upload included), the 1M-point 8-d HDBSCAN trees per
core-distance method (auto, learned radius, radix select) with their identity,
and the CPU baseline (scikit-learn's Prim MST, one thread, on a prefix).

    python bench.py [--steps K] [--warmup W] [--n N] [--plain-steps P] [--no-c3]
"""

import argparse
import json
import os
import pathlib
import sys
import time

import numpy as np

ROOT = pathlib.Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

METRIC = "s per full tree @1M pts + dist-tile MFMA util; mixed/learned vs plain FP32"
FP16_MFMA_PEAK_TFLOPS = 2500.0  # MI355X dense FP16 MFMA
HBM_PEAK_GBS = 8000.0


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def make_points(n, d, seed):
    return np.random.default_rng(seed).standard_normal((n, d), dtype=np.float32)


def tree_on_device(lib, ptr, n, d, tree):
    from cuhc_amd import _lib
    _lib.check(lib.cuhc_linkage_device(ptr, n, d, 0, 2.0, 0, 1, _lib.double_ptr(tree)))


PROFILE_C4 = ROOT / "profiles" / "c4_single_linkage_1M_d32_mixed.md"


def profiled_traffic():
    """HBM bytes per launch of the candidate kernel, from the committed rocprofv3
    FETCH_SIZE/WRITE_SIZE summary (its 'traffic_bytes_per_launch' line), or None."""
    try:
        for line in PROFILE_C4.read_text().splitlines():
            if line.startswith("traffic_bytes_per_launch:"):
                return float(line.split(":", 1)[1])
    except OSError:
        pass
    return None


def calibrate_copy(mib):
    """One device-to-device copy of mib MiB (a plain copy kernel of known size,
    run under rocprofv3 --pmc to scale FETCH_SIZE / WRITE_SIZE)."""
    import ctypes
    from cuhc_amd import _lib
    nbytes = mib << 20
    src, dst = ctypes.c_void_p(), ctypes.c_void_p()
    _lib.check(_lib.lib.cuhc_device_alloc(nbytes, ctypes.byref(src)))
    _lib.check(_lib.lib.cuhc_device_alloc(nbytes, ctypes.byref(dst)))
    try:
        for _ in range(2):
            _lib.check(_lib.lib.cuhc_copy_device(dst, src, nbytes))
    finally:
        _lib.lib.cuhc_device_free(src)
        _lib.lib.cuhc_device_free(dst)
    print(json.dumps({"copy_kib_per_launch": nbytes >> 10, "launches": 2}), flush=True)


# MFMA probe shapes (cuhc_mfma_probe): what the candidate pass's instruction
# mix can issue without memory traffic, beside the data-sheet peak.
MFMA_PROBE_SHAPES = {0: "32x32x16 f16, accumulate in place",
                     1: "16x16x32 f16, accumulate in place",
                     2: "candidate loop: seeded 32x32x16, LDS operands, min epilogue",
                     3: "candidate loop on 16x16x32: seeded, LDS operands, min epilogue",
                     4: "candidate loop on 32x32x16, 2 row blocks, 4 waves per SIMD",
                     5: "candidate loop on 16x16x32, 4 row blocks, 4 waves per SIMD",
                     6: "symmetric candidate loop on 32x32x16 (row and column tests), "
                        "flops counted twice"}


def mfma_ceiling(workgroups=2048, iters=4096):
    """TFLOP/s of each probe shape (three timed launches each, HIP events)."""
    import ctypes
    from cuhc_amd import _lib
    out = {}
    for shape, name in MFMA_PROBE_SHAPES.items():
        ms, flops = ctypes.c_float(), ctypes.c_double()
        _lib.check(_lib.lib.cuhc_mfma_probe(shape, workgroups, iters, ctypes.byref(ms),
                                            ctypes.byref(flops)))
        tf = flops.value / (ms.value * 1e-3) / 1e12
        out[name] = {"ms": round(ms.value, 3), "tflops": round(tf, 1),
                     "frac_of_peak": round(tf / FP16_MFMA_PEAK_TFLOPS, 3)}
    return out


def timed_steps(fn, steps):
    # Each call is blocking: it returns after the tree is in host memory and
    # the library's stream is idle, so wall time brackets exactly the steps.
    t0 = time.perf_counter()
    for _ in range(steps):
        fn()
    return (time.perf_counter() - t0) / steps


def host_info():
    """Cores, CPU model and memory of the machine the CPU baseline runs on."""
    info = {"nproc": os.cpu_count()}
    try:
        import subprocess
        out = subprocess.run(["lscpu"], capture_output=True, text=True, timeout=10).stdout
        for line in out.splitlines():
            if line.startswith("Model name:"):
                info["cpu_model"] = line.split(":", 1)[1].strip()
    except (OSError, subprocess.SubprocessError):
        pass
    try:
        for line in pathlib.Path("/proc/meminfo").read_text().splitlines():
            if line.startswith("MemTotal:"):
                info["ram_gib"] = round(int(line.split()[1]) / 2**20, 1)
    except OSError:
        pass
    return info


def sklearn_single_linkage(X64):
    """scikit-learn's CPU single-linkage tree: Prim's MST over the data matrix
    (O(n) memory), edges stably sorted by weight, then make_single_linkage."""
    from sklearn.cluster._hdbscan._linkage import make_single_linkage, mst_from_data_matrix
    from sklearn.metrics import DistanceMetric
    mst = mst_from_data_matrix(X64, np.zeros(len(X64)), DistanceMetric.get_metric("euclidean"))
    mst = mst[np.argsort(mst["distance"], kind="mergesort")]
    h = make_single_linkage(mst)
    return np.column_stack([h["left_node"], h["right_node"], h["value"], h["cluster_size"]])


def cpu_baseline(d, seed, sample_n, full_n):
    """The reference's CPU comparison path for this workload (scikit-learn's
    Prim MST + single-linkage tree, one thread: BLAS pinned by threadpoolctl)
    on a bounded prefix of the same data, extrapolated to full_n by n^2 (the
    cost of Prim's algorithm on a dense graph). The prefix's GPU tree is
    compared with scikit-learn's FP64 tree."""
    from threadpoolctl import threadpool_limits

    import cuhc_amd as cuhc
    X = make_points(full_n, d, seed)[:sample_n].copy()
    with threadpool_limits(1):
        t0 = time.perf_counter()
        Zs = sklearn_single_linkage(X.astype(np.float64))
        t = time.perf_counter() - t0
    Zg = cuhc.linkage(X, method="single")
    ours, theirs = tree_clusters(Zg), tree_clusters(Zs)
    common = sorted(ours.keys() & theirs.keys())
    a = np.array([ours[c] for c in common])
    b = np.array([theirs[c] for c in common])
    out = {"value": t * (full_n / sample_n) ** 2, "unit": "s", "cores": 1, "kind": "reference",
           "sample": f"scikit-learn {_sklearn_version()} mst_from_data_matrix + "
                     f"make_single_linkage (FP64, threadpool_limits(1)) on the first "
                     f"{sample_n} points (d={d}) in {t:.2f} s, extrapolated to n={full_n} by n^2",
           "host": host_info(),
           "prefix_tree_vs_sklearn": {
               "clusters_differing_by_tie_order": len(ours.keys() - theirs.keys()),
               "max_rel_height_error": float((np.abs(a - b) / b).max())}}
    return out


def tree_clusters(Z):
    """{(smallest point, size): height} for every cluster a merge tree creates
    (tied merges may come in either order; the clusters and heights agree)."""
    n = len(Z) + 1
    smallest = np.empty(2 * n - 1, dtype=np.int64)
    smallest[:n] = np.arange(n)
    out = {}
    for i, (a, b, h, size) in enumerate(np.asarray(Z).tolist()):
        smallest[n + i] = min(smallest[int(a)], smallest[int(b)])
        out[(int(smallest[n + i]), int(size))] = h
    return out


def _sklearn_version():
    import sklearn
    return sklearn.__version__


def make_c3(seed=2):
    """20 isotropic Gaussian blobs in 8-d (centres uniform in [-10, 10], std 1,
    45,000 points each) plus 100,000 uniform noise points over their bounding box."""
    rng = np.random.default_rng(seed)
    centres = rng.uniform(-10, 10, (20, 8))
    blobs = [c + rng.standard_normal((45_000, 8)) for c in centres]
    X = np.vstack(blobs)
    noise = rng.uniform(X.min(0), X.max(0), (100_000, 8))
    return np.vstack([X, noise]).astype(np.float32)


C3_WORKLOAD = "HDBSCAN mcs=min_samples=25, n=1,000,000, d=8, 20 blobs + 10% noise"
CORE_METHODS = {"auto": 0, "fused": 1, "radix": 2, "learned": 3}


def _large_k_block(lib, dev, n, d, k=100):
    """The same points with min_samples = 100, above the fused top-k's k <= 64:
    auto hands the core distances to the learned radius; the method without
    it is the full-row radix select. One tree each, identical trees."""
    from cuhc_amd import _lib
    tree = np.empty((n - 1, 4))
    core = np.empty(n, np.float32)
    out, results = {"min_samples": k}, []
    for method in ["auto", "radix"]:
        _lib.learned_stats()
        t = timed_steps(lambda: _lib.check(lib.cuhc_hdbscan_tree_device(  # noqa: B023
            dev.ptr, n, d, 0, 2.0, k, 1.0, CORE_METHODS[method], _lib.double_ptr(tree),
            _lib.float_ptr(core))), 1)
        out[method] = {"tree_s": t}
        if method == "auto":
            out[method]["learned_stats"] = _lib.learned_stats()
        results.append((tree.copy(), core.copy()))
        log(f"c3 k={k} {method}: {t:.3f} s")
    out["identical"] = bool(np.array_equal(results[0][0], results[1][0]) and
                            np.array_equal(results[0][1], results[1][1]))
    out["radix_vs_auto"] = out["radix"]["tree_s"] / out["auto"]["tree_s"]
    return out


def c3_block(methods, steps, radix_steps=1, large_k=True):
    """HDBSCAN trees of the 1M-point 8-d workload from points resident in HBM,
    one entry per core-distance method (auto: fused k-d top-k + pruned
    Boruvka; learned: the radius network's proposals, verified; radix: the
    full-row radix select), with the identity of their trees and core
    distances and the extraction time; large_k adds the same points at
    min_samples = 100 (auto, which then uses the learned radius, against the
    radix select)."""
    from cuhc_amd import _lib
    from cuhc_amd._hdbscan import tree_to_labels
    lib = _lib.lib
    X = make_c3()
    n, d, k = len(X), 8, 25
    dev = _lib.DeviceArray(X)
    tree = np.empty((n - 1, 4))
    core = np.empty(n, np.float32)
    out, trees = {"workload": C3_WORKLOAD}, {}
    try:
        if large_k:
            out["k100"] = _large_k_block(lib, dev, n, d)
        for method in methods:
            call = lambda: _lib.check(lib.cuhc_hdbscan_tree_device(  # noqa: E731
                dev.ptr, n, d, 0, 2.0, k, 1.0, CORE_METHODS[method], _lib.double_ptr(tree),
                _lib.float_ptr(core)))
            reps = radix_steps if method == "radix" else steps
            if method != "radix":
                call()
            _lib.set_timing(True)
            _lib.reset_timings()
            _lib.prune_stats()
            _lib.learned_stats()
            t = timed_steps(call, reps)
            out[method] = {"tree_s": t, "steps": reps, "phases_ms": {
                k2: round(v[0] / reps, 2) for k2, v in _lib.read_timings().items() if v[1]},
                "prune_stats": _lib.prune_stats()}
            if method == "learned":
                out[method]["learned_stats_per_call"] = {
                    k2: v / reps for k2, v in _lib.learned_stats().items()}
            _lib.set_timing(False)
            trees[method] = (tree.copy(), core.copy())
            log(f"c3 {method}: {out[method]}")
    finally:
        dev.free()
    first = next(iter(trees.values()))
    out["identical_across_methods"] = all(
        np.array_equal(a, first[0]) and np.array_equal(c, first[1]) for a, c in trees.values())
    if "auto" in out and "learned" in out:
        out["learned_vs_auto"] = out["learned"]["tree_s"] / out["auto"]["tree_s"]
    t0 = time.perf_counter()
    labels, _ = tree_to_labels(first[0], 25)
    out.update(extraction_s=time.perf_counter() - t0,
               clusters=int(labels.max() + 1), noise=int((labels < 0).sum()))
    return out


def other_config(name, args):
    """Seconds per call for the other configurations (parity-test cases of the
    same path; printed as one JSON line, not the headline)."""
    from cuhc_amd import _lib
    lib = _lib.lib
    out = {"config": name}
    if name == "c3":
        out.update(c3_block(args.core_methods.split(","), args.steps))
    else:
        spec = {"c1": (20_000, 32, 0, 0, 1, 0, "single/mst, n=20k, d=32"),
                "c2": (100_000, 32, 1, 3, 3, 0, "ward/rnn, n=100k, d=32"),
                "c5": (100_000, 64, 4, 2, 0, 5, "average/dense cosine, n=100k, d=64")}[name]
        n, d, seed, linkage_id, algorithm, metric, label = spec
        X = make_points(n, d, seed)
        if name == "c5":
            X /= np.linalg.norm(X, axis=1, keepdims=True)
        dev = _lib.DeviceArray(X)
        tree = np.empty((n - 1, 4))
        call = lambda: _lib.check(lib.cuhc_linkage_device(  # noqa: E731
            dev.ptr, n, d, metric, 2.0, linkage_id, algorithm, _lib.double_ptr(tree)))
        call()
        _lib.set_timing(True)
        _lib.reset_timings()
        t = timed_steps(call, args.steps)
        out["phases_ms"] = {k2: round(v[0] / args.steps, 2) for k2, v in
                            _lib.read_timings().items() if v[1]}
        _lib.set_timing(False)
        dev.free()
        out.update(tree_s=t, workload=label)
    print(json.dumps(out), flush=True)


def main():
    parser = argparse.ArgumentParser(description=__doc__)
    parser.add_argument("--config", default="c4", choices=["c1", "c2", "c3", "c4", "c5"])
    parser.add_argument("--gpus", type=int, default=1)
    parser.add_argument("--steps", type=int, default=3)
    parser.add_argument("--warmup", type=int, default=1)
    parser.add_argument("--n", type=int, default=1_000_000)
    parser.add_argument("--d", type=int, default=32)
    parser.add_argument("--seed", type=int, default=3)
    parser.add_argument("--plain-steps", type=int, default=1)
    parser.add_argument("--cpu-sample", type=int, default=20_000)
    parser.add_argument("--no-cpu", action="store_true")
    parser.add_argument("--core-methods", default="auto,learned,radix",
                        help="c3: comma-separated core-distance methods to time")
    parser.add_argument("--no-c3", action="store_true",
                        help="leave out the nested HDBSCAN (1M, d=8) block of the default run")
    parser.add_argument("--host-steps", type=int, default=2,
                        help="trees timed from host memory (cuhc_linkage, PCIe included)")
    parser.add_argument("--print-maps", action="store_true",
                        help="print the memory-map lines of libcuhc.so and the HIP runtime "
                             "after loading (to attribute a crash address)")
    parser.add_argument("--mfma-ceiling", action="store_true",
                        help="only run the MFMA issue-rate probe (attainable FP16 ceiling)")
    parser.add_argument("--calibrate-copy-mib", type=int, default=0,
                        help="only run a device-to-device copy of this size (counter scaling)")
    args = parser.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py runs on one GPU only (there is no multi-GPU path); "
              "use --gpus 1", file=sys.stderr)
        sys.exit(2)

    import cuhc_amd  # noqa: F401
    from cuhc_amd import _lib

    if args.print_maps:
        with open("/proc/self/maps") as maps:
            for line in maps:
                if any(k in line for k in ("libcuhc", "libamdhip64", "libhsa-runtime")):
                    log(f"maps: {line.rstrip()}")

    if args.calibrate_copy_mib > 0:
        calibrate_copy(args.calibrate_copy_mib)
        return
    if args.mfma_ceiling:
        print(json.dumps({"mfma_ceiling": mfma_ceiling()}), flush=True)
        return
    if args.config != "c4":
        other_config(args.config, args)
        return
    n, d = args.n, args.d
    log(f"generating n={n} d={d} seed={args.seed}")
    X = make_points(n, d, args.seed)
    dev = _lib.DeviceArray(X)  # inputs resident in HBM before the timed region
    tree = np.empty((n - 1, 4), dtype=np.float64)
    lib = _lib.lib

    # Mixed form: warm-up, then K timed trees with per-phase HIP-event timing.
    _lib.set_precision("mixed")
    for i in range(args.warmup):
        log(f"warmup {i + 1}/{args.warmup}")
        tree_on_device(lib, dev.ptr, n, d, tree)
    _lib.set_timing(True)
    _lib.reset_timings()
    _lib.mixed_stats()
    log(f"timing {args.steps} steps (mixed)")
    s_mixed = timed_steps(lambda: tree_on_device(lib, dev.ptr, n, d, tree), args.steps)
    phases = _lib.read_timings()
    stats = _lib.mixed_stats()
    _lib.set_timing(False)
    z_mixed = tree.copy()
    log(f"mixed: {s_mixed:.3f} s/tree; phases {phases}; stats {stats}")

    ms, launches, pairs = phases["mfma_candidates"]
    kernel_ms = ms / max(launches, 1)
    flops_per_launch = 2.0 * d * pairs / max(launches, 1)
    achieved = flops_per_launch / (kernel_ms * 1e-3) / 1e12 if kernel_ms > 0 else 0.0

    result = {
        "metric": METRIC, "value": s_mixed, "unit": "s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": s_mixed * 1e3, "higher_is_better": False,
        "scaling": "strong", "vs_baseline": None, "dtype": "fp16 mfma + fp32",
        "data": "synthetic N(0,1) float32, seed 3",
        "config": {"workload": f"single linkage (Boruvka MST), n={n}, d={d}, euclidean, "
                               "points resident in HBM", "n": n, "d": d,
                   "form": "mixed: FP16 MFMA candidate tiles + certified FP32 recheck",
                   "parallelism": "1 GPU"},
        "roofline": {"bound": "mfma", "kernel": "mfmaCandidateKernel<2>",
                     "achieved": round(achieved, 2), "peak": FP16_MFMA_PEAK_TFLOPS,
                     "unit": "TFLOP/s", "frac": round(achieved / FP16_MFMA_PEAK_TFLOPS, 4),
                     "traffic": profiled_traffic(), "avg_kernel_ms": round(kernel_ms, 3),
                     "launches_per_tree": launches / max(args.steps, 1),
                     "algorithmic_flops_per_launch": flops_per_launch},
        "phase_ms_per_tree": {k: round(v[0] / max(args.steps, 1), 3) for k, v in phases.items()
                              if v[1]},
        "mixed_stats_per_tree": {k: v / max(args.steps, 1) for k, v in stats.items()},
    }

    # The same instruction mix without memory traffic or records (the probe's
    # candidate-loop shape) on this GPU: an attainable ceiling beside the
    # data-sheet peak.
    ceiling = mfma_ceiling()
    attainable = ceiling[MFMA_PROBE_SHAPES[2]]["tflops"]
    result["roofline"]["attainable"] = attainable
    result["roofline"]["frac_of_attainable"] = round(achieved / attainable, 4)
    result["mfma_probe"] = ceiling
    log(f"mfma probe: {ceiling}")

    if args.plain_steps > 0:
        _lib.set_precision("fp32")
        log(f"timing {args.plain_steps} steps (plain fp32)")
        s_plain = timed_steps(lambda: tree_on_device(lib, dev.ptr, n, d, tree),
                              args.plain_steps)
        result["plain_fp32_s_per_tree"] = s_plain
        result["speedup_vs_plain"] = s_plain / s_mixed
        result["tree_identical_to_plain"] = bool(np.array_equal(tree, z_mixed))
        log(f"plain: {s_plain:.3f} s/tree, identical={result['tree_identical_to_plain']}")

    if args.host_steps > 0:
        # The drop-in entry point: points in host memory, so every tree adds
        # the upload of n*d*4 bytes over PCIe. Reported beside value, not as it.
        _lib.set_precision("mixed")
        log(f"timing {args.host_steps} steps (mixed, host-array cuhc_linkage)")
        call = lambda: _lib.check(lib.cuhc_linkage(  # noqa: E731
            _lib.float_ptr(X), n, d, 0, 2.0, 0, 1, _lib.double_ptr(tree)))
        s_host = timed_steps(call, args.host_steps)
        result["host_array"] = {
            "entry_point": "cuhc_linkage", "s_per_tree": s_host,
            "upload_bytes": n * d * 4, "tree_identical": bool(np.array_equal(tree, z_mixed))}
        log(f"host array: {s_host:.3f} s/tree")
    _lib.set_precision("auto")
    dev.free()

    if not args.no_c3:
        log("c3 block")
        result["c3"] = c3_block(args.core_methods.split(","), args.steps)

    if not args.no_cpu:
        log("cpu baseline")
        result["cpu_baseline"] = cpu_baseline(d, args.seed, min(args.cpu_sample, n), n)
    print(json.dumps(result), flush=True)


if __name__ == "__main__":
    main()
