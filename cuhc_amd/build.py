"""Builds libcuhc.so for gfx950 from cuhc_amd/csrc with hipcc.

Every translation unit compiles in parallel into cuhc_amd/_build/, then links
into cuhc_amd/libcuhc.so next to the Python package (where _lib.py loads it).

    python cuhc_amd/build.py [--jobs N]

(run by path: importing the cuhc_amd package would load the library being built)
"""

import argparse
import concurrent.futures
import os
import pathlib
import shutil
import subprocess
import sys

PACKAGE = pathlib.Path(__file__).resolve().parent
CSRC = PACKAGE / "csrc"
BUILD = PACKAGE / "_build"
LIBRARY = PACKAGE / "libcuhc.so"

ARCH = "gfx950"
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

# -ffp-contract=off: the only fused multiply-adds are the explicit fmaf calls,
# so FP32 distances are the exact fold the tests and the oracle restate.
# -packed-fp32-ops: no v_pk_*_f32 instructions in any kernel.
COMMON = [
    "-O3", "-fPIC", "-std=c++17", "-ffp-contract=off", "-fno-fast-math",
    "-Wall", "-Wno-unused-function", "-Wno-unused-variable",
]
DEVICE = [
    f"--offload-arch={ARCH}", "-fno-gpu-flush-denormals-to-zero",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
]

SOURCES = [
    "runtime.hip", "primitives.hip", "distance.hip", "boruvka.hip", "core.hip", "dense.hip", "ward.hip",
    "mfma.hip", "mfma_probe.hip", "mlp.hip", "learned.hip", "edge_sort.hip", "contract.hip", "capi.hip", "noalloc.hip", "host_tree.cpp", "extract.cpp",
]

# Per-file extra flags. mfma.hip: inputs are validated finite, so min/max need
# no NaN canonicalisation (the epilogue is one v_min per value).
EXTRA = {"mfma.hip": ["-fno-honor-nans"], "mfma_probe.hip": ["-fno-honor-nans"]}

NOISE = "is not a recognized feature for this target"


def _compile(source):
    src = CSRC / source
    obj = BUILD / (source + ".o")
    if source.endswith(".cpp"):
        cmd = ["g++", *COMMON, "-I", str(CSRC), "-c", str(src), "-o", str(obj)]
    else:
        cmd = [HIPCC, *COMMON, *DEVICE, *EXTRA.get(source, []), "-I", str(CSRC), "-c", str(src),
               "-o", str(obj)]
    proc = subprocess.run(cmd, capture_output=True, text=True)
    out = "\n".join(line for line in (proc.stdout + proc.stderr).splitlines() if NOISE not in line)
    return source, proc.returncode, out, obj


def build(jobs=None, verbose=False):
    """Compiles every source and links libcuhc.so; raises RuntimeError on failure."""
    if shutil.which(HIPCC) is None and not pathlib.Path(HIPCC).exists():
        raise RuntimeError(f"hipcc not found at {HIPCC}")
    BUILD.mkdir(exist_ok=True)
    jobs = jobs or min(len(SOURCES), os.cpu_count() or 4, 16)
    objects, errors = [], []
    with concurrent.futures.ThreadPoolExecutor(max_workers=jobs) as pool:
        for source, code, out, obj in pool.map(_compile, SOURCES):
            if out.strip() and (verbose or code):
                print(f"[{source}]\n{out}", file=sys.stderr)
            if code:
                errors.append(source)
            objects.append(str(obj))
    if errors:
        raise RuntimeError(f"compilation failed: {', '.join(errors)}")
    tmp = LIBRARY.with_suffix(".so.tmp")
    cmd = [HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", str(tmp), *objects]
    proc = subprocess.run(cmd, capture_output=True, text=True)
    if proc.returncode:
        raise RuntimeError(f"link failed:\n{proc.stdout}{proc.stderr}")
    tmp.replace(LIBRARY)
    return LIBRARY


def main():
    parser = argparse.ArgumentParser(description=__doc__)
    parser.add_argument("--jobs", type=int, default=None)
    parser.add_argument("--verbose", action="store_true")
    args = parser.parse_args()
    print(build(args.jobs, args.verbose))


if __name__ == "__main__":
    main()
