"""Builds the CPU oracle (oracle.c) into oracle/_build/liboracle.so with gcc.

-ffp-contract=off keeps every FP32 operation as written (the explicit fmaf
calls are the only fused operations); -mavx2 -mfma lets gcc vectorise the
row loops without changing any result (IEEE single operations and fmaf).

The reference implementation is CUDA (nvcc, CUB, cooperative groups) and is
not buildable in this image, so there is no oracle/_ref build.

    python -m oracle.build
"""

import pathlib
import subprocess

HERE = pathlib.Path(__file__).resolve().parent
OUT = HERE / "_build" / "liboracle.so"
FLAGS = ["-O3", "-std=c11", "-fPIC", "-shared", "-ffp-contract=off", "-fno-fast-math",
         "-fno-math-errno", "-mavx2", "-mfma", "-Wall"]


def build():
    OUT.parent.mkdir(exist_ok=True)
    tmp = OUT.with_suffix(".so.tmp")
    cmd = ["gcc", *FLAGS, str(HERE / "oracle.c"), "-o", str(tmp), "-lm"]
    proc = subprocess.run(cmd, capture_output=True, text=True)
    if proc.returncode:
        raise RuntimeError(f"oracle build failed:\n{proc.stdout}{proc.stderr}")
    tmp.replace(OUT)
    return OUT


if __name__ == "__main__":
    print(build())
