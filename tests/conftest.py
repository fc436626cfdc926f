"""Test configuration: the repository root on sys.path and the `gpu` marker.

Tests marked `gpu` run the HIP library on a real MI355X; the rest run on the
CPU (oracle, host C++ in libcuhc.so, argument checks). CUHC_TEST_PRUNING=always
(or never) forces spatial pruning for the whole run; results must not change.
"""

import os
import pathlib
import sys

ROOT = pathlib.Path(__file__).resolve().parents[1]
sys.path.insert(0, str(ROOT))
sys.path.insert(0, str(ROOT / "tests"))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs HIP kernels)")
    if os.environ.get("CUHC_TEST_PRUNING"):
        from cuhc_amd import _lib

        _lib.set_pruning(os.environ["CUHC_TEST_PRUNING"])
