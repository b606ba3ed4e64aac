"""Shared pytest setup: the ``gpu`` marker and import paths.

GPU tests are marked ``@pytest.mark.gpu`` and run only on a machine with a
ROCm device; everything else runs on the CPU.
"""
import os
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
for p in (ROOT, os.path.join(ROOT, "oracle")):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (ROCm device) to run")


def _ensure_built():
    """Builds the library and the oracle when they are missing, and brings
    them up to date when their object files are here (a build tree): make
    then rebuilds whatever changed.  A tree that carries only the built .so
    files (no objects) is used as it is."""
    import glob
    import subprocess

    csrc = os.path.join(ROOT, "reed-solomon_amd", "csrc")
    lib = os.path.join(ROOT, "reed-solomon_amd", "libtrs_gpu.so")
    orc = os.path.join(ROOT, "oracle", "liboracle.so")
    if not os.path.exists(lib) or glob.glob(os.path.join(csrc, "*.o")):
        subprocess.run(["make", "-j8", "-C", csrc], check=True)
    if not os.path.exists(orc) or os.path.exists(os.path.join(ROOT, "oracle", "simd_fft_gen.c")):
        subprocess.run(["make", "-C", os.path.join(ROOT, "oracle")], check=True)


_ensure_built()

import pytest  # noqa: E402


@pytest.fixture(scope="session")
def trs():
    import __graft_entry__

    return __graft_entry__.load()
