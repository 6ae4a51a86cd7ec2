"""MI355X-native Chebyshev-filtered subspace iteration engine for TTDFT.

The engine is the C-ABI library ``libttdft_chfsi.so`` (HIP kernels for gfx950
plus host orchestration, declared in ``include/ttdft_chfsi.h``).  This package
holds its sources (``csrc/``), the C++ mirror of the reference's
ChebyShevFilter / EigenSolver / AX / MultTM interface (``host/``) and a ctypes
binding (``engine``) used by the tests and the benchmark.
"""
import os
import subprocess

PKG_DIR = os.path.dirname(os.path.abspath(__file__))


def build(jobs=8):
    """Compile the HIP library and the C++ host layer for gfx950 (always rebuilds)."""
    subprocess.run(["make", "-B", f"-j{jobs}", "-C", os.path.join(PKG_DIR, "csrc")], check=True)
    host = os.path.join(PKG_DIR, "host")
    if os.path.exists(os.path.join(host, "Makefile")):
        subprocess.run(["make", "-B", f"-j{jobs}", "-C", host], check=True)


from . import engine  # noqa: E402,F401
