"""Build recipe for ``tensormesh_amd/libtensormesh_amd.so`` (gfx950).

HIP sources are compiled by ``hipcc --offload-arch=gfx950``; the host-only
setup code by ``g++``.  The shared library links ``libamdhip64.so.7``: when
torch is imported first, the process already holds torch's copy of that
soname and the library binds to it, so both share one HIP runtime.

Usage: ``python -m tensormesh_amd.csrc.build`` (or ``build()`` from Python).
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
PKG = os.path.dirname(HERE)
LIB = os.path.join(PKG, "libtensormesh_amd.so")
OBJ = os.path.join(PKG, "_build")

HIP_SOURCES = ["tm_assemble.hip", "tm_krylov.hip", "tm_hyper.hip", "tm_batched.hip", "tm_gnn.hip",
               "tm_condense.hip", "tm_pattern.hip", "tm_facet.hip"]
CXX_SOURCES = ["tm_host.cpp"]

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
HIP_FLAGS = [
    "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-Wall", "-Wno-unused-function",
    # no packed-FP32 instructions anywhere in device code
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
]
CXX_FLAGS = ["-O3", "-std=c++17", "-fPIC", "-pthread", "-Wall"]


def _run(cmd):
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"build step failed ({r.returncode}): {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    return r


def build(verbose: bool = False) -> str:
    """Compile every source and link the shared library; returns its path."""
    os.makedirs(OBJ, exist_ok=True)
    jobs = []
    for s in HIP_SOURCES:
        o = os.path.join(OBJ, s.replace(".hip", ".o"))
        jobs.append(([HIPCC, *HIP_FLAGS, "-c", os.path.join(HERE, s), "-o", o], o))
    for s in CXX_SOURCES:
        o = os.path.join(OBJ, s.replace(".cpp", ".o"))
        jobs.append((["g++", *CXX_FLAGS, "-c", os.path.join(HERE, s), "-o", o], o))
    with ThreadPoolExecutor(max_workers=min(8, len(jobs))) as ex:
        list(ex.map(lambda j: _run(j[0]), jobs))
    tmp = LIB + ".tmp"
    _run([HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-pthread", "-o", tmp,
          *[o for _, o in jobs]])
    os.replace(tmp, LIB)
    if verbose:
        print(f"built {LIB}")
    return LIB


def clean():
    shutil.rmtree(OBJ, ignore_errors=True)


if __name__ == "__main__":
    build(verbose=True)
    sys.exit(0)
