"""Build recipes for the native libraries (all in-tree).

* ``shci_amd/libshci_hip.so``  -- gfx950 HIP kernels + the C-ABI of include/shci_hip.h
* ``shci_amd/libshci_host.so`` -- C++ host side (integrals, CPU heat-bath selection,
  the drop-in Hamiltonian/Davidson adapter and variational driver); links libshci_hip.so
* ``oracle/liboracle.so``      -- the CPU oracle (test infrastructure, built by oracle/Makefile)
"""
from __future__ import annotations

import os
import shutil
import subprocess
from pathlib import Path

ROOT = Path(__file__).resolve().parent.parent
PKG = ROOT / "shci_amd"
CSRC = PKG / "csrc"
HOST = PKG / "host"

HIP_LIB = PKG / "libshci_hip.so"
HOST_LIB = PKG / "libshci_host.so"
ORACLE_LIB = ROOT / "oracle" / "liboracle.so"

HIP_SOURCES = ["shci_hip.hip"]
HOST_SOURCES = ["chem_host.cc", "host_capi.cc", "variation.cc", "rdm_io.cc", "wf_io.cc"]

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
HIP_FLAGS = [
    "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-shared",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-munsafe-fp-atomics", "-Wall", "-Wno-unused-result",
    f"-I{ROOT / 'include'}",
]


def _run(cmd, cwd=None):
    r = subprocess.run(cmd, cwd=cwd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"build failed: {' '.join(map(str, cmd))}\n{r.stdout}\n{r.stderr}")
    return r


def build_hip(verbose: bool = False) -> Path:
    srcs = [str(CSRC / s) for s in HIP_SOURCES]
    tmp = HIP_LIB.with_suffix(".so.tmp")
    _run([HIPCC, *HIP_FLAGS, *srcs, "-o", str(tmp)])
    os.replace(tmp, HIP_LIB)
    return HIP_LIB


def build_host(verbose: bool = False) -> Path:
    srcs = [str(HOST / s) for s in HOST_SOURCES if (HOST / s).exists()]
    tmp = HOST_LIB.with_suffix(".so.tmp")
    cmd = ["g++", "-std=c++17", "-O3", "-fopenmp", "-fPIC", "-shared", "-Wall",
           f"-I{ROOT / 'include'}", f"-I{HOST / 'include'}", *srcs,
           f"-L{PKG}", "-lshci_hip", "-Wl,-rpath,$ORIGIN", "-o", str(tmp)]
    _run(cmd)
    os.replace(tmp, HOST_LIB)
    return HOST_LIB


def build_oracle(verbose: bool = False) -> Path:
    make = shutil.which("make") or "make"
    _run([make, "-B", "-C", str(ROOT / "oracle")])
    return ORACLE_LIB


def build_all(verbose: bool = False):
    build_oracle(verbose)
    build_hip(verbose)
    build_host(verbose)
