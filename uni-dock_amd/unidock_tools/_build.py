"""Build the HIP library (libunidock_amd.so) in-tree for gfx950.

Compiles every time it is called; objects go to uni-dock_amd/_build/ and the
shared library next to this file so that it ships with the source tree.
"""
from __future__ import annotations

import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
PKG = os.path.dirname(HERE)                 # uni-dock_amd/
CSRC = os.path.join(PKG, "csrc")
BUILD = os.path.join(PKG, "_build")
LIB = os.path.join(HERE, "libunidock_amd.so")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

# gfx950 only; packed-FP32 instructions are disabled for every HIP source
HIP_FLAGS = ["--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC",
             "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
             "-Wno-unused-command-line-argument"]
SOURCES = ["ud_kernels.hip", "ud_capi.cpp", "ud_post.cpp", "ud_pdbqt.cpp", "synth_host.cpp"]


def _compile(src: str) -> str:
    obj = os.path.join(BUILD, os.path.splitext(src)[0] + ".o")
    cmd = [HIPCC, *HIP_FLAGS, "-c", os.path.join(CSRC, src), "-o", obj]
    if src.endswith(".cpp"):
        cmd = [HIPCC, *HIP_FLAGS, "-x", "hip", "-c", os.path.join(CSRC, src), "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError(f"hipcc failed for {src}")
    return obj


def build(verbose: bool = False) -> str:
    os.makedirs(BUILD, exist_ok=True)
    with ThreadPoolExecutor(max_workers=5) as ex:
        objs = list(ex.map(_compile, SOURCES))
    tmp = LIB + ".tmp"
    cmd = [HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", tmp, *objs]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError("link of libunidock_amd.so failed")
    os.replace(tmp, LIB)
    if verbose:
        print("built", LIB)
    build_host(verbose)
    return LIB


HOST_DIR = os.path.join(os.path.dirname(CSRC), "host")
HOST_LIB = os.path.join(HOST_DIR, "libunidock_amd_vina.so")


def build_host(verbose: bool = False) -> str:
    """The Vina-style C++ host (host/unidock_vina.cpp) as a shared library
    linked against libunidock_amd.so: host code, built with g++."""
    include = os.path.join(os.path.dirname(os.path.dirname(CSRC)), "include")
    libdir = os.path.dirname(LIB)
    cmd = ["g++", "-std=c++17", "-O2", "-fPIC", "-shared", "-Wall", "-Werror", "-I", include, "-I", HOST_DIR,
           os.path.join(HOST_DIR, "unidock_vina.cpp"), "-o", HOST_LIB + ".tmp", "-L", libdir, "-lunidock_amd",
           "-Wl,-rpath,$ORIGIN/../unidock_tools"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError("build of libunidock_amd_vina.so failed")
    os.replace(HOST_LIB + ".tmp", HOST_LIB)
    if verbose:
        print("built", HOST_LIB)
    return HOST_LIB


if __name__ == "__main__":
    build(verbose=True)
