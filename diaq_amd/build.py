"""Build the native parts in-tree: libdiaq_hip.so (HIP kernels + C ABI, gfx950) and the diaq_gpu pybind11 module.

    python diaq_amd/build.py          # or __graft_entry__.build()

Objects go to diaq_amd/_build/ (git-ignored); the shared libraries land next to this file so they travel with the
source tree.
"""
from __future__ import annotations

import concurrent.futures as cf
import glob
import os
import subprocess
import sys
import sysconfig

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
OUT = os.path.join(HERE, "_build")
INCLUDE = os.path.join(ROOT, "include")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"

HIP_FLAGS = ["-O3", "-std=c++17", "-fPIC", f"--offload-arch={ARCH}", "-Wall", "-Wno-unused-result",
             # packed FP32 instructions stay off (they are generated from plain float code otherwise)
             "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
             "-I", CSRC, "-I", INCLUDE]


def _run(cmd):
    p = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    if p.returncode != 0:
        raise RuntimeError("build failed:\n" + " ".join(cmd) + "\n" + p.stdout)
    return p.stdout


def build_hip(jobs: int = 8) -> str:
    os.makedirs(OUT, exist_ok=True)
    srcs = sorted(glob.glob(os.path.join(CSRC, "*.hip")))
    objs = [os.path.join(OUT, os.path.basename(s)[:-4] + ".o") for s in srcs]
    with cf.ThreadPoolExecutor(max_workers=jobs) as ex:
        list(ex.map(_run, [[HIPCC, *HIP_FLAGS, "-c", s, "-o", o] for s, o in zip(srcs, objs)]))
    lib = os.path.join(HERE, "libdiaq_hip.so")
    tmp = lib + ".tmp"
    _run([HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp, *objs])
    os.replace(tmp, lib)
    return lib


def build_pybind() -> str:
    import pybind11
    src = os.path.join(CSRC, "pybind", "diaq_gpu.cpp")
    ext = sysconfig.get_config_var("EXT_SUFFIX")
    out = os.path.join(HERE, "diaq_gpu" + ext)
    tmp = out + ".tmp"
    py_inc = sysconfig.get_paths()["include"]
    _run(["g++", "-O3", "-std=c++17", "-fPIC", "-shared", "-fopenmp", "-Wall",
          "-I", pybind11.get_include(), "-I", py_inc, "-I", INCLUDE, "-I", os.path.join(CSRC, "pybind"),
          src, "-o", tmp, "-L", HERE, "-ldiaq_hip", "-Wl,-rpath,$ORIGIN"])
    os.replace(tmp, out)
    return out


def build_all() -> None:
    build_hip()
    build_pybind()


if __name__ == "__main__":
    build_hip()
    if "--hip-only" not in sys.argv:
        build_pybind()
    print("built", os.path.join(HERE, "libdiaq_hip.so"))
