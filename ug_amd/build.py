"""Build recipe for the native parts (called by __graft_entry__.build()).

  ug_amd/libugmg.so           HIP kernels + C ABI (hipcc, gfx950) + numproc layer
  ug_amd/libugmg_problems.so  synthetic UG-layout problems (g++, host only)
  oracle/liboracle.so         CPU restatement used by tests (gcc, test only)
"""
from __future__ import annotations

import os
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "ug_amd")
CSRC = os.path.join(PKG, "csrc")
BUILD = os.path.join(PKG, "_build")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"
# packed FP32 instructions disabled: hipcc must not emit v_pk_{add,mul,fma}_f32
HIP_FLAGS = ["-O3", "-fPIC", "-std=c++17", f"--offload-arch={ARCH}",
             "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
             "-Wall", "-Wno-unused-result", "-Wno-unused-function", "-Wno-unused-value"]
CXX_FLAGS = ["-O3", "-fPIC", "-std=c++17", "-Wall"]


def _run(cmd, cwd=None):
    print("+", " ".join(cmd), flush=True)
    subprocess.run(cmd, check=True, cwd=cwd)


def build_problems():
    _run(["g++", *CXX_FLAGS, "-fopenmp", "-shared", "-o", os.path.join(PKG, "libugmg_problems.so"),
          os.path.join(CSRC, "problems.cc")])


def build_oracle():
    _run(["make", "-s", "-C", os.path.join(ROOT, "oracle")])


def build_ugmg():
    os.makedirs(BUILD, exist_ok=True)
    objs = []
    for src in ("ugmg_kernels.hip", "ugmg_base.hip", "ugmg_assemble.hip", "ugmg_devsetup.hip",
                "ugmg_api.hip"):
        obj = os.path.join(BUILD, src.replace(".hip", ".o"))
        _run([HIPCC, *HIP_FLAGS, "-x", "hip", "-c", os.path.join(CSRC, src), "-o", obj])
        objs.append(obj)
    obj = os.path.join(BUILD, "ugmg_setup.o")
    _run(["g++", *CXX_FLAGS, "-fopenmp", "-c", os.path.join(CSRC, "ugmg_setup.cc"), "-o", obj])
    objs.append(obj)
    obj = os.path.join(BUILD, "ugmg_np.o")   # numproc layer over the C ABI (host only)
    _run(["g++", *CXX_FLAGS, "-c", os.path.join(CSRC, "ugmg_np.cc"), "-o", obj])
    objs.append(obj)
    obj = os.path.join(BUILD, "ugmg_dio.o")  # UG data files (host only)
    _run(["g++", *CXX_FLAGS, "-c", os.path.join(CSRC, "ugmg_dio.cc"), "-o", obj])
    objs.append(obj)
    obj = os.path.join(BUILD, "ugmg_mgio.o")  # UG multigrid files (host only)
    _run(["g++", *CXX_FLAGS, "-c", os.path.join(CSRC, "ugmg_mgio.cc"), "-o", obj])
    objs.append(obj)
    out = os.path.join(PKG, "libugmg.so")
    tmp = out + ".tmp"
    _run([HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp, *objs, "-lgomp"])
    os.replace(tmp, out)


def build_all():
    build_problems()
    build_oracle()
    build_ugmg()


if __name__ == "__main__":
    what = sys.argv[1:] or ["all"]
    for w in what:
        {"all": build_all, "problems": build_problems, "oracle": build_oracle, "ugmg": build_ugmg}[w]()
