"""Build recipe for the native parts of seqtree_amd (run by ``__graft_entry__.build()``).

Produces, in-tree next to this file:

* ``libseqtree_hip.so`` -- the gfx950 HIP kernels and the C-ABI of ``include/seqtree_hip.h``,
  compiled by ``hipcc --offload-arch=gfx950``;
* ``libseqtree.so`` -- the C++ API (``namespace seqtree``, the reference's signatures) on top
  of that C-ABI, compiled by ``g++``;
* ``_core<ext>`` -- the pybind11 module mirroring the reference's ``seqtree._core`` hot-path
  bindings, linked against ``libseqtree.so``.

Every library finds its neighbours through ``$ORIGIN`` so the tree can be copied anywhere.
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys
import sysconfig
from concurrent.futures import ThreadPoolExecutor
from pathlib import Path

HERE = Path(__file__).resolve().parent
ROOT = HERE.parent
CSRC = HERE / "csrc"
HOST = CSRC / "host"
BUILD = HERE / "_build"

HIP_SOURCES = ["prep.hip", "gapblock.hip", "batch.hip", "gotoh.hip", "traceback.hip", "distance.hip",
               "single.hip", "probe.hip", "prefilter.hip", "capi.hip"]
HOST_SOURCES = ["codec.cpp", "substitution_matrix.cpp", "matrix_data.cpp", "api.cpp"]

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
CXX = os.environ.get("CXX", "g++")

HIP_FLAGS = [
    "-O3", "-std=c++17", "-fPIC", "--offload-arch=gfx950",
    # no packed FP32 instructions anywhere (the kernels are integer code in any case)
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-result",
]
CXX_FLAGS = ["-O3", "-std=c++17", "-fPIC", "-Wall", "-Wextra", "-Wno-unused-parameter"]


def _run(cmd: list[str]) -> None:
    res = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"build step failed ({res.returncode}): {' '.join(cmd)}\n{res.stdout}")


def ext_suffix() -> str:
    return sysconfig.get_config_var("EXT_SUFFIX") or ".so"


def build_hip(jobs: int = 8) -> Path:
    """Compile the HIP sources to objects in parallel, then link libseqtree_hip.so."""
    BUILD.mkdir(exist_ok=True)
    objs = []

    def compile_one(src: str) -> Path:
        obj = BUILD / (Path(src).stem + ".o")
        _run([HIPCC, *HIP_FLAGS, "-I", str(ROOT / "include"), "-c", str(CSRC / src), "-o", str(obj)])
        return obj

    with ThreadPoolExecutor(max_workers=jobs) as ex:
        objs = list(ex.map(compile_one, HIP_SOURCES))
    out = HERE / "libseqtree_hip.so"
    tmp = out.with_suffix(".so.tmp")
    _run([HIPCC, "-shared", "--offload-arch=gfx950", "-fPIC", *map(str, objs), "-o", str(tmp),
          "-Wl,-soname,libseqtree_hip.so"])
    os.replace(tmp, out)
    return out


def build_host() -> Path:
    """Compile the C++ API layer (libseqtree.so) against libseqtree_hip.so."""
    out = HERE / "libseqtree.so"
    tmp = out.with_suffix(".so.tmp")
    _run([CXX, *CXX_FLAGS, "-shared", "-I", str(ROOT / "include"), "-I", str(HOST),
          *[str(HOST / s) for s in HOST_SOURCES], "-o", str(tmp),
          "-L", str(HERE), "-lseqtree_hip", "-Wl,-rpath,$ORIGIN", "-Wl,-soname,libseqtree.so",
          "-lpthread"])
    os.replace(tmp, out)
    return out


def build_core() -> Path:
    """Compile the pybind11 module _core against libseqtree.so."""
    import pybind11

    out = HERE / ("_core" + ext_suffix())
    tmp = HERE / ("_core.tmp" + ext_suffix())
    py_inc = sysconfig.get_paths()["include"]
    _run([CXX, *CXX_FLAGS, "-shared", "-fvisibility=hidden", "-I", pybind11.get_include(),
          "-I", py_inc, "-I", str(ROOT / "include"), "-I", str(HOST), str(HOST / "bindings.cpp"),
          "-o", str(tmp), "-L", str(HERE), "-lseqtree", "-Wl,-rpath,$ORIGIN"])
    os.replace(tmp, out)
    return out


def build_all() -> None:
    build_hip()
    build_host()
    build_core()


def clean() -> None:
    shutil.rmtree(BUILD, ignore_errors=True)


if __name__ == "__main__":
    what = sys.argv[1:] or ["all"]
    for w in what:
        {"all": build_all, "hip": build_hip, "host": build_host, "core": build_core,
         "clean": clean}[w]()
