"""Builds the native libraries of cloverleaf_amd in-tree.

* ``libcloverleaf_ep.so``    HIP kernels + host runtime (C ABI of include/cloverleaf_ep.h),
                             compiled for gfx950 with hipcc.
* ``libcloverleaf_synth.so`` host-only seeded graph/feature generators (g++).

Objects go to ``cloverleaf_amd/_build/``; the libraries sit next to this file so they
travel with the source tree.
"""
from __future__ import annotations

import os
import shutil
import subprocess
from concurrent.futures import ThreadPoolExecutor
from pathlib import Path

PKG = Path(__file__).resolve().parent
CSRC = PKG / "csrc"
BUILD = PKG / "_build"
ROOT = PKG.parent

HIP_SOURCES = ["ep_kernels.hip", "ep_attention.hip", "ep_capi.hip"]
HOST_SOURCES = ["ep_io.cpp"]  # host-only parts of libcloverleaf_ep.so (g++, zlib)
SYNTH_SOURCES = ["synth.cpp"]

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("CLOVERLEAF_OFFLOAD_ARCH", "gfx950")

HIP_FLAGS = [
    f"--offload-arch={ARCH}",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-munsafe-fp-atomics",
    # no packed-FP32 VALU ops (v_pk_*_f32) in any generated code
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    f"-I{ROOT / 'include'}",
    f"-I{CSRC}",
]


def _run(cmd):
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    if r.returncode != 0:
        raise RuntimeError("command failed (%d): %s\n%s" % (r.returncode, " ".join(cmd), r.stdout))
    return r.stdout


def build_ep(verbose: bool = False) -> Path:
    BUILD.mkdir(exist_ok=True)
    objs = [BUILD / (Path(s).stem + ".o") for s in HIP_SOURCES]

    def compile_one(src_obj):
        src, obj = src_obj
        out = _run([HIPCC, *HIP_FLAGS, "-c", str(CSRC / src), "-o", str(obj)])
        if verbose and out.strip():
            print(out)

    with ThreadPoolExecutor(max_workers=len(HIP_SOURCES)) as ex:
        list(ex.map(compile_one, zip(HIP_SOURCES, objs)))
    cxx = shutil.which("g++") or "g++"
    for src in HOST_SOURCES:
        obj = BUILD / (Path(src).stem + ".o")
        _run([cxx, "-O3", "-std=c++17", "-fPIC", f"-I{ROOT / 'include'}", "-c", str(CSRC / src), "-o", str(obj)])
        objs.append(obj)
    lib = PKG / "libcloverleaf_ep.so"
    tmp = PKG / "libcloverleaf_ep.so.tmp"
    _run([HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", str(tmp), *map(str, objs), "-lz"])
    os.replace(tmp, lib)
    return lib


def build_synth(verbose: bool = False) -> Path:
    lib = PKG / "libcloverleaf_synth.so"
    tmp = PKG / "libcloverleaf_synth.so.tmp"
    cxx = shutil.which("g++") or "g++"
    _run([cxx, "-O3", "-std=c++17", "-fPIC", "-shared", "-pthread", "-o", str(tmp),
          *[str(CSRC / s) for s in SYNTH_SOURCES]])
    os.replace(tmp, lib)
    return lib


def build_all(verbose: bool = False):
    return build_ep(verbose), build_synth(verbose)


if __name__ == "__main__":
    for p in build_all(verbose=True):
        print("built", p)
