"""Builds the CPU restatement (oracle/ep_cpu.cpp) into oracle/_build/ep_cpu.

Test infrastructure only: used by bench.py's CPU-baseline leg and the AUC comparison.
"""
import os
import shutil
import subprocess
from pathlib import Path

HERE = Path(__file__).resolve().parent


def build() -> Path:
    """oracle/_build/ep_cpu (command line: AUC runs, explicit-list batches) and
    oracle/_build/libep_cpu.so (ep_cpu_bench, the in-process CPU baseline)."""
    out = HERE / "_build"
    out.mkdir(exist_ok=True)
    exe = out / "ep_cpu"
    cxx = shutil.which("g++") or "g++"
    tmp = out / "ep_cpu.tmp"
    subprocess.run([cxx, "-O3", "-std=c++17", "-pthread", "-o", str(tmp), str(HERE / "ep_cpu.cpp")], check=True)
    os.replace(tmp, exe)
    lib = out / "libep_cpu.so"
    tmp = out / "libep_cpu.so.tmp"
    subprocess.run([cxx, "-O3", "-std=c++17", "-pthread", "-fPIC", "-shared", "-DEP_CPU_NO_MAIN", "-o", str(tmp),
                    str(HERE / "ep_cpu.cpp")], check=True)
    os.replace(tmp, lib)
    return exe


if __name__ == "__main__":
    print(build())
