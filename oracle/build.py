"""Builds the CPU restatement (oracle/ep_cpu.cpp) into oracle/_build/ep_cpu.

Test infrastructure only: used by bench.py's CPU-baseline leg and the AUC comparison.
"""
import os
import shutil
import subprocess
from pathlib import Path

HERE = Path(__file__).resolve().parent


def build() -> Path:
    out = HERE / "_build"
    out.mkdir(exist_ok=True)
    exe = out / "ep_cpu"
    cxx = shutil.which("g++") or "g++"
    tmp = out / "ep_cpu.tmp"
    subprocess.run([cxx, "-O3", "-std=c++17", "-pthread", "-o", str(tmp), str(HERE / "ep_cpu.cpp")], check=True)
    os.replace(tmp, exe)
    return exe


if __name__ == "__main__":
    print(build())
