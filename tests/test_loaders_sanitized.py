"""The host file loaders (cloverleaf_amd/csrc/ep_io.cpp) built with g++ under
AddressSanitizer and UndefinedBehaviorSanitizer and driven by tests/native/loaders_harness.cpp
over graph, feature and embedding files, including malformed ones: every call must
return its documented result with no sanitizer report."""
import os
import shutil
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def write_inputs(d):
    rng = np.random.default_rng(1)
    with open(os.path.join(d, "g.tsv"), "w") as fh:
        for a, b in zip(rng.integers(0, 30, 500), rng.integers(0, 30, 500)):
            fh.write(f"u\tn{a}\tu\tn{b}\t{rng.uniform(0, 3):.3f}\n")
    with open(os.path.join(d, "f.tsv"), "w", encoding="utf-8") as fh:
        for _ in range(300):
            a = int(rng.integers(0, 14))
            toks = [f"x{int(rng.integers(0, 9))}" + (":" + str(int(rng.integers(0, 3))) if rng.random() < 0.5 else "")
                    for _ in range(int(rng.integers(0, 5)))]
            fh.write(f"u\tn{a}\t" + str(rng.choice([" ", "\u3000", "\xa0 "])).join(toks) + "\n")
    with open(os.path.join(d, "e.tsv"), "w", encoding="utf-8") as fh:
        fh.write("header line\n")
        for i in range(40):
            vals = ", ".join(f"{x:.6g}" for x in rng.standard_normal(5))
            fh.write(f"{'u' if i % 2 else 'i'}\tn{i}\t[{vals}\u3000]\n")
    with open(os.path.join(d, "bad1.tsv"), "w") as fh:
        fh.write("u\ta\t[1,2]\nu\ta\t[1,2]\n")
    with open(os.path.join(d, "bad2.tsv"), "w") as fh:
        fh.write("u\ta\t[\n")
    with open(os.path.join(d, "bad3.tsv"), "w") as fh:
        fh.write("u\ta\t\n")


@pytest.mark.skipif(shutil.which("g++") is None, reason="g++ is needed to build the sanitized harness")
def test_loaders_under_asan_ubsan(tmp_path):
    exe = str(tmp_path / "loaders_harness")
    build = subprocess.run(
        ["g++", "-std=c++17", "-O1", "-g", "-fsanitize=address,undefined", "-fno-sanitize-recover=all",
         "-fno-omit-frame-pointer", "-I", os.path.join(ROOT, "include"),
         os.path.join(ROOT, "cloverleaf_amd", "csrc", "ep_io.cpp"),
         os.path.join(ROOT, "tests", "native", "loaders_harness.cpp"), "-lz", "-o", exe],
        capture_output=True, text=True, timeout=300)
    assert build.returncode == 0, build.stderr[-2000:]
    write_inputs(str(tmp_path))
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:verify_asan_link_order=0", UBSAN_OPTIONS="print_stacktrace=1")
    run = subprocess.run([exe, str(tmp_path)], capture_output=True, text=True, timeout=120, env=env)
    assert run.returncode == 0, run.stderr[-4000:]
    assert "ERROR: AddressSanitizer" not in run.stderr and "runtime error" not in run.stderr
    out = run.stdout.splitlines()
    assert out[0] == "graph rc=0 n=30 e=600"
    assert all(line.startswith(f"features kind {k} rc=0 set=10 ") for k, line in zip(range(3), out[1:4]))
    assert out[4] == "embeddings filter 0 rc=0 rows=40 dims=5"
    assert out[5] == "embeddings filter 1 rc=0 rows=20 dims=5"
    assert out[6] == "bad bad1.tsv rc=-6 msg=found duplicate node at 1!"
    assert out[7].startswith("bad bad2.tsv rc=-1 msg=Error parsing line")
    assert out[8].startswith("bad bad3.tsv rc=-1 msg=Error parsing line")
