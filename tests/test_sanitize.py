"""The host code under AddressSanitizer and UndefinedBehaviorSanitizer: the
graph library (acctricnt_amd/csrc/graph_host.cpp) and the oracle
(oracle/oracle.c) are compiled with g++/gcc and -fsanitize=address,undefined
(host code only; nothing here is GPU code) into a driver that runs every
entry point on small graphs, and the run must end with no sanitizer report."""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SAN = ["-fsanitize=address,undefined", "-fno-sanitize-recover=all", "-fno-omit-frame-pointer", "-g", "-O1"]


@pytest.mark.skipif(shutil.which("g++") is None or shutil.which("gcc") is None, reason="no host compiler")
def test_host_code_under_sanitizers(tmp_path):
    inc = ["-I" + os.path.join(ROOT, "include"), "-I" + os.path.join(ROOT, "oracle")]
    obj = tmp_path / "oracle.o"
    subprocess.run(["gcc", *SAN, "-fopenmp", "-std=gnu11", "-c", os.path.join(ROOT, "oracle", "oracle.c"),
                    "-o", str(obj), *inc], check=True, capture_output=True, text=True)
    exe = tmp_path / "driver"
    subprocess.run(["g++", *SAN, "-fopenmp", "-std=c++17", os.path.join(ROOT, "tests", "sanitize", "driver.cpp"),
                    os.path.join(ROOT, "acctricnt_amd", "csrc", "graph_host.cpp"), str(obj), "-o", str(exe), *inc,
                    "-lm"], check=True, capture_output=True, text=True)
    work = tmp_path / "work"
    work.mkdir()
    env = {**os.environ, "ASAN_OPTIONS": "detect_leaks=1:abort_on_error=0", "UBSAN_OPTIONS": "print_stacktrace=1",
           "OMP_NUM_THREADS": "4"}
    env.pop("LD_PRELOAD", None)
    r = subprocess.run([str(exe), str(work)], capture_output=True, text=True, timeout=300, env=env)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "0 failures" in r.stdout
    assert "runtime error" not in r.stderr and "AddressSanitizer" not in r.stderr, r.stderr
