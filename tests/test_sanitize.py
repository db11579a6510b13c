"""The host code under AddressSanitizer and UndefinedBehaviorSanitizer: the
graph library (acctricnt_amd/csrc/graph_host.cpp) and the oracle
(oracle/oracle.c) are compiled with g++/gcc and -fsanitize=address,undefined
(host code only; nothing here is GPU code) into a driver that runs every
entry point on small graphs, and the run must end with no sanitizer report.
The serving core (acctricnt_amd/csrc/serve_core.h) runs under
ThreadSanitizer over its fake device with concurrent clients."""
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
    env = {**os.environ, "ASAN_OPTIONS": "detect_leaks=1:abort_on_error=0:verify_asan_link_order=0",
           "UBSAN_OPTIONS": "print_stacktrace=1", "OMP_NUM_THREADS": "4"}
    r = subprocess.run([str(exe), str(work)], capture_output=True, text=True, timeout=300, env=env)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "0 failures" in r.stdout
    assert "runtime error" not in r.stderr and "AddressSanitizer" not in r.stderr, r.stderr


@pytest.mark.skipif(shutil.which("g++") is None, reason="no host compiler")
def test_serving_core_under_thread_sanitizer(tmp_path):
    """Six client threads submit, wait (some with timeouts), cancel and
    release queries while a seventh registers and removes graphs under a
    budget of three graphs, so admissions, evictions and re-admissions race
    with batches on four streams; ThreadSanitizer must report nothing."""
    exe = tmp_path / "serve_tsan"
    src = os.path.join(ROOT, "tests", "serve")
    subprocess.run(["g++", "-fsanitize=thread", "-g", "-O1", "-std=c++17", "-pthread",
                    "-I" + os.path.join(ROOT, "acctricnt_amd", "csrc"), "-I" + os.path.join(ROOT, "include"),
                    os.path.join(src, "serve_stress.cpp"), os.path.join(src, "serve_driver.cpp"), "-o", str(exe)],
                   check=True, capture_output=True, text=True)
    env = {**os.environ, "TSAN_OPTIONS": "halt_on_error=1:second_deadlock_stack=1"}
    r = subprocess.run([str(exe)], capture_output=True, text=True, timeout=600, env=env)
    assert r.returncode == 0, r.stdout + r.stderr[-4000:]
    assert "ThreadSanitizer" not in r.stderr, r.stderr[-4000:]
    assert "stress ok" in r.stdout
