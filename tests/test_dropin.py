"""The drop-in executables: argument layouts of the reference's mains, log
tags, exit behaviour and the SCAN result file / triangle total against the
oracle.  The GPU cases run the executables as child processes."""
import os
import re
import subprocess

import numpy as np
import pytest

import bruteforce as BF
import oracle
from acctricnt_amd import graph as G

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(ROOT, "acctricnt_amd", "bin")


def run(exe, *args, timeout=120, env=None):
    return subprocess.run([os.path.join(BIN, exe), *map(str, args)], capture_output=True, text=True,
                          timeout=timeout, env=None if env is None else {**os.environ, **env})


@pytest.fixture(scope="module", autouse=True)
def _built(built):
    assert os.path.exists(os.path.join(BIN, "main_scanxp")), "drop-in executables are not built"
    return built


def test_usage_without_arguments():
    r = run("main_scanxp")
    assert r.returncode == 0 and "Usage" in r.stdout


def test_missing_graph_exits_1(tmp_path):
    r = run("main_scanxp", tmp_path / "nope", "0.2", "5", "2", "org", tmp_path / "log.txt")
    assert r.returncode == 1
    assert "cannot open" in r.stderr


def test_self_loop_exits_1(tmp_path):
    # b_degree.bin / b_adj.bin with a self-loop at vertex 0
    deg = np.array([4, 2, 2, 2], np.int32)
    hdr = np.array([4, 3, 6], np.int32)
    (tmp_path / "b_degree.bin").write_bytes(hdr.tobytes() + deg[1:].tobytes())
    (tmp_path / "b_adj.bin").write_bytes(np.array([0, 1, 0, 2, 1, 2], np.int32).tobytes())
    r = run("main_triangle_count", tmp_path, "2", "org", tmp_path / "log.txt")
    assert r.returncode == 1
    assert "Self loop" in r.stdout


def harness_intersection_time(text):
    """The reference harness's "O(E) intersection count time"
    (python_experiments/data_analysis/common/varying_datasets_cuda_exec.py:110-116):
    Step1 minus the similarity cost minus the task-range init, each parsed as
    get_time_mem_info_updated does (read_file_utils_updated.py:13-15, 47-71):
    the message after the log header's 10 fields, the time before 's,'."""
    header = len("2019-03-23 11:05:57 INFO  (ts: 1553310357.245065 s,  et: 0.000756 s) main_scanxp.cpp:40:".split())

    def tag_time(tag):
        for line in text.splitlines():
            rest = " ".join(line.split()[header:])
            if rest.startswith(tag):
                return float(rest.split("s,")[0].split(":")[-1].split()[0])
        return None

    t = [tag_time(x) for x in ("Step1 - CheckCore:", "core-checking sim-core-comp cost",
                               "task range init by filtered accumulated degree cost")]
    return None if None in t else t[0] - t[1] - t[2]


def _count(text, tag):
    m = re.search(r"\b" + tag + r": (\d[\d,.]*)", text)
    return int(re.sub(r"[^0-9]", "", m.group(1))) if m else None


@pytest.mark.gpu
def test_k4_pendant_reference_run(tmp_path):
    """The run recorded from the reference binaries: Core 4, Cluster 1, Hub 0,
    Outlier 1, result file c 0..3 -> 0 and n 4 0; Triangle Cnt 4."""
    G.write_bins(BF.k4_pendant(), str(tmp_path))
    log = tmp_path / "log.txt"
    r = run("main_scanxp", tmp_path, "0.6", "2", "4", "org", log)
    assert r.returncode == 0, r.stderr
    text = log.read_text()
    assert "Core: 4" in text and re.search(r"Cluster: 1, Hub: 0, Outlier: 1", text)
    for tag in ("Step1 - CheckCore:", "Step2 - ClusterCore", "Step3 - LabelNonCore", "Step4 - output to the disk cost",
                "All-Edge Computation", "CUDA Kernel Time", "CUDA kernel lauch cost in GPU 0", "bin-search cost",
                "task range init by filtered accumulated degree cost", "core-checking sim-core-comp cost"):
        assert tag in text, tag
    assert "Verify" not in text  # opt-in
    kernel = re.search(r"CUDA Kernel Time: ([0-9.]+) ms", text)
    assert kernel and float(kernel.group(1)) > 0
    t = harness_intersection_time(text)
    assert t is not None and t >= 0
    res = (tmp_path / "scanxp-result-0.6-2.txt").read_text().splitlines()
    assert res == ["c/n vertex_id cluster_id", "c 0 0", "c 1 0", "c 2 0", "c 3 0", "n 4 0"]
    r = run("main_triangle_count", tmp_path, "4", "deg", log)
    assert r.returncode == 0, r.stderr
    assert _count(log.read_text(), "Triangle Cnt") == 4


@pytest.mark.gpu
@pytest.mark.parametrize("layout", ["cpu-deg", "cpu-org", "gpu", "gpu-experimental"])
def test_scan_result_file_vs_oracle(tmp_path, layout):
    """Raw (not relabelled) R-MAT ids; with the deg reorder the executable works
    on relabelled ids and must still print the original ids."""
    g = G.gen_rmat(14, 16, seed=51, relabel=False)
    G.write_bins(g, str(tmp_path))
    eps, mu = "0.3", 4
    ref = oracle.scan(g, oracle.cn_merge(g), eps, mu)
    expect = tmp_path / "expect.txt"
    oracle.write_result(str(expect), g, ref)
    log = tmp_path / "log.txt"
    args = {"cpu-deg": [eps, mu, 4, "deg", log], "cpu-org": [eps, mu, 4, "org", log], "gpu": [eps, mu, 4, log],
            "gpu-experimental": [eps, mu, 4, 1, 16, log]}[layout]
    r = run("main_scanxp", tmp_path, *args)
    assert r.returncode == 0, r.stderr
    out = tmp_path / f"scanxp-result-{oracle.eps_string(float(eps))}-{mu}.txt"
    assert out.read_bytes() == expect.read_bytes()
    text = log.read_text()
    assert _count(text, "Core") == ref["core_num"]


@pytest.mark.gpu
def test_triangle_count_vs_oracle(tmp_path):
    g = G.gen_chung_lu(1 << 15, 2.1, 24, 4096, seed=52)
    G.write_bins(g, str(tmp_path))
    log = tmp_path / "log.txt"
    for reorder in ("deg", "org"):
        r = run("main_triangle_count", tmp_path, 4, reorder, log)
        assert r.returncode == 0, r.stderr
    counts = [int(re.sub(r"[^0-9]", "", m)) for m in re.findall(r"Triangle Cnt: ([0-9][0-9,.]*)", log.read_text())]
    assert counts == [oracle.triangle_count(g)] * 2


@pytest.mark.gpu
def test_scan_result_file_full_size_rmat22(tmp_path):
    """R-MAT scale 22 as generated, through the drop-in with the deg reorder
    (relabelled on the GPU): the result file equals the oracle's SCAN of the
    graph in its original ids, byte for byte (counts from the oracle's
    bit-vector path, independent of the GPU)."""
    g = G.gen_rmat(22, 16, seed=1, relabel=False)
    print(f"[drop-in rmat22] n={g.n} m={g.m}", flush=True)
    G.write_bins(g, str(tmp_path))
    eps, mu = "0.2", 5
    log = tmp_path / "log.txt"
    r = run("main_scanxp", tmp_path, eps, mu, 16, "deg", log, timeout=280, env={"SCANXP_VERIFY": "1"})
    assert r.returncode == 0, r.stderr
    ref = oracle.scan(g, oracle.cn_bitmap(g), eps, mu)
    expect = tmp_path / "expect.txt"
    oracle.write_result(str(expect), g, ref)
    out = tmp_path / f"scanxp-result-{oracle.eps_string(float(eps))}-{mu}.txt"
    assert out.read_bytes() == expect.read_bytes()
    text = log.read_text()
    assert _count(text, "Core") == ref["core_num"]
    assert "Verify with BitVec" in text and "Verify Pass, Correct!" in text
    assert "Verify Fail" not in text
    print(f"[drop-in rmat22] harness intersection time {harness_intersection_time(text):.3f} s", flush=True)
