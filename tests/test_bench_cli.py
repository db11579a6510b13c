"""bench.py's command-line contract and its reading of the committed profiles
(CPU only: nothing here initialises a GPU)."""
import os
import re
import subprocess
import sys

import pytest

from acctricnt_amd import graph as G

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def run_bench(*args):
    return subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), *args], cwd=ROOT,
                          capture_output=True, text=True, timeout=120)


@pytest.mark.parametrize("gpus", ["2", "8"])
def test_more_than_one_gpu_is_refused(gpus):
    r = run_bench("--gpus", gpus)
    assert r.returncode != 0
    assert "one GPU" in r.stderr and r.stdout == ""


def test_help_lists_the_contract_flags():
    r = run_bench("--help")
    assert r.returncode == 0
    for flag in ("--gpus", "--steps", "--warmup", "--workload"):
        assert flag in r.stdout


@pytest.mark.parametrize("workload", ["rmat22", "chunglu23", "batch512"])
def test_traffic_comes_from_the_committed_profile(workload):
    """Each bench workload's dominant kernel has a committed rocprofv3 summary whose
    traffic line bench.py reads, and the figure equals the calibrated FETCH + WRITE the
    same file states in GB."""
    sys.path.insert(0, ROOT)
    import bench

    t = bench.profiled_traffic("k_fwd_count<4>", workload)
    assert t is not None and t > 0
    text = open(os.path.join(ROOT, "profiles", f"k_fwd_count-{workload}.md")).read()
    fetch, write, total = map(float, re.search(
        r"fetch ([0-9.]+) GB \+ write ([0-9.]+) GB = \*\*([0-9.]+) GB\*\*", text).groups())
    assert abs(fetch + write - total) < 0.02
    assert abs(t / 1e9 - total) < 0.01
    # every profile names the command that made it, on committed files only
    assert "rocprofv3 --kernel-trace --stats" in text and "bench.py" in text and "_work" not in text


@pytest.mark.parametrize("name", sorted(f for f in os.listdir(os.path.join(ROOT, "profiles")) if f.endswith(".md")))
def test_every_profile_names_its_command(name):
    """Each committed profile summary gives the rocprofv3 command that made it, running
    committed files only (bench.py from the repository root)."""
    text = open(os.path.join(ROOT, "profiles", name)).read()
    cmds = [ln for ln in text.splitlines() if ln.startswith("rocprofv3 ")]
    assert cmds, name
    for c in cmds:
        assert " -- python3 bench.py" in c and "_work" not in c, c


@pytest.mark.parametrize("maker", [lambda: G.gen_rmat(9, 8, seed=31), lambda: G.gen_er(300, 6, seed=32),
                                   lambda: G.gen_chung_lu(400, 2.1, 8, 100, seed=33)])
def test_forward_bytes_by_brute_force(maker):
    """The forward kernel's algorithmic bytes: the probe count P is the sum over
    upper slots (b, c) of |{x in N(c) : x < b}|, counted here from the lists;
    the other terms follow from n, m and T."""
    import oracle

    g = maker()
    t = oracle.triangle_count(g)
    off, adj = g.node_off, g.edge_dst
    p = 0
    for b in range(g.n):
        for c in adj[off[b]:off[b + 1]]:
            if c > b:
                p += int((adj[off[c]:off[c + 1]] < b).sum())
    fb = g.forward_bytes(t)
    assert fb["probe_elements"] == p
    half = g.m // 2
    assert fb["terms"] == {"offsets": 4 * (g.n + 1), "lower_lists": 4 * half, "upper_slots": 16 * half,
                           "probes": 4 * p, "updates": 4 * t + 4 * half}
    assert fb["bytes"] == sum(fb["terms"].values())
