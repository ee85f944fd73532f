"""CPU tests of bench.py's measurement arithmetic: the algorithmic bytes and
flops its figures divide by, the traffic lines it reads from the committed
rocprofv3 summaries, its refusal of more than one GPU, and the CPU leg of the
pooled workload on a small sample (no GPU needed)."""
from __future__ import annotations

import math
import os
import subprocess
import sys

import numpy as np
import pytest

import bench
from oracle import oracle as O

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.mark.parametrize("d,n", [(6, 6), (3, 8), (6, 4), (6, 8), (4, 5)])
def test_flops_follow_the_reference_flop_model(d, n):
    # square factors: kron_flops = 2 * sum over the d stages of n * n^d
    assert bench.flops_per_call(d, n, 1) == O.kron_flops([n] * d, [n] * d)
    assert bench.flops_per_call(d, n, 7) == 7 * 2 * d * n ** (d + 1)


def test_algorithmic_bytes_of_the_bench_workloads():
    # X and A read once, every distinct output read and written once
    b = bench.algorithmic_bytes(6, 6, 4096, 4)
    assert b == 8 * (4096 * 46656 + 4096 * 6 * 36 + 2 * 1024 * 46656)
    assert math.isclose(b, 2.3003e9, rel_tol=1e-4)
    assert math.isclose(bench.algorithmic_bytes(3, 8, 262144, 4), 2.0133e9, rel_tol=1e-4)
    assert math.isclose(bench.algorithmic_bytes(6, 8, 1024, 2), 4.2981e9, rel_tol=1e-4)
    # overlap 1: every entry has its own output
    assert bench.algorithmic_bytes(2, 3, 5, 1) == 8 * 5 * (9 + 2 * 9 + 2 * 9)


# kron_slab d6n6 visits the output run it starts inside twice (staggered run
# switches), one extra Y read and write per workgroup: 1.08x
@pytest.mark.parametrize("name,ratio_lo,ratio_hi", [
    ("kron_slab_d6n6_b4096.txt", 0.98, 1.10),
    ("kron_reg_d3n8_b262144.txt", 0.98, 1.10),
    ("kron_reg_d6n4_b65536.txt", 0.90, 1.05),
])
def test_profiled_traffic_is_near_the_algorithmic_bytes(name, ratio_lo, ratio_hi):
    workload = {"kron_slab_d6n6_b4096.txt": "d6n6", "kron_reg_d3n8_b262144.txt": "d3n8",
                "kron_reg_d6n4_b65536.txt": "d6n4"}[name]
    d, n, batch, overlap = bench.WORKLOADS[workload]
    t = bench.parse_traffic(os.path.join(ROOT, "profiles", name))
    assert t is not None
    assert ratio_lo <= t / bench.algorithmic_bytes(d, n, batch, overlap) <= ratio_hi


def test_profiled_traffic_of_the_pooled_workload():
    c = bench.XPOOL
    s = c["n"] ** c["d"]
    t = bench.parse_traffic(os.path.join(ROOT, "profiles", f"kron_reg_xpool_d{c['d']}n{c['n']}_b{c['batch']}.txt"))
    assert t is not None
    # every entry reads its X vector from HBM; A and Y once
    streamed = 8 * (c["batch"] * s + c["pool"] * c["d"] * c["n"] ** 2 + 2 * c["ny"] * s)
    assert 1.0 <= t / streamed <= 1.1


def test_missing_profile_gives_no_traffic(tmp_path):
    assert bench.parse_traffic(str(tmp_path / "absent.txt")) is None


def test_more_than_one_gpu_is_refused():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 2
    assert "one GPU" in r.stderr


def test_pooled_cpu_leg_on_a_small_sample(monkeypatch):
    torch = pytest.importorskip("torch")
    monkeypatch.setattr(bench, "XPOOL_CPU_SAMPLE", 64)
    c = bench.XPOOL
    d, n = c["d"], c["n"]
    rng = np.random.default_rng(0)
    pools = [rng.uniform(-1, 1, (c["pool"], n, n)) for _ in range(d)]
    X = torch.tensor(rng.uniform(-1, 1, (300, n ** d)))
    batch = 256
    aidx = rng.integers(0, c["pool"], (batch, d))
    xidx = rng.integers(0, 300, batch)
    yidx = rng.permutation(batch) // 16
    r = bench.xpool_cpu_baseline(pools, aidx, X, xidx, yidx, reps=1)
    assert r["value"] > 0 and r["unit"] == "GFLOP/s" and r["cores"] >= 1
    assert r["kind"] == ("reference" if O.ref_available("recursive") else "port")
    assert r["sample"].startswith("first 64 of ")


def test_profiled_traffic_of_the_grouped_workload():
    c = bench.GROUPED
    rng = np.random.default_rng(bench.SEED)
    ns = rng.choice(c["sizes"], c["batch"])
    d = c["d"]
    alg = 8.0 * (sum(int(n) ** d + d * int(n) ** 2 for n in ns) + 2 * c["groups"] * sum(n ** d for n in c["sizes"]))
    t = bench.parse_traffic(os.path.join(ROOT, "profiles", f"grouped_d{d}_b{c['batch']}.txt"))
    assert t is not None
    # reads match the algorithmic bytes; shared runs add Y writes
    assert 1.0 <= t / alg <= 1.4


def test_workload_names():
    assert set(bench.WORKLOADS) >= {"d6n6", "d3n8", "d6n4", "d6n8", "d6n4-interleaved"}
    assert bench.INTERLEAVED <= set(bench.WORKLOADS)
