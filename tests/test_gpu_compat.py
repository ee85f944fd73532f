"""GPU run of the C++ caller of the reference-named interface
(tests/compat/caller.cpp, built by __graft_entry__.build()): every overload,
kronmult{1..6}_{batched,pbatched,xbatched,vbatched} (shared outputs, a pooled
and padded A, rectangular vbatched factors) and kgemm_{nn,nt,tt}_batched
(beta 0, 1 with a shared C, 0.5), each against an explicit Kronecker matrix or
product on the host."""
from __future__ import annotations

import os
import subprocess

import pytest

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CALLER = os.path.join(ROOT, "tests", "compat", "_bin", "caller")


def test_cpp_caller_runs():
    assert os.path.exists(CALLER), "tests/compat/_bin/caller is missing: run __graft_entry__.build()"
    r = subprocess.run([CALLER], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    lines = r.stdout.strip().splitlines()
    assert lines[-1] == "ok", r.stdout
    for d in range(1, 7):
        for what in ("batched", "pbatched", "xbatched", "vbatched"):
            assert sum(line.startswith(f"kronmult{d}_{what} ") for line in lines) == 1, (d, what, r.stdout)
    assert sum(line.startswith("kgemm_") for line in lines) == 9, r.stdout
