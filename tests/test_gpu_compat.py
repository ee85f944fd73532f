"""GPU run of the C++ caller of the reference-named interface
(tests/compat/caller.cpp, built by __graft_entry__.build()): kronmult3_batched,
kronmult3_pbatched with shared outputs and kronmult3_xbatched with a pooled,
padded A, each against an explicit Kronecker matrix on the host."""
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
    assert sum(line.startswith("kronmult3_") for line in lines) == 3
