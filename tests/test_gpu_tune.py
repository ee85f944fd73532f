"""Kernel-variant tuning: the autotuner measures the candidates of a shape and
records the fastest; the table survives a save / clear / load round trip and
calls honour it.  Every variant the table can select is checked against the
oracle, in atomic and plan mode, FP64 and float (the split-pair kron_half is
used only through the table)."""
from __future__ import annotations

import json

import numpy as np
import pytest

from _helpers import RTOL_FP32, RTOL_FP64, gpu_pbatched, rel_err, rng_inputs
from oracle import oracle as O

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def km():
    import torch
    import kronmult_amd
    assert torch.cuda.is_available()
    kronmult_amd.lib()
    kronmult_amd.tune_clear()
    yield kronmult_amd
    kronmult_amd.tune_clear()


def _family(km, v):
    return {km.KERNEL_SLAB: km.FAMILY_SLAB, km.KERNEL_HALF1: km.FAMILY_HALF,
            km.KERNEL_HALF2: km.FAMILY_HALF, km.KERNEL_REG: km.FAMILY_REG,
            km.KERNEL_LDS: km.FAMILY_LDS, km.KERNEL_LDS_WIDE: km.FAMILY_LDS}[v]


def test_tune_shape_save_load_honoured(km, tmp_path):
    d, n, batch = 5, 8, 256
    km.tune_clear()
    default = km.query_kernel_for(d, n, batch, owned=True)
    v = km.tune_shape(d, n, batch, overlap=2, owned=True)
    assert v in (km.KERNEL_SLAB, km.KERNEL_HALF1, km.KERNEL_HALF2)
    assert km.tune_get(d, n, batch, owned=True) == v
    assert km.query_kernel_for(d, n, batch, owned=True)[0] == _family(km, v)
    path = str(tmp_path / "tune.json")
    km.tune_save(path)
    doc = json.load(open(path))
    assert doc["arch"].startswith("gfx")
    (e,) = [e for e in doc["entries"] if (e["d"], e["n"], e["owned"]) == (d, n, 1)]
    assert e["bin"] == int(np.log2(batch)) and set(e["ms"]) == {"slab", "half1", "half2"}
    assert all(t > 0 for t in e["ms"].values())
    km.tune_clear()
    assert km.query_kernel_for(d, n, batch, owned=True) == default
    km.tune_load(path)
    assert km.tune_get(d, n, batch, owned=True) == v
    assert km.query_kernel_for(d, n, batch, owned=True)[0] == _family(km, v)
    # a call in that batch bin runs the tuned kernel and matches the oracle
    A, X, Y0 = rng_inputs(1200, d, n, batch, batch // 2)
    yidx = np.arange(batch) // 2
    got = gpu_pbatched(A, X, Y0, yidx, plan_kw={})
    assert rel_err(got, O.pbatched(A, X, Y0, yidx, variant=O.BACKWARD)) <= RTOL_FP64
    km.tune_clear()


def test_tune_two_pass_shape(km):
    """d6n7 (one slab above LDS): the autotuner times the two-kernel form and
    the fused cooperative kernel, each call with a workspace per entry, and
    records the faster; the call then reports that family."""
    km.tune_clear()
    v = km.tune_shape(6, 7, 32, overlap=2)
    assert v in (km.KERNEL_TWOPASS, km.KERNEL_XCD)
    fam = km.FAMILY_TWOPASS if v == km.KERNEL_TWOPASS else km.FAMILY_XCD
    assert km.query_kernel_for(6, 7, 32)[0] == fam
    km.tune_clear()
    assert km.query_kernel_for(6, 7, 32)[0] == km.FAMILY_TWOPASS


VARIANTS = [(6, 6, "HALF1"), (6, 6, "HALF2"), (5, 8, "HALF1"), (5, 8, "HALF2"), (6, 6, "SLAB"),
            (6, 4, "LDS"), (4, 5, "LDS_WIDE"), (3, 8, "LDS"), (4, 4, "REG"),
            # kron_reg with a partly filled last wave (125 and 343 threads)
            (4, 5, "REG"), (4, 7, "REG")]


@pytest.mark.parametrize("d,n,name", VARIANTS)
@pytest.mark.parametrize("mode", ["atomic", "plan"])
@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_forced_variant_vs_oracle(km, d, n, name, mode, dtype):
    v = getattr(km, "KERNEL_" + name)
    batch = 24 if n ** d > 20000 else 96
    eb = 8 if dtype == np.float64 else 4
    owned = mode == "plan"
    km.tune_set(d, n, batch, v, owned=owned, elem_bytes=eb)
    try:
        assert km.query_kernel_for(d, n, batch, owned=owned, elem_bytes=eb)[0] == _family(km, v)
        A, X, Y0 = rng_inputs(1300 + d * n + v, d, n, batch, batch // 4, dtype=dtype)
        yidx = (np.arange(batch) // 4) if owned else np.random.default_rng(v).integers(0, batch // 4, batch)
        got = gpu_pbatched(A, X, Y0, yidx, plan_kw={} if owned else None)
        want = O.pbatched(A, X, Y0, yidx, variant=O.BACKWARD)
        assert rel_err(got, want) <= (RTOL_FP64 if dtype == np.float64 else RTOL_FP32)
    finally:
        km.tune_set(d, n, batch, km.KERNEL_DEFAULT, owned=owned, elem_bytes=eb)


def test_half_variant_unaligned_outputs_and_inputs(km):
    """kron_half with X vectors aligned to one element only (element loads
    instead of pair loads) and outputs aligned to one element."""
    import torch
    d, n, batch = 6, 6, 16
    km.tune_set(d, n, batch, km.KERNEL_HALF2, owned=False)
    try:
        A, X, Y0 = rng_inputs(1400, d, n, batch, batch // 4)
        yidx = np.arange(batch) // 4
        dev = torch.device("cuda:0")
        s = n ** d
        xb = torch.zeros(X.size + 1, dtype=torch.float64, device=dev)
        tX = xb[1:].view(batch, s)
        tX.copy_(torch.tensor(X, device=dev))
        yb = torch.zeros(Y0.size + 1, dtype=torch.float64, device=dev)
        tY = yb[1:].view(batch // 4, s)
        tY.copy_(torch.tensor(Y0, device=dev))
        assert tX.data_ptr() % 16 and tY.data_ptr() % 16
        km.kronmult_pbatched(d, n, torch.tensor(A, device=dev), km.pointer_array(tX, np.arange(batch)),
                             km.pointer_array(tY, yidx), None, batch)
        torch.cuda.synchronize()
        assert rel_err(tY.cpu().numpy(), O.pbatched(A, X, Y0, yidx, variant=O.BACKWARD)) <= RTOL_FP64
    finally:
        km.tune_set(d, n, batch, km.KERNEL_DEFAULT, owned=False)
