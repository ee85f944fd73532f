"""Precision study at d = 6, n = 8 (outputs shared 2:1; A with
entries N(0,1)/sqrt(8) and one row per matrix scaled by 1e3, X ~ U[-1,1]),
on a batch the CPU oracle finishes in seconds.

Every arithmetic path the library offers is compared with the FP64 oracle
(the reference's rounding, backward order):
  * FP64, fused multiply-add (default);
  * FP64, KM_FLAG_EXACT (separate multiply and add);
  * float entry points (inputs rounded to float, accumulation in double).
Reported per path: normwise relative error ||dY|| / ||Y|| over all outputs,
and the max and median over outputs of the elementwise relative error
|dY_i| / max_i |Y_i|.  The bars asserted are the path's documented ones
(1e-13 FP64; bitwise for EXACT where each output gets one entry; 1e-5
float).  Run with -s to see the figures.
"""
from __future__ import annotations

import json

import numpy as np
import pytest

from _helpers import RTOL_FP32, RTOL_FP64, gpu_pbatched, rel_err
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def scaled_row_inputs(batch: int, seed: int = 4):
    d, n = 6, 8
    rng = np.random.default_rng(seed)
    A = rng.standard_normal((batch, d, n, n)) / np.sqrt(8.0)
    rows = rng.integers(0, n, (batch, d))
    for b in range(batch):
        for k in range(d):
            A[b, k, :, rows[b, k]] *= 1e3     # row i of A_k is A[b, k, :, i]
    X = rng.uniform(-1, 1, (batch, n ** d))
    Y0 = np.zeros((batch // 2, n ** d))
    yidx = np.arange(batch) // 2
    return A, X, Y0, yidx


def _stats(got, want):
    scale = np.abs(want).max(axis=1, keepdims=True)
    el = (np.abs(got - want) / scale).max(axis=1)
    return {"normwise": float(rel_err(got, want)), "max": float(el.max()), "median": float(np.median(el))}


@pytest.fixture(scope="module")
def km():
    import torch
    import kronmult_amd
    assert torch.cuda.is_available()
    kronmult_amd.lib()
    return kronmult_amd


def test_precision_study_d6n8_scaled_rows(km):
    batch = 16
    A, X, Y0, yidx = scaled_row_inputs(batch)
    want = O.pbatched(A, X, Y0, yidx, variant=O.BACKWARD)
    out = {}
    fma = gpu_pbatched(A, X, Y0, yidx, with_w=True)
    out["fp64_fma"] = _stats(fma, want)
    assert out["fp64_fma"]["normwise"] <= RTOL_FP64
    exact = gpu_pbatched(A, X, Y0, yidx, with_w=True, opt=km.options(exact=True))
    out["fp64_exact"] = _stats(exact, want)
    assert out["fp64_exact"]["normwise"] <= RTOL_FP64
    f32 = gpu_pbatched(A.astype(np.float32), X.astype(np.float32), Y0.astype(np.float32), yidx, with_w=True)
    out["float_inputs"] = _stats(f32.astype(np.float64), want)
    assert out["float_inputs"]["normwise"] <= RTOL_FP32
    # single-entry outputs: EXACT reproduces the oracle bitwise
    A1, X1, Y1, _ = scaled_row_inputs(4, seed=5)
    y1 = np.arange(4)
    Y1 = np.zeros((4, X1.shape[1]))
    w1 = O.pbatched(A1, X1, Y1, y1, variant=O.BACKWARD)
    g1 = gpu_pbatched(A1, X1, Y1, y1, with_w=True, opt=km.options(exact=True))
    assert np.array_equal(g1, w1)
    print("precision study d=6 n=8 (16 entries, 2:1):", json.dumps(out))
