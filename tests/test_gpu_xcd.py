"""The fused two-pass kernel (kron_xcd, family 7, variant KM_KERNEL_XCD): d = 6
with n = 7 or 8, where one n^(d-1) slab exceeds LDS.  One cooperative launch
per call; each XCD runs its share of the entries with the intermediate W in
its own L2, separated by XCD-local barriers.  It is selected through the
variant table (the two kernels stay the built-in choice: they measured
faster).

Checks:
- against the oracle at 1e-13 (FMA rounding; atomics reorder sums);
- bitwise against the two-kernel form (kron_sub + kron_pair) in plan mode,
  where both use the same arithmetic and the same summation order for every
  output;
- batches that leave XCDs without entries;
- output runs that straddle the even split of the batch over the eight XCDs;
- interleaved outputs;
- the batched entry point;
- float;
- small-integer inputs, bitwise;
- repeated calls, bitwise equal (W read through L1 gave stale values here);
- the give-up counter (km_cooperative_faults) unchanged by every call."""
from __future__ import annotations

import contextlib

import numpy as np
import pytest

from _helpers import (RTOL_FP32, RTOL_FP64, gpu_batched, gpu_pbatched, int_inputs, rel_err,
                      rng_inputs)
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


@pytest.fixture
def faults_unchanged(km):
    before = km.cooperative_faults()
    yield
    assert km.cooperative_faults() == before


@contextlib.contextmanager
def variant(km, n, batch, v, owned=False, elem_bytes=8):
    km.tune_set(6, n, batch, v, owned=owned, elem_bytes=elem_bytes)
    try:
        yield
    finally:
        km.tune_clear()


def fused(km, A, X, Y0, yidx, mode="atomic", calls=1):
    batch, _, n, _ = A.shape
    with variant(km, n, batch, km.KERNEL_XCD, owned=(mode == "plan"), elem_bytes=A.itemsize):
        assert km.query_kernel_for(6, n, batch, owned=(mode == "plan"),
                                   elem_bytes=A.itemsize)[0] == km.FAMILY_XCD
        return gpu_pbatched(A, X, Y0, yidx, with_w=True,
                            plan_kw={} if mode == "plan" else None, calls=calls)


def test_family_reported(km):
    assert km.query_kernel_for(6, 8, 1024)[0] == km.FAMILY_TWOPASS
    with variant(km, 8, 1024, km.KERNEL_XCD):
        assert km.query_kernel_for(6, 8, 1024)[0] == km.FAMILY_XCD
        # EXACT keeps the two kernels
        assert km.query_kernel_for(6, 8, 1024, opt=km.options(exact=True))[0] == km.FAMILY_TWOPASS


@pytest.mark.parametrize("n", [7, 8])
@pytest.mark.parametrize("batch,ov", [(3, 1), (20, 2), (37, 3)])
@pytest.mark.parametrize("mode", ["atomic", "plan"])
def test_vs_oracle(km, faults_unchanged, n, batch, ov, mode):
    """3 entries: five XCDs idle.  20 entries 2:1 and 37 entries 3:1: the even
    split over eight XCDs cuts output runs, which plan mode moves whole to one
    XCD."""
    nY = (batch + ov - 1) // ov
    A, X, Y0 = rng_inputs(2100 + 10 * n + batch, 6, n, batch, nY)
    yidx = np.arange(batch) // ov
    got = fused(km, A, X, Y0, yidx, mode)
    assert rel_err(got, O.pbatched(A, X, Y0, yidx, variant=O.BACKWARD)) <= RTOL_FP64


@pytest.mark.parametrize("mode", ["atomic", "plan"])
def test_interleaved_outputs(km, faults_unchanged, mode):
    """pY[k] = Y + (k mod 5) n^d: every XCD's share touches every output."""
    batch, nY = 30, 5
    A, X, Y0 = rng_inputs(2150, 6, 8, batch, nY)
    yidx = np.arange(batch) % nY
    got = fused(km, A, X, Y0, yidx, mode)
    assert rel_err(got, O.pbatched(A, X, Y0, yidx, variant=O.BACKWARD)) <= RTOL_FP64


@pytest.mark.parametrize("n", [7, 8])
def test_bitwise_vs_two_kernels(km, faults_unchanged, n):
    batch, ov = 24, 2
    A, X, Y0 = rng_inputs(2200 + n, 6, n, batch, batch // ov)
    yidx = np.arange(batch) // ov
    one = fused(km, A, X, Y0, yidx, "plan")
    assert km.query_kernel_for(6, n, batch, owned=True)[0] == km.FAMILY_TWOPASS
    two = gpu_pbatched(A, X, Y0, yidx, with_w=True, plan_kw={})
    assert np.array_equal(one, two)
    assert rel_err(one, O.pbatched(A, X, Y0, yidx, variant=O.BACKWARD)) <= RTOL_FP64


@pytest.mark.parametrize("mode", ["atomic", "plan"])
def test_integer_inputs_bitwise(km, faults_unchanged, mode):
    """Small integers: every partial sum is exact, so a lost or doubly counted
    sub-slab, position or entry shows bitwise (runs of 3 entries, d=6 n=8)."""
    batch = 27
    A, X, Y0 = int_inputs(2250, 6, 8, batch, 9)
    yidx = np.arange(batch) // 3
    got = fused(km, A, X, Y0, yidx, mode)
    assert np.array_equal(got, O.pbatched(A, X, Y0, yidx, variant=O.BACKWARD))


def test_batched_entry_point(km, faults_unchanged):
    batch = 11
    A, X, Y0 = rng_inputs(2300, 6, 7, batch, batch)
    with variant(km, 7, batch, km.KERNEL_XCD):
        got = gpu_batched(A, X, Y0, with_w=True)
    want = O.pbatched(A, X, Y0, np.arange(batch), variant=O.BACKWARD)
    assert rel_err(got, want) <= RTOL_FP64


@pytest.mark.parametrize("which", ["fused", "two-kernel"])
def test_float(km, faults_unchanged, which):
    """Float W is half the bytes, so a workgroup's positions of W fit its L1:
    a read of W that could hit a line kept from the previous entry on the XCD
    (the second entry of a run) gave stale values here, every few calls."""
    batch = 10
    A, X, Y0 = rng_inputs(2350, 6, 7, batch, 5, dtype=np.float32)
    yidx = np.arange(batch) // 2
    want = O.pbatched(A, X, Y0, yidx, variant=O.BACKWARD)
    if which == "fused":
        outs = fused(km, A, X, Y0, yidx, calls=12)
    else:
        assert km.query_kernel_for(6, 7, batch, elem_bytes=4)[0] == km.FAMILY_TWOPASS
        outs = gpu_pbatched(A, X, Y0, yidx, with_w=True, calls=12)
    for g in outs:
        assert rel_err(g, want) <= RTOL_FP32


def test_repeated_calls_reuse_control_blocks(km, faults_unchanged):
    """More calls than the ring of control blocks holds, each bitwise equal
    (plan mode is deterministic)."""
    batch = 9
    A, X, Y0 = rng_inputs(2400, 6, 7, batch, 3)
    yidx = np.arange(batch) // 3
    outs = fused(km, A, X, Y0, yidx, "plan", calls=70)
    want = O.pbatched(A, X, Y0, yidx, variant=O.BACKWARD)
    for g in outs:
        assert np.array_equal(g, outs[0])
    assert rel_err(outs[0], want) <= RTOL_FP64
