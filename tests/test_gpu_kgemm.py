"""GPU parity of the batched slender GEMM (km_kgemm_{nn,nt,tt}_batched_*,
called through the C ABI) against the CPU oracle on the same inputs.

Bars:
  * FMA (default): normwise relative error <= 1e-13 (double), 1e-5 (float);
  * KM_FLAG_EXACT: bitwise equal to the oracle (the reference CPU build's
    rounding) wherever no two entries share a C;
  * small-integer inputs: bitwise in every mode, shared C included;
  * the reference tests' closed-form setup: within their own max-abs bound.
Shapes cover both kernels: the row kernel (mm >= 64, kk <= 16) and the flat
one (everything else), padded leading dimensions, beta in {0, 1, other}.
"""
from __future__ import annotations

import numpy as np
import pytest

from _helpers import (KGEMM_OP, KGEMM_REF_TOL, RTOL_FP32, RTOL_FP64, gpu_kgemm, kgemm_inputs,
                      rel_err)
from oracle import oracle as O

pytestmark = pytest.mark.gpu

OPS = ("nn", "nt", "tt")
SHAPES = [(1, 1, 1), (8, 8, 8), (15, 1, 22), (29, 36, 8), (64, 8, 8), (1, 43, 57), (200, 8, 8),
          (4096, 8, 8), (300, 5, 16), (130, 3, 17), (64, 64, 64)]


@pytest.fixture(scope="module")
def km():
    import torch
    import kronmult_amd
    assert torch.cuda.is_available()
    kronmult_amd.lib()
    return kronmult_amd


def _logical(c, mm):
    return np.asarray(c)[:mm, :]


@pytest.mark.parametrize("op", OPS)
@pytest.mark.parametrize("mm,nn,kk", SHAPES)
def test_fma_vs_oracle(km, op, mm, nn, kk):
    for beta in (0.0, 1.0, 1.2):
        batch = 3 if mm * nn < 10000 else 2
        As, Bs, Cs = kgemm_inputs(op, mm, nn, kk, batch, seed=mm + 3 * nn + kk, pad=(mm % 3),
                                  shared_c=beta == 1.0)
        want = O.kgemm_batched(KGEMM_OP[op], mm, nn, kk, 1.3, As, Bs, beta, Cs)
        got = gpu_kgemm(op, mm, nn, kk, 1.3, As, Bs, beta, Cs)
        for g, w, c0 in zip(got, want, Cs):
            assert rel_err(_logical(g, mm), _logical(w, mm)) <= RTOL_FP64, (beta,)
            assert np.array_equal(g[mm:, :], c0[mm:, :])   # padding rows untouched


@pytest.mark.parametrize("op", OPS)
@pytest.mark.parametrize("mm,nn,kk", [(1, 1, 1), (15, 1, 22), (64, 8, 8), (130, 3, 17), (4096, 8, 8)])
def test_exact_bitwise(km, op, mm, nn, kk):
    opt = km.options(exact=True)
    for beta in (0.0, 1.0, -0.75):
        As, Bs, Cs = kgemm_inputs(op, mm, nn, kk, 2, seed=7 + kk, pad=1)
        want = O.kgemm_batched(KGEMM_OP[op], mm, nn, kk, 1.3, As, Bs, beta, Cs)
        got = gpu_kgemm(op, mm, nn, kk, 1.3, As, Bs, beta, Cs, opt=opt)
        for g, w in zip(got, want):
            assert np.array_equal(g, w), beta


@pytest.mark.parametrize("op", OPS)
def test_reference_test_setup(km, op):
    """The reference tests' inputs and acceptance bound (alpha 1.3, beta 1.2,
    C = 1, closed-form A and B; test_kgemm_nt_batched.cpp:64-200, 389-433)."""
    for mm in (1, 22, 64):
        for nn in (1, 15, 64):
            for kk in (8, 57):
                for batch in (1, 15):
                    As, Bs, Cs = kgemm_inputs(op, mm, nn, kk, batch, kind="reference")
                    want = O.kgemm_batched(KGEMM_OP[op], mm, nn, kk, 1.3, As, Bs, 1.2, Cs)
                    got = gpu_kgemm(op, mm, nn, kk, 1.3, As, Bs, 1.2, Cs)
                    err = max(np.abs(_logical(g, mm) - _logical(w, mm)).max() for g, w in zip(got, want))
                    assert err < KGEMM_REF_TOL[op], (mm, nn, kk, batch, err)


@pytest.mark.parametrize("op", OPS)
def test_integer_inputs_shared_c_bitwise(km, op):
    """Every entry adds into one C with beta = 1 (hardware atomics across
    workgroups); small-integer inputs make every order exact."""
    for mm, nn, kk in [(9, 7, 5), (256, 8, 8), (70, 4, 20)]:
        As, Bs, Cs = kgemm_inputs(op, mm, nn, kk, 12, kind="integer", seed=mm, shared_c=True)
        want = O.kgemm_batched(KGEMM_OP[op], mm, nn, kk, 2.0, As, Bs, 1.0, Cs)
        got = gpu_kgemm(op, mm, nn, kk, 2.0, As, Bs, 1.0, Cs)
        assert np.array_equal(got[0], want[0])


@pytest.mark.parametrize("op", OPS)
def test_float(km, op):
    for mm, nn, kk in [(8, 8, 8), (300, 6, 6), (5, 33, 40)]:
        for beta in (0.0, 1.2):
            As, Bs, Cs = kgemm_inputs(op, mm, nn, kk, 3, seed=kk, dtype=np.float32)
            want = O.kgemm_batched(KGEMM_OP[op], mm, nn, kk, 1.3, As, Bs, beta, Cs)
            got = gpu_kgemm(op, mm, nn, kk, 1.3, As, Bs, beta, Cs)
            for g, w in zip(got, want):
                assert rel_err(_logical(g, mm), _logical(w, mm)) <= RTOL_FP32
            got = gpu_kgemm(op, mm, nn, kk, 1.3, As, Bs, beta, Cs, opt=km.options(exact=True))
            for g, w in zip(got, want):
                assert np.array_equal(g, w)


def test_beta_zero_never_reads_c_and_kk_zero(km):
    mm, nn, kk = 100, 4, 6
    As, Bs, Cs = kgemm_inputs("nt", mm, nn, kk, 2, seed=3)
    for c in Cs:
        c[:, :] = np.nan
    got = gpu_kgemm("nt", mm, nn, kk, 1.3, As, Bs, 0.0, Cs)
    assert all(np.isfinite(g).all() for g in got)
    # kk = 0: C = beta*C + alpha*0
    As, Bs, Cs = kgemm_inputs("nn", 70, 3, 1, 2, seed=4)
    want = O.kgemm_batched(0, 70, 3, 0, 1.3, As, Bs, 0.5, Cs)
    got = gpu_kgemm("nn", 70, 3, 0, 1.3, As, Bs, 0.5, Cs)
    for g, w in zip(got, want):
        assert np.array_equal(g, w)


def test_empty_and_invalid(km):
    L = km.lib()
    f = L.km_kgemm_nn_batched_f64
    assert f(0, 4, 4, 1.0, None, None, None, None, 0.0, None, None, 5, None, None) == km.SUCCESS
    assert f(4, 4, 4, 1.0, None, None, None, None, 0.0, None, None, 0, None, None) == km.SUCCESS
    assert f(-1, 4, 4, 1.0, None, None, None, None, 0.0, None, None, 1, None, None) == km.ERROR_INVALID_VALUE
    assert f(4, 4, 4, 1.0, None, None, None, None, 0.0, None, None, 1, None, None) == km.ERROR_INVALID_VALUE
