"""CPU tests that pin the batched slender-GEMM oracle
(oracle_kgemm_batched_*, the restatement of kgemm_{nn,nt,tt}_batched).

1. Golden vectors written by the reference's own CPU build
   (tests/golden/make_golden.py -> ref_kgemm.npz): reproduced bitwise.
2. Live against the reference build where it was compiled here (oracle/_ref).
3. Mathematics: numpy matmul; the reference tests' closed-form inputs with
   their own acceptance bounds; op identities (nt(A,B) = nn(A,B^T),
   tt(A^T,B) = nt(A,B), bitwise); small-integer inputs (exact); beta == 0
   never reads C; a shared C with beta == 1 receives the sum of the entries.
A transposed operand is shown to fail the numpy check, so that check can
catch the plausible indexing mistakes.
"""
from __future__ import annotations

import os

import numpy as np
import pytest

from _helpers import KGEMM_OP, KGEMM_REF_TOL, kgemm_inputs, kgemm_shapes
from oracle import oracle as O

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "ref_kgemm.npz")
OPS = ("nn", "nt", "tt")


def _logical(buf, rows):
    return np.asarray(buf)[:rows, :]


def _numpy(op, mm, nn, kk, alpha, As, Bs, beta, Cs):
    """alpha*op(A)*op(B) + beta*C per entry in float64 (entries in order)."""
    (ra, _), (rb, _) = kgemm_shapes(op, mm, nn, kk)
    work = {}
    outs = []
    for a, b, c in zip(As, Bs, Cs):
        if id(c) not in work:
            work[id(c)] = _logical(c, mm).astype(np.float64)
        A = _logical(a, ra).astype(np.float64)
        B = _logical(b, rb).astype(np.float64)
        opA = A.T if op == "tt" else A
        opB = B if op == "nn" else B.T
        work[id(c)] = alpha * (opA @ opB) + beta * work[id(c)]
        outs.append(work[id(c)])
    return [work[id(c)] for c in Cs]


def test_golden_bitwise():
    g = np.load(GOLD)
    n = int(g["ncases"][0])
    assert n >= 100
    for c in range(n):
        key = f"c{c}"
        op, mm, nn, kk, shared = (int(v) for v in g[key + "_meta"])
        alpha = g[key + "_alpha"][0]
        beta = g[key + "_beta"][0]
        As = [np.asfortranarray(g[f"{key}_A{b}"]) for b in range(2)]
        Bs = [np.asfortranarray(g[f"{key}_B{b}"]) for b in range(2)]
        if shared:
            c0 = np.asfortranarray(g[f"{key}_C00"])
            Cs = [c0, c0]
            want = [g[f"{key}_C0"]] * 2
        else:
            Cs = [np.asfortranarray(g[f"{key}_C0{b}"]) for b in range(2)]
            want = [g[f"{key}_C{b}"] for b in range(2)]
        got = O.kgemm_batched(op, mm, nn, kk, alpha, As, Bs, beta, Cs)
        for x, y in zip(got, want):
            assert x.dtype == y.dtype
            assert np.array_equal(x, y), (c, OPS[op], mm, nn, kk, beta)


@pytest.mark.skipif(not O.ref_available("recursive"), reason="reference CPU build not present")
@pytest.mark.parametrize("op", OPS)
@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_live_against_reference_build(op, dtype):
    rng = np.random.default_rng(11)
    for t in range(12):
        mm, nn, kk = (int(v) for v in rng.integers(1, 40, 3))
        beta = [0.0, 1.0, 1.2, -0.7][t % 4]
        As, Bs, Cs = kgemm_inputs(op, mm, nn, kk, 4, seed=t, dtype=dtype, pad=t % 3, shared_c=beta == 1.0)
        got = O.kgemm_batched(KGEMM_OP[op], mm, nn, kk, 1.3, As, Bs, beta, Cs)
        ref = O.kgemm_batched(KGEMM_OP[op], mm, nn, kk, 1.3, As, Bs, beta, Cs, use_ref=True)
        for x, y in zip(got, ref):
            assert np.array_equal(x, y), (op, mm, nn, kk, beta)


@pytest.mark.parametrize("op", OPS)
@pytest.mark.parametrize("beta", [0.0, 1.0, 1.2])
def test_matches_numpy(op, beta):
    for mm, nn, kk in [(1, 1, 1), (7, 5, 3), (64, 8, 8), (3, 40, 17)]:
        As, Bs, Cs = kgemm_inputs(op, mm, nn, kk, 3, seed=mm + nn, pad=2, shared_c=beta == 1.0)
        got = O.kgemm_batched(KGEMM_OP[op], mm, nn, kk, 1.3, As, Bs, beta, Cs)
        want = _numpy(op, mm, nn, kk, 1.3, As, Bs, beta, Cs)
        for x, y in zip(got, want):
            assert np.abs(_logical(x, mm) - y).max() <= 1e-13 * max(1.0, np.abs(y).max())
    # the padding rows of C are never written
    As, Bs, Cs = kgemm_inputs(op, 5, 4, 3, 2, seed=1, pad=3)
    for c in Cs:
        c[5:, :] = 7.0
    got = O.kgemm_batched(KGEMM_OP[op], 5, 4, 3, 1.3, As, Bs, beta, Cs)
    for x in got:
        assert np.all(x[5:, :] == 7.0)


@pytest.mark.parametrize("op", OPS)
def test_reference_test_setup_within_its_bound(op):
    """The reference tests' closed-form inputs, alpha = 1.3, beta = 1.2,
    C = 1, and their own max-abs-error bound (test_kgemm_nt_batched.cpp:
    389-433, 584)."""
    for mm in (1, 15, 64):
        for nn in (1, 22, 64):
            for kk in (1, 29, 64):
                As, Bs, Cs = kgemm_inputs(op, mm, nn, kk, 8, kind="reference")
                got = O.kgemm_batched(KGEMM_OP[op], mm, nn, kk, 1.3, As, Bs, 1.2, Cs)
                want = _numpy(op, mm, nn, kk, 1.3, As, Bs, 1.2, Cs)
                err = max(np.abs(_logical(x, mm) - y).max() for x, y in zip(got, want))
                assert err < KGEMM_REF_TOL[op], (mm, nn, kk, err)


def test_op_identities_bitwise():
    rng = np.random.default_rng(5)
    mm, nn, kk = 9, 6, 11
    for beta in (0.0, 1.0, 1.2):
        A = rng.uniform(-1, 1, (mm, kk))
        B = rng.uniform(-1, 1, (kk, nn))
        C0 = rng.uniform(-1, 1, (mm, nn))
        nn_ = O.kgemm_batched(0, mm, nn, kk, 0.7, [O.colmajor(A)], [O.colmajor(B)], beta, [O.colmajor(C0)])[0]
        nt_ = O.kgemm_batched(1, mm, nn, kk, 0.7, [O.colmajor(A)], [O.colmajor(B.T)], beta, [O.colmajor(C0)])[0]
        tt_ = O.kgemm_batched(2, mm, nn, kk, 0.7, [O.colmajor(A.T)], [O.colmajor(B.T)], beta, [O.colmajor(C0)])[0]
        assert np.array_equal(nn_, nt_)
        assert np.array_equal(nt_, tt_)


@pytest.mark.parametrize("op", OPS)
def test_integer_inputs_exact(op):
    mm, nn, kk = 13, 7, 9
    As, Bs, Cs = kgemm_inputs(op, mm, nn, kk, 4, kind="integer", seed=3, shared_c=True)
    got = O.kgemm_batched(KGEMM_OP[op], mm, nn, kk, 2.0, As, Bs, 1.0, Cs)
    want = _numpy(op, mm, nn, kk, 2.0, As, Bs, 1.0, Cs)
    assert np.array_equal(_logical(got[0], mm), want[0])


def test_beta_zero_does_not_read_c_and_shared_c_sums_entries():
    mm, nn, kk = 6, 5, 4
    As, Bs, Cs = kgemm_inputs("nt", mm, nn, kk, 3, seed=9)
    for c in Cs:
        c[:, :] = np.nan
    got = O.kgemm_batched(1, mm, nn, kk, 1.3, As, Bs, 0.0, Cs)
    assert all(np.isfinite(x).all() for x in got)
    # beta == 1 with one shared C: C0 + sum of the entries' alpha*op(A)op(B)
    As, Bs, Cs = kgemm_inputs("nn", mm, nn, kk, 5, seed=4, shared_c=True)
    got = O.kgemm_batched(0, mm, nn, kk, 1.3, As, Bs, 1.0, Cs)[0]
    each = [O.kgemm_batched(0, mm, nn, kk, 1.3, [a], [b], 0.0, [np.zeros_like(Cs[0])])[0] for a, b in zip(As, Bs)]
    assert np.abs(got - (Cs[0] + sum(each))).max() < 1e-14 * 10


def test_transposed_operand_is_caught():
    """A plausible mistake (B used untransposed in nt) fails the numpy check."""
    mm, nn, kk = 5, 4, 4
    As, Bs, Cs = kgemm_inputs("nt", mm, nn, kk, 1, seed=2)
    wrong = O.kgemm_batched(0, mm, nn, kk, 1.0, As, Bs, 0.0, Cs)[0]   # nn instead of nt
    want = _numpy("nt", mm, nn, kk, 1.0, As, Bs, 0.0, Cs)[0]
    assert np.abs(wrong - want).max() > 1e-3


def test_float_accumulates_in_double():
    """Float data: products and sums in double, one rounding to float at the
    store (kgemm_nn.hpp:214-221); a float accumulator would lose the small
    terms below."""
    kk = 3
    A = np.array([[1.0, 1e-8, -1.0]], dtype=np.float32)
    B = np.ones((kk, 1), dtype=np.float32)
    got = O.kgemm_batched(0, 1, 1, kk, 1.0, [O.colmajor(A)], [O.colmajor(B)], 0.0,
                          [O.colmajor(np.zeros((1, 1), np.float32))])[0]
    assert got[0, 0] == np.float32(1e-8)
