"""CPU tests that pin the oracle (oracle/kron_oracle.c) before it is trusted.

1. Golden vectors written by the reference's OWN CPU path (both builds:
   default recursive and -DUSE_ALG993), tests/golden/make_golden.py: the
   restatement must reproduce every one bitwise.
2. Where the reference CPU build is present in this process's tree
   (oracle/_ref), the restatement against it live on fresh inputs.
3. Mathematics that fixes the answer independently of any implementation:
   explicit Kronecker sums (brute force, numpy.kron), identity and
   permutation matrices, linearity, the mixed-product property
   kron(A)kron(B) = kron(AB), d = 1 (GEMV), d = 2 (vec(A2 X A1^T)),
   and small-integer inputs where every summation order is exact.
Each check is chosen so that a plausible mistake (transposed A, wrong mode
stride, reversed mode order, dropped term) fails it.
"""
from __future__ import annotations

import os

import numpy as np
import pytest

from _helpers import int_inputs, reference_inputs, rel_err
from oracle import oracle as O

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
VARIANT = {"recursive": O.RECURSIVE, "alg993": O.ALG993}


def _load(name):
    return np.load(os.path.join(GOLD, f"ref_{name}.npz"))


@pytest.fixture(scope="module", autouse=True)
def one_thread():
    # overlapping outputs are then accumulated in entry order, as in the
    # single-threaded reference runs that wrote the golden vectors
    old = O.num_threads()
    O.set_num_threads(1)
    yield
    O.set_num_threads(old)


# ---------------------------------------------------------------------------
# 1. golden vectors from the reference build
# ---------------------------------------------------------------------------
def test_golden_pbatched_bitwise():
    g = _load("pbatched")
    keys = sorted({k.rsplit("_", 1)[0] for k in g.files if k.endswith("_A")})
    assert len(keys) >= 20
    for key in keys:
        A, X, Y0, yidx = g[key + "_A"], g[key + "_X"], g[key + "_Y0"], g[key + "_yidx"]
        for b, v in VARIANT.items():
            got = O.pbatched(A, X, Y0, yidx, variant=v)
            assert np.array_equal(got, g[key + f"_Y_{b}"]), (key, b)


def test_golden_batched_bitwise():
    g = _load("batched")
    keys = sorted({k.rsplit("_", 1)[0] for k in g.files if k.endswith("_A")})
    for key in keys:
        for b, v in VARIANT.items():
            got = O.batched(g[key + "_A"], g[key + "_X"], g[key + "_Y0"], variant=v)
            assert np.array_equal(got, g[key + f"_Y_{b}"]), (key, b)


def test_golden_xbatched_reference_setup_bitwise():
    g = _load("xbatched")
    keys = sorted({k.rsplit("_", 1)[0] for k in g.files if k.endswith("_A")})
    assert len(keys) >= 10
    for key in keys:
        A, X = g[key + "_A"], g[key + "_X"]
        batch, d = A.shape[0], A.shape[1]
        aidx = np.stack([np.arange(batch)] * d, axis=1)
        for b, v in VARIANT.items():
            got = O.xbatched([A[:, k] for k in range(d)], aidx, X, np.arange(batch),
                             np.zeros((1, X.shape[1])), np.zeros(batch, dtype=np.int64), variant=v)
            assert np.array_equal(got, g[key + f"_Y_{b}"]), (key, b)
        # the golden inputs are the reference test's closed forms
        n = A.shape[2]
        A2, X2 = reference_inputs(d, n, batch, A.shape[3])
        assert np.array_equal(A2, A) and np.array_equal(X2, X)


def test_golden_vbatched_bitwise():
    g = _load("vbatched")
    keys = sorted({k.rsplit("_", 1)[0] for k in g.files if k.endswith("_X")})
    for key in keys:
        m, n = list(g[key + "_m"]), list(g[key + "_n"])
        d = len(m)
        X, Y0 = g[key + "_X"], g[key + "_Y0"]
        batch = X.shape[0]
        mats = [[g[key + f"_A{k}"][b] for k in range(d)] for b in range(batch)]
        for b, v in VARIANT.items():
            got = O.vbatched(m, n, mats, X, Y0, np.arange(batch), variant=v)
            assert np.array_equal(got, g[key + f"_Y_{b}"]), (key, b)


def test_golden_float_bitwise():
    g = _load("pbatched_f32")
    keys = sorted({k.rsplit("_", 1)[0] for k in g.files if k.endswith("_A")})
    for key in keys:
        A, X, Y0, yidx = g[key + "_A"], g[key + "_X"], g[key + "_Y0"], g[key + "_yidx"]
        assert A.dtype == np.float32
        for b, v in VARIANT.items():
            got = O.pbatched(A, X, Y0, yidx, variant=v)
            assert np.array_equal(got, g[key + f"_Y_{b}"]), (key, b)


def test_golden_kron_flops():
    g = _load("kron_flops")
    for m, n, d, f in zip(g["shapes_m"], g["shapes_n"], g["d"], g["flops"]):
        assert O.kron_flops(m[:d], n[:d]) == f
    # the reference harness's d = 6 convention 12 n^7 (test_kronmult6_batched.cpp:270)
    assert O.kron_flops([6] * 6, [6] * 6) == 12 * 6 ** 7


# ---------------------------------------------------------------------------
# 2. live against the reference build (where it was built)
# ---------------------------------------------------------------------------
@pytest.mark.skipif(not (O.ref_available("recursive") and O.ref_available("alg993")),
                    reason="reference CPU build not present (oracle/_ref)")
@pytest.mark.parametrize("d,n", [(1, 7), (2, 6), (3, 5), (4, 4), (5, 3), (6, 4), (6, 6)])
def test_live_against_reference_build(d, n):
    rng = np.random.default_rng(d * 100 + n)
    batch = 3
    A = rng.uniform(-1, 1, (batch, d, n, n))
    X = rng.uniform(-1, 1, (batch, n ** d))
    Y0 = rng.uniform(-1, 1, (2, n ** d))
    yidx = np.array([0, 1, 0])
    for b, v in VARIANT.items():
        assert np.array_equal(O.pbatched(A, X, Y0, yidx, variant=v), O.pbatched(A, X, Y0, yidx, use_ref=b))
    # xbatched with pooled, padded A
    lda = n + 2
    pools = [rng.uniform(-1, 1, (4, n, lda)) for _ in range(d)]
    aidx = rng.integers(0, 4, (batch, d))
    for b, v in VARIANT.items():
        got = O.xbatched(pools, aidx, X, [0, 1, 2], Y0, yidx, variant=v)
        want = O.xbatched(pools, aidx, X, [0, 1, 2], Y0, yidx, use_ref=b)
        assert np.array_equal(got, want)


# ---------------------------------------------------------------------------
# 3. mathematics
# ---------------------------------------------------------------------------
def _kron_np(mats):
    K = np.ones((1, 1))
    for a in mats:
        K = np.kron(K, a)
    return K


@pytest.mark.parametrize("m,n", [((2,), (3,)), ((2, 3), (3, 2)), ((2, 2, 3), (3, 1, 2)), ((2,) * 4, (3, 2, 1, 2)),
                                 ((2,) * 5, (2,) * 5), ((2, 1, 2, 1, 2, 2), (1, 2, 2, 2, 1, 2))])
@pytest.mark.parametrize("variant", [O.RECURSIVE, O.BACKWARD, O.FORWARD, O.ALG993])
def test_staged_equals_numpy_kron(m, n, variant):
    """Index order: the first factor is outermost (j_d fastest)."""
    rng = np.random.default_rng(sum(m) + 10 * sum(n))
    mats = [rng.uniform(-1, 1, (m[k], n[k])) for k in range(len(m))]
    x = rng.uniform(-1, 1, int(np.prod(n)))
    want = _kron_np(mats) @ x
    assert rel_err(O.entry(mats, x, variant), want) <= 1e-14
    assert rel_err(O.brute(mats, x), want) <= 1e-14


def test_variants_differ_only_in_rounding_and_order_is_as_stated():
    """Recursive and forward contract mode 1 first, backward mode d first: on
    random inputs recursive == forward bitwise and backward differs in the
    last bits only."""
    rng = np.random.default_rng(3)
    mats = [rng.uniform(-1, 1, (4, 4)) for _ in range(5)]
    x = rng.uniform(-1, 1, 4 ** 5)
    r = O.entry(mats, x, O.RECURSIVE)
    f = O.entry(mats, x, O.FORWARD)
    b = O.entry(mats, x, O.BACKWARD)
    assert np.array_equal(r, f)
    assert not np.array_equal(r, b)
    assert rel_err(r, b) <= 1e-14


def test_identity_and_permutation():
    rng = np.random.default_rng(4)
    d, n = 4, 3
    x = rng.uniform(-1, 1, n ** d)
    eye = [np.eye(n)] * d
    for v in (O.RECURSIVE, O.BACKWARD):
        assert np.array_equal(O.entry(eye, x, v), x)
    perms = [np.eye(n)[rng.permutation(n)] for _ in range(d)]
    want = _kron_np(perms) @ x  # a permutation of x: exact
    for v in (O.RECURSIVE, O.BACKWARD):
        assert np.array_equal(O.entry(perms, x, v), want)


def test_linearity():
    rng = np.random.default_rng(5)
    mats = [rng.uniform(-1, 1, (3, 3)) for _ in range(4)]
    x1, x2 = rng.uniform(-1, 1, (2, 81))
    a, b = 0.75, -1.25
    lhs = O.entry(mats, a * x1 + b * x2, O.BACKWARD)
    rhs = a * O.entry(mats, x1, O.BACKWARD) + b * O.entry(mats, x2, O.BACKWARD)
    assert rel_err(lhs, rhs) <= 1e-14


def test_mixed_product():
    """kron(A) (kron(B) x) = kron(AB) x."""
    rng = np.random.default_rng(6)
    A = [rng.uniform(-1, 1, (3, 3)) for _ in range(3)]
    B = [rng.uniform(-1, 1, (3, 3)) for _ in range(3)]
    x = rng.uniform(-1, 1, 27)
    lhs = O.entry(A, O.entry(B, x, O.RECURSIVE), O.RECURSIVE)
    rhs = O.entry([a @ b for a, b in zip(A, B)], x, O.RECURSIVE)
    assert rel_err(lhs, rhs) <= 1e-14


def test_d1_is_gemv_and_d2_is_vec_identity():
    rng = np.random.default_rng(8)
    A1 = rng.uniform(-1, 1, (5, 5))
    x = rng.uniform(-1, 1, 5)
    # d = 1: one dot product per row, same order as the reference -> exact
    want = np.array([sum(A1[i, j] * x[j] for j in range(5)) for i in range(5)])
    assert np.array_equal(O.entry([A1], x, O.RECURSIVE), want)
    # d = 2: kron(A1, A2) vec_r(X) = vec_r(A1 X A2^T) with X row-major (j2 fastest)
    A1 = rng.uniform(-1, 1, (3, 4))
    A2 = rng.uniform(-1, 1, (2, 5))
    Xm = rng.uniform(-1, 1, (4, 5))
    want = (A1 @ Xm @ A2.T).reshape(-1)
    assert rel_err(O.entry([A1, A2], Xm.reshape(-1), O.BACKWARD), want) <= 1e-14


def test_transposed_operand_is_caught():
    """A_k(i, j) is stored column-major; feeding A^T must change the answer."""
    rng = np.random.default_rng(9)
    mats = [rng.uniform(-1, 1, (3, 3)) for _ in range(3)]
    x = rng.uniform(-1, 1, 27)
    good = O.entry(mats, x, O.BACKWARD)
    bad = O.entry([a.T for a in mats], x, O.BACKWARD)
    assert rel_err(good, bad) > 1e-3
    assert rel_err(good, _kron_np(mats) @ x) <= 1e-14


def test_integer_inputs_exact_in_every_order():
    """Small integers: all partial sums exact, so every variant and any number
    of threads (any atomic order) gives the identical result."""
    d, n, batch = 4, 3, 12
    A, X, Y0 = int_inputs(10, d, n, batch, 3)
    yidx = np.arange(batch) % 3
    ys = [O.pbatched(A, X, Y0, yidx, variant=v) for v in (O.RECURSIVE, O.BACKWARD, O.FORWARD)]
    O.set_num_threads(4)
    ys.append(O.pbatched(A, X, Y0, yidx, variant=O.BACKWARD))
    O.set_num_threads(1)
    for y in ys[1:]:
        assert np.array_equal(ys[0], y)
    # and against a float64 brute force (exact for integers)
    want = Y0.copy()
    for b in range(batch):
        want[yidx[b]] += O.brute([A[b, k].T for k in range(d)], X[b])
    assert np.array_equal(ys[0], want)


def test_overlap_is_sum_of_entries():
    rng = np.random.default_rng(12)
    d, n, batch = 3, 4, 9
    A = rng.uniform(-1, 1, (batch, d, n, n))
    X = rng.uniform(-1, 1, (batch, n ** d))
    Y0 = np.zeros((1, n ** d))
    got = O.pbatched(A, X, Y0, np.zeros(batch, dtype=np.int64), variant=O.BACKWARD)
    want = sum(O.entry([A[b, k].T for k in range(d)], X[b], O.BACKWARD) for b in range(batch))
    assert rel_err(got[0], want) <= 1e-14


def test_alg993_choice():
    """kronm.hpp:89: forward only when strictly cheaper; square => backward."""
    assert not O.alg993_uses_forward([4] * 6, [4] * 6)
    assert O.alg993_uses_forward([1, 4], [4, 4]) != O.alg993_uses_forward([4, 1], [4, 4])


def test_vbatched_refuses_too_small_workspace():
    mats = [[np.eye(2)] * 2]
    with pytest.raises(ValueError):
        O.vbatched([2, 2], [2, 2], mats, np.ones((1, 4)), np.zeros((1, 4)), [0], Wcap_bytes=8)


def test_float_accumulates_in_double():
    """Float entries: each dot product is summed in double and rounded once
    (kgemm_nt.hpp:213-221) -- closer to the double result than float sums."""
    rng = np.random.default_rng(13)
    mats = [rng.uniform(-1, 1, (6, 6)).astype(np.float32) for _ in range(3)]
    x = rng.uniform(-1, 1, 216).astype(np.float32)
    y32 = O.entry(mats, x, O.BACKWARD)
    y64 = O.entry([m.astype(np.float64) for m in mats], x.astype(np.float64), O.BACKWARD)
    assert y32.dtype == np.float32
    assert rel_err(y32, y64) <= 1e-6
