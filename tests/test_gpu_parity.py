"""GPU parity: the HIP kernels (called through the C ABI) against the CPU
oracle on the same seeded inputs.

Bars (stated per test):
  * FP64, fused multiply-add (default): normwise relative error <= 1e-13;
  * KM_FLAG_EXACT: bitwise equal to the oracle's restatement of the matching
    reference build (backward order -> -DUSE_ALG993 build, forward order ->
    default recursive build) wherever each output receives one entry;
  * small-integer inputs: bitwise equal in every mode (all sums exact);
  * float entry points: relative error <= 1e-5.
"""
from __future__ import annotations

import numpy as np
import pytest

from _helpers import (RTOL_FP32, RTOL_FP64, gpu_batched, gpu_pbatched, gpu_vbatched,
                      gpu_xbatched, int_inputs, ref_err, reference_inputs, rel_err,
                      rng_inputs)
from oracle import oracle as O

pytestmark = pytest.mark.gpu

# (d, n) covering every kernel family: LDS-resident (n^d <= 19200), slab
# (6,6) and (5,8), and the workspace-staged fallback (6,7), (6,8)
SHAPES = [(d, n) for d in range(1, 7) for n in range(1, 9) if n ** d <= 20000] + [(6, 6), (5, 8), (6, 7), (6, 8)]


def _batch_for(d, n):
    s = n ** d
    return 12 if s <= 5000 else (6 if s <= 50000 else 3)


@pytest.fixture(scope="module")
def km():
    import torch
    import kronmult_amd
    assert torch.cuda.is_available()
    kronmult_amd.lib()
    return kronmult_amd


@pytest.mark.parametrize("d,n", SHAPES)
def test_pbatched_overlap_vs_oracle(km, d, n):
    batch = _batch_for(d, n)
    nY = max(1, batch // 3)
    A, X, Y0 = rng_inputs(100 + 10 * d + n, d, n, batch, nY)
    # consecutive runs of 3 and a scrambled tail: exercises on-chip run
    # merging and atomics between workgroups
    yidx = (np.arange(batch) // 3) % nY
    yidx[-2:] = yidx[:2]
    want = O.pbatched(A, X, Y0, yidx, variant=O.BACKWARD)
    need_w = (d, n) in [(6, 7), (6, 8)]
    got = gpu_pbatched(A, X, Y0, yidx, with_w=need_w)
    assert rel_err(got, want) <= RTOL_FP64, (d, n, rel_err(got, want))


@pytest.mark.parametrize("d,n", [(1, 5), (2, 7), (3, 8), (4, 4), (4, 6), (4, 8), (5, 3), (5, 4), (5, 7), (6, 2), (6, 4), (6, 5), (6, 6),
                                 (5, 8), (6, 7), (6, 8)])
@pytest.mark.parametrize("variant", ["backward", "forward"])
def test_exact_mode_bitwise(km, d, n, variant):
    batch = _batch_for(d, n)
    A, X, Y0 = rng_inputs(7 + d * n, d, n, batch, batch)
    yidx = np.arange(batch)
    if variant == "backward":
        opt = km.options(exact=True, variant=km.VARIANT_BACKWARD)
        want = O.pbatched(A, X, Y0, yidx, variant=O.BACKWARD)
    else:
        opt = km.options(exact=True, variant=km.VARIANT_FORWARD)
        want = O.pbatched(A, X, Y0, yidx, variant=O.RECURSIVE)
    with_w = km.query_kernel(d, n, 8, opt)[0] in (km.FAMILY_GLOBAL, km.FAMILY_TWOPASS)
    got = gpu_pbatched(A, X, Y0, yidx, opt=opt, with_w=with_w)
    assert np.array_equal(got, want), (d, n, variant, rel_err(got, want))


@pytest.mark.parametrize("d,n", [(2, 5), (4, 6), (3, 8), (6, 4), (6, 6), (5, 8)])
@pytest.mark.parametrize("variant", ["backward", "forward"])
def test_exact_mode_bitwise_float(km, d, n, variant):
    """Float data (every stage's result rounded to float, dot products in
    double, as the reference's float templates): EXACT is bitwise to the
    oracle for single-entry outputs in kron_lds, kron_reg, kron_slab and
    kron_slab_fwd.  kron_slab at n <= 6 moves the widened A to SGPRs; a stale
    low half of those doubles would show here as a 1e-6 error."""
    if variant == "backward":
        opt, ovar = km.options(exact=True, variant=km.VARIANT_BACKWARD), O.BACKWARD
    else:
        opt, ovar = km.options(exact=True, variant=km.VARIANT_FORWARD), O.RECURSIVE
    batch = 4
    A, X, Y0 = rng_inputs(830 + d * n, d, n, batch, batch, dtype=np.float32)
    yidx = np.arange(batch)
    got = gpu_pbatched(A, X, Y0, yidx, opt=opt)
    want = O.pbatched(A, X, Y0, yidx, variant=ovar)
    assert np.array_equal(got, want), rel_err(got, want)


@pytest.mark.parametrize("d,n", [(6, 6), (5, 8)])
@pytest.mark.parametrize("mode", ["atomic", "plan"])
def test_exact_forward_slab(km, d, n, mode):
    """EXACT with the forward order above the LDS size runs kron_slab_fwd
    (mode 1 accumulated in registers first).  With a plan every output is
    owned and updated once per entry in entry order, which is the oracle's
    sequential order: bitwise even for shared outputs (three entries each,
    interleaved).  Atomic: single-entry outputs bitwise."""
    import torch
    opt0 = km.options(exact=True, variant=km.VARIANT_FORWARD)
    assert km.query_kernel(d, n, 8, opt0)[0] == km.FAMILY_SLAB
    batch = 6
    A, X, Y0 = rng_inputs(820 + d * n, d, n, batch, batch)
    if mode == "plan":
        yidx = np.array([0, 0, 1, 1, 0, 1])
        got = gpu_pbatched(A, X, Y0, yidx, plan_kw=dict(exact=True, variant=km.VARIANT_FORWARD))
    else:
        yidx = np.arange(batch)
        got = gpu_pbatched(A, X, Y0, yidx, opt=opt0)
    # one oracle thread: entries reach a shared output in entry order
    threads = O.num_threads()
    O.set_num_threads(1)
    try:
        want = O.pbatched(A, X, Y0, yidx, variant=O.RECURSIVE)
    finally:
        O.set_num_threads(threads)
    assert np.array_equal(got, want), rel_err(got, want)


@pytest.mark.parametrize("d,n", [(6, 7), (6, 8)])
@pytest.mark.parametrize("mode", ["atomic", "plan"])
def test_exact_forward_twopass(km, d, n, mode):
    """EXACT with the forward order where one slab overflows LDS runs the
    forward two-pass pair (modes 1, 2 per position, then modes 3..d per
    sub-slab).  With a plan each part of an owned output is updated by one
    workgroup in entry order: bitwise to a single-thread oracle even for
    shared outputs.  Atomic: single-entry outputs bitwise."""
    import torch
    opt0 = km.options(exact=True, variant=km.VARIANT_FORWARD)
    assert km.query_kernel(d, n, 8, opt0)[0] == km.FAMILY_TWOPASS
    batch = 6 if n == 7 else 4
    A, X, Y0 = rng_inputs(840 + d * n, d, n, batch, batch)
    if mode == "plan":
        yidx = np.array([0, 0, 1, 1, 0, 1])[:batch]
        got = gpu_pbatched(A, X, Y0, yidx, with_w=True,
                           plan_kw=dict(exact=True, variant=km.VARIANT_FORWARD))
    else:
        yidx = np.arange(batch)
        got = gpu_pbatched(A, X, Y0, yidx, opt=opt0, with_w=True)
    threads = O.num_threads()
    O.set_num_threads(1)
    try:
        want = O.pbatched(A, X, Y0, yidx, variant=O.RECURSIVE)
    finally:
        O.set_num_threads(threads)
    assert np.array_equal(got, want), rel_err(got, want)


@pytest.mark.parametrize("d,n", [(2, 3), (3, 4), (4, 5), (6, 3), (6, 4), (6, 6)])
def test_exact_mode_forced_global_bitwise(km, d, n):
    batch = 4
    A, X, Y0 = rng_inputs(300 + d * n, d, n, batch, batch)
    yidx = np.arange(batch)
    for variant, ovar in ((km.VARIANT_BACKWARD, O.BACKWARD), (km.VARIANT_FORWARD, O.RECURSIVE)):
        opt = km.options(exact=True, variant=variant, force_global=True)
        got = gpu_pbatched(A, X, Y0, yidx, opt=opt, with_w=True)
        want = O.pbatched(A, X, Y0, yidx, variant=ovar)
        assert np.array_equal(got, want), (d, n, variant)


@pytest.mark.parametrize("d,n", [(3, 8), (6, 4), (6, 6), (6, 8)])
def test_integer_inputs_bitwise_with_overlap(km, d, n):
    batch = 8 if n ** d < 100000 else 4
    nY = 2
    A, X, Y0 = int_inputs(5 + d + n, d, n, batch, nY)
    yidx = np.arange(batch) % nY  # non-adjacent overlap: atomics between workgroups
    want = O.pbatched(A, X, Y0, yidx, variant=O.BACKWARD)
    got = gpu_pbatched(A, X, Y0, yidx, with_w=(n ** d > 50000))
    assert np.array_equal(got, want)


@pytest.mark.parametrize("d,n", [(1, 4), (3, 3), (6, 2), (6, 4), (4, 8), (6, 6)])
def test_batched_contiguous(km, d, n):
    batch = 7
    A, X, Y0 = rng_inputs(40 + d + n, d, n, batch, batch)
    want = O.batched(A, X, Y0, variant=O.BACKWARD)
    got = gpu_batched(A, X, Y0)
    assert rel_err(got, want) <= RTOL_FP64


@pytest.mark.parametrize("d,n", [(1, 3), (2, 2), (3, 4), (4, 3), (5, 2), (6, 4), (6, 6)])
def test_xbatched_reference_inputs_all_outputs_shared(km, d, n):
    """The reference test's own setup: closed-form inputs, lda = n + 3 and
    every pY[k] pointing at the same output (test_kronmult6_xbatched.cpp:
    65, 189-262)."""
    batch = 16 if n ** d <= 5000 else 4
    lda = n + 3
    A, X = reference_inputs(d, n, batch, lda)
    pools = [A[:, k] for k in range(d)]
    aidx = np.stack([np.arange(batch)] * d, axis=1)
    Y0 = np.zeros((1, n ** d))
    yidx = np.zeros(batch, dtype=np.int64)
    want = O.xbatched(pools, aidx, X, np.arange(batch), Y0, yidx, variant=O.BACKWARD)
    got = gpu_xbatched(pools, aidx, X, np.arange(batch), Y0, yidx)
    assert rel_err(got, want) <= RTOL_FP64
    # the reference's own acceptance measure and threshold
    assert ref_err(got, want) < 3e-7


@pytest.mark.parametrize("d,n", [(3, 4), (6, 4), (4, 8)])
def test_xbatched_pooled(km, d, n):
    """Pooled A (pointers repeat), repeated X, random overlapping Y."""
    rng = np.random.default_rng(99 + d * n)
    batch, P, nX, nY, lda = 64, 5, 9, 7, n + 1
    pools = [rng.uniform(-1, 1, (P, n, lda)) for _ in range(d)]
    aidx = rng.integers(0, P, (batch, d))
    X = rng.uniform(-1, 1, (nX, n ** d))
    xidx = rng.integers(0, nX, batch)
    Y0 = rng.uniform(-1, 1, (nY, n ** d))
    yidx = rng.integers(0, nY, batch)
    want = O.xbatched(pools, aidx, X, xidx, Y0, yidx, variant=O.BACKWARD)
    got = gpu_xbatched(pools, aidx, X, xidx, Y0, yidx, with_w=False)
    assert rel_err(got, want) <= RTOL_FP64


@pytest.mark.parametrize("m,n", [((3,), (5,)), ((2, 3), (4, 2)), ((1, 2, 3), (3, 2, 1)),
                                 ((2, 2, 2, 2), (3, 1, 2, 2)), ((4, 4, 4), (4, 4, 4)), ((2,) * 6, (3,) * 6)])
def test_vbatched_rectangular(km, m, n):
    rng = np.random.default_rng(sum(m) * 7 + sum(n))
    batch, nY = 6, 3
    d = len(m)
    mats = [[rng.uniform(-1, 1, (n[k], m[k])) for k in range(d)] for _ in range(batch)]
    X = rng.uniform(-1, 1, (batch, int(np.prod(n))))
    Y0 = rng.uniform(-1, 1, (nY, int(np.prod(m))))
    yidx = np.arange(batch) % nY
    # oracle takes matrices as (n_k, m_k) storage too
    want_fwd = O.vbatched(m, n, mats, X, Y0, yidx, variant=O.RECURSIVE)
    got = gpu_vbatched(m, n, mats, X, Y0, yidx)
    assert rel_err(got, want_fwd) <= RTOL_FP64
    # exact + ALG993 choice reproduces the reference's ALG993 build bitwise
    # (non-overlapping outputs)
    Y1 = rng.uniform(-1, 1, (batch, int(np.prod(m))))
    opt = km.options(exact=True, variant=km.VARIANT_ALG993)
    got1 = gpu_vbatched(m, n, mats, X, Y1, np.arange(batch), opt=opt)
    want1 = O.vbatched(m, n, mats, X, Y1, np.arange(batch), variant=O.ALG993)
    assert np.array_equal(got1, want1)


def test_vbatched_insufficient_workspace(km):
    m = n = (3, 3, 3)
    mats = [[np.eye(3) for _ in range(3)] for _ in range(2)]
    X = np.ones((2, 27))
    Y0 = np.zeros((2, 27))
    with pytest.raises(km.KronmultError) as ei:
        gpu_vbatched((2, 3, 3), n, [[np.ones((3, 2)), np.eye(3), np.eye(3)]] * 2, X, np.zeros((2, 18)),
                     np.arange(2), Wcap=8)
    assert ei.value.status == km.ERROR_INSUFFICIENT_WORKSPACE
    del mats, Y0


@pytest.mark.parametrize("d,n", [(2, 5), (3, 8), (6, 4), (5, 8), (6, 6)])
def test_float_entry_points(km, d, n):
    batch = 6
    A, X, Y0 = rng_inputs(500 + d * n, d, n, batch, 3, dtype=np.float32)
    yidx = np.arange(batch) % 3
    want = O.pbatched(A, X, Y0, yidx, variant=O.BACKWARD)
    got = gpu_pbatched(A, X, Y0, yidx)
    assert rel_err(got, want) <= RTOL_FP32
    want64 = O.pbatched(A.astype(np.float64), X.astype(np.float64), Y0.astype(np.float64), yidx,
                        variant=O.BACKWARD)
    assert rel_err(got, want64) <= RTOL_FP32


@pytest.mark.parametrize("d,n", [(4, 5), (6, 6), (5, 8), (6, 4)])
def test_float_subnormal_and_scaled_factors(km, d, n):
    """Float factors are widened to double on their bit patterns.  Of the
    three last modes, the middle one is scaled by 2^-140 (every entry a float
    subnormal) and its neighbours by 2^70, so the intermediates stay normal
    floats in either mode order and the net scale is 1: a mis-widened
    subnormal or exponent shows as an O(1) error against the oracle, which
    widens the same floats."""
    batch = 4
    A, X, Y0 = rng_inputs(810 + d * n, d, n, batch, 2, dtype=np.float32)
    A = A.astype(np.float64)
    A[:, d - 1] *= 2.0 ** 70
    A[:, d - 2] *= 2.0 ** -140
    A[:, d - 3] *= 2.0 ** 70
    A = A.astype(np.float32)
    assert np.all(np.abs(A[:, d - 2]) < np.finfo(np.float32).tiny)
    yidx = np.arange(batch) % 2
    want = O.pbatched(A, X, Y0, yidx, variant=O.BACKWARD)
    got = gpu_pbatched(A, X, Y0, yidx, with_w=True)
    assert rel_err(got, want) <= RTOL_FP32


def test_float_zero_and_infinite_factors(km):
    """+-0 and +-inf float factors (d = 1, n = 3): infinities and NaNs land
    where the oracle puts them, finite outputs match."""
    A = np.array([[[[np.inf, 0.0, 1.0], [-0.0, 1e-40, 2.0], [1.0, -np.inf, 3.0]]]], dtype=np.float32)
    X = np.array([[1.0, 2.0, 0.5]], dtype=np.float32)
    Y0 = np.zeros((1, 3), dtype=np.float32)
    want = O.pbatched(A, X, Y0, np.zeros(1, dtype=np.int64), variant=O.BACKWARD)
    got = gpu_pbatched(A, X, Y0, np.zeros(1, dtype=np.int64))
    np.testing.assert_array_equal(np.isnan(got), np.isnan(want))
    np.testing.assert_allclose(got[~np.isnan(got)], want[~np.isnan(want)], rtol=1e-6)


@pytest.mark.parametrize("d,n,pattern", [(6, 4, "mod"), (6, 4, "block"), (3, 8, "block"),
                                         (6, 6, "block"), (6, 6, "mod"), (4, 5, "random")])
def test_plan_deterministic_and_correct(km, d, n, pattern):
    import torch
    batch = 48 if n ** d <= 5000 else 16
    nY = 8 if pattern != "random" else 5
    A, X, Y0 = rng_inputs(700 + d + n, d, n, batch, nY)
    if pattern == "mod":
        yidx = np.arange(batch) % nY           # interleaved outputs: k mod nY
    elif pattern == "block":
        yidx = np.arange(batch) * nY // batch  # consecutive blocks of equal outputs
    else:
        yidx = np.random.default_rng(3).integers(0, nY, batch)
    dev = torch.device("cuda:0")
    tY = torch.tensor(Y0, device=dev)
    pY = km.pointer_array(tY, yidx)
    plan = km.Plan(d, n, 8, pY, batch)
    nout, ngroups = plan.info()
    assert nout == len(set(yidx.tolist()))
    plan.close()
    del tY, pY
    got1, got2 = gpu_pbatched(A, X, Y0, yidx, plan_kw={}, calls=2)
    assert np.array_equal(got1, got2), "plan mode must be bitwise reproducible"
    want = O.pbatched(A, X, Y0, yidx, variant=O.BACKWARD)
    assert rel_err(got1, want) <= RTOL_FP64


def test_plan_mismatch_rejected(km):
    import torch
    dev = torch.device("cuda:0")
    tY = torch.zeros((2, 64), dtype=torch.float64, device=dev)
    pY = km.pointer_array(tY, [0, 1, 0, 1])
    plan = km.Plan(3, 4, 8, pY, 4)
    A, X, Y0 = rng_inputs(1, 3, 4, 6, 2)
    with pytest.raises(km.KronmultError) as ei:
        gpu_pbatched(A, X, Y0, np.arange(6) % 2, opt=km.options(plan=plan))
    assert ei.value.status == km.ERROR_PLAN_MISMATCH
    plan.close()


@pytest.mark.parametrize("d,n", [(6, 6), (6, 4), (4, 5)])
@pytest.mark.parametrize("dtype", [np.float64, np.float32])
@pytest.mark.parametrize("mode", ["atomic", "plan"])
def test_outputs_aligned_to_one_element(km, d, n, dtype, mode):
    """Output vectors only aligned to their element size (the reference's
    pointers need no more), 4:1: kron_slab (owned runs move Y two elements
    per lane and fall back to single elements here), kron_reg (wide stores
    of consecutive elements), kron_lds."""
    import torch
    batch = 16
    A, X, Y0 = rng_inputs(960, d, n, batch, batch // 4, dtype=dtype)
    yidx = np.arange(batch) // 4
    dev = torch.device("cuda:0")
    s = n ** d
    buf = torch.zeros(Y0.size + 1, dtype=torch.float64 if dtype == np.float64 else torch.float32, device=dev)
    tY = buf[1:].view(batch // 4, s)
    tY.copy_(torch.tensor(Y0, device=dev))
    assert tY.data_ptr() % 16 != 0
    pY = km.pointer_array(tY, yidx)
    plan = km.Plan(d, n, X.itemsize, pY, batch) if mode == "plan" else None
    tA = torch.tensor(A, device=dev)
    tX = torch.tensor(X, device=dev)
    km.kronmult_pbatched(d, n, tA, km.pointer_array(tX, np.arange(batch)), pY, None, batch,
                         dtype="f64" if dtype == np.float64 else "f32", opt=km.options(plan=plan))
    torch.cuda.synchronize()
    want = O.pbatched(A, X, Y0, yidx, variant=O.BACKWARD)
    assert rel_err(tY.cpu().numpy(), want) <= (RTOL_FP64 if dtype == np.float64 else RTOL_FP32)
    if plan is not None:
        plan.close()


@pytest.mark.parametrize("d,n,batch", [(6, 4, 2048), (5, 8, 256), (6, 5, 540), (3, 8, 16384)])
@pytest.mark.parametrize("mode", ["atomic", "plan"])
def test_streamed_inputs_vs_oracle(km, d, n, batch, mode):
    """Calls whose X totals at least 64 MiB load X with the non-temporal
    hint (kron_reg, kron_lds; kron_slab with shared outputs), in separately
    compiled bodies: each against the oracle, with 4:1 output overlap."""
    import torch
    A, X, Y0 = rng_inputs(950 + d * n, d, n, batch, batch // 4)
    assert X.nbytes >= 64 << 20
    yidx = np.arange(batch) // 4
    got = gpu_pbatched(A, X, Y0, yidx, plan_kw={} if mode == "plan" else None)
    want = O.pbatched(A, X, Y0, yidx, variant=O.BACKWARD)
    assert rel_err(got, want) <= RTOL_FP64


@pytest.mark.parametrize("mode", ["atomic", "plan"])
def test_identity_full_size_d6n6_b4096(km, mode):
    """A = I at the full d=6, n=6, 4096-entry shape with 4:1 output overlap:
    Y must equal Y0 + sum of the X of its four entries exactly (each Y
    element receives four additions of exactly representable values).  With
    atomic outputs each workgroup starts its range part-way into an output
    run, so that run is flushed in two pieces."""
    import torch
    d, n, batch = 6, 6, 4096
    dev = torch.device("cuda:0")
    s = n ** d
    A = torch.zeros((batch, d, n, n), dtype=torch.float64, device=dev)
    idx = torch.arange(n, device=dev)
    A[:, :, idx, idx] = 1.0
    g = torch.Generator(device=dev)
    g.manual_seed(11)
    # multiples of 2^-10 in [-1, 1]: sums of four stay exact
    X = (torch.randint(-1024, 1025, (batch, s), generator=g, device=dev).double() / 1024)
    Y = torch.zeros((batch // 4, s), dtype=torch.float64, device=dev)
    pX = km.pointer_array(X, np.arange(batch))
    pY = km.pointer_array(Y, np.arange(batch) // 4)
    plan = km.Plan(d, n, 8, pY, batch) if mode == "plan" else None
    km.kronmult_pbatched(d, n, A, pX, pY, None, batch, opt=km.options(plan=plan))
    torch.cuda.synchronize()
    want = X.view(batch // 4, 4, s).sum(dim=1)
    assert torch.equal(Y, want)
    if plan is not None:
        plan.close()


def test_linearity_full_size_d3n8_b262144(km):
    """d=3, n=8, 262144 entries, 4:1 overlap: kron(A)(X1 + X2) equals
    kron(A)X1 + kron(A)X2 to rounding (size-independent property)."""
    import torch
    d, n, batch = 3, 8, 262144
    dev = torch.device("cuda:0")
    s = n ** d
    g = torch.Generator(device=dev)
    g.manual_seed(5)
    A = torch.rand((batch, d, n, n), generator=g, device=dev, dtype=torch.float64) * 2 - 1
    X1 = torch.rand((batch, s), generator=g, device=dev, dtype=torch.float64) * 2 - 1
    X2 = torch.rand((batch, s), generator=g, device=dev, dtype=torch.float64) * 2 - 1
    X12 = X1 + X2
    outs = []
    for X in (X1, X2, X12):
        Y = torch.zeros((batch // 4, s), dtype=torch.float64, device=dev)
        km.kronmult_pbatched(d, n, A, km.pointer_array(X, np.arange(batch)),
                             km.pointer_array(Y, np.arange(batch) // 4), None, batch)
        outs.append(Y)
    torch.cuda.synchronize()
    diff = (outs[0] + outs[1] - outs[2]).abs().max().item()
    scale = outs[2].abs().max().item()
    assert diff <= 1e-13 * scale
    # and spot-check 64 entries' outputs against the oracle
    sel = np.arange(0, 256)
    An = A[sel].cpu().numpy()
    Xn = X1[sel].cpu().numpy()
    want = O.pbatched(An, Xn, np.zeros((64, s)), sel // 4, variant=O.BACKWARD)
    assert rel_err(outs[0][:64].cpu().numpy(), want) <= RTOL_FP64


def test_empty_batch_is_noop(km):
    import torch
    dev = torch.device("cuda:0")
    Y = torch.ones((1, 16), dtype=torch.float64, device=dev)
    km.kronmult_pbatched(2, 4, Y, Y, Y, None, 0)
    torch.cuda.synchronize()
    assert torch.equal(Y, torch.ones_like(Y))


@pytest.mark.parametrize("d,n", [(6, 7), (6, 8)])
def test_twopass_family_selected(km, d, n):
    assert km.query_kernel(d, n, 8)[0] == km.FAMILY_TWOPASS
    assert km.query_kernel(d, n, 4)[0] == km.FAMILY_TWOPASS


@pytest.mark.parametrize("d,n,batch", [(6, 7, 150), (6, 8, 70)])
def test_twopass_large_batch(km, d, n, batch):
    """Two-pass family over a batch whose intermediates (one n^d tensor per
    entry in W, 140 MB at n = 7) far exceed the caches, outputs shared 2:1."""
    A, X, Y0 = rng_inputs(900 + n, d, n, batch, batch // 2)
    yidx = np.arange(batch) % (batch // 2)
    want = O.pbatched(A, X, Y0, yidx, variant=O.BACKWARD)
    got = gpu_pbatched(A, X, Y0, yidx, with_w=True)
    assert rel_err(got, want) <= RTOL_FP64


def test_twopass_xbatched_chunked_by_subbatch(km):
    """d=6, n=7 through xbatched with subbatchCount=2: the two passes run in
    chunks of two entries, each using the workspace pair of its slot."""
    rng = np.random.default_rng(77)
    d, n, batch, lda = 6, 7, 5, 8
    pools = [rng.uniform(-1, 1, (3, n, lda)) for _ in range(d)]
    aidx = rng.integers(0, 3, (batch, d))
    X = rng.uniform(-1, 1, (3, n ** d))
    xidx = np.array([0, 1, 2, 0, 1])
    Y0 = rng.uniform(-1, 1, (2, n ** d))
    yidx = np.array([0, 1, 1, 0, 0])
    want = O.xbatched(pools, aidx, X, xidx, Y0, yidx, variant=O.BACKWARD)
    got = gpu_xbatched(pools, aidx, X, xidx, Y0, yidx, subbatch=2)
    assert rel_err(got, want) <= RTOL_FP64


def test_twopass_vbatched_square_chunked(km):
    """Square vbatched at d=6, n=7 with room for two slot pairs only."""
    rng = np.random.default_rng(78)
    d, n, batch = 6, 7, 3
    m = (n,) * d
    mats = [[rng.uniform(-1, 1, (n, n)) for _ in range(d)] for _ in range(batch)]
    X = rng.uniform(-1, 1, (batch, n ** d))
    Y0 = rng.uniform(-1, 1, (2, n ** d))
    yidx = np.array([0, 1, 0])
    want = O.vbatched(m, m, mats, X, Y0, yidx, variant=O.BACKWARD)
    got = gpu_vbatched(m, m, mats, X, Y0, yidx, Wcap=2 * 2 * n ** d * 8)
    assert rel_err(got, want) <= RTOL_FP64


@pytest.mark.parametrize("mode", ["atomic", "plan"])
def test_twopass_merged_runs_integer_bitwise(km, mode):
    """Pass 2 sums each output run in registers (runs of 3 entries here, d=6,
    n=7): with small-integer inputs every sum is exact, so a dropped or
    doubly counted entry, or a run flushed twice, shows bitwise."""
    import torch
    d, n, batch = 6, 7, 12
    A, X, Y0 = int_inputs(81, d, n, batch, 4)
    yidx = np.arange(batch) // 3
    got = gpu_pbatched(A, X, Y0, yidx, with_w=True, plan_kw={} if mode == "plan" else None)
    want = O.pbatched(A, X, Y0, yidx, variant=O.BACKWARD)
    assert np.array_equal(got, want)


def test_twopass_plan_deterministic(km):
    import torch
    d, n, batch = 6, 8, 6
    A, X, Y0 = rng_inputs(79, d, n, batch, 2)
    yidx = np.array([0, 0, 1, 1, 0, 1])
    g1, g2 = gpu_pbatched(A, X, Y0, yidx, with_w=True, plan_kw={}, calls=2)
    assert np.array_equal(g1, g2)
    want = O.pbatched(A, X, Y0, yidx, variant=O.BACKWARD)
    assert rel_err(g1, want) <= RTOL_FP64


def test_input_ordered_plan_pooled_xbatched(km):
    """Input-ordered plan (entries sharing an X become neighbours, atomic
    flush) on pooled xbatched inputs."""
    import torch
    rng = np.random.default_rng(123)
    d, n, batch, P, nX, nY, lda = 6, 4, 200, 7, 11, 9, 4
    pools = [rng.uniform(-1, 1, (P, n, lda)) for _ in range(d)]
    aidx = rng.integers(0, P, (batch, d))
    X = rng.uniform(-1, 1, (nX, n ** d))
    xidx = rng.integers(0, nX, batch)
    Y0 = rng.uniform(-1, 1, (nY, n ** d))
    yidx = rng.integers(0, nY, batch)
    want = O.xbatched(pools, aidx, X, xidx, Y0, yidx, variant=O.BACKWARD)
    dev = torch.device("cuda:0")
    tpools = [torch.tensor(p, device=dev) for p in pools]
    table = km.matrix_pointer_table(tpools, aidx)
    tX = torch.tensor(X, device=dev)
    tY = torch.tensor(Y0, device=dev)
    pX = km.pointer_array(tX, xidx)
    pY = km.pointer_array(tY, yidx)
    plan = km.Plan(d, n, 8, pY, batch, order=km.ORDER_BY_INPUT, pX=pX)
    nout, _ = plan.info()
    assert nout == len(set(yidx.tolist()))
    km.kronmult_xbatched(d, n, table, lda, pX, pY, None, batch, opt=km.options(plan=plan))
    torch.cuda.synchronize()
    assert rel_err(tY.cpu().numpy(), want) <= RTOL_FP64
    plan.close()


@pytest.mark.parametrize("d,n", [(3, 8), (6, 6)])
def test_graph_capture_replays(km, d, n):
    """A pbatched call captured in a HIP graph (through torch.cuda.graph) and
    replayed several times, with eager work in between, matches the eager
    call every time."""
    import torch
    dev = torch.device("cuda:0")
    batch = 64 if n ** d < 5000 else 16
    A, X, Y0 = rng_inputs(900 + d * n, d, n, batch, batch // 4)
    tA = torch.tensor(A, device=dev)
    tX = torch.tensor(X, device=dev)
    tY0 = torch.tensor(Y0, device=dev)
    tY = tY0.clone()
    pX = km.pointer_array(tX, np.arange(batch))
    pY = km.pointer_array(tY, np.arange(batch) // 4)
    km.kronmult_pbatched(d, n, tA, pX, pY, None, batch)  # eager (also warms caches)
    torch.cuda.synchronize()
    eager = (tY - tY0).clone()
    g = torch.cuda.CUDAGraph()
    with torch.cuda.graph(g):
        km.kronmult_pbatched(d, n, tA, pX, pY, None, batch)
    for rep in range(4):
        tY.copy_(tY0)
        junk = torch.ones(1 << 20, device=dev, dtype=torch.float64).sum()  # eager work between replays
        g.replay()
        torch.cuda.synchronize()
        assert junk.item() == float(1 << 20)
        assert torch.allclose(tY - tY0, eager, rtol=0, atol=1e-12 * float(eager.abs().max())), rep
    want = O.pbatched(A, X, Y0, np.arange(batch) // 4, variant=O.BACKWARD)
    assert rel_err(tY.cpu().numpy(), want) <= RTOL_FP64


@pytest.mark.parametrize("d,n", [(4, 5), (6, 6), (6, 8)])
def test_explicit_stream(km, d, n):
    """A call on a caller-created stream (default priority) runs there: the
    result is complete once that stream alone is synchronized, and equals the
    oracle's."""
    import torch
    dev = torch.device("cuda:0")
    batch = 12 if n ** d < 50000 else 4
    A, X, Y0 = rng_inputs(950 + d * n, d, n, batch, batch // 2)
    yidx = np.arange(batch) // 2
    s = torch.cuda.Stream(device=dev)
    with torch.cuda.stream(s):
        tA = torch.tensor(A, device=dev)
        tX = torch.tensor(X, device=dev)
        tY = torch.tensor(Y0, device=dev)
        tW = torch.zeros_like(tX)
        pX = km.pointer_array(tX, np.arange(batch))
        pY = km.pointer_array(tY, yidx)
        pW = km.pointer_array(tW, np.arange(batch))
    km.kronmult_pbatched(d, n, tA, pX, pY, pW, batch, stream=s)
    s.synchronize()
    want = O.pbatched(A, X, Y0, yidx, variant=O.BACKWARD)
    assert rel_err(tY.cpu().numpy(), want) <= RTOL_FP64
