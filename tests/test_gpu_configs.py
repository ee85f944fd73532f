"""Parity at the benchmark configurations' own sizes, with random factors.

C1 runs whole on the CPU oracle.  The larger configurations run whole on the
GPU; afterwards 64 outputs are sampled (seeded), every entry that feeds a
sampled output is recomputed on the CPU oracle from the same inputs, and the
sampled outputs are compared.  Random, non-symmetric factors make a transposed
A, a wrong lda, a wrong mode order or a dropped entry visible at full size
(identity factors cannot show those).

  C1  kronmult6_pbatched, d=6, n=4, 512 entries, Y = 0, pY[k] = Y + (k mod 64)*4096
      (test_kronmult6_pbatched.cpp:216-223): atomic and plan <= 1e-13;
      KM_FLAG_EXACT + forward order with a plan bitwise equal to the
      one-thread recursive oracle (the reference's default build).
  C2a d=6, n=6, 4096 entries, pY[k] = Y + (k/4)*n^d, atomic and plan.
  C3  grouped d=4, 20000 entries, n in 2..8, ~10 entries per output, every
      launch mode.
  C4  d=6, n=8, 1024 entries (all run), outputs shared 2:1, A ~ N(0,1)/sqrt(8)
      with one row per matrix scaled by 1e3, atomic and plan.
  C5  pooled xbatched d=6, n=4, 1,048,576 entries, 512 A per mode, 65536 X
      and 65536 Y (16 entries each), atomic and output-ordered plan.
Bar: normwise relative error 1e-13 over the sampled outputs (atomics and
fused multiply-adds reorder the rounding; the values agree to ~1e-16)."""
from __future__ import annotations

import numpy as np
import pytest

from _helpers import RTOL_FP64, rel_err
from oracle import oracle as O

pytestmark = pytest.mark.gpu

NSAMPLE = 64


@pytest.fixture(scope="module")
def km():
    import torch
    import kronmult_amd
    assert torch.cuda.is_available()
    kronmult_amd.lib()
    return kronmult_amd


def _dev():
    import torch
    return torch.device("cuda:0")


def _urand(shape, gen):
    import torch
    return torch.rand(shape, generator=gen, device=_dev(), dtype=torch.float64) * 2 - 1


def _gen(seed):
    import torch
    g = torch.Generator(device=_dev())
    g.manual_seed(seed)
    return g


def _sample(n_out, seed):
    rng = np.random.default_rng(seed)
    return np.sort(rng.choice(n_out, size=min(NSAMPLE, n_out), replace=False))


# ---------------------------------------------------------------------------
# C1: the reference's CPU configuration, whole
# ---------------------------------------------------------------------------
def _c1_inputs():
    d, n, batch, nY = 6, 4, 512, 64
    rng = np.random.default_rng(101)
    A = rng.uniform(-1, 1, (batch, d, n, n))
    X = rng.uniform(-1, 1, (batch, n ** d))
    Y0 = np.zeros((nY, n ** d))
    yidx = np.arange(batch) % nY
    return A, X, Y0, yidx


def _run_pbatched(km, A, X, Y0, yidx, opt_fn, with_w=False):
    """Runs pbatched on the GPU; opt_fn(pY) -> (options, plan or None)."""
    import torch
    dev = _dev()
    batch, d, n, _ = A.shape
    tA = torch.tensor(A, device=dev)
    tX = torch.tensor(X, device=dev)
    tY = torch.tensor(Y0, device=dev)
    pX = km.pointer_array(tX, np.arange(batch))
    pY = km.pointer_array(tY, yidx)
    pW = None
    if with_w:
        tW = torch.zeros((batch, n ** d), dtype=torch.float64, device=dev)
        pW = km.pointer_array(tW, np.arange(batch))
    opt, plan = opt_fn(pY)
    try:
        km.kronmult_pbatched(d, n, tA, pX, pY, pW, batch, opt=opt)
        torch.cuda.synchronize()
    finally:
        if plan is not None:
            plan.close()
    return tY.cpu().numpy()


@pytest.mark.parametrize("mode", ["atomic", "plan"])
def test_c1_d6n4_b512_interleaved(km, mode):
    A, X, Y0, yidx = _c1_inputs()
    batch, d, n, _ = A.shape

    def opt_fn(pY):
        if mode == "atomic":
            return None, None
        plan = km.Plan(d, n, 8, pY, batch)
        return km.options(plan=plan), plan

    got = _run_pbatched(km, A, X, Y0, yidx, opt_fn)
    want = O.pbatched(A, X, Y0, yidx, variant=O.BACKWARD)
    assert rel_err(got, want) <= RTOL_FP64


def test_c1_exact_forward_bitwise_one_thread_oracle(km):
    """KM_FLAG_EXACT + forward order with an output-ordered plan: each output's
    8 entries are summed in entry order starting from Y = 0, which is the
    order of the reference's default build run on one thread."""
    A, X, Y0, yidx = _c1_inputs()
    batch, d, n, _ = A.shape

    def opt_fn(pY):
        plan = km.Plan(d, n, 8, pY, batch)
        return km.options(exact=True, variant=km.VARIANT_FORWARD, plan=plan), plan

    got = _run_pbatched(km, A, X, Y0, yidx, opt_fn)
    saved = O.num_threads()
    O.set_num_threads(1)
    try:
        want = O.pbatched(A, X, Y0, yidx, variant=O.RECURSIVE)
    finally:
        O.set_num_threads(saved)
    assert np.array_equal(got, want)


# ---------------------------------------------------------------------------
# helpers for the sampled full-size checks
# ---------------------------------------------------------------------------
def _entries_of(yidx, outs):
    """Entries (ascending) whose output is in outs, and their local output
    index into outs."""
    pos = {int(o): i for i, o in enumerate(outs)}
    ents = np.array([e for e in range(len(yidx)) if int(yidx[e]) in pos], dtype=np.int64)
    loc = np.array([pos[int(yidx[e])] for e in ents], dtype=np.int64)
    return ents, loc


def _check_sampled_pbatched(tA, tX, Y0s, Ygot, yidx, outs):
    """tA (batch,d,n,n) and tX (batch,n^d) on the device; Y0s / Ygot the
    sampled outputs before / after (host); recomputes them on the oracle."""
    import torch
    ents, loc = _entries_of(yidx, outs)
    it = torch.as_tensor(ents, device=tA.device)
    A = tA.index_select(0, it).cpu().numpy()
    X = tX.index_select(0, it).cpu().numpy()
    want = O.pbatched(A, X, Y0s, loc, variant=O.BACKWARD)
    err = rel_err(Ygot, want)
    assert err <= RTOL_FP64, err
    return err


# ---------------------------------------------------------------------------
# C2a: d=6, n=6, 4096 entries, 4:1
# ---------------------------------------------------------------------------
@pytest.mark.parametrize("mode", ["atomic", "plan"])
def test_c2a_d6n6_b4096_random_sampled(km, mode):
    import torch
    d, n, batch, ov = 6, 6, 4096, 4
    s = n ** d
    g = _gen(202)
    tA = _urand((batch, d, n, n), g)
    tX = _urand((batch, s), g)
    Y0 = _urand((batch // ov, s), g)
    tY = Y0.clone()
    yidx = np.arange(batch) // ov
    pX = km.pointer_array(tX, np.arange(batch))
    pY = km.pointer_array(tY, yidx)
    plan = km.Plan(d, n, 8, pY, batch) if mode == "plan" else None
    try:
        km.kronmult_pbatched(d, n, tA, pX, pY, None, batch, opt=km.options(plan=plan))
        torch.cuda.synchronize()
    finally:
        if plan is not None:
            plan.close()
    outs = _sample(batch // ov, 2020)
    it = torch.as_tensor(outs, device=_dev())
    _check_sampled_pbatched(tA, tX, Y0.index_select(0, it).cpu().numpy(),
                            tY.index_select(0, it).cpu().numpy(), yidx, outs)


# ---------------------------------------------------------------------------
# C4: d=6, n=8, 1024 entries (the precision study's inputs), 2:1
# ---------------------------------------------------------------------------
@pytest.mark.parametrize("mode", ["atomic", "plan"])
def test_c4_d6n8_b1024_scaled_rows_sampled(km, mode):
    import torch
    d, n, batch, ov = 6, 8, 1024, 2
    s = n ** d
    dev = _dev()
    g = _gen(404)
    tA = torch.randn((batch, d, n, n), generator=g, device=dev, dtype=torch.float64) / np.sqrt(8.0)
    rows = torch.randint(0, n, (batch, d), generator=g, device=dev)
    # row i of A_k is A[b, k, :, i] (column-major storage, A_k(i,j) = A[b,k,j,i])
    scale = torch.ones((batch, d, n), dtype=torch.float64, device=dev)
    scale.scatter_(2, rows.unsqueeze(2), 1e3)
    tA *= scale.unsqueeze(2)
    tX = _urand((batch, s), g)
    Y0 = torch.zeros((batch // ov, s), dtype=torch.float64, device=dev)
    tY = Y0.clone()
    yidx = np.arange(batch) // ov
    tW = torch.zeros((batch, s), dtype=torch.float64, device=dev)
    pX = km.pointer_array(tX, np.arange(batch))
    pY = km.pointer_array(tY, yidx)
    pW = km.pointer_array(tW, np.arange(batch))
    plan = km.Plan(d, n, 8, pY, batch) if mode == "plan" else None
    try:
        km.kronmult_pbatched(d, n, tA, pX, pY, pW, batch, opt=km.options(plan=plan))
        torch.cuda.synchronize()
    finally:
        if plan is not None:
            plan.close()
    del tW
    outs = _sample(batch // ov, 4040)
    it = torch.as_tensor(outs, device=dev)
    err = _check_sampled_pbatched(tA, tX, Y0.index_select(0, it).cpu().numpy(),
                                  tY.index_select(0, it).cpu().numpy(), yidx, outs)
    print(f"C4 {mode}: normwise relative error over {len(outs)} sampled outputs {err:.2e}")


# ---------------------------------------------------------------------------
# C3: grouped d=4, 20000 entries, n in 2..8
# ---------------------------------------------------------------------------
@pytest.mark.parametrize("mode", ["atomic", "deterministic", "persistent"])
def test_c3_grouped_d4_b20000_random_sampled(km, mode):
    import torch
    dev = _dev()
    d, batch, sizes, groups = 4, 20000, (2, 3, 4, 5, 6, 7, 8), 285
    rng = np.random.default_rng(303)
    ns = rng.choice(sizes, batch)
    ys = rng.integers(0, groups, batch)
    g = _gen(303)
    idx = {n: np.flatnonzero(ns == n) for n in sizes}
    A = {n: _urand((len(idx[n]), d, n, n), g) for n in sizes}
    X = {n: _urand((len(idx[n]), n ** d), g) for n in sizes}
    Y0 = {n: _urand((groups, n ** d), g) for n in sizes}
    Y = {n: v.clone() for n, v in Y0.items()}
    slot = np.empty(batch, dtype=np.int64)
    for n in sizes:
        slot[idx[n]] = np.arange(len(idx[n]))
    i64 = dict(dtype=torch.int64, device=dev)
    pX = torch.tensor([X[n].data_ptr() + int(slot[e]) * X[n].stride(0) * 8 for e, n in enumerate(ns)], **i64)
    pY = torch.tensor([Y[n].data_ptr() + int(ys[e]) * Y[n].stride(0) * 8 for e, n in enumerate(ns)], **i64)
    table = torch.tensor([A[n].data_ptr() + (int(slot[e]) * d + k) * n * n * 8
                          for e, n in enumerate(ns) for k in range(d)], **i64)
    plan = km.GroupPlan(d, [int(v) for v in ns], pY)
    opt = km.options(deterministic=(mode == "deterministic"), persistent=(mode == "persistent"))
    try:
        km.kronmult_grouped(plan, table, pX, pY, opt=opt)
        torch.cuda.synchronize()
    finally:
        plan.close()
    # 64 sampled outputs spread over the size bins
    per = NSAMPLE // len(sizes) + 1
    for n in sizes:
        outs = _sample(groups, 3030 + n)[:per]
        outs = np.sort(outs)
        local_y = ys[idx[n]]
        it = torch.as_tensor(outs, device=dev)
        _check_sampled_pbatched(A[n], X[n], Y0[n].index_select(0, it).cpu().numpy(),
                                Y[n].index_select(0, it).cpu().numpy(), local_y, outs)


# ---------------------------------------------------------------------------
# C5: pooled xbatched, 1,048,576 entries
# ---------------------------------------------------------------------------
@pytest.mark.parametrize("mode", ["atomic", "plan"])
def test_c5_pooled_xbatched_d6n4_random_sampled(km, mode):
    import torch
    dev = _dev()
    d, n, batch, pool, nx, ny = 6, 4, 1 << 20, 512, 65536, 65536
    s = n ** d
    g = _gen(505)
    pools = [_urand((pool, n, n), g) for _ in range(d)]
    X = _urand((nx, s), g)
    Y0 = _urand((ny, s), g)
    Y = Y0.clone()
    rng = np.random.default_rng(505)
    aidx = rng.integers(0, pool, (batch, d))
    xidx = rng.integers(0, nx, batch)
    yidx = rng.permutation(batch) // (batch // ny)
    table = km.matrix_pointer_table(pools, aidx)
    pX = km.pointer_array(X, xidx)
    pY = km.pointer_array(Y, yidx)
    plan = km.Plan(d, n, 8, pY, batch) if mode == "plan" else None
    try:
        km.kronmult_xbatched(d, n, table, n, pX, pY, None, batch, opt=km.options(plan=plan))
        torch.cuda.synchronize()
    finally:
        if plan is not None:
            plan.close()
    outs = _sample(ny, 5050)
    sel = np.isin(yidx, outs)
    ents = np.flatnonzero(sel)
    pos = {int(o): i for i, o in enumerate(outs)}
    loc = np.array([pos[int(yidx[e])] for e in ents])
    xs_u, xs_loc = np.unique(xidx[ents], return_inverse=True)
    Xh = X.index_select(0, torch.as_tensor(xs_u, device=dev)).cpu().numpy()
    poolh = [p.cpu().numpy() for p in pools]  # [p, j, i]: column-major, lda = n
    it = torch.as_tensor(outs, device=dev)
    want = O.xbatched(poolh, aidx[ents], Xh, xs_loc, Y0.index_select(0, it).cpu().numpy(), loc,
                      variant=O.BACKWARD)
    got = Y.index_select(0, it).cpu().numpy()
    err = rel_err(got, want)
    assert err <= RTOL_FP64, err


def test_c5_chunks_captured_in_a_graph(km):
    """The pooled apply in chunks of entries, each with its own output-ordered
    plan, captured once in a HIP graph: every replay (with eager work between
    replays) equals the eager chunked call bitwise and the single call within
    1e-13.  65,536 entries, 512 A per mode, 4096 X and 4096 Y (16 entries
    each), 4 chunks."""
    import torch
    dev = _dev()
    d, n, batch, pool, nx, ny, k = 6, 4, 1 << 16, 512, 4096, 4096, 4
    s = n ** d
    g = _gen(606)
    pools = [_urand((pool, n, n), g) for _ in range(d)]
    X = _urand((nx, s), g)
    Y0 = _urand((ny, s), g)
    Y = Y0.clone()
    rng = np.random.default_rng(606)
    aidx = rng.integers(0, pool, (batch, d))
    xidx = rng.integers(0, nx, batch)
    yidx = rng.permutation(batch) // (batch // ny)
    table = km.matrix_pointer_table(pools, aidx)
    pX = km.pointer_array(X, xidx)
    pY = km.pointer_array(Y, yidx)
    bounds = [batch * i // k for i in range(k + 1)]
    plans = [km.Plan(d, n, 8, pY[bounds[i]:bounds[i + 1]], bounds[i + 1] - bounds[i]) for i in range(k)]

    def chunked():
        for i in range(k):
            b0, b1 = bounds[i], bounds[i + 1]
            km.kronmult_xbatched(d, n, table[d * b0:d * b1], n, pX[b0:b1], pY[b0:b1], None, b1 - b0,
                                 opt=km.options(plan=plans[i]))

    try:
        # single call (atomic) as the reference value
        km.kronmult_xbatched(d, n, table, n, pX, pY, None, batch)
        torch.cuda.synchronize()
        single = Y - Y0
        Y.copy_(Y0)
        chunked()
        torch.cuda.synchronize()
        eager = Y - Y0
        assert _rel_t(eager, single) <= RTOL_FP64
        cg = torch.cuda.CUDAGraph()
        Y.copy_(Y0)
        with torch.cuda.graph(cg):
            chunked()
        for rep in range(3):
            Y.copy_(Y0)
            junk = torch.ones(1 << 20, device=dev, dtype=torch.float64).sum()
            cg.replay()
            torch.cuda.synchronize()
            assert junk.item() == float(1 << 20)
            assert torch.equal(Y - Y0, eager), rep
        del cg
    finally:
        for p in plans:
            p.close()


def _rel_t(a, b):
    return float((a - b).abs().max() / max(1.0, float(b.abs().max())))
