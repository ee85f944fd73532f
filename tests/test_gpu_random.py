"""Seeded randomized parity sweep: pbatched and xbatched calls with random
shapes (every d <= 6, n <= 8 up to n^d = 262144), random X and Y index maps
(repeats and collisions anywhere), FMA or exact mode, every variant, with or
without a plan, FP64 or float, and for xbatched a padded lda, a small A pool
and a random subbatchCount; every call against the CPU oracle."""
from __future__ import annotations

import os

import numpy as np
import pytest

from _helpers import RTOL_FP32, RTOL_FP64, rel_err
from oracle import oracle as O

pytestmark = pytest.mark.gpu

SHAPES = [(d, n) for d in range(1, 7) for n in range(1, 9) if n ** d <= 262144]

# KM_RANDOM_EXTRA_SEEDS=K adds K more seeds to the call and grouped sweeps
# (a longer run of the same generators: 200 calls or 15 grouped calls per seed)
EXTRA = int(os.environ.get("KM_RANDOM_EXTRA_SEEDS", "0"))


def _case(km, rng):
    import torch
    dev = torch.device("cuda:0")
    d, n = SHAPES[rng.integers(len(SHAPES))]
    s = n ** d
    bmax = 64 if s <= 5000 else (8 if s <= 50000 else 3)
    batch = int(rng.integers(1, bmax + 1))
    f32 = bool(rng.random() < 0.15)
    dt = np.float32 if f32 else np.float64
    exact = bool(rng.random() < 0.3)
    variant = int(rng.integers(0, 3))
    use_plan = bool(rng.random() < 0.4)
    xb = bool(rng.random() < 0.3)
    nY = int(rng.integers(1, batch + 1))
    yidx = rng.integers(0, nY, batch)
    nX = int(rng.integers(1, batch + 1))
    xidx = rng.integers(0, nX, batch)
    X = rng.uniform(-1, 1, (nX, s)).astype(dt)
    Y0 = rng.uniform(-1, 1, (nY, s)).astype(dt)
    desc = dict(xbatched=xb, d=d, n=n, batch=batch, f32=f32, exact=exact, variant=variant, plan=use_plan)
    tX = torch.tensor(X, device=dev)
    tY = torch.tensor(Y0, device=dev)
    pX = km.pointer_array(tX, xidx)
    pY = km.pointer_array(tY, yidx)
    plan = km.Plan(d, n, 4 if f32 else 8, pY, batch) if use_plan else None
    opt = km.options(plan=plan, exact=exact, variant=variant)
    dts = "f32" if f32 else "f64"
    ov = O.FORWARD if variant == km.VARIANT_FORWARD else O.BACKWARD
    # every device buffer a call reads stays referenced until the
    # synchronize below: a freed tensor's memory goes to the next allocation
    # while the kernel may still read it
    try:
        if not xb:
            A = rng.uniform(-1, 1, (batch, d, n, n)).astype(dt)
            tA = torch.tensor(A, device=dev)
            tW = torch.zeros((batch, s), dtype=tX.dtype, device=dev)
            pW = km.pointer_array(tW, np.arange(batch))
            km.kronmult_pbatched(d, n, tA, pX, pY, pW, batch, dtype=dts, opt=opt)
            want = O.pbatched(A, X, Y0, yidx, xidx=xidx, variant=ov)
        else:
            lda = n + int(rng.integers(0, 4))
            P = int(rng.integers(1, 6))
            pools = [rng.uniform(-1, 1, (P, n, lda)).astype(dt) for _ in range(d)]
            aidx = rng.integers(0, P, (batch, d))
            sub = int(rng.integers(0, batch + 1))
            tpools = [torch.tensor(p, device=dev) for p in pools]
            table = km.matrix_pointer_table(tpools, aidx)
            nW = max(2 * (sub if sub else max(1, batch // 2)), batch)
            tW = torch.zeros((nW, s), dtype=tX.dtype, device=dev)
            pW = km.pointer_array(tW, np.arange(nW))
            try:
                km.kronmult_xbatched(d, n, table, lda, pX, pY, pW, batch, sub, dtype=dts, opt=opt)
            except km.KronmultError as e:
                # a plan cannot cover a workspace that is reused chunk by chunk
                assert use_plan and e.status == km.ERROR_NOT_SUPPORTED and sub < batch, desc
                return None
            want = O.xbatched(pools, aidx, X, xidx, Y0, yidx, subbatch=sub, variant=ov)
        torch.cuda.synchronize()
        assert np.array_equal(tX.cpu().numpy(), X), ("X was modified", desc)
        return desc, rel_err(tY.cpu().numpy(), want), (RTOL_FP32 if f32 else RTOL_FP64)
    finally:
        if plan is not None:
            plan.close()


@pytest.fixture(scope="module")
def km():
    import kronmult_amd
    kronmult_amd.lib()
    return kronmult_amd


@pytest.mark.parametrize("seed", [11, 23, 37] + list(range(1000, 1000 + EXTRA)))
def test_random_calls_vs_oracle(km, seed):
    rng = np.random.default_rng(seed)
    ran = 0
    for _ in range(200):
        r = _case(km, rng)
        if r is None:
            continue
        desc, err, tol = r
        assert err <= tol, (desc, err)
        ran += 1
    assert ran >= 190


@pytest.mark.parametrize("seed", [5, 17])
def test_random_vbatched_vs_oracle(km, seed):
    """Uniform rectangular shapes (m_k, n_k in 1..6, d <= 6), random output
    maps, FMA or exact, every variant, FP64 or float."""
    from _helpers import gpu_vbatched
    rng = np.random.default_rng(seed)
    for _ in range(60):
        d = int(rng.integers(1, 7))
        while True:
            m = tuple(int(v) for v in rng.integers(1, 7, d))
            n = tuple(int(v) for v in rng.integers(1, 7, d))
            if np.prod(m) <= 5000 and np.prod(n) <= 5000:
                break
        batch = int(rng.integers(1, 17))
        nY = int(rng.integers(1, batch + 1))
        yidx = rng.integers(0, nY, batch)
        f32 = bool(rng.random() < 0.2)
        dt = np.float32 if f32 else np.float64
        exact = bool(rng.random() < 0.3)
        variant = int(rng.integers(0, 3))
        mats = [[rng.uniform(-1, 1, (n[k], m[k])).astype(dt) for k in range(d)] for _ in range(batch)]
        X = rng.uniform(-1, 1, (batch, int(np.prod(n)))).astype(dt)
        Y0 = rng.uniform(-1, 1, (nY, int(np.prod(m)))).astype(dt)
        got = gpu_vbatched(m, n, mats, X, Y0, yidx, opt=km.options(exact=exact, variant=variant))
        ov = {km.VARIANT_BACKWARD: O.BACKWARD, km.VARIANT_FORWARD: O.RECURSIVE, km.VARIANT_ALG993: O.ALG993}[variant]
        want = O.vbatched(m, n, mats, X, Y0, yidx, variant=ov)
        tol = RTOL_FP32 if f32 else RTOL_FP64
        assert rel_err(got, want) <= tol, (m, n, batch, f32, exact, variant, rel_err(got, want))


@pytest.mark.parametrize("seed", [3, 29] + list(range(2000, 2000 + EXTRA)))
def test_random_grouped_vs_oracle(km, seed):
    """Grouped calls with random d, random per-entry sizes (within each d's
    supported range) and random output groups, in every launch mode."""
    import test_gpu_grouped as G
    rng = np.random.default_rng(seed)
    nmax = {1: 8, 2: 8, 3: 8, 4: 8, 5: 6, 6: 4}
    for _ in range(15):
        d = int(rng.integers(1, 7))
        k = int(rng.integers(1, 5))
        sizes = sorted({int(v) for v in rng.integers(1, nmax[d] + 1, k)})
        batch = int(rng.integers(1, 120))
        ngroups = int(rng.integers(1, 8))
        mode = list(G.MODES)[int(rng.integers(len(G.MODES)))]
        ns, yg, mats, X, Y0 = G._make(int(rng.integers(1 << 30)), d, sizes, batch, ngroups)
        want = G._oracle(d, ns, yg, mats, X, Y0)
        outs, _ = G._gpu(km, d, ns, yg, mats, X, Y0, mode=mode)
        for nn in want:
            assert rel_err(outs[0][nn], want[nn]) <= RTOL_FP64, (d, sizes, batch, ngroups, mode, nn)


@pytest.mark.parametrize("d,sizes", [(4, [2, 3, 4, 5, 6, 7, 8]), (3, [1, 7, 8]), (5, [3, 4, 5]), (6, [2, 3, 4])])
@pytest.mark.parametrize("mode", ["atomic", "deterministic", "persistent"])
def test_grouped_float(km, d, sizes, mode):
    """The float grouped entry point (float data, double accumulation) against
    the FP64 oracle on the same float inputs, in every launch mode."""
    import torch
    import test_gpu_grouped as G
    dev = torch.device("cuda:0")
    ns, yg, mats, X, Y0 = G._make(7 * d + len(sizes), d, sizes, 240, 4)
    mats = [[m.astype(np.float32) for m in ms] for ms in mats]
    X = [x.astype(np.float32) for x in X]
    Y0 = {n: v.astype(np.float32) for n, v in Y0.items()}
    want = G._oracle(d, ns, yg, [[m.astype(np.float64) for m in ms] for ms in mats],
                     [x.astype(np.float64) for x in X], {n: v.astype(np.float64) for n, v in Y0.items()})
    batch = len(ns)
    tY = {n: torch.tensor(v, device=dev) for n, v in Y0.items()}
    tX = [torch.tensor(x, device=dev) for x in X]
    tA = [torch.tensor(np.ascontiguousarray(mats[b][k].T), device=dev) for b in range(batch) for k in range(d)]
    table = torch.tensor([t.data_ptr() for t in tA], dtype=torch.int64, device=dev)
    pX = torch.tensor([t.data_ptr() for t in tX], dtype=torch.int64, device=dev)
    pY = torch.tensor([tY[n][yg[b]].data_ptr() for b, n in enumerate(ns)], dtype=torch.int64, device=dev)
    plan = km.GroupPlan(d, list(ns), pY, elem_bytes=4)
    try:
        km.kronmult_grouped(plan, table, pX, pY, dtype="f32", opt=km.options(**G.MODES[mode]))
        torch.cuda.synchronize()
    finally:
        plan.close()
    for n in want:
        assert rel_err(tY[n].cpu().numpy(), want[n]) <= RTOL_FP32, (d, n, mode)
