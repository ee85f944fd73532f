"""Shared helpers for the parity tests: seeded inputs, GPU runners that call
the library through its C ABI, and the error measures used by the tests."""
from __future__ import annotations

import numpy as np

RTOL_FP64 = 1e-13   # plain FP64 path vs the FP64 oracle (atomics reorder sums)
RTOL_FP32 = 1e-5    # float entry points (test_kronmult6_xbatched.cpp:730)


def rel_err(got: np.ndarray, want: np.ndarray) -> float:
    """Normwise relative error max|got-want| / max(1, max|want|)."""
    got = np.asarray(got, dtype=np.float64)
    want = np.asarray(want, dtype=np.float64)
    if want.size == 0:
        return 0.0
    return float(np.abs(got - want).max() / max(1.0, np.abs(want).max()))


def ref_err(got: np.ndarray, want: np.ndarray) -> float:
    """The reference tests' elementwise measure abs/(1+|y|)
    (test_kronmult6_xbatched.cpp:522-523)."""
    got = np.asarray(got, dtype=np.float64)
    want = np.asarray(want, dtype=np.float64)
    return float((np.abs(got - want) / (1.0 + np.abs(want))).max()) if want.size else 0.0


def rng_inputs(seed: int, d: int, n: int, batch: int, nY: int, dtype=np.float64):
    rng = np.random.default_rng(seed)
    A = rng.uniform(-1, 1, (batch, d, n, n)).astype(dtype)
    X = rng.uniform(-1, 1, (batch, n ** d)).astype(dtype)
    Y0 = rng.uniform(-1, 1, (nY, n ** d)).astype(dtype)
    return A, X, Y0


def int_inputs(seed: int, d: int, n: int, batch: int, nY: int, dtype=np.float64):
    """Small-integer inputs: every partial sum is an integer far below 2^53
    (2^24 for float), so every summation order gives the same exact result."""
    rng = np.random.default_rng(seed)
    A = rng.integers(-2, 3, (batch, d, n, n)).astype(dtype)
    X = rng.integers(-2, 3, (batch, n ** d)).astype(dtype)
    Y0 = rng.integers(-4, 5, (nY, n ** d)).astype(dtype)
    return A, X, Y0


def reference_inputs(d: int, n: int, batch: int, lda: int, dtype=np.float64):
    """The reference tests' closed-form inputs (test_kronmult6_xbatched.cpp:
    189-225): X(i,b) = (i + (b-1)*Xsize)/(Xsize*batch),
    A(i,j,k,b) = (i + (j-1)*n + (k-1)*n*n + (b-1)*batch)/(n*n*d*batch),
    1-based i, j, k, b; A stored with leading dimension lda."""
    xs = n ** d
    i = np.arange(1, xs + 1)
    b = np.arange(1, batch + 1)[:, None]
    X = ((i[None, :] + (b - 1) * xs) / (xs * batch)).astype(dtype)
    A = np.zeros((batch, d, n, lda), dtype=dtype)
    for bb in range(1, batch + 1):
        for k in range(1, d + 1):
            for j in range(1, n + 1):
                for ii in range(1, n + 1):
                    r1 = ii + (j - 1) * n + (k - 1) * n * n + (bb - 1) * batch
                    A[bb - 1, k - 1, j - 1, ii - 1] = dtype(r1) / dtype(n * n * d * batch)
    return A, X


# ---------------------------------------------------------------------------
# GPU runners (through the C ABI)
# ---------------------------------------------------------------------------
def _torch():
    import torch
    return torch


def gpu_pbatched(A, X, Y0, yidx, xidx=None, opt=None, with_w=False, check_x=True,
                 plan_kw=None, calls=1):
    """pbatched through the C ABI.  plan_kw (a dict of km.options keywords)
    builds an output-ordered plan from this call's own pY and passes it;
    calls > 1 repeats the call from Y0 and returns every result."""
    import kronmult_amd as km
    torch = _torch()
    dev = torch.device("cuda:0")
    batch, d, n, _ = A.shape
    dt = "f64" if A.dtype == np.float64 else "f32"
    xidx = np.arange(batch) if xidx is None else np.asarray(xidx)
    tA = torch.tensor(A, device=dev)
    tX = torch.tensor(X, device=dev)
    tY0 = torch.tensor(Y0, device=dev)
    tY = tY0.clone()
    pX = km.pointer_array(tX, xidx)
    pY = km.pointer_array(tY, yidx)
    tW = pW = None
    if with_w:
        tW = torch.zeros((batch, n ** d), dtype=tX.dtype, device=dev)
        pW = km.pointer_array(tW, np.arange(batch))
    plan = None
    if plan_kw is not None:
        plan = km.Plan(d, n, X.itemsize, pY, batch)
        opt = km.options(plan=plan, **plan_kw)
    outs = []
    try:
        for _ in range(calls):
            tY.copy_(tY0)
            km.kronmult_pbatched(d, n, tA, pX, pY, pW, batch, dtype=dt, opt=opt)
            torch.cuda.synchronize()
            outs.append(tY.cpu().numpy())
    finally:
        if plan is not None:
            plan.close()
    if check_x:
        assert np.array_equal(tX.cpu().numpy(), X), "X was modified"
    return outs[0] if calls == 1 else outs


def gpu_batched(A, X, Y0, opt=None, with_w=False):
    import kronmult_amd as km
    torch = _torch()
    dev = torch.device("cuda:0")
    batch, d, n, _ = A.shape
    tA = torch.tensor(A, device=dev)
    tX = torch.tensor(X, device=dev)
    tY = torch.tensor(Y0, device=dev)
    tW = torch.zeros_like(tX) if with_w else None
    km.kronmult_batched(d, n, tA, tX, tY, tW, batch, opt=opt)
    torch.cuda.synchronize()
    assert np.array_equal(tX.cpu().numpy(), X), "X was modified"
    return tY.cpu().numpy()


def gpu_xbatched(pools, aidx, X, xidx, Y0, yidx, subbatch=0, opt=None, with_w=True):
    """pools[k]: numpy (P_k, n, lda) [p, j, i]; aidx (batch, d)."""
    import kronmult_amd as km
    torch = _torch()
    dev = torch.device("cuda:0")
    aidx = np.asarray(aidx)
    batch, d = aidx.shape
    n, lda = pools[0].shape[1], pools[0].shape[2]
    dt = "f64" if X.dtype == np.float64 else "f32"
    tpools = [torch.tensor(p, device=dev) for p in pools]
    table = km.matrix_pointer_table(tpools, aidx)
    tX = torch.tensor(X, device=dev)
    tY = torch.tensor(Y0, device=dev)
    pX = km.pointer_array(tX, xidx)
    pY = km.pointer_array(tY, yidx)
    pW = tW = None
    if with_w:
        sub = subbatch if subbatch else max(1, batch // 2)
        nW = max(2 * sub, batch)
        tW = torch.zeros((nW, n ** d), dtype=tX.dtype, device=dev)
        pW = km.pointer_array(tW, np.arange(nW))
    km.kronmult_xbatched(d, n, table, lda, pX, pY, pW, batch, subbatch, dtype=dt, opt=opt)
    torch.cuda.synchronize()
    assert np.array_equal(tX.cpu().numpy(), X), "X was modified"
    return tY.cpu().numpy()


def gpu_vbatched(m, n, mats, X, Y0, yidx, Wcap=None, opt=None):
    """mats[b][k]: numpy (n_k, m_k) column-major storage of A_k (ld = m_k)."""
    import kronmult_amd as km
    torch = _torch()
    dev = torch.device("cuda:0")
    batch = len(mats)
    d = len(m)
    dt = "f64" if X.dtype == np.float64 else "f32"
    flat = [torch.tensor(np.ascontiguousarray(mats[b][k]), device=dev) for b in range(batch) for k in range(d)]
    table = torch.tensor([t.data_ptr() for t in flat], dtype=torch.int64, device=dev)
    tX = torch.tensor(X, device=dev)
    tY = torch.tensor(Y0, device=dev)
    pX = km.pointer_array(tX, np.arange(batch))
    pY = km.pointer_array(tY, yidx)
    sx = int(np.prod(n))
    sizes = [sx, int(np.prod(m))]
    for L in range(1, d):
        sizes.append(int(np.prod(m[:L])) * int(np.prod(n[L:])))
        sizes.append(int(np.prod(n[: d - L])) * int(np.prod(m[d - L:])))
    sxw = max(sizes)
    if Wcap is None:
        Wcap = 2 * sxw * X.itemsize * batch
    tW = torch.zeros(max(1, Wcap // X.itemsize), dtype=tX.dtype, device=dev)
    km.kronmult_vbatched(list(m), list(n), table, pX, pY, tW, Wcap, batch, dtype=dt, opt=opt)
    torch.cuda.synchronize()
    assert np.array_equal(tX.cpu().numpy(), X), "X was modified"
    return tY.cpu().numpy()


# ---------------------------------------------------------------------------
# batched slender GEMM (kgemm_{nn,nt,tt}_batched)
# ---------------------------------------------------------------------------
KGEMM_OP = {"nn": 0, "nt": 1, "tt": 2}
# the reference tests' acceptance bounds on max |error| (FP64):
# test_kgemm_nn_batched.cpp:570, test_kgemm_nt_batched.cpp:584,
# test_kgemm_tt_batched.cpp:588
KGEMM_REF_TOL = {"nn": 1e-9, "nt": 1e-6, "tt": 1e-6}


def kgemm_shapes(op: str, mm: int, nn: int, kk: int):
    """(rows, cols) of A and B for op (C is mm x nn)."""
    a = (kk, mm) if op == "tt" else (mm, kk)
    b = (kk, nn) if op == "nn" else (nn, kk)
    return a, b


def kgemm_inputs(op: str, mm: int, nn: int, kk: int, batch: int, kind: str = "random", seed: int = 0,
                 dtype=np.float64, pad: int = 0, shared_c: bool = False):
    """Per-entry column-major buffers (oracle.colmajor) As, Bs, Cs.

    kind: "random" U[-1,1); "reference" the reference tests' closed form
    A(i,j) = 1 + i + j*nrowA + b, B(i,j) = 1/(1 + i + j*nrowB + b), C = 1
    (1-based i, j; test_kgemm_nt_batched.cpp:174-197); "integer" values in
    {-2..2} (every sum exact).  pad adds to every leading dimension.
    shared_c: all entries update one C buffer (the same object)."""
    from oracle import oracle as O
    rng = np.random.default_rng(seed)
    (ra, ca), (rb, cb) = kgemm_shapes(op, mm, nn, kk)

    def make(rows, cols, b, which):
        if kind == "reference":
            i = np.arange(1, rows + 1)[:, None]
            j = np.arange(1, cols + 1)[None, :]
            if which == "A":
                return 1.0 + i + j * rows + b
            if which == "B":
                return 1.0 / (1.0 + i + j * rows + b)
            return np.ones((rows, cols))
        if kind == "integer":
            return rng.integers(-2, 3, (rows, cols)).astype(np.float64)
        return rng.uniform(-1, 1, (rows, cols))

    As = [O.colmajor(make(ra, ca, b, "A"), ld=ra + pad, dtype=dtype) for b in range(batch)]
    Bs = [O.colmajor(make(rb, cb, b, "B"), ld=rb + pad, dtype=dtype) for b in range(batch)]
    if shared_c:
        c = O.colmajor(make(mm, nn, 0, "C"), ld=mm + pad, dtype=dtype)
        Cs = [c] * batch
    else:
        Cs = [O.colmajor(make(mm, nn, b, "C"), ld=mm + pad, dtype=dtype) for b in range(batch)]
    return As, Bs, Cs


def gpu_kgemm(op: str, mm: int, nn: int, kk: int, alpha: float, As, Bs, beta: float, Cs, opt=None):
    """Run kgemm_{op}_batched on the GPU through the C ABI; returns the C
    buffers per entry (shared inputs stay shared)."""
    import kronmult_amd as km
    torch = _torch()
    dev = torch.device("cuda:0")
    dt = "f64" if As[0].dtype == np.float64 else "f32"
    tA = [torch.tensor(np.ascontiguousarray(a.T), device=dev) for a in As]
    tB = [torch.tensor(np.ascontiguousarray(b.T), device=dev) for b in Bs]
    uniq = {}
    for c in Cs:
        if id(c) not in uniq:
            uniq[id(c)] = (c, torch.tensor(np.ascontiguousarray(c.T), device=dev))
    tC = [uniq[id(c)][1] for c in Cs]
    i64 = dict(dtype=torch.int64, device=dev)
    i32 = dict(dtype=torch.int32, device=dev)
    pA = torch.tensor([t.data_ptr() for t in tA], **i64)
    pB = torch.tensor([t.data_ptr() for t in tB], **i64)
    pC = torch.tensor([t.data_ptr() for t in tC], **i64)
    lA = torch.tensor([a.shape[0] for a in As], **i32)
    lB = torch.tensor([b.shape[0] for b in Bs], **i32)
    lC = torch.tensor([c.shape[0] for c in Cs], **i32)
    km.kgemm_batched(op, mm, nn, kk, alpha, pA, lA, pB, lB, beta, pC, lC, len(As), dtype=dt, opt=opt)
    torch.cuda.synchronize()
    out = {k: np.asfortranarray(t.cpu().numpy().T) for k, (c, t) in uniq.items()}
    return [out[id(c)] for c in Cs]
