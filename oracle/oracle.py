"""Python face of the CPU oracle -- TEST INFRASTRUCTURE ONLY.

Only tests/, ``__graft_entry__.smoke()`` and ``bench.py``'s cpu_baseline leg
import this module, and only to check or to time the CPU baseline; the
product path (``kronmult_amd``) never imports it.

Two CPU libraries sit behind the same functions:

* ``liboracle.so`` -- the clean-room C restatement (``kron_oracle.c``) of the
  reference's CPU/OpenMP path; built on every machine by ``oracle/Makefile``.
* ``_ref/libkronref_{recursive,alg993}.so`` -- the reference's own CPU path
  compiled from its headers (``ref_driver.cpp``); only where the reference
  sources were present at build time.

Array conventions (all numpy, C-contiguous):

* square A for ``batched``/``pbatched`` -- shape ``(batch, d, n, n)`` indexed
  ``[b, k, j, i]``, i.e. the reference's column-major ``(n, n, d, batch)``
  block (kronmult6_pbatched.hpp:42-45);
* X -- ``(rows, prod(n))``; Y -- ``(rows, prod(m))``; an index vector maps each
  entry to its X / Y row, so overlapping outputs are expressed as repeated Y
  indices (kronmult6_pbatched.hpp:59-61 reads them through pointer arrays).

Every function returns the updated Y (a copy); the inputs are not modified
(the reference clobbers X in ``batched``/``pbatched``, kronm_backward.hpp:128,
so the wrappers hand it private copies).
"""
from __future__ import annotations

import ctypes as C
import os
import time
from typing import Optional, Sequence

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))

RECURSIVE, BACKWARD, FORWARD, ALG993 = 0, 1, 2, 3

_P = C.c_void_p
_I = C.c_int
_SZ = C.c_size_t

_lib = None
_ref = {}


def _declare(lib, prefix):
    for suf in ("f64", "f32"):
        getattr(lib, f"{prefix}_batched_{suf}").argtypes = [_I, _I, _P, _P, _P, _P, _I] + ([_I] if prefix == "oracle" else [])
        getattr(lib, f"{prefix}_pbatched_{suf}").argtypes = [_I, _I, _P, _P, _P, _P, _I] + ([_I] if prefix == "oracle" else [])
        getattr(lib, f"{prefix}_xbatched_{suf}").argtypes = [_I, _I, _P, _I, _P, _P, _P, _I, _I] + ([_I] if prefix == "oracle" else [])
        getattr(lib, f"{prefix}_vbatched_{suf}").argtypes = [_I, _P, _P, _P, _P, _P, _P, _SZ, _I] + ([_I] if prefix == "oracle" else [])
        getattr(lib, f"{prefix}_vbatched_{suf}").restype = _I
    return lib


def _declare_kgemm(lib, prefix):
    for suf, real in (("f64", C.c_double), ("f32", C.c_float)):
        f = getattr(lib, f"{prefix}_kgemm_batched_{suf}")
        f.argtypes = [_I, _I, _I, _I, real, _P, _P, _P, _P, real, _P, _P, _I]


def lib():
    """The restatement library (raises if it has not been built)."""
    global _lib
    if _lib is None:
        path = os.path.join(HERE, "liboracle.so")
        if not os.path.exists(path):
            raise RuntimeError(f"oracle library missing: {path} (run `make -C oracle`)")
        L = C.CDLL(path)
        _declare(L, "oracle")
        for suf in ("f64", "f32"):
            getattr(L, f"oracle_entry_{suf}").argtypes = [_I, _P, _P, _P, _P, _I, _P, _P, _P]
            getattr(L, f"oracle_brute_{suf}").argtypes = [_I, _P, _P, _P, _P, _P, _P]
        L.oracle_kron_flops.argtypes = [_I, _P, _P]
        L.oracle_kron_flops.restype = C.c_double
        L.oracle_alg993_uses_forward.argtypes = [_I, _P, _P]
        L.oracle_num_threads.restype = _I
        L.oracle_set_num_threads.argtypes = [_I]
        _declare_kgemm(L, "oracle")
        _lib = L
    return _lib


def ref_available(variant: str = "recursive") -> bool:
    return os.path.exists(os.path.join(HERE, "_ref", f"libkronref_{variant}.so"))


def ref_lib(variant: str = "recursive"):
    """The reference's own CPU path (``recursive`` = stock build, ``alg993`` =
    built with -DUSE_ALG993).  Raises if it was not built here."""
    if variant not in _ref:
        path = os.path.join(HERE, "_ref", f"libkronref_{variant}.so")
        if not os.path.exists(path):
            raise RuntimeError(f"reference CPU library not built: {path}")
        L = C.CDLL(path)
        _declare(L, "ref")
        L.ref_kron_flops.argtypes = [_I, _P, _P]
        L.ref_kron_flops.restype = C.c_double
        L.ref_variant.restype = _I
        _declare_kgemm(L, "ref")
        _ref[variant] = L
    return _ref[variant]


def set_num_threads(t: int) -> None:
    lib().oracle_set_num_threads(int(t))


def num_threads() -> int:
    return lib().oracle_num_threads()


# ---------------------------------------------------------------------------
# helpers
# ---------------------------------------------------------------------------
def _suffix(dtype) -> str:
    dt = np.dtype(dtype)
    if dt == np.float64:
        return "f64"
    if dt == np.float32:
        return "f32"
    raise TypeError(f"unsupported dtype {dt}")


def _ptr(a: np.ndarray) -> int:
    return a.ctypes.data


def _rowptrs(arr: np.ndarray, idx: Sequence[int]) -> np.ndarray:
    """Pointer array to rows ``arr[idx[b]]`` (uintp, keeps ``arr`` alive via caller)."""
    base = _ptr(arr)
    stride = arr.strides[0]
    return (base + np.asarray(idx, dtype=np.int64) * stride).astype(np.uintp)


def _check_index(idx, batch: int, rows: int, name: str) -> None:
    """Index vectors become raw pointers in C: reject a wrong shape or an
    out-of-range entry here instead of reading out of bounds."""
    idx = np.asarray(idx)
    if idx.shape != (batch,) or not np.issubdtype(idx.dtype, np.integer):
        raise ValueError(f"{name} must be {batch} integers, got shape {idx.shape} dtype {idx.dtype}")
    if batch and (idx.min() < 0 or idx.max() >= rows):
        raise ValueError(f"{name} out of range [0, {rows})")


def _ivec(v: Sequence[int]) -> np.ndarray:
    return np.ascontiguousarray(np.asarray(v, dtype=np.int32))


def kron_flops(m: Sequence[int], n: Sequence[int]) -> float:
    """Reference flop model (kroncommon.hpp:154-185)."""
    mm, nn = _ivec(m), _ivec(n)
    return lib().oracle_kron_flops(len(mm), _ptr(mm), _ptr(nn))


def alg993_uses_forward(m: Sequence[int], n: Sequence[int]) -> bool:
    mm, nn = _ivec(m), _ivec(n)
    return bool(lib().oracle_alg993_uses_forward(len(mm), _ptr(mm), _ptr(nn)))


# ---------------------------------------------------------------------------
# batch drivers
# ---------------------------------------------------------------------------
def batched(A: np.ndarray, X: np.ndarray, Y: np.ndarray, variant: int = RECURSIVE,
            use_ref: Optional[str] = None) -> np.ndarray:
    """kronmultN_batched: A (batch,d,n,n), X (batch,n^d), Y (batch,n^d) -> new Y."""
    batch, d, n, _ = A.shape
    A = np.ascontiguousarray(A)
    Xw = np.array(X, copy=True, order="C")
    Yo = np.array(Y, copy=True, order="C")
    W = np.zeros_like(Xw)
    suf = _suffix(A.dtype)
    if use_ref:
        rc = getattr(ref_lib(use_ref), f"ref_batched_{suf}")(d, n, _ptr(A), _ptr(Xw), _ptr(Yo), _ptr(W), batch)
        assert rc == 0
    else:
        getattr(lib(), f"oracle_batched_{suf}")(d, n, _ptr(A), _ptr(Xw), _ptr(Yo), _ptr(W), batch, variant)
    return Yo


def pbatched(A: np.ndarray, X: np.ndarray, Y: np.ndarray, yidx: Sequence[int],
             xidx: Optional[Sequence[int]] = None, variant: int = RECURSIVE,
             use_ref: Optional[str] = None, timings: Optional[list] = None) -> np.ndarray:
    """kronmultN_pbatched: entry b reads X[xidx[b]] and accumulates into
    Y[yidx[b]].  Every entry gets a private X copy (the reference clobbers X).
    When ``timings`` is a list, the wall time of the library call alone
    (without the copies) is appended to it."""
    batch, d, n, _ = A.shape
    A = np.ascontiguousarray(A)
    xidx = np.arange(batch) if xidx is None else np.asarray(xidx)
    _check_index(xidx, batch, X.shape[0], "xidx")
    _check_index(yidx, batch, Y.shape[0], "yidx")
    Xw = np.ascontiguousarray(X[xidx]).copy()
    Yo = np.array(Y, copy=True, order="C")
    W = np.zeros_like(Xw)
    pX = _rowptrs(Xw, np.arange(batch))
    pY = _rowptrs(Yo, yidx)
    pW = _rowptrs(W, np.arange(batch))
    suf = _suffix(A.dtype)
    t0 = time.perf_counter()
    if use_ref:
        rc = getattr(ref_lib(use_ref), f"ref_pbatched_{suf}")(d, n, _ptr(A), _ptr(pX), _ptr(pY), _ptr(pW), batch)
        assert rc == 0
    else:
        getattr(lib(), f"oracle_pbatched_{suf}")(d, n, _ptr(A), _ptr(pX), _ptr(pY), _ptr(pW), batch, variant)
    if timings is not None:
        timings.append(time.perf_counter() - t0)
    return Yo


def xbatched(Apool: Sequence[np.ndarray], aidx: np.ndarray, X: np.ndarray, xidx: Sequence[int],
             Y: np.ndarray, yidx: Sequence[int], subbatch: int = 0, variant: int = RECURSIVE,
             use_ref: Optional[str] = None, timings: Optional[list] = None) -> np.ndarray:
    """kronmultN_xbatched with pooled A.

    ``Apool[k]`` has shape ``(P_k, n, lda)`` indexed ``[p, j, i]`` (column
    major, leading dimension lda >= n); ``aidx`` is ``(batch, d)`` and entry b
    uses ``Apool[k][aidx[b, k]]`` as A_k (the reference's (d,batch) pointer
    table, kronmult_xbatched.hpp:67-69).  X is preserved (copied into the
    workspace pairs, kronmult_xbatched.hpp:110-119).  When ``timings`` is a
    list, the wall time of the library call alone is appended to it."""
    aidx = np.asarray(aidx)
    batch, d = aidx.shape
    n = Apool[0].shape[1]
    lda = Apool[0].shape[2]
    dt = Apool[0].dtype
    for k in range(d):
        _check_index(aidx[:, k], batch, Apool[k].shape[0], f"aidx[:, {k}]")
    _check_index(xidx, batch, X.shape[0], "xidx")
    _check_index(yidx, batch, Y.shape[0], "yidx")
    pools = [np.ascontiguousarray(p) for p in Apool]
    Aptr = np.empty((batch, d), dtype=np.uintp)
    for k in range(d):
        Aptr[:, k] = _rowptrs(pools[k], aidx[:, k])
    Xw = np.array(X, copy=True, order="C")
    Yo = np.array(Y, copy=True, order="C")
    xs = Xw.shape[1]
    sub = subbatch if subbatch else max(1, batch // 2)
    nW = max(2 * sub, batch)
    W = np.zeros((nW, xs), dtype=dt)
    pX = _rowptrs(Xw, xidx)
    pY = _rowptrs(Yo, yidx)
    pW = _rowptrs(W, np.arange(nW))
    suf = _suffix(dt)
    t0 = time.perf_counter()
    if use_ref:
        rc = getattr(ref_lib(use_ref), f"ref_xbatched_{suf}")(d, n, _ptr(Aptr), lda, _ptr(pX), _ptr(pY), _ptr(pW), batch, subbatch)
        assert rc == 0
    else:
        getattr(lib(), f"oracle_xbatched_{suf}")(d, n, _ptr(Aptr), lda, _ptr(pX), _ptr(pY), _ptr(pW), batch, subbatch, variant)
    if timings is not None:
        timings.append(time.perf_counter() - t0)
    return Yo


def vbatched(m: Sequence[int], n: Sequence[int], A: Sequence[Sequence[np.ndarray]], X: np.ndarray,
             Y: np.ndarray, yidx: Sequence[int], Wcap_bytes: Optional[int] = None,
             variant: int = RECURSIVE, use_ref: Optional[str] = None) -> np.ndarray:
    """kronmultN_vbatched: uniform rectangular shapes; ``A[b][k]`` is an
    ``m_k x n_k`` column-major matrix given as a numpy array of shape
    ``(n_k, m_k)`` (ld = m_k, kronmult_vbatched.hpp:184-189)."""
    d = len(m)
    batch = len(A)
    dt = X.dtype
    _check_index(yidx, batch, Y.shape[0], "yidx")
    if X.shape[0] < batch:
        raise ValueError(f"X needs {batch} rows, has {X.shape[0]}")
    mats = [[np.ascontiguousarray(A[b][k], dtype=dt) for k in range(d)] for b in range(batch)]
    Aptr = np.array([[_ptr(mats[b][k]) for k in range(d)] for b in range(batch)], dtype=np.uintp).reshape(batch, d)
    Xw = np.array(X, copy=True, order="C")
    Yo = np.array(Y, copy=True, order="C")
    mm, nn = _ivec(m), _ivec(n)
    sx = int(np.prod(n))
    sw = 0
    for L in range(1, d):
        sw = max(sw, int(np.prod(m[:L])) * int(np.prod(n[L:])),
                 int(np.prod(n[: d - L])) * int(np.prod(m[d - L:])))
    sxw = max(sx, sw)
    if Wcap_bytes is None:
        Wcap_bytes = 2 * sxw * dt.itemsize * batch
    W = np.zeros(max(1, Wcap_bytes // dt.itemsize), dtype=dt)
    pX = _rowptrs(Xw, np.arange(batch))
    pY = _rowptrs(Yo, yidx)
    suf = _suffix(dt)
    if use_ref:
        rc = getattr(ref_lib(use_ref), f"ref_vbatched_{suf}")(d, _ptr(mm), _ptr(nn), _ptr(Aptr), _ptr(pX), _ptr(pY), _ptr(W), Wcap_bytes, batch)
    else:
        rc = getattr(lib(), f"oracle_vbatched_{suf}")(d, _ptr(mm), _ptr(nn), _ptr(Aptr), _ptr(pX), _ptr(pY), _ptr(W), Wcap_bytes, batch, variant)
    if rc < 0:
        raise ValueError("vbatched workspace too small for one (X copy, W) slot pair")
    return Yo


def brute(mats: Sequence[np.ndarray], x: np.ndarray) -> np.ndarray:
    """Explicit Kronecker sum for one entry (test_kronmult6_xbatched.cpp:447-499).

    ``mats[k]`` is A_k as a numpy (m_k, n_k) array (row i, column j);
    returns y = kron(A_1..A_d) x."""
    d = len(mats)
    m = _ivec([a.shape[0] for a in mats])
    n = _ivec([a.shape[1] for a in mats])
    dt = x.dtype
    cols = [np.asfortranarray(a, dtype=dt) for a in mats]  # column-major, ld = m_k
    ld = _ivec([a.shape[0] for a in mats])
    Aptr = np.array([_ptr(c) for c in cols], dtype=np.uintp)
    xx = np.ascontiguousarray(x, dtype=dt)
    y = np.zeros(int(np.prod(m)), dtype=dt)
    getattr(lib(), f"oracle_brute_{_suffix(dt)}")(d, _ptr(m), _ptr(n), _ptr(Aptr), _ptr(ld), _ptr(xx), _ptr(y))
    return y


def entry(mats: Sequence[np.ndarray], x: np.ndarray, variant: int = RECURSIVE) -> np.ndarray:
    """One entry through the restated staged algorithm: returns kron(A) x."""
    d = len(mats)
    m = _ivec([a.shape[0] for a in mats])
    n = _ivec([a.shape[1] for a in mats])
    dt = x.dtype
    cols = [np.asfortranarray(a, dtype=dt) for a in mats]
    ld = _ivec([a.shape[0] for a in mats])
    Aptr = np.array([_ptr(c) for c in cols], dtype=np.uintp)
    xs = int(np.prod(n))
    # stage sizes of both orders: forward (modes 1..d) and backward (d..1)
    ws = max([xs, int(np.prod(m))] + [int(np.prod(m[:L])) * int(np.prod(n[L:])) for L in range(1, d)]
             + [int(np.prod(n[:d - L])) * int(np.prod(m[d - L:])) for L in range(1, d)])
    xx = np.zeros(ws, dtype=dt)
    xx[:xs] = x
    w = np.zeros(ws, dtype=dt)
    y = np.zeros(int(np.prod(m)), dtype=dt)
    getattr(lib(), f"oracle_entry_{_suffix(dt)}")(d, _ptr(m), _ptr(n), _ptr(Aptr), _ptr(ld), variant, _ptr(xx), _ptr(y), _ptr(w))
    return y


# ---------------------------------------------------------------------------
# batched slender GEMM (kgemm_{nn,nt,tt}_batched)
# ---------------------------------------------------------------------------
KGEMM_NN, KGEMM_NT, KGEMM_TT = 0, 1, 2


def colmajor(M: np.ndarray, ld: Optional[int] = None, dtype=None) -> np.ndarray:
    """Column-major storage of matrix M with leading dimension ``ld`` >= rows:
    a Fortran-ordered ``(ld, cols)`` array whose first ``rows`` rows hold M."""
    M = np.asarray(M)
    rows, cols = M.shape
    ld = rows if ld is None else int(ld)
    assert ld >= max(rows, 1)
    buf = np.zeros((ld, cols), dtype=dtype or M.dtype, order="F")
    buf[:rows, :] = M
    return buf


def kgemm_batched(op: int, mm: int, nn: int, kk: int, alpha: float, As: Sequence[np.ndarray],
                  Bs: Sequence[np.ndarray], beta: float, Cs: Sequence[np.ndarray],
                  use_ref: bool = False) -> list:
    """C_b = alpha*op(A_b)*op(B_b) + beta*C_b for every entry b in order.

    ``As``, ``Bs``, ``Cs`` hold one column-major buffer per entry (see
    :func:`colmajor`; leading dimension = first extent).  Entries may share
    a C buffer (the same array object), as callers of the reference may.
    Returns the updated C buffers per entry (copies; shared inputs stay
    shared).  op: KGEMM_NN (A mm x kk, B kk x nn), KGEMM_NT (B nn x kk),
    KGEMM_TT (A kk x mm, B nn x kk)."""
    batch = len(As)
    dt = np.dtype(As[0].dtype) if batch else np.dtype(np.float64)
    suf = _suffix(dt)
    work = {}
    outs = []
    for c in Cs:
        if id(c) not in work:
            work[id(c)] = np.array(c, dtype=dt, order="F", copy=True)
        outs.append(work[id(c)])
    Ak = [np.asfortranarray(a, dtype=dt) for a in As]
    Bk = [np.asfortranarray(b, dtype=dt) for b in Bs]
    pa = np.array([_ptr(a) for a in Ak], dtype=np.uintp)
    pb = np.array([_ptr(b) for b in Bk], dtype=np.uintp)
    pc = np.array([_ptr(c) for c in outs], dtype=np.uintp)
    la = _ivec([a.shape[0] for a in Ak])
    lb = _ivec([b.shape[0] for b in Bk])
    lc = _ivec([c.shape[0] for c in outs])
    L = ref_lib() if use_ref else lib()
    f = getattr(L, f"{'ref' if use_ref else 'oracle'}_kgemm_batched_{suf}")
    f(int(op), int(mm), int(nn), int(kk), float(alpha), _ptr(pa), _ptr(la), _ptr(pb), _ptr(lb),
      float(beta), _ptr(pc), _ptr(lc), int(batch))
    return outs
