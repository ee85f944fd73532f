"""kronmult_amd -- MI355X (gfx950) batched Kronecker matrix-vector products.

    Y(:,k) += kron(A_1(k), ..., A_d(k)) X(:,k),   k = 0..batchCount-1,  d = 1..6

Python mirror of the C ABI in ``include/kronmult_amd.h`` (the product is the
shared library ``libkronmult_amd.so`` built from ``kronmult_amd/csrc``).  The
functions keep the reference's entry-point names and argument meaning
(``kronmultN_batched / _pbatched / _xbatched / _vbatched``,
kronmult6_batched.cpp:4-18, kronmult6_pbatched.cpp:4-19,
kronmult6_xbatched.cpp:4-28, kronmult6_vbatched.hpp:10-34) with ``N`` given as
the argument ``d``.  Device memory and streams come from PyTorch; every call is
asynchronous on the current stream, like the reference's launches.

The library must be built (``__graft_entry__.build()`` or
``make -C kronmult_amd/csrc``); importing works without it, but every compute
call raises ``RuntimeError`` -- there is no CPU fallback.
"""
from __future__ import annotations

import ctypes as C
import os
from typing import Optional, Sequence

HERE = os.path.dirname(os.path.abspath(__file__))
# KRONMULT_AMD_LIB selects another build of the library (e.g. a tuning variant)
LIB_PATH = os.environ.get("KRONMULT_AMD_LIB", os.path.join(HERE, "libkronmult_amd.so"))

# km_status
SUCCESS = 0
ERROR_INVALID_VALUE = 1
ERROR_INSUFFICIENT_WORKSPACE = 2
ERROR_NOT_SUPPORTED = 3
ERROR_HIP = 4
ERROR_PLAN_MISMATCH = 5

# km_variant
VARIANT_BACKWARD = 0
VARIANT_FORWARD = 1
VARIANT_ALG993 = 2

# km_precision
PRECISION_FP64 = 0
PRECISION_FP64_MFMA = 1
PRECISION_F16X4 = 2

# flags
FLAG_EXACT = 0x1
FLAG_FORCE_GLOBAL = 0x2
FLAG_PERSISTENT = 0x4
FLAG_DETERMINISTIC = 0x8
FLAG_PLAN_OUTPUTS = 0x10

# kernel families reported by query_kernel
FAMILY_LDS, FAMILY_SLAB, FAMILY_GLOBAL, FAMILY_TWOPASS, FAMILY_REG = 0, 1, 2, 3, 4
FAMILY_GROUPED, FAMILY_HALF, FAMILY_XCD = 5, 6, 7

# every symbol the C ABI exports (include/kronmult_amd.h)
C_ABI_SYMBOLS = (
    "km_options_init", "km_status_string", "km_last_error",
    "km_kronmult_batched_f64", "km_kronmult_batched_f32",
    "km_kronmult_pbatched_f64", "km_kronmult_pbatched_f32",
    "km_kronmult_xbatched_f64", "km_kronmult_xbatched_f32",
    "km_kronmult_vbatched_f64", "km_kronmult_vbatched_f32",
    "km_plan_create", "km_plan_create_ordered", "km_plan_destroy", "km_plan_info",
    "km_query_kernel", "km_kron_flops",
    "km_group_plan_create", "km_group_plan_destroy", "km_group_plan_info",
    "km_kronmult_grouped_f64", "km_kronmult_grouped_f32",
    "km_kgemm_nn_batched_f64", "km_kgemm_nt_batched_f64", "km_kgemm_tt_batched_f64",
    "km_kgemm_nn_batched_f32", "km_kgemm_nt_batched_f32", "km_kgemm_tt_batched_f32",
    "km_xbatched_tables_f64", "km_xbatched_tables_f32",
    "km_kronmult_pbatched_noalloc_f64",
    "km_tune_shape", "km_tune_set", "km_tune_get", "km_tune_clear", "km_tune_save", "km_tune_load",
    "km_query_kernel_for", "km_cooperative_faults",
)

# km_kernel_variant
(KERNEL_DEFAULT, KERNEL_SLAB, KERNEL_HALF1, KERNEL_HALF2, KERNEL_REG, KERNEL_LDS, KERNEL_LDS_WIDE,
 KERNEL_TWOPASS, KERNEL_XCD) = range(9)

# largest batchCount of kronmult_pbatched_noalloc (KM_NOALLOC_MAX_BATCH)
NOALLOC_MAX_BATCH = 262144


class KronmultError(RuntimeError):
    def __init__(self, status: int, where: str):
        self.status = status
        detail = ""
        if status == ERROR_HIP:
            try:
                detail = " [" + lib().km_last_error().decode() + "]"
            except Exception:
                pass
        super().__init__(f"{where}: {status_string(status)} (status {status}){detail}")


class Options(C.Structure):
    """``km_options``."""
    _fields_ = [
        ("flags", C.c_uint),
        ("variant", C.c_int),
        ("precision", C.c_int),
        ("plan", C.c_void_p),
        ("reserved", C.c_int * 4),
    ]


_P, _I, _SZ = C.c_void_p, C.c_int, C.c_size_t
_lib = None


def library_loaded() -> bool:
    return _lib is not None


def lib():
    """Load ``libkronmult_amd.so``; raise if it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(
                f"kronmult_amd: native library not built ({LIB_PATH}); run "
                "`python -c 'import __graft_entry__ as g; g.build()'`")
        # One HIP runtime per process: PyTorch ships its own libamdhip64.so.7
        # and names it differently in its DT_NEEDED entries, so if this
        # library were loaded first (binding the system runtime) torch would
        # bring up a second runtime, and device pointers and streams would not
        # be shared.  Loading torch first makes this library bind to torch's.
        try:
            import torch  # noqa: F401
        except ImportError:
            pass
        L = C.CDLL(LIB_PATH)
        L.km_options_init.argtypes = [C.POINTER(Options)]
        L.km_options_init.restype = None
        L.km_status_string.argtypes = [_I]
        L.km_status_string.restype = C.c_char_p
        L.km_last_error.argtypes = []
        L.km_last_error.restype = C.c_char_p
        for suf in ("f64", "f32"):
            f = getattr(L, f"km_kronmult_batched_{suf}")
            f.argtypes = [_I, _I, _P, _P, _P, _P, _I, _P, C.POINTER(Options)]
            f.restype = _I
            f = getattr(L, f"km_kronmult_pbatched_{suf}")
            f.argtypes = [_I, _I, _P, _P, _P, _P, _I, _P, C.POINTER(Options)]
            f.restype = _I
            f = getattr(L, f"km_kronmult_xbatched_{suf}")
            f.argtypes = [_I, _I, _P, _I, _P, _P, _P, _I, _I, _P, C.POINTER(Options)]
            f.restype = _I
            f = getattr(L, f"km_kronmult_vbatched_{suf}")
            f.argtypes = [_I, _P, _P, _P, _P, _P, _P, _SZ, _I, _P, C.POINTER(Options)]
            f.restype = _I
        L.km_kronmult_pbatched_noalloc_f64.argtypes = [_I, _I, _P, _P, _P, _P, _I, _P, C.POINTER(Options)]
        L.km_kronmult_pbatched_noalloc_f64.restype = _I
        L.km_plan_create.argtypes = [C.POINTER(C.c_void_p), _I, _I, _I, _P, _I, _P]
        L.km_plan_create.restype = _I
        L.km_plan_create_ordered.argtypes = [C.POINTER(C.c_void_p), _I, _I, _I, _P, _P, _I, _I, _P]
        L.km_plan_create_ordered.restype = _I
        L.km_plan_destroy.argtypes = [_P]
        L.km_plan_destroy.restype = _I
        L.km_plan_info.argtypes = [_P, C.POINTER(_I), C.POINTER(_I)]
        L.km_plan_info.restype = _I
        L.km_query_kernel.argtypes = [_I, _I, _I, C.POINTER(Options), C.POINTER(_I), C.POINTER(_I)]
        L.km_query_kernel.restype = _I
        L.km_query_kernel_for.argtypes = [_I, _I, _I, _I, _I, C.POINTER(Options), C.POINTER(_I), C.POINTER(_I)]
        L.km_query_kernel_for.restype = _I
        L.km_cooperative_faults.argtypes = [C.POINTER(C.c_ulonglong)]
        L.km_cooperative_faults.restype = _I
        L.km_tune_shape.argtypes = [_I, _I, _I, _I, _I, _I, _P, C.POINTER(_I)]
        L.km_tune_set.argtypes = [_I, _I, _I, _I, _I, _I]
        L.km_tune_get.argtypes = [_I, _I, _I, _I, _I, C.POINTER(_I)]
        L.km_tune_clear.argtypes = []
        L.km_tune_save.argtypes = [C.c_char_p]
        L.km_tune_load.argtypes = [C.c_char_p]
        for nm in ("km_tune_shape", "km_tune_set", "km_tune_get", "km_tune_clear", "km_tune_save", "km_tune_load"):
            getattr(L, nm).restype = _I
        L.km_kron_flops.argtypes = [_I, _P, _P]
        L.km_kron_flops.restype = C.c_double
        L.km_group_plan_create.argtypes = [C.POINTER(C.c_void_p), _I, _I, _P, _P, _I, _P]
        L.km_group_plan_create.restype = _I
        L.km_group_plan_destroy.argtypes = [_P]
        L.km_group_plan_destroy.restype = _I
        L.km_group_plan_info.argtypes = [_P, C.POINTER(_I), C.POINTER(_I), C.POINTER(_I)]
        for suf in ("f64", "f32"):
            f = getattr(L, f"km_xbatched_tables_{suf}")
            f.argtypes = [_I, _I, _P, _P, _P, _P, _P, C.c_longlong, _I, _P, _P, C.c_longlong, _I, _P,
                          _P, _P, _P, _P, _P]
            f.restype = _I
        for op in ("nn", "nt", "tt"):
            for suf, real in (("f64", C.c_double), ("f32", C.c_float)):
                f = getattr(L, f"km_kgemm_{op}_batched_{suf}")
                f.argtypes = [_I, _I, _I, real, _P, _P, _P, _P, real, _P, _P, _I, _P, C.POINTER(Options)]
                f.restype = _I
        L.km_group_plan_info.restype = _I
        for suf in ("f64", "f32"):
            f = getattr(L, f"km_kronmult_grouped_{suf}")
            f.argtypes = [_P, _P, _P, _P, _P, C.POINTER(Options)]
            f.restype = _I
        _lib = L
    return _lib


def status_string(status: int) -> str:
    try:
        return lib().km_status_string(int(status)).decode()
    except RuntimeError:
        return f"status {status}"


def options(exact: bool = False, variant: int = VARIANT_BACKWARD, plan: "Optional[Plan]" = None,
            force_global: bool = False, precision: int = PRECISION_FP64,
            persistent: bool = False, deterministic: bool = False,
            plan_outputs: bool = False) -> Options:
    o = Options()
    lib().km_options_init(C.byref(o))
    o.flags = ((FLAG_EXACT if exact else 0) | (FLAG_FORCE_GLOBAL if force_global else 0)
               | (FLAG_PERSISTENT if persistent else 0) | (FLAG_DETERMINISTIC if deterministic else 0)
               | (FLAG_PLAN_OUTPUTS if plan_outputs else 0))
    o.variant = int(variant)
    o.precision = int(precision)
    o.plan = plan.handle if plan is not None else None
    return o


def _check(status: int, where: str) -> None:
    if status != SUCCESS:
        raise KronmultError(status, where)


def _stream(stream) -> int:
    if stream is None:
        import torch
        return torch.cuda.current_stream().cuda_stream
    if isinstance(stream, int):
        return stream
    return stream.cuda_stream


def _addr(t) -> Optional[int]:
    """Device address of a torch tensor, an int address, or None."""
    if t is None:
        return None
    if isinstance(t, int):
        return t
    return t.data_ptr()


def _suffix(t) -> str:
    import torch
    if t.dtype == torch.float64:
        return "f64"
    if t.dtype == torch.float32:
        return "f32"
    raise TypeError(f"kronmult_amd: unsupported dtype {t.dtype}")


def _opt_ref(opt: Optional[Options]):
    return C.byref(opt) if opt is not None else None


# ---------------------------------------------------------------------------
# entry points (mirror of the reference's host wrappers)
# ---------------------------------------------------------------------------
def kronmult_batched(d: int, n: int, A, X, Y, W=None, batchCount: Optional[int] = None,
                     stream=None, opt: Optional[Options] = None) -> None:
    """kronmultN_batched: ``A`` holds the reference's (n,n,d,batch) block
    (a tensor of shape (batch, d, n, n) indexed [b, k, j, i]); ``X``, ``Y``
    (and optional ``W``) hold (n^d, batch) column blocks, i.e. tensors of
    shape (batch, n^d).  Y is accumulated; X is not modified."""
    if batchCount is None:
        batchCount = Y.shape[0]
    f = getattr(lib(), f"km_kronmult_batched_{_suffix(Y)}")
    _check(f(d, n, _addr(A), _addr(X), _addr(Y), _addr(W), batchCount, _stream(stream), _opt_ref(opt)),
           f"kronmult{d}_batched")


def kronmult_pbatched(d: int, n: int, A, pX, pY, pW, batchCount: int, dtype="f64",
                      stream=None, opt: Optional[Options] = None) -> None:
    """kronmultN_pbatched: ``pX``, ``pY``, ``pW`` are device arrays of
    ``batchCount`` device pointers (int64 tensors, see :func:`pointer_array`);
    ``pW`` may be None when the shape has an on-chip kernel."""
    f = getattr(lib(), f"km_kronmult_pbatched_{dtype}")
    _check(f(d, n, _addr(A), _addr(pX), _addr(pY), _addr(pW), batchCount, _stream(stream), _opt_ref(opt)),
           f"kronmult{d}_pbatched")


def kronmult_pbatched_noalloc(d: int, n: int, A, pX, pY, pW, batchCount: int, stream=None,
                              opt: Optional[Options] = None) -> None:
    """Deterministic (output-ordered) FP64 pbatched that allocates nothing:
    the entry order is built on the device in the library's static
    per-device buffer (km_kronmult_pbatched_noalloc_f64).  batchCount <=
    NOALLOC_MAX_BATCH; calls that overlap on one device race."""
    _check(lib().km_kronmult_pbatched_noalloc_f64(d, n, _addr(A), _addr(pX), _addr(pY), _addr(pW), batchCount,
                                                  _stream(stream), _opt_ref(opt)),
           f"kronmult{d}_pbatched_noalloc")


def kronmult_xbatched(d: int, n: int, Aarray, lda: int, pX, pY, pW, batchCount: int,
                      subbatchCount: int = 0, dtype="f64", stream=None,
                      opt: Optional[Options] = None) -> None:
    """kronmultN_xbatched: ``Aarray`` is a device array of d*batchCount
    pointers (A_k of entry b at index k + d*b) to n x n matrices with leading
    dimension ``lda``."""
    f = getattr(lib(), f"km_kronmult_xbatched_{dtype}")
    _check(f(d, n, _addr(Aarray), lda, _addr(pX), _addr(pY), _addr(pW), batchCount, subbatchCount,
             _stream(stream), _opt_ref(opt)), f"kronmult{d}_xbatched")


def kronmult_vbatched(m: Sequence[int], n: Sequence[int], Aarray, pX, pY, W, Wcapacity_bytes: int,
                      batchCount: int, dtype="f64", stream=None, opt: Optional[Options] = None) -> None:
    """kronmultN_vbatched with N = len(m): A_k is m[k] x n[k] (leading
    dimension m[k]); shapes are uniform over the batch."""
    d = len(m)
    mm = (C.c_int * d)(*m)
    nn = (C.c_int * d)(*n)
    f = getattr(lib(), f"km_kronmult_vbatched_{dtype}")
    _check(f(d, C.cast(mm, C.c_void_p), C.cast(nn, C.c_void_p), _addr(Aarray), _addr(pX), _addr(pY),
             _addr(W), int(Wcapacity_bytes), batchCount, _stream(stream), _opt_ref(opt)),
           f"kronmult{d}_vbatched")


def xbatched_tables(pools: Sequence, aidx, X, xidx, Y, yidx, bad=None, stream=None):
    """Device-built pooled-xbatched tables (km_xbatched_tables_*).

    ``pools[k]``: device tensor holding pool k's matrices, first index =
    matrix; ``aidx``: device int32 ``(batch, d)``; ``X``/``Y``: device tensors
    whose first index is the vector; ``xidx``/``yidx``: device int32
    ``(batch,)``.  Returns ``(Aarray, pX, pY)`` as device int64 tensors of
    pointers; ``bad`` (device int32, one element) counts out-of-pool indices."""
    import torch
    d = len(pools)
    batch = int(aidx.shape[0])
    dev = X.device
    dt = _suffix(X)
    P = (C.c_void_p * d)(*[p.data_ptr() for p in pools])
    S = (C.c_longlong * d)(*[p.stride(0) for p in pools])
    N = (C.c_int * d)(*[p.shape[0] for p in pools])
    Aarray = torch.empty(d * batch, dtype=torch.int64, device=dev)
    pX = torch.empty(batch, dtype=torch.int64, device=dev)
    pY = torch.empty(batch, dtype=torch.int64, device=dev)
    f = getattr(lib(), f"km_xbatched_tables_{dt}")
    _check(f(d, batch, C.cast(P, C.c_void_p), C.cast(S, C.c_void_p), C.cast(N, C.c_void_p), _addr(aidx),
             _addr(X), X.stride(0), X.shape[0], _addr(xidx), _addr(Y), Y.stride(0), Y.shape[0], _addr(yidx),
             _addr(Aarray), _addr(pX), _addr(pY), _addr(bad), _stream(stream)), "xbatched_tables")
    return Aarray, pX, pY


def kgemm_batched(op: str, mm: int, nn: int, kk: int, alpha: float, Aarray, ldAarray, Barray, ldBarray,
                  beta: float, Carray, ldCarray, batchCount: int, dtype="f64", stream=None,
                  opt: Optional[Options] = None) -> None:
    """kgemm_{nn,nt,tt}_batched: C_b = alpha*op(A_b)*op(B_b) + beta*C_b.
    ``op`` is "nn", "nt" or "tt"; the six arrays are device arrays of
    ``batchCount`` pointers / leading dimensions (int32)."""
    if op not in ("nn", "nt", "tt"):
        raise ValueError(f"op must be nn, nt or tt, not {op!r}")
    f = getattr(lib(), f"km_kgemm_{op}_batched_{dtype}")
    _check(f(mm, nn, kk, alpha, _addr(Aarray), _addr(ldAarray), _addr(Barray), _addr(ldBarray), beta,
             _addr(Carray), _addr(ldCarray), batchCount, _stream(stream), _opt_ref(opt)),
           f"kgemm_{op}_batched")


# ---------------------------------------------------------------------------
# plans (deterministic output-sorted reduction)
# ---------------------------------------------------------------------------
ORDER_BY_OUTPUT, ORDER_BY_INPUT = 0, 1


class Plan:
    """Reduction plan for one output pointer array ``pY``: by default the
    deterministic output-ordered plan; ``order=ORDER_BY_INPUT`` (with ``pX``)
    orders the entries by input instead (atomic flush)."""

    def __init__(self, d: int, n: int, elem_bytes: int, pY, batchCount: int, stream=None,
                 order: int = ORDER_BY_OUTPUT, pX=None):
        h = C.c_void_p()
        if order == ORDER_BY_OUTPUT and pX is None:
            st = lib().km_plan_create(C.byref(h), d, n, elem_bytes, _addr(pY), batchCount, _stream(stream))
        else:
            st = lib().km_plan_create_ordered(C.byref(h), d, n, elem_bytes, _addr(pX), _addr(pY), batchCount,
                                              order, _stream(stream))
        _check(st, "km_plan_create")
        self.handle = h.value

    def info(self):
        no, ng = C.c_int(), C.c_int()
        _check(lib().km_plan_info(self.handle, C.byref(no), C.byref(ng)), "km_plan_info")
        return no.value, ng.value

    def close(self):
        if self.handle:
            lib().km_plan_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class GroupPlan:
    """Plan of a grouped variable-size batch: per-entry sizes ``n_host``
    (host integers) and the device output pointer array ``pY``."""

    def __init__(self, d: int, n_host, pY, elem_bytes: int = 8, stream=None):
        nn = (C.c_int * len(n_host))(*[int(v) for v in n_host])
        h = C.c_void_p()
        self.batch = len(n_host)
        _check(lib().km_group_plan_create(C.byref(h), d, self.batch, C.cast(nn, C.c_void_p), _addr(pY),
                                          elem_bytes, _stream(stream)), "km_group_plan_create")
        self.handle = h.value
        self.d = d

    def info(self):
        no, ng, nm = C.c_int(), C.c_int(), C.c_int()
        _check(lib().km_group_plan_info(self.handle, C.byref(no), C.byref(ng), C.byref(nm)), "km_group_plan_info")
        return no.value, ng.value, nm.value

    def close(self):
        if self.handle:
            lib().km_group_plan_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def kronmult_grouped(plan: GroupPlan, Aarray, pX, pY, dtype="f64", stream=None,
                     opt: Optional[Options] = None) -> None:
    """Grouped variable-size batch in one launch (see km_kronmult_grouped_f64)."""
    f = getattr(lib(), f"km_kronmult_grouped_{dtype}")
    _check(f(plan.handle, _addr(Aarray), _addr(pX), _addr(pY), _stream(stream), _opt_ref(opt)),
           "kronmult_grouped")


def query_kernel(d: int, n: int, elem_bytes: int = 8, opt: Optional[Options] = None):
    """(family, threads) of the kernel a square (d, n) call uses."""
    fam, nt = C.c_int(), C.c_int()
    _check(lib().km_query_kernel(d, n, elem_bytes, _opt_ref(opt), C.byref(fam), C.byref(nt)),
           "km_query_kernel")
    return fam.value, nt.value


def query_kernel_for(d: int, n: int, batchCount: int, owned: bool = False, elem_bytes: int = 8,
                     opt: Optional[Options] = None):
    """(family, threads) of the kernel a call of batchCount entries uses, with
    (owned) or without an output-ordered plan; tuned variants apply."""
    fam, nt = C.c_int(), C.c_int()
    _check(lib().km_query_kernel_for(d, n, elem_bytes, batchCount, int(owned), _opt_ref(opt),
                                     C.byref(fam), C.byref(nt)), "km_query_kernel_for")
    return fam.value, nt.value


def cooperative_faults() -> int:
    """Launches of the fused two-pass kernel on the current device whose
    bounded inter-workgroup waits gave up (km_cooperative_faults)."""
    v = C.c_ulonglong()
    _check(lib().km_cooperative_faults(C.byref(v)), "km_cooperative_faults")
    return v.value


def tune_shape(d: int, n: int, batchCount: int, overlap: int = 1, owned: bool = False,
               elem_bytes: int = 8, stream=None) -> int:
    """Time every kernel variant for this shape on synthetic device data and
    record the fastest (km_tune_shape); returns the chosen KERNEL_* value."""
    v = C.c_int()
    _check(lib().km_tune_shape(d, n, elem_bytes, batchCount, overlap, int(owned), _stream(stream),
                               C.byref(v)), "km_tune_shape")
    return v.value


def tune_set(d: int, n: int, batchCount: int, variant: int, owned: bool = False, elem_bytes: int = 8) -> None:
    _check(lib().km_tune_set(d, n, elem_bytes, batchCount, int(owned), variant), "km_tune_set")


def tune_get(d: int, n: int, batchCount: int, owned: bool = False, elem_bytes: int = 8) -> int:
    v = C.c_int()
    _check(lib().km_tune_get(d, n, elem_bytes, batchCount, int(owned), C.byref(v)), "km_tune_get")
    return v.value


def tune_clear() -> None:
    _check(lib().km_tune_clear(), "km_tune_clear")


def tune_save(path: str) -> None:
    _check(lib().km_tune_save(path.encode()), "km_tune_save")


def tune_load(path: str) -> None:
    _check(lib().km_tune_load(path.encode()), "km_tune_load")


def kron_flops(m: Sequence[int], n: Sequence[int]) -> float:
    """Reference flop model for one entry (kroncommon.hpp:154-185)."""
    d = len(m)
    mm = (C.c_int * d)(*m)
    nn = (C.c_int * d)(*n)
    return lib().km_kron_flops(d, C.cast(mm, C.c_void_p), C.cast(nn, C.c_void_p))


# ---------------------------------------------------------------------------
# helpers for building device pointer arrays from torch tensors
# ---------------------------------------------------------------------------
def pointer_array(base, idx, device=None):
    """Device int64 tensor of row addresses ``base[idx[b]]`` (base 2-D, rows
    contiguous).  ``idx`` is any integer sequence or tensor."""
    import torch
    idx_t = torch.as_tensor(idx, dtype=torch.int64, device="cpu")
    addrs = base.data_ptr() + idx_t * (base.stride(0) * base.element_size())
    return addrs.to(device if device is not None else base.device)


def matrix_pointer_table(pools: Sequence, aidx, device=None):
    """Device int64 tensor of d*batch matrix addresses for xbatched/vbatched:
    entry b, mode k -> ``pools[k][aidx[b, k]]`` (each pool 3-D, matrices
    contiguous)."""
    import torch
    aidx_t = torch.as_tensor(aidx, dtype=torch.int64, device="cpu")
    batch, d = aidx_t.shape
    cols = []
    for k in range(d):
        p = pools[k]
        cols.append(p.data_ptr() + aidx_t[:, k] * (p.stride(0) * p.element_size()))
    table = torch.stack(cols, dim=1).reshape(-1)
    return table.to(device if device is not None else pools[0].device)
