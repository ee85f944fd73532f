"""CPU tests of the C ABI boundary: the shared library loads without a GPU,
exports every function include/kronmult_amd.h declares (and the
reference-named C++ overloads of include/kronmult_compat.hpp), and rejects
invalid arguments with the documented status before touching any device."""
from __future__ import annotations

import ctypes as C
import os
import re
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB = os.path.join(ROOT, "kronmult_amd", "libkronmult_amd.so")

pytestmark = pytest.mark.skipif(not os.path.exists(LIB), reason="library not built")


def _header_functions():
    text = open(os.path.join(ROOT, "include", "kronmult_amd.h")).read()
    return sorted(set(re.findall(r"\b(km_[a-z0-9_]+)\s*\(", text)))


def _exported():
    out = subprocess.run(["nm", "-D", "--defined-only", LIB], capture_output=True, text=True, check=True).stdout
    return {line.split()[-1] for line in out.splitlines() if " T " in line}


def test_every_declared_function_is_exported():
    names = _header_functions()
    assert len(names) >= 20
    exported = _exported()
    missing = [n for n in names if n not in exported]
    assert not missing, missing


def test_python_mirror_noalloc_limit_matches_header():
    import re
    import kronmult_amd as km
    hdr = open(os.path.join(ROOT, "include", "kronmult_amd.h")).read()
    assert int(re.search(r"#define KM_NOALLOC_MAX_BATCH (\d+)", hdr).group(1)) == km.NOALLOC_MAX_BATCH


def test_python_mirror_lists_the_abi():
    import kronmult_amd as km
    assert set(km.C_ABI_SYMBOLS) == set(_header_functions())


def test_compat_overloads_exported():
    out = subprocess.run(["nm", "-DC", "--defined-only", LIB], capture_output=True, text=True, check=True).stdout
    for kind in ("batched", "pbatched", "xbatched", "vbatched"):
        for n in range(1, 7):
            assert re.search(rf"\bkronmult{n}_{kind}\(", out), f"kronmult{n}_{kind}"
    hdr = open(os.path.join(ROOT, "include", "kronmult_compat.hpp")).read()
    assert hdr.count("void kronmult") == 24


def test_invalid_arguments_rejected_on_host():
    import kronmult_amd as km
    L = km.lib()
    f = L.km_kronmult_pbatched_f64
    assert f(0, 4, None, None, None, None, 1, None, None) == km.ERROR_INVALID_VALUE   # d < 1
    assert f(7, 4, None, None, None, None, 1, None, None) == km.ERROR_INVALID_VALUE   # d > 6
    assert f(3, 0, None, None, None, None, 1, None, None) == km.ERROR_INVALID_VALUE   # n < 1
    assert f(3, 4, None, None, None, None, -1, None, None) == km.ERROR_INVALID_VALUE  # batch < 0
    assert f(3, 4, None, None, None, None, 5, None, None) == km.ERROR_INVALID_VALUE   # null arrays
    assert f(6, 40, 1, 1, 1, None, 1, None, None) == km.ERROR_INVALID_VALUE           # n^d > 2^31-1
    assert f(3, 4, None, None, None, None, 0, None, None) == km.SUCCESS               # empty batch
    x = L.km_kronmult_xbatched_f64
    assert x(3, 4, 1, 3, 1, 1, None, 4, 0, None, None) == km.ERROR_INVALID_VALUE      # lda < n
    assert x(3, 4, 1, 4, 1, 1, None, 4, -1, None, None) == km.ERROR_INVALID_VALUE     # subbatch < 0
    v = L.km_kronmult_vbatched_f64
    m = (C.c_int * 2)(2, 0)
    assert v(2, C.cast(m, C.c_void_p), C.cast(m, C.c_void_p), 1, 1, 1, None, 0, 1, None, None) == \
        km.ERROR_INVALID_VALUE
    o = km.Options()
    L.km_options_init(C.byref(o))
    o.precision = 7
    assert f(3, 4, 1, 1, 1, None, 2, None, C.byref(o)) == km.ERROR_NOT_SUPPORTED
    o.precision = 0
    o.variant = 9
    assert f(3, 4, 1, 1, 1, None, 2, None, C.byref(o)) == km.ERROR_INVALID_VALUE
    # the allocation-free entry point: its size limit and arguments are
    # checked before any device work
    na = L.km_kronmult_pbatched_noalloc_f64
    assert na(6, 4, 1, 1, 1, None, km.NOALLOC_MAX_BATCH + 1, None, None) == km.ERROR_INVALID_VALUE
    assert b"KM_NOALLOC_MAX_BATCH" in L.km_last_error()
    assert na(0, 4, 1, 1, 1, None, 4, None, None) == km.ERROR_INVALID_VALUE
    assert na(3, 4, None, None, None, None, 4, None, None) == km.ERROR_INVALID_VALUE
    assert na(3, 4, None, None, None, None, 0, None, None) == km.SUCCESS
    o.variant = 0
    o.plan = 1  # a plan is not accepted: the call builds its own order
    assert na(3, 4, 1, 1, 1, None, 4, None, C.byref(o)) == km.ERROR_INVALID_VALUE


def test_options_defaults_and_status_strings():
    import kronmult_amd as km
    o = km.Options()
    km.lib().km_options_init(C.byref(o))
    assert (o.flags, o.variant, o.precision, o.plan) == (0, km.VARIANT_BACKWARD, km.PRECISION_FP64, None)
    for s in range(6):
        assert km.status_string(s) and "unknown" not in km.status_string(s)


def test_kron_flops_matches_reference_model():
    import kronmult_amd as km
    from oracle import oracle as O
    for m, n in (((6,) * 6, (6,) * 6), ((2, 3, 4), (4, 3, 2)), ((8,) * 3, (8,) * 3), ((5,), (7,))):
        assert km.kron_flops(m, n) == O.kron_flops(m, n)


def test_query_kernel_families():
    import kronmult_amd as km
    assert km.query_kernel(6, 4)[0] == km.FAMILY_REG
    assert km.query_kernel(3, 8)[0] == km.FAMILY_REG
    assert km.query_kernel(6, 5)[0] == km.FAMILY_LDS
    assert km.query_kernel(6, 4, opt=km.options(exact=True))[0] == km.FAMILY_LDS   # backward order
    assert km.query_kernel(6, 6)[0] == km.FAMILY_SLAB
    assert km.query_kernel(5, 8)[0] == km.FAMILY_SLAB
    assert km.query_kernel(6, 8)[0] == km.FAMILY_TWOPASS
    assert km.query_kernel(3, 12)[0] == km.FAMILY_GLOBAL
    assert km.query_kernel(6, 4, opt=km.options(force_global=True))[0] == km.FAMILY_GLOBAL


def test_cpp_caller_of_the_reference_interface_links(tmp_path):
    """A C++ program that includes only kronmult_compat.hpp and the HIP
    runtime (tests/compat/caller.cpp) compiles and links against the library
    with the reference's function names and argument lists."""
    out = tmp_path / "caller"
    r = subprocess.run(["make", "-B", "-C", os.path.join(ROOT, "tests", "compat"), f"OUT={out}"],
                       capture_output=True, text=True)
    assert r.returncode == 0, r.stderr[-2000:]
    assert out.exists()


C_PROGRAM = r"""
#include <stdio.h>
#include "kronmult_amd.h"
int main(void)
{
  km_options o;
  int m[6] = {6, 6, 6, 6, 6, 6};
  km_options_init(&o);
  printf("%d %d %s %.0f\n", (int)o.precision, (int)o.variant,
         km_status_string(KM_ERROR_INVALID_VALUE), km_kron_flops(6, m, m));
  return 0;
}
"""


def test_c_abi_header_is_plain_c(tmp_path):
    """kronmult_amd.h compiles as pedantic C99 (what cgo, JNI and other FFI
    hosts see), links with gcc and its host-only functions run without a GPU."""
    src = tmp_path / "c_abi.c"
    src.write_text(C_PROGRAM)
    exe = tmp_path / "c_abi"
    r = subprocess.run(["gcc", "-std=c99", "-Wall", "-Wextra", "-pedantic", "-Werror",
                        "-I", os.path.join(ROOT, "include"), str(src), "-o", str(exe),
                        "-L", os.path.dirname(LIB), "-lkronmult_amd", "-L/opt/rocm/lib", "-lamdhip64",
                        f"-Wl,-rpath,{os.path.dirname(LIB)}"], capture_output=True, text=True)
    assert r.returncode == 0, r.stderr[-2000:]
    out = subprocess.run([str(exe)], capture_output=True, text=True, timeout=60)
    assert out.returncode == 0, out.stderr
    # defaults FP64 and backward order, the status text, the flop model
    assert out.stdout.strip() == f"0 0 invalid value {12 * 6 ** 7}"


def test_tune_table_json_round_trip(tmp_path):
    """The variant table's JSON file round trip (host-only: no GPU is needed
    to set, save, clear and load entries)."""
    import kronmult_amd as km
    km.tune_clear()
    try:
        km.tune_set(6, 6, 4096, km.KERNEL_HALF2, owned=True)
        km.tune_set(3, 8, 262144, km.KERNEL_LDS)
        with pytest.raises(km.KronmultError):
            km.tune_set(6, 6, 4096, km.KERNEL_REG)      # no kron_reg for d6n6
        path = str(tmp_path / "t.json")
        km.tune_save(path)
        km.tune_clear()
        assert km.tune_get(6, 6, 4096, owned=True) == km.KERNEL_DEFAULT
        km.tune_load(path)
        assert km.tune_get(6, 6, 4096, owned=True) == km.KERNEL_HALF2
        assert km.tune_get(6, 6, 5000, owned=True) == km.KERNEL_HALF2   # same log2 bin
        assert km.tune_get(6, 6, 8192, owned=True) == km.KERNEL_DEFAULT
        assert km.tune_get(6, 6, 4096, owned=False) == km.KERNEL_DEFAULT
        assert km.tune_get(3, 8, 262144) == km.KERNEL_LDS
        import json
        doc = json.load(open(path))
        assert len(doc["entries"]) == 2
    finally:
        km.tune_clear()


_TWO_RUNTIMES = r'''
import ctypes, os, sys
lib = ctypes.CDLL(sys.argv[1])          # binds the system HIP runtime first
if sys.argv[2] == "second":
    ctypes.CDLL(sys.argv[3])            # a second copy of the runtime library
f = lib.km_kronmult_pbatched_f64
f.argtypes = [ctypes.c_int] * 2 + [ctypes.c_void_p] * 4 + [ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p]
lib.km_last_error.restype = ctypes.c_char_p
st = f(3, 4, 8, 8, 8, None, 2, None, None)
print(st, lib.km_last_error().decode())
'''


def test_second_hip_runtime_is_reported(tmp_path):
    """A process holding two HIP runtime libraries (this library loaded
    before PyTorch's own copy) gets KM_ERROR_HIP with an explanation instead
    of mixing device state across runtimes; one runtime is not reported."""
    import torch
    second = os.path.join(os.path.dirname(torch.__file__), "lib", "libamdhip64.so")
    if not os.path.exists(second):
        pytest.skip("no second HIP runtime library to load")
    script = tmp_path / "two.py"
    script.write_text(_TWO_RUNTIMES)
    outs = {}
    for mode in ("one", "second"):
        r = subprocess.run([sys.executable, str(script), LIB, mode, second], capture_output=True, text=True,
                           timeout=120, env={**os.environ, "HIP_VISIBLE_DEVICES": os.environ.get("HIP_VISIBLE_DEVICES", "")})
        assert r.returncode == 0, r.stderr
        outs[mode] = r.stdout.strip().splitlines()[-1]
    assert outs["second"].startswith("4 ") and "two HIP runtimes" in outs["second"], outs
    assert "two HIP runtimes" not in outs["one"], outs
