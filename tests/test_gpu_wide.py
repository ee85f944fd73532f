"""Modes wider than 32 (the workspace-staged kernel caches such fibers in LDS
instead of registers; the reference accepts any n): square shapes, which
update their workspace in place, and rectangular vbatched shapes, against the
oracle -- FMA within 1e-13, KM_FLAG_EXACT bitwise (one entry per output)."""
from __future__ import annotations

import numpy as np
import pytest

from _helpers import RTOL_FP64, gpu_pbatched, gpu_vbatched, rel_err, rng_inputs
from oracle import oracle as O

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def km():
    import torch
    import kronmult_amd
    assert torch.cuda.is_available()
    kronmult_amd.lib()
    return kronmult_amd


@pytest.mark.parametrize("d,n", [(1, 100), (2, 40), (3, 33), (1, 300)])
@pytest.mark.parametrize("mode", ["atomic", "plan"])
def test_wide_square_vs_oracle(km, d, n, mode):
    batch = 6
    A, X, Y0 = rng_inputs(1600 + d * n, d, n, batch, batch // 2)
    yidx = np.arange(batch) // 2
    assert km.query_kernel(d, n)[0] == km.FAMILY_GLOBAL
    got = gpu_pbatched(A, X, Y0, yidx, with_w=True, plan_kw={} if mode == "plan" else None)
    assert rel_err(got, O.pbatched(A, X, Y0, yidx, variant=O.BACKWARD)) <= RTOL_FP64


@pytest.mark.parametrize("d,n", [(2, 40), (3, 33)])
@pytest.mark.parametrize("variant", ["backward", "forward"])
def test_wide_square_exact_bitwise(km, d, n, variant):
    batch = 3
    A, X, Y0 = rng_inputs(1700 + d * n, d, n, batch, batch)
    yidx = np.arange(batch)
    kv = km.VARIANT_BACKWARD if variant == "backward" else km.VARIANT_FORWARD
    ov = O.BACKWARD if variant == "backward" else O.RECURSIVE
    got = gpu_pbatched(A, X, Y0, yidx, with_w=True, opt=km.options(exact=True, variant=kv))
    assert np.array_equal(got, O.pbatched(A, X, Y0, yidx, variant=ov))


def test_wide_rectangular_vbatched(km):
    m, n = [50, 3, 34], [37, 45, 2]
    batch = 4
    rng = np.random.default_rng(1800)
    mats = [[rng.uniform(-1, 1, (n[k], m[k])) for k in range(3)] for _ in range(batch)]
    X = rng.uniform(-1, 1, (batch, int(np.prod(n))))
    Y0 = rng.uniform(-1, 1, (2, int(np.prod(m))))
    yidx = np.arange(batch) % 2
    got = gpu_vbatched(m, n, mats, X, Y0, yidx)
    want = O.vbatched(m, n, [[mats[b][k] for k in range(3)] for b in range(batch)], X, Y0, yidx,
                      variant=O.BACKWARD)
    assert rel_err(got, want) <= RTOL_FP64


def test_mode_above_300_not_supported(km):
    import torch
    t = torch.zeros(8, dtype=torch.float64, device=torch.device("cuda:0"))
    with pytest.raises(km.KronmultError) as ei:
        km.kronmult_pbatched(1, 301, t, t, t, t, 1)
    assert ei.value.status == km.ERROR_NOT_SUPPORTED
