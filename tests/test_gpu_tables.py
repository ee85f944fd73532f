"""GPU tests of the device-built pooled xbatched tables
(km_xbatched_tables_*): the pointers equal the host-built ones bitwise, an
xbatched call through them gives the same result as through host-built tables
and matches the oracle, and out-of-pool indices are counted and left NULL."""
from __future__ import annotations

import numpy as np
import pytest

from _helpers import RTOL_FP64, gpu_xbatched, rel_err
from oracle import oracle as O

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def km():
    import torch
    import kronmult_amd
    assert torch.cuda.is_available()
    kronmult_amd.lib()
    return kronmult_amd


@pytest.mark.parametrize("d,n,dtype", [(6, 4, np.float64), (3, 5, np.float64), (2, 8, np.float32)])
def test_tables_match_host_and_drive_xbatched(km, d, n, dtype):
    import torch
    dev = torch.device("cuda:0")
    rng = np.random.default_rng(d * 10 + n)
    npool, nx, ny, batch = 7, 40, 9, 300
    lda = n + 1
    pools_np = [rng.uniform(-1, 1, (npool, n, lda)).astype(dtype) for _ in range(d)]
    X = rng.uniform(-1, 1, (nx, n ** d)).astype(dtype)
    Y0 = rng.uniform(-1, 1, (ny, n ** d)).astype(dtype)
    aidx = rng.integers(0, npool, (batch, d))
    xidx = rng.integers(0, nx, batch)
    yidx = rng.integers(0, ny, batch)

    pools = [torch.tensor(p, device=dev) for p in pools_np]
    tX = torch.tensor(X, device=dev)
    tY = torch.tensor(Y0, device=dev)
    i32 = dict(dtype=torch.int32, device=dev)
    bad = torch.zeros(1, **i32)
    Aarray, pX, pY = km.xbatched_tables(pools, torch.tensor(aidx, **i32), tX, torch.tensor(xidx, **i32), tY,
                                        torch.tensor(yidx, **i32), bad=bad)
    torch.cuda.synchronize()
    assert int(bad.item()) == 0
    assert torch.equal(Aarray, km.matrix_pointer_table(pools, aidx))
    assert torch.equal(pX, km.pointer_array(tX, xidx))
    assert torch.equal(pY, km.pointer_array(tY, yidx))

    # an xbatched call through the device-built tables
    sub = max(1, batch // 2)
    tW = torch.zeros((max(2 * sub, batch), n ** d), dtype=tX.dtype, device=dev)
    pW = km.pointer_array(tW, np.arange(tW.shape[0]))
    km.kronmult_xbatched(d, n, Aarray, lda, pX, pY, pW, batch, sub, dtype="f64" if dtype == np.float64 else "f32")
    torch.cuda.synchronize()
    got = tY.cpu().numpy()
    host = gpu_xbatched(pools_np, aidx, X, xidx, Y0, yidx)
    want = O.xbatched(pools_np, aidx, X, xidx, Y0, yidx)
    tol = RTOL_FP64 if dtype == np.float64 else 1e-5
    assert rel_err(got, want) <= tol
    assert rel_err(got, host) <= tol


def test_out_of_pool_indices_are_counted(km):
    import torch
    dev = torch.device("cuda:0")
    d, n = 2, 3
    pools = [torch.zeros((4, n, n), dtype=torch.float64, device=dev) for _ in range(d)]
    X = torch.zeros((5, n ** d), dtype=torch.float64, device=dev)
    Y = torch.zeros((5, n ** d), dtype=torch.float64, device=dev)
    i32 = dict(dtype=torch.int32, device=dev)
    aidx = torch.tensor([[0, 1], [4, 0], [-1, 3]], **i32)   # 4 and -1 outside the pool of 4
    xidx = torch.tensor([0, 5, 1], **i32)                     # 5 outside
    yidx = torch.tensor([4, 0, 0], **i32)
    bad = torch.zeros(1, **i32)
    Aarray, pX, pY = km.xbatched_tables(pools, aidx, X, xidx, Y, yidx, bad=bad)
    torch.cuda.synchronize()
    assert int(bad.item()) == 3
    a = Aarray.cpu().numpy().reshape(3, d)
    assert a[1, 0] == 0 and a[2, 0] == 0 and a[0, 0] != 0
    assert pX.cpu().numpy()[1] == 0


def test_invalid_arguments(km):
    L = km.lib()
    f = L.km_xbatched_tables_f64
    assert f(0, 1, None, None, None, None, None, 0, 0, None, None, 0, 0, None, None, None, None, None,
             None) == km.ERROR_INVALID_VALUE
    assert f(2, 0, None, None, None, None, None, 0, 0, None, None, 0, 0, None, None, None, None, None,
             None) == km.SUCCESS
