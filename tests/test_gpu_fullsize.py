"""Full-size parity through a size-independent property: with every factor
the identity, each output must equal its initial value plus the sum of its
entries' X vectors.  The expected result is computed independently on the
GPU with torch's index_add (float64), so the oracle is not needed at sizes it
cannot finish in seconds.

  * pooled xbatched: d=6, n=4, 1,048,576 entries, 512 pooled A per mode,
    65536 X drawn uniformly, 65536 Y with 16 entries each, atomic and
    output-ordered plan;
  * grouped: d=4, 20000 entries with n in 2..8, 285 output groups per n, in
    every launch mode.
Sums of up to ~40 terms are taken in different orders, so the bar is a
normwise relative error of 1e-14."""
from __future__ import annotations

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

TOL = 1e-14


@pytest.fixture(scope="module")
def km():
    import kronmult_amd
    kronmult_amd.lib()
    return kronmult_amd


def _rel(got, want):
    return float((got - want).abs().max() / max(1.0, float(want.abs().max())))


@pytest.mark.parametrize("mode", ["atomic", "plan"])
def test_pooled_xbatched_identity_full_size(km, mode):
    import torch
    dev = torch.device("cuda:0")
    d, n, batch, pool, nx, ny = 6, 4, 1 << 20, 512, 65536, 65536
    s = n ** d
    g = torch.Generator(device=dev)
    g.manual_seed(61)
    eye = torch.eye(n, dtype=torch.float64, device=dev)
    pools = [eye.expand(pool, n, n).contiguous() for _ in range(d)]
    X = torch.rand((nx, s), generator=g, device=dev, dtype=torch.float64) * 2 - 1
    Y0 = torch.rand((ny, s), generator=g, device=dev, dtype=torch.float64) * 2 - 1
    Y = Y0.clone()
    rng = np.random.default_rng(61)
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
    want = Y0.index_add(0, torch.as_tensor(yidx, device=dev), X[torch.as_tensor(xidx, device=dev)])
    assert _rel(Y, want) <= TOL


@pytest.mark.parametrize("mode", ["atomic", "deterministic", "persistent"])
def test_grouped_identity_full_size(km, mode):
    import torch
    dev = torch.device("cuda:0")
    d, batch, sizes, groups = 4, 20000, (2, 3, 4, 5, 6, 7, 8), 285
    rng = np.random.default_rng(62)
    ns = rng.choice(sizes, batch)
    ys = rng.integers(0, groups, batch)
    g = torch.Generator(device=dev)
    g.manual_seed(62)
    idx = {n: np.flatnonzero(ns == n) for n in sizes}
    X = {n: torch.rand((len(idx[n]), n ** d), generator=g, device=dev, dtype=torch.float64) * 2 - 1 for n in sizes}
    Y0 = {n: torch.rand((groups, n ** d), generator=g, device=dev, dtype=torch.float64) * 2 - 1 for n in sizes}
    Y = {n: v.clone() for n, v in Y0.items()}
    eye = {n: torch.eye(n, dtype=torch.float64, device=dev) for n in sizes}
    slot = np.empty(batch, dtype=np.int64)
    for n in sizes:
        slot[idx[n]] = np.arange(len(idx[n]))
    pX = torch.tensor([X[n].data_ptr() + int(slot[e]) * X[n].stride(0) * 8 for e, n in enumerate(ns)],
                      dtype=torch.int64, device=dev)
    pY = torch.tensor([Y[n].data_ptr() + int(ys[e]) * Y[n].stride(0) * 8 for e, n in enumerate(ns)],
                      dtype=torch.int64, device=dev)
    table = torch.tensor([eye[n].data_ptr() for n in ns for _ in range(d)], dtype=torch.int64, device=dev)
    plan = km.GroupPlan(d, [int(v) for v in ns], pY)
    opt = km.options(deterministic=(mode == "deterministic"), persistent=(mode == "persistent"))
    try:
        km.kronmult_grouped(plan, table, pX, pY, opt=opt)
        torch.cuda.synchronize()
    finally:
        plan.close()
    for n in sizes:
        want = Y0[n].index_add(0, torch.as_tensor(ys[idx[n]], device=dev), X[n])
        assert _rel(Y[n], want) <= TOL, n
