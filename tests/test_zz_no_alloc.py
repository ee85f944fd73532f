"""km_kronmult_pbatched_noalloc_f64 against the existing entry point it
replaces for callers that cannot allocate: km_kronmult_pbatched_f64 with an
output-ordered plan built from the same pY.  Both run the same kernels over
the same entry order, so the outputs must be bitwise equal; they are also
checked against the CPU oracle (normwise 1e-13).

Every case makes two calls in a row, the second with a smaller batch and a
different output grouping, so stale contents of the static per-device buffer
from the first call would show.  Shapes cover the register-blocked (d6n4),
LDS (d4n5), slab (d6n6) and two-pass (d6n8, with pW) kernels, one output for
the whole batch (one long run, most workgroup ranges empty), and the largest
batch the entry point accepts."""
from __future__ import annotations

import numpy as np
import pytest

from _helpers import RTOL_FP64, rel_err, rng_inputs
from oracle import oracle as O

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def km():
    import torch
    import kronmult_amd
    assert torch.cuda.is_available()
    kronmult_amd.lib()
    return kronmult_amd


def _run(km, A, X, Y0, yidx, noalloc):
    import torch
    dev = torch.device("cuda:0")
    batch, d, n, _ = A.shape
    tA = torch.tensor(A, device=dev)
    tX = torch.tensor(X, device=dev)
    tY = torch.tensor(Y0, device=dev)
    pX = km.pointer_array(tX, np.arange(batch))
    pY = km.pointer_array(tY, yidx)
    tW = pW = None
    if n ** d > 150 * 1024 // 8 and n >= 7:
        tW = torch.zeros((batch, n ** d), dtype=torch.float64, device=dev)
        pW = km.pointer_array(tW, np.arange(batch))
    if noalloc:
        km.kronmult_pbatched_noalloc(d, n, tA, pX, pY, pW, batch)
        torch.cuda.synchronize()
    else:
        plan = km.Plan(d, n, 8, pY, batch)
        try:
            km.kronmult_pbatched(d, n, tA, pX, pY, pW, batch, opt=km.options(plan=plan))
            torch.cuda.synchronize()
        finally:
            plan.close()
    return tY.cpu().numpy()


CASES = [
    # (d, n, first batch, first outputs, second batch, second outputs)
    (6, 4, 4096, 256, 1000, 37),
    (4, 5, 3000, 100, 700, 700),
    (6, 6, 512, 128, 96, 5),
    (6, 8, 16, 8, 6, 2),
    (4, 4, 5000, 1, 300, 1),
]


@pytest.mark.parametrize("d,n,b1,o1,b2,o2", CASES)
def test_noalloc_matches_plan_entry_point(km, d, n, b1, o1, b2, o2):
    rng = np.random.default_rng(1000 + d * n)
    for call, (batch, nout) in enumerate(((b1, o1), (b2, o2))):
        A, X, Y0 = rng_inputs(1100 + 10 * d + n + call, d, n, batch, nout)
        yidx = rng.integers(0, nout, batch) if call else (np.arange(batch) % nout)
        got = _run(km, A, X, Y0, yidx, noalloc=True)
        ref = _run(km, A, X, Y0, yidx, noalloc=False)
        assert np.array_equal(got, ref), (call, rel_err(got, ref))
        sel = np.unique(np.concatenate([yidx[:8], yidx[-8:]]))
        ents = np.flatnonzero(np.isin(yidx, sel))
        pos = {int(o): i for i, o in enumerate(sel)}
        want = O.pbatched(A[ents], X[ents], Y0[sel], np.array([pos[int(yidx[e])] for e in ents]),
                          variant=O.BACKWARD)
        assert rel_err(got[sel], want) <= RTOL_FP64, call


def test_noalloc_largest_batch(km):
    """KM_NOALLOC_MAX_BATCH entries (d=3, n=8, 4:1 outputs), then a smaller
    call; and one entry more is refused."""
    import torch
    dev = torch.device("cuda:0")
    d, n = 3, 8
    for call, batch in enumerate((km.NOALLOC_MAX_BATCH, 50000)):
        g = torch.Generator(device=dev)
        g.manual_seed(77 + call)
        tA = torch.rand((batch, d, n, n), generator=g, device=dev, dtype=torch.float64) * 2 - 1
        tX = torch.rand((batch, n ** d), generator=g, device=dev, dtype=torch.float64) * 2 - 1
        Y0 = torch.rand((batch // 4, n ** d), generator=g, device=dev, dtype=torch.float64) * 2 - 1
        yidx = np.random.default_rng(call).permutation(batch) // 4
        pX = km.pointer_array(tX, np.arange(batch))
        outs = []
        for noalloc in (True, False):
            tY = Y0.clone()
            pY = km.pointer_array(tY, yidx)
            if noalloc:
                km.kronmult_pbatched_noalloc(d, n, tA, pX, pY, None, batch)
            else:
                plan = km.Plan(d, n, 8, pY, batch)
                km.kronmult_pbatched(d, n, tA, pX, pY, None, batch, opt=km.options(plan=plan))
                torch.cuda.synchronize()
                plan.close()
            torch.cuda.synchronize()
            outs.append(tY)
        assert torch.equal(outs[0], outs[1]), call
    with pytest.raises(km.KronmultError) as ei:
        km.kronmult_pbatched_noalloc(d, n, tA, pX, pY, None, km.NOALLOC_MAX_BATCH + 1)
    assert ei.value.status == km.ERROR_INVALID_VALUE
