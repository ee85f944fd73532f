"""64-bit entry offsets: kronmultN_batched with contiguous X and Y
(kronmult6_batched.hpp:12-75 addresses entry k at k*n^d) on batches whose
X and Y hold more than 2^31 elements, so an entry offset computed in 32 bits
would wrap.  Entries on both sides of the 2^31-element boundary, the first
and the last entry and a seeded random sample are recomputed on the CPU
oracle from the same inputs (random factors, FMA order: normwise relative
error 1e-13)."""
from __future__ import annotations

import numpy as np
import pytest

from _helpers import RTOL_FP64, rel_err
from oracle import oracle as O

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def km():
    import torch
    import kronmult_amd
    assert torch.cuda.is_available()
    kronmult_amd.lib()
    return kronmult_amd


# (d, n, batch): X and Y hold batch*n^d > 2^31 elements each (about 20 GB)
SHAPES = [(6, 4, 600_000),   # kron_reg, 2.46e9 elements
          (6, 6, 48_000)]    # kron_slab, 2.24e9 elements


@pytest.mark.parametrize("d,n,batch", SHAPES)
def test_batched_offsets_beyond_2e31_elements(km, d, n, batch):
    import torch
    s = n ** d
    assert batch * s > 2 ** 31
    dev = torch.device("cuda:0")
    g = torch.Generator(device=dev)
    g.manual_seed(404 + n)
    try:
        tA = torch.empty((batch, d, n, n), dtype=torch.float64, device=dev).uniform_(-1, 1, generator=g)
        tX = torch.empty((batch, s), dtype=torch.float64, device=dev).uniform_(-1, 1, generator=g)
        tY = torch.empty((batch, s), dtype=torch.float64, device=dev).uniform_(-1, 1, generator=g)
        edge = 2 ** 31 // s  # first entry whose offset is at least 2^31 elements
        rng = np.random.default_rng(505 + n)
        pick = {0, 1, edge - 1, edge, edge + 1, batch - 1}
        pick |= {int(v) for v in rng.choice(batch, 40, replace=False)}
        ents = np.array(sorted(pick))
        it = torch.as_tensor(ents, device=dev)
        Y0 = tY.index_select(0, it).cpu().numpy()
        km.kronmult_batched(d, n, tA, tX, tY, None, batch)
        torch.cuda.synchronize()
        got = tY.index_select(0, it).cpu().numpy()
        want = O.batched(tA.index_select(0, it).cpu().numpy(), tX.index_select(0, it).cpu().numpy(),
                         Y0, variant=O.BACKWARD)
        err = rel_err(got, want)
        assert err <= RTOL_FP64, err
        # the boundary entries one by one
        for k in (edge - 1, edge, batch - 1):
            j = int(np.searchsorted(ents, k))
            assert rel_err(got[j], want[j]) <= RTOL_FP64, k
    finally:
        tA = tX = tY = None
        torch.cuda.empty_cache()
