"""Matrix-core stage arithmetic on the precision-study configuration:
d = 6, n = 8, 1024 entries, outputs shared 2:1, A with entries N(0,1)/sqrt(8)
and one row per matrix scaled by 1e3, X ~ U[-1,1].  Every entry runs; 64
sampled outputs are recomputed by the FP64 CPU oracle.

  KM_PRECISION_FP64_MFMA  FP64 matrix instructions: the FP64 bar, 1e-13;
  KM_PRECISION_F16X4      mixed (FP16 slices, exact FP32 sums, FP64
                          recombination): the stated bar, 1e-10.
Reported per path (run with -s): the normwise relative error over the
sampled outputs, and the max and median over outputs of max_i |dY_i| /
max_i |Y_i| (the error relative to the output's largest element)."""
from __future__ import annotations

import numpy as np
import pytest

from _helpers import RTOL_FP64, rel_err, rng_inputs
from oracle import oracle as O

pytestmark = pytest.mark.gpu

BAR = {"fp64_mfma": RTOL_FP64, "f16x4": 1e-10}


@pytest.fixture(scope="module")
def km():
    import torch
    import kronmult_amd
    assert torch.cuda.is_available()
    kronmult_amd.lib()
    return kronmult_amd


@pytest.fixture(scope="module")
def c4():
    """C4 inputs on the device and the oracle's result for 64 sampled outputs."""
    import torch
    dev = torch.device("cuda:0")
    d, n, batch, ov = 6, 8, 1024, 2
    s = n ** d
    g = torch.Generator(device=dev)
    g.manual_seed(404)
    tA = torch.randn((batch, d, n, n), generator=g, device=dev, dtype=torch.float64) / np.sqrt(8.0)
    rows = torch.randint(0, n, (batch, d), generator=g, device=dev)
    scale = torch.ones((batch, d, n), dtype=torch.float64, device=dev)
    scale.scatter_(2, rows.unsqueeze(2), 1e3)
    tA *= scale.unsqueeze(2)      # row i of A_k is A[b, k, :, i]
    tX = torch.rand((batch, s), generator=g, device=dev, dtype=torch.float64) * 2 - 1
    yidx = np.arange(batch) // ov
    outs = np.sort(np.random.default_rng(4040).choice(batch // ov, 64, replace=False))
    ents = np.flatnonzero(np.isin(yidx, outs))
    pos = {int(o): i for i, o in enumerate(outs)}
    it = torch.as_tensor(ents, device=dev)
    want = O.pbatched(tA.index_select(0, it).cpu().numpy(), tX.index_select(0, it).cpu().numpy(),
                      np.zeros((len(outs), s)), np.array([pos[int(yidx[e])] for e in ents]),
                      variant=O.BACKWARD)
    return dict(d=d, n=n, batch=batch, tA=tA, tX=tX, yidx=yidx, outs=outs, want=want)


def _stats(got, want):
    scale = np.abs(want).max(axis=1, keepdims=True)
    el = (np.abs(got - want) / scale).max(axis=1)
    return rel_err(got, want), float(el.max()), float(np.median(el))


@pytest.mark.parametrize("prec", ["fp64_mfma", "f16x4"])
@pytest.mark.parametrize("mode", ["atomic", "plan"])
def test_c4_matrix_core_stages(km, c4, prec, mode):
    import torch
    dev = torch.device("cuda:0")
    d, n, batch = c4["d"], c4["n"], c4["batch"]
    s = n ** d
    tY = torch.zeros((batch // 2, s), dtype=torch.float64, device=dev)
    tW = torch.zeros((batch, s), dtype=torch.float64, device=dev)
    pX = km.pointer_array(c4["tX"], np.arange(batch))
    pY = km.pointer_array(tY, c4["yidx"])
    pW = km.pointer_array(tW, np.arange(batch))
    plan = km.Plan(d, n, 8, pY, batch) if mode == "plan" else None
    precision = {"fp64_mfma": km.PRECISION_FP64_MFMA, "f16x4": km.PRECISION_F16X4}[prec]
    try:
        km.kronmult_pbatched(d, n, c4["tA"], pX, pY, pW, batch, opt=km.options(plan=plan, precision=precision))
        torch.cuda.synchronize()
    finally:
        if plan is not None:
            plan.close()
    got = tY.index_select(0, torch.as_tensor(c4["outs"], device=dev)).cpu().numpy()
    nw, mx, med = _stats(got, c4["want"])
    print(f"C4 {prec} {mode}: normwise {nw:.2e}  max over outputs {mx:.2e}  median {med:.2e}")
    assert nw <= BAR[prec]


@pytest.mark.parametrize("prec", ["fp64_mfma", "f16x4"])
def test_matrix_core_stages_small_whole_batch(km, prec):
    """Every output of a small batch (outputs shared 2:1, plan and atomic)."""
    d, n, batch = 6, 8, 6
    A, X, Y0 = rng_inputs(1500, d, n, batch, batch // 2)
    yidx = np.arange(batch) // 2
    want = O.pbatched(A, X, Y0, yidx, variant=O.BACKWARD)
    precision = {"fp64_mfma": km.PRECISION_FP64_MFMA, "f16x4": km.PRECISION_F16X4}[prec]
    from _helpers import gpu_pbatched
    got = gpu_pbatched(A, X, Y0, yidx, with_w=True, opt=km.options(precision=precision))
    assert rel_err(got, want) <= BAR[prec]
    got = gpu_pbatched(A, X, Y0, yidx, with_w=True, plan_kw=dict(precision=precision))
    assert rel_err(got, want) <= BAR[prec]


def test_matrix_core_stages_unsupported(km):
    import torch
    dev = torch.device("cuda:0")
    t = torch.zeros(8, dtype=torch.float64, device=dev)
    for d, n, extra in ((6, 6, {}), (3, 8, {}), (6, 8, {"exact": True})):
        with pytest.raises(km.KronmultError) as ei:
            km.kronmult_pbatched(d, n, t, t, t, t, 2, opt=km.options(precision=km.PRECISION_F16X4, **extra))
        assert ei.value.status == km.ERROR_NOT_SUPPORTED
