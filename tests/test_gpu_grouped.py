"""GPU parity of the grouped variable-size entry point (per-entry n; per-size
launches with atomic or owned outputs, or one persistent launch) against the
CPU oracle, entry by entry."""
from __future__ import annotations

import numpy as np
import pytest

from _helpers import RTOL_FP64, rel_err
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def _make(seed, d, sizes, batch, ngroups_per_n):
    """Entries with n drawn from `sizes`; outputs form groups of equal n and
    every entry targets a random group of its own n."""
    rng = np.random.default_rng(seed)
    ns = rng.choice(sizes, batch)
    groups = {n: list(range(ngroups_per_n)) for n in sizes}
    yg = np.array([rng.integers(0, ngroups_per_n) for _ in range(batch)])
    mats = [[rng.uniform(-1, 1, (n, n)) for _ in range(d)] for n in ns]
    X = [rng.uniform(-1, 1, n ** d) for n in ns]
    Y0 = {n: rng.uniform(-1, 1, (ngroups_per_n, n ** d)) for n in sizes}
    del groups
    return ns, yg, mats, X, Y0


def _oracle(d, ns, yg, mats, X, Y0):
    Y = {n: v.copy() for n, v in Y0.items()}
    for b, n in enumerate(ns):
        Y[n][yg[b]] += O.entry(mats[b], X[b], O.BACKWARD)
    return Y


MODES = {"atomic": {}, "deterministic": {"deterministic": True}, "persistent": {"persistent": True}}


def _gpu(km, d, ns, yg, mats, X, Y0, exact=False, runs=1, mode="atomic"):
    import torch
    dev = torch.device("cuda:0")
    batch = len(ns)
    tY = {n: torch.tensor(v, device=dev) for n, v in Y0.items()}
    tX = [torch.tensor(x, device=dev) for x in X]
    # A_k of entry b: column-major n x n (mats hold A as (row, col))
    tA = [torch.tensor(np.ascontiguousarray(mats[b][k].T), device=dev) for b in range(batch) for k in range(d)]
    table = torch.tensor([t.data_ptr() for t in tA], dtype=torch.int64, device=dev)
    pX = torch.tensor([t.data_ptr() for t in tX], dtype=torch.int64, device=dev)
    pY = torch.tensor([tY[n][yg[b]].data_ptr() for b, n in enumerate(ns)], dtype=torch.int64, device=dev)
    plan = km.GroupPlan(d, list(ns), pY)
    outs = []
    for _ in range(runs):
        for n in tY:
            tY[n].copy_(torch.tensor(Y0[n], device=dev))
        km.kronmult_grouped(plan, table, pX, pY, opt=km.options(exact=exact, **MODES[mode]))
        torch.cuda.synchronize()
        outs.append({n: t.cpu().numpy() for n, t in tY.items()})
    info = plan.info()
    plan.close()
    return outs, info


@pytest.fixture(scope="module")
def km():
    import kronmult_amd
    kronmult_amd.lib()
    return kronmult_amd


@pytest.mark.parametrize("d,sizes", [(4, [2, 3, 4, 5, 6, 7, 8]), (3, [1, 2, 5, 8]), (2, [3, 8]),
                                     (5, [2, 3, 6]), (6, [2, 3, 4]), (1, [4, 7])])
@pytest.mark.parametrize("mode", list(MODES))
def test_grouped_vs_oracle(km, d, sizes, mode):
    ns, yg, mats, X, Y0 = _make(d * 10 + len(sizes), d, sizes, 300, 5)
    want = _oracle(d, ns, yg, mats, X, Y0)
    outs, (nout, ngroups, nmax) = _gpu(km, d, ns, yg, mats, X, Y0, runs=2, mode=mode)
    assert nmax == max(ns)
    assert nout == len({(n, g) for n, g in zip(ns, yg)})
    for n in want:
        assert rel_err(outs[0][n], want[n]) <= RTOL_FP64, n
        for r in outs[1:]:
            assert rel_err(r[n], want[n]) <= RTOL_FP64, n
        if mode != "atomic":
            # owned outputs, fixed order: bitwise run to run
            assert np.array_equal(outs[0][n], outs[1][n])


def test_grouped_persistent_long_runs_both_paths(km):
    """One launch over long output runs of both kinds of entry: d = 4 with
    n = 8 (a whole workgroup per entry, next X copied in behind it) and n = 3
    (entries dealt out to the wavefronts, partial sums combined in LDS), 40
    entries per output, so every wave takes several entries of one run;
    bitwise equal from call to call."""
    d, sizes = 4, [3, 8]
    ns, yg, mats, X, Y0 = _make(91, d, sizes, 160, 2)
    want = _oracle(d, ns, yg, mats, X, Y0)
    outs, (nout, _, _) = _gpu(km, d, ns, yg, mats, X, Y0, runs=2, mode="persistent")
    assert nout == 4
    for n in want:
        assert rel_err(outs[0][n], want[n]) <= RTOL_FP64, n
        assert np.array_equal(outs[0][n], outs[1][n])


@pytest.mark.parametrize("fp32", [False, True])
def test_grouped_persistent_split_runs(km, fp32):
    """One output per size, about 180 entries of n = 7 into one: the one-launch
    mode cuts the run into parts over several workgroups and a second launch
    adds the parts in order, so the result is still bitwise equal from call to
    call; float data with double accumulation as well."""
    import torch
    d, sizes = 4, [2, 7]
    ns, yg, mats, X, Y0 = _make(313, d, sizes, 360, 1)
    want = _oracle(d, ns, yg, mats, X, Y0)
    if not fp32:
        outs, (nout, ngroups, _) = _gpu(km, d, ns, yg, mats, X, Y0, runs=2, mode="persistent")
        assert nout == 2 and ngroups > 2
        for n in want:
            assert rel_err(outs[0][n], want[n]) <= RTOL_FP64, n
            assert np.array_equal(outs[0][n], outs[1][n])
        return
    dev = torch.device("cuda:0")
    tY = {n: torch.tensor(v, device=dev, dtype=torch.float32) for n, v in Y0.items()}
    tX = [torch.tensor(x, device=dev, dtype=torch.float32) for x in X]
    tA = [torch.tensor(np.ascontiguousarray(mats[b][k].T), device=dev, dtype=torch.float32)
          for b in range(len(ns)) for k in range(d)]
    table = torch.tensor([t.data_ptr() for t in tA], dtype=torch.int64, device=dev)
    pX = torch.tensor([t.data_ptr() for t in tX], dtype=torch.int64, device=dev)
    pY = torch.tensor([tY[n][yg[b]].data_ptr() for b, n in enumerate(ns)], dtype=torch.int64, device=dev)
    plan = km.GroupPlan(d, list(ns), pY, elem_bytes=4)
    try:
        km.kronmult_grouped(plan, table, pX, pY, opt=km.options(persistent=True), dtype="f32")
        torch.cuda.synchronize()
    finally:
        plan.close()
    for n in want:
        assert rel_err(tY[n].cpu().numpy().astype(np.float64), want[n]) <= 1e-5, n


def test_grouped_large_bins_vs_oracle(km):
    """Bins larger than one resident wave of workgroups, so the even split
    cuts output runs and the atomic flush of a shared run is exercised."""
    d, sizes = 3, [2, 4]
    ns, yg, mats, X, Y0 = _make(77, d, sizes, 6000, 7)
    want = _oracle(d, ns, yg, mats, X, Y0)
    for mode in MODES:
        outs, _ = _gpu(km, d, ns, yg, mats, X, Y0, mode=mode)
        for n in want:
            assert rel_err(outs[0][n], want[n]) <= RTOL_FP64, (mode, n)


def test_grouped_exact_single_entry_outputs_bitwise(km):
    d = 4
    rng = np.random.default_rng(4)
    ns = np.array([2, 8, 5, 3, 7, 6, 4, 8])
    yg = np.arange(len(ns))
    mats = [[rng.uniform(-1, 1, (n, n)) for _ in range(d)] for n in ns]
    X = [rng.uniform(-1, 1, n ** d) for n in ns]
    Y0 = {}
    for b, n in enumerate(ns):
        Y0.setdefault(n, np.zeros((len(ns), n ** d)))
    want = _oracle(d, ns, yg, mats, X, Y0)
    for mode in MODES:
        outs, _ = _gpu(km, d, ns, yg, mats, X, Y0, exact=True, mode=mode)
        for n in want:
            assert np.array_equal(outs[0][n], want[n])


def test_grouped_rejects_oversize(km):
    import torch
    dev = torch.device("cuda:0")
    Y = torch.zeros(4096 * 2, dtype=torch.float64, device=dev)
    pY = torch.tensor([Y.data_ptr()] * 2, dtype=torch.int64, device=dev)
    with pytest.raises(km.KronmultError) as ei:
        km.GroupPlan(6, [4, 5], pY)
    assert ei.value.status == km.ERROR_NOT_SUPPORTED


@pytest.mark.parametrize("mode", ["atomic", "deterministic", "persistent"])
def test_grouped_graph_capture_replays(km, mode):
    """The per-size launches fork onto the plan's streams and join back into
    the caller's stream; captured in a HIP graph (torch.cuda.graph) and
    replayed with eager work in between, every replay matches the eager call
    (bitwise with owned outputs) and the oracle."""
    import torch
    dev = torch.device("cuda:0")
    d, sizes = 3, [2, 4, 6, 8, 5]
    ns, yg, mats, X, Y0 = _make(31, d, sizes, 400, 6)
    want = _oracle(d, ns, yg, mats, X, Y0)
    tY0 = {n: torch.tensor(v, device=dev) for n, v in Y0.items()}
    tY = {n: t.clone() for n, t in tY0.items()}
    tX = [torch.tensor(x, device=dev) for x in X]
    tA = [torch.tensor(np.ascontiguousarray(mats[b][k].T), device=dev) for b in range(len(ns)) for k in range(d)]
    table = torch.tensor([t.data_ptr() for t in tA], dtype=torch.int64, device=dev)
    pX = torch.tensor([t.data_ptr() for t in tX], dtype=torch.int64, device=dev)
    pY = torch.tensor([tY[n][yg[b]].data_ptr() for b, n in enumerate(ns)], dtype=torch.int64, device=dev)
    plan = km.GroupPlan(d, list(ns), pY)
    opt = km.options(**MODES[mode])
    km.kronmult_grouped(plan, table, pX, pY, opt=opt)  # eager
    torch.cuda.synchronize()
    eager = {n: (tY[n] - tY0[n]).clone() for n in tY}
    g = torch.cuda.CUDAGraph()
    with torch.cuda.graph(g):
        km.kronmult_grouped(plan, table, pX, pY, opt=opt)
    for rep in range(4):
        for n in tY:
            tY[n].copy_(tY0[n])
        junk = torch.ones(1 << 20, device=dev, dtype=torch.float64).sum()  # eager work between replays
        g.replay()
        torch.cuda.synchronize()
        assert junk.item() == float(1 << 20)
        for n in tY:
            got = tY[n] - tY0[n]
            if mode != "atomic":
                assert torch.equal(got, eager[n]), (rep, n)
            else:
                assert torch.allclose(got, eager[n], rtol=0, atol=1e-12 * float(eager[n].abs().max())), (rep, n)
            assert rel_err(tY[n].cpu().numpy(), want[n]) <= RTOL_FP64, (rep, n)
    del g
    plan.close()
