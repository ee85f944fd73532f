"""Benchmark: d=6 batched Kronecker mat-vec on one MI355X.

Workload (the plain-FP64 GPU configuration): kronmult6_pbatched, FP64,
d = 6, n = 6 (vectors of 6^6 = 46656 doubles), 4096 entries, every entry with
its own six 6x6 A ~ U[-1,1) and its own X ~ U[-1,1); outputs overlap 4:1,
pY[k] = Y + (k/4)*n^d.  One step = one kronmult6_pbatched call over the whole
batch, inputs resident in HBM.

Printed: one JSON line with GFLOP/s (reference flop convention 2*d*n^(d+1) per
entry, test_kronmult6_batched.cpp:270), the HBM roofline of the kernel
(algorithmic bytes = X + A read once, Y read and written once), and the
reference's CPU path timed on this machine's host cores.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--mode atomic|plan]
                    [--workload d6n6|d3n8|d6n4|d6n8|d6n4-interleaved|grouped|xpool]

--workload selects another configuration for profiling (default d6n6, the one
above): d3n8 = d=3 n=8 x262144 4:1, d6n4 = d=6 n=4 x65536 16:1, d6n8 = d=6
n=8 x1024 2:1 (workspace kernels); d6n4-interleaved = d=6 n=4 x512 with
pY[k] = Y + (k mod 64)*n^d (8 entries per output, spread over the batch);
grouped = d=4, 20000 entries with n drawn from 2..8, 285 output groups per n,
each entry adding into a random group of its n, through the grouped entry
point (compared with the deterministic and persistent modes and with one
kronmult4_vbatched call per n); xpool = the pooled sparse-grid apply
through kronmult6_xbatched: d=6, n=4, 1,048,576 entries, 512 pooled 4x4 A per
mode, 65536 X and 65536 Y vectors, X drawn uniformly, 16 entries per output.
"""
from __future__ import annotations

import argparse
import json
import os
import statistics
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0  # MI355X published HBM3E bandwidth
FP64_PEAK_TFLOPS = 78.6  # MI355X published dense FP64 (vector and matrix)
SEED = 20261203
# name -> (d, n, batch, overlap): pY[k] = Y + (k/overlap)*n^d
WORKLOADS = {
    "d6n6": (6, 6, 4096, 4),
    "d3n8": (3, 8, 262144, 4),
    "d6n4": (6, 4, 65536, 16),
    "d6n8": (6, 8, 1024, 2),
    "d6n4-interleaved": (6, 4, 512, 8),
}
# workloads whose outputs interleave: pY[k] = Y + (k mod (batch/overlap))*n^d
INTERLEAVED = {"d6n4-interleaved"}
# grouped: d, entries, sizes, output groups per size
GROUPED = dict(d=4, batch=20000, sizes=(2, 3, 4, 5, 6, 7, 8), groups=285)
GROUPED_CPU_SAMPLE = 20000
# pooled xbatched: entries, A matrices per mode, X vectors, Y vectors, entries per Y
XPOOL = dict(d=6, n=4, batch=1 << 20, pool=512, nx=65536, ny=65536, fan_in=16)
XPOOL_CPU_SAMPLE = 16384


def flops_per_call(d, n, batch):
    return 2.0 * d * n ** (d + 1) * batch


def algorithmic_bytes(d, n, batch, overlap, elem=8):
    s = n ** d
    n_y = batch // overlap
    return elem * (batch * s + batch * d * n * n + 2 * n_y * s)


def parse_traffic(path):
    """HBM bytes per launch from a committed rocprofv3 PMC summary: lines
    'traffic_bytes_per_launch: <int>'."""
    try:
        with open(path) as f:
            for line in f:
                if line.startswith("traffic_bytes_per_launch:"):
                    return int(float(line.split(":", 1)[1].strip()))
    except OSError:
        pass
    return None


def cpu_baseline(A, X, yidx, n_y, overlap, reps=3):
    """The reference's own CPU/OpenMP path (oracle/_ref, default recursive
    build) when it was built, else the oracle restatement; timed on the host
    cores over the full workload (median of `reps`)."""
    import numpy as np

    from oracle import oracle as O
    use_ref = "recursive" if O.ref_available("recursive") else None
    threads = O.num_threads()
    batch, d, n, _ = A.shape
    s = n ** d
    Y = np.zeros((n_y, s))
    times = []
    for _ in range(reps):
        # the reference uses X as scratch: every rep gets fresh copies, and
        # only the library call itself is timed
        if use_ref:
            O.pbatched(A, X, Y, yidx, use_ref=use_ref, timings=times)
        else:
            O.pbatched(A, X, Y, yidx, variant=O.RECURSIVE, timings=times)
    t = statistics.median(times)
    return {
        "value": round(flops_per_call(d, n, batch) / t / 1e9, 3),
        "unit": "GFLOP/s",
        "cores": threads,
        "kind": "reference" if use_ref else "port",
        "sample": f"full workload ({batch} entries, d={d}, n={n}, {overlap}:1 overlap), median of {reps} "
                  f"calls of kronmult{d}_pbatched (default recursive build), OpenMP threads={threads}",
    }


PRECISIONS = {"fp64": 0, "fp64_mfma": 1, "f16x4": 2}


def precision_study(dev, steps=10, warmup=2, nsample=16):
    """The precision-study configuration: d=6, n=8, 1024 entries, outputs
    shared 2:1, A ~ N(0,1)/sqrt(8) with one seeded row per matrix scaled by
    1e3, X ~ U[-1,1], run with plain FP64 stages, FP64 matrix instructions and
    the mixed FP16-slice stages (plan reduction).  Time per call, GFLOP/s and
    the error against the FP64 CPU oracle on `nsample` sampled outputs."""
    import numpy as np
    import torch

    import kronmult_amd as km
    from oracle import oracle as O

    d, n, batch, ov = 6, 8, 1024, 2
    s = n ** d
    g = torch.Generator(device=dev)
    g.manual_seed(SEED + 4)
    A = torch.randn((batch, d, n, n), generator=g, device=dev, dtype=torch.float64) / np.sqrt(8.0)
    rows = torch.randint(0, n, (batch, d), generator=g, device=dev)
    scale = torch.ones((batch, d, n), dtype=torch.float64, device=dev)
    scale.scatter_(2, rows.unsqueeze(2), 1e3)
    A *= scale.unsqueeze(2)  # row i of A_k is A[b, k, :, i]
    X = torch.rand((batch, s), generator=g, device=dev, dtype=torch.float64) * 2 - 1
    Y = torch.zeros((batch // ov, s), device=dev, dtype=torch.float64)
    W = torch.empty((batch, s), device=dev, dtype=torch.float64)
    yidx = np.arange(batch) // ov
    pX = km.pointer_array(X, np.arange(batch))
    pY = km.pointer_array(Y, yidx)
    pW = km.pointer_array(W, np.arange(batch))
    plan = km.Plan(d, n, 8, pY, batch)
    outs = np.sort(np.random.default_rng(SEED).choice(batch // ov, nsample, replace=False))
    ents = np.flatnonzero(np.isin(yidx, outs))
    pos = {int(o): i for i, o in enumerate(outs)}
    it = torch.as_tensor(ents, device=dev)
    want = O.pbatched(A.index_select(0, it).cpu().numpy(), X.index_select(0, it).cpu().numpy(),
                      np.zeros((len(outs), s)), np.array([pos[int(yidx[e])] for e in ents]),
                      variant=O.BACKWARD)
    scale_w = np.abs(want).max(axis=1, keepdims=True)
    res = {"ms": {}, "gflops": {}, "speed_vs_fp64": {}, "rel_err_normwise": {},
           "rel_err_max": {}, "rel_err_median": {}}
    stream = torch.cuda.current_stream()
    for name, prec in PRECISIONS.items():
        opt = km.options(plan=plan, precision=prec)
        Y.zero_()
        km.kronmult_pbatched(d, n, A, pX, pY, pW, batch, stream=stream, opt=opt)
        torch.cuda.synchronize()
        got = Y.index_select(0, torch.as_tensor(outs, device=dev)).cpu().numpy()
        el = (np.abs(got - want) / scale_w).max(axis=1)
        res["rel_err_normwise"][name] = float(np.abs(got - want).max() / np.abs(want).max())
        res["rel_err_max"][name] = float(el.max())
        res["rel_err_median"][name] = float(np.median(el))
        for _ in range(warmup):
            km.kronmult_pbatched(d, n, A, pX, pY, pW, batch, stream=stream, opt=opt)
        ev0 = torch.cuda.Event(enable_timing=True)
        ev1 = torch.cuda.Event(enable_timing=True)
        ev0.record(stream)
        for _ in range(steps):
            km.kronmult_pbatched(d, n, A, pX, pY, pW, batch, stream=stream, opt=opt)
        ev1.record(stream)
        torch.cuda.synchronize()
        ms = ev0.elapsed_time(ev1) / steps
        res["ms"][name] = round(ms, 4)
        res["gflops"][name] = round(flops_per_call(d, n, batch) / (ms * 1e-3) / 1e9, 1)
    for name in PRECISIONS:
        res["speed_vs_fp64"][name] = round(res["ms"]["fp64"] / res["ms"][name], 3)
    plan.close()
    res["workload"] = ("kronmult6_pbatched d=6, n=8, batch=1024, 2:1 outputs, A ~ N(0,1)/sqrt(8) with one "
                       "row per matrix x1e3, X ~ U[-1,1], plan reduction; fp64 = FP64 vector stages, "
                       "fp64_mfma = FP64 matrix instructions, f16x4 = mixed FP16 slices (first pass; "
                       "last two modes and Y in FP64)")
    res["error_sample"] = (f"{len(outs)} outputs ({len(ents)} entries) recomputed by the FP64 CPU oracle; "
                           "rel_err_max/median: over outputs of max|dY|/max|Y|")
    return res


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--mode", choices=["best", "atomic", "plan", "persistent"], default="best",
                    help="output reduction: atomic adds (drop-in), a deterministic plan, or "
                         "time both and report the faster; grouped only: persistent times the "
                         "one-launch call (KM_FLAG_PERSISTENT) alone")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="also run a plain 1 GiB device copy (PMC byte-counter calibration)")
    ap.add_argument("--workload", choices=sorted(WORKLOADS) + ["grouped", "xpool"], default="d6n6",
                    help="configuration (default d6n6: d=6, n=6, 4096 entries, 4:1)")
    ap.add_argument("--precision", choices=sorted(PRECISIONS), default="fp64",
                    help="stage arithmetic of the timed calls (fp64_mfma and f16x4: d6n8 only)")
    ap.add_argument("--no-precision-study", action="store_true",
                    help="skip the d=6, n=8 precision comparison added to the default line")
    ap.add_argument("--chunks", type=int, default=8,
                    help="xpool: entries split into this many chunks, each with its own plan, "
                         "captured in one HIP graph and replayed (0 = skip)")
    ap.add_argument("--xpool-variant", choices=["uniform", "x-const", "x-const-a1"], default="uniform",
                    help="xpool counter runs: every entry reads X[0] (x-const), and also one A per "
                         "mode (x-const-a1); the L2 requests and misses of the A pool are the "
                         "difference of the two")
    args = ap.parse_args()

    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py: this build measures one GPU only (the batched kernel on one "
              "GPU's share); --gpus must be 1", file=sys.stderr)
        sys.exit(2)
    if args.mode == "persistent" and args.workload != "grouped":
        print("bench.py: --mode persistent applies to --workload grouped only", file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import torch

    import kronmult_amd as km

    dev = torch.device("cuda:0")
    torch.cuda.set_device(dev)
    if args.workload == "xpool":
        bench_xpool(args, dev)
        return
    if args.workload == "grouped":
        bench_grouped(args, dev)
        return
    d, n, batch, overlap = WORKLOADS[args.workload]
    s = n ** d
    n_y = batch // overlap

    g = torch.Generator(device=dev)
    g.manual_seed(SEED)
    A = torch.rand((batch, d, n, n), generator=g, device=dev, dtype=torch.float64) * 2 - 1
    X = torch.rand((batch, s), generator=g, device=dev, dtype=torch.float64) * 2 - 1
    Y = torch.zeros((n_y, s), device=dev, dtype=torch.float64)
    yidx = (np.arange(batch) % n_y) if args.workload in INTERLEAVED else (np.arange(batch) // overlap)
    pX = km.pointer_array(X, np.arange(batch))
    pY = km.pointer_array(Y, yidx)
    fam, nt = km.query_kernel(d, n, 8, None)
    # the workspace kernels (two-pass, generic) use one n^d slot per entry
    W = torch.empty((batch, s), device=dev, dtype=torch.float64) if fam in (2, 3, 7) else None
    pW = km.pointer_array(W, np.arange(batch)) if W is not None else None

    # a plan is set-up work (once per output pointer array), outside the
    # timed region; every timed call does the full computation
    plan = km.Plan(d, n, 8, pY, batch) if args.mode in ("best", "plan") else None
    prec = PRECISIONS[args.precision]
    modes = {"atomic": km.options(precision=prec),
             "plan": km.options(plan=plan, precision=prec) if plan else None}
    todo = ["atomic", "plan"] if args.mode == "best" else [args.mode]

    kernel_name = {0: "kron_lds", 1: "kron_slab", 2: "kron_global", 3: "kron_sub+kron_pair",
                   4: "kron_reg", 6: "kron_half", 7: "kron_xcd"}[fam]

    stream = torch.cuda.current_stream()

    def timed(opt):
        def step():
            km.kronmult_pbatched(d, n, A, pX, pY, pW, batch, stream=stream, opt=opt)

        for _ in range(args.warmup):
            step()
        torch.cuda.synchronize()
        ev0 = torch.cuda.Event(enable_timing=True)
        ev1 = torch.cuda.Event(enable_timing=True)
        t0 = time.perf_counter()
        ev0.record(stream)
        for _ in range(args.steps):
            step()
        ev1.record(stream)
        torch.cuda.synchronize()
        return ev0.elapsed_time(ev1) / args.steps, time.perf_counter() - t0

    results = {m: timed(modes[m]) for m in todo}
    best = min(results, key=lambda m: results[m][0])
    ms_per_step, wall = results[best]

    fl = flops_per_call(d, n, batch)
    by = algorithmic_bytes(d, n, batch, overlap)
    gflops = fl / (ms_per_step * 1e-3) / 1e9
    gbs = by / (ms_per_step * 1e-3) / 1e9

    if args.copy_calibration:
        a = torch.empty(1 << 27, dtype=torch.float64, device=dev)  # 1 GiB
        b = torch.empty_like(a)
        a.fill_(1.0)
        for _ in range(3):
            b.copy_(a)
        torch.cuda.synchronize()

    traffic = parse_traffic(os.path.join(ROOT, "profiles", f"{kernel_name}_d{d}n{n}_b{batch}.txt"))

    cpu = None
    if not args.no_cpu:
        # same seeded inputs, host copies
        cpu = cpu_baseline(A.cpu().numpy(), X.cpu().numpy(), yidx, n_y, overlap)
    study = None
    if args.workload == "d6n6" and not args.no_precision_study:
        del A, X, Y, W
        torch.cuda.empty_cache()
        study = precision_study(dev)

    out = {
        "metric": "d=6 batched kronmult GFLOP/s + % HBM roofline; mixed vs FP64 at rel err <=1e-10",
        "value": round(gflops, 2),
        "unit": "GFLOP/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 5),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": {"fp64": "f64", "fp64_mfma": "f64", "f16x4": "f16 slices, f32 sums, f64 recombination"}[
            args.precision],
        "data": "synthetic (seeded U[-1,1) A and X on device)",
        "config": {
            "workload": f"kronmult{d}_pbatched FP64, d={d}, n={n}, batch={batch}, {overlap}:1 output "
                        + (f"overlap (pY[k] = Y + (k mod {n_y})*n^d)" if args.workload in INTERLEAVED
                           else f"overlap (pY[k] = Y + (k/{overlap})*n^d)"),
            "d": d, "n": n, "batch": batch, "overlap": overlap,
            "reduction": {"plan": "plan (deterministic output-sorted, no atomics)",
                          "atomic": "atomic adds (drop-in)"}[best],
            "ms_per_step_by_reduction": {m: round(v[0], 5) for m, v in results.items()},
            "kernel": f"{kernel_name} ({nt} threads/workgroup)",
            "parallelism": "single GPU",
        },
        "roofline": {
            "bound": "hbm",
            "achieved": round(gbs, 1),
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": round(gbs / HBM_PEAK_GBS, 4),
            "traffic": traffic,
        },
        "cpu_baseline": cpu,
        "wall_s": round(wall, 4),
    }
    if args.precision != "fp64":
        out["config"]["precision"] = args.precision
    if study is not None:
        out["precision_study"] = study
    print(json.dumps(out))
    if plan is not None:
        plan.close()


def xpool_cpu_baseline(pools, aidx, X, xidx, yidx, reps=3):
    """The reference's CPU kronmult6_xbatched (oracle/_ref, default recursive
    build) when it was built, else the oracle restatement, on the first
    XPOOL_CPU_SAMPLE entries of the pooled workload (their X and Y vectors
    gathered into compact host arrays); median of `reps` calls.  ``X`` is the
    device X pool; only the sampled rows are copied to the host."""
    import numpy as np
    import torch

    from oracle import oracle as O
    use_ref = "recursive" if O.ref_available("recursive") else None
    threads = O.num_threads()
    m = XPOOL_CPU_SAMPLE
    d, n = XPOOL["d"], XPOOL["n"]
    xs, xmap = np.unique(xidx[:m], return_inverse=True)
    ys, ymap = np.unique(yidx[:m], return_inverse=True)
    Xs = X[torch.as_tensor(xs, device=X.device)].cpu().numpy()
    Ys = np.zeros((len(ys), X.shape[1]))
    times = []
    for _ in range(reps):
        O.xbatched(pools, aidx[:m], Xs, xmap, Ys, ymap, use_ref=use_ref,
                   variant=O.RECURSIVE, timings=times)
    t = statistics.median(times)
    return {
        "value": round(flops_per_call(d, n, m) / t / 1e9, 3),
        "unit": "GFLOP/s",
        "cores": threads,
        "kind": "reference" if use_ref else "port",
        "sample": f"first {m} of {XPOOL['batch']} entries ({len(xs)} distinct X, {len(ys)} distinct Y), "
                  f"median of {reps} calls of kronmult{d}_xbatched (default recursive build), "
                  f"OpenMP threads={threads}",
    }


def bench_xpool(args, dev):
    """Pooled sparse-grid apply through kronmult6_xbatched (lda = n).

    Entry e uses A_k = pool_k[aidx[e, k]], reads X[xidx[e]] (uniform over the
    X vectors) and accumulates into Y[yidx[e]], yidx = a seeded permutation of
    the entries // fan_in, so every output receives fan_in entries.  The
    compulsory bytes (every X and Y vector once) put this workload above the
    FP64 ridge (32 flop/B against 78.6 TF / 8 TB/s = 9.8), so its roofline is
    the FP64 rate; the bytes with X read once per entry are reported too."""
    import numpy as np
    import torch

    import kronmult_amd as km

    c = XPOOL
    d, n, batch = c["d"], c["n"], c["batch"]
    s = n ** d
    g = torch.Generator(device=dev)
    g.manual_seed(SEED)
    pools = [torch.rand((c["pool"], n, n), generator=g, device=dev, dtype=torch.float64) * 2 - 1
             for _ in range(d)]
    X = torch.rand((c["nx"], s), generator=g, device=dev, dtype=torch.float64) * 2 - 1
    Y = torch.zeros((c["ny"], s), device=dev, dtype=torch.float64)
    rng = np.random.default_rng(SEED)
    aidx = rng.integers(0, c["pool"], (batch, d))
    xidx = rng.integers(0, c["nx"], batch)
    yidx = rng.permutation(batch) // c["fan_in"]
    if args.xpool_variant != "uniform":
        xidx[:] = 0
    if args.xpool_variant == "x-const-a1":
        aidx[:] = 0
    table = km.matrix_pointer_table(pools, aidx)
    pX = km.pointer_array(X, xidx)
    pY = km.pointer_array(Y, yidx)
    plan = km.Plan(d, n, 8, pY, batch) if args.mode in ("best", "plan") else None
    prec = PRECISIONS[args.precision]
    modes = {"atomic": km.options(precision=prec),
             "plan": km.options(plan=plan, precision=prec) if plan else None}
    todo = ["atomic", "plan"] if args.mode == "best" else [args.mode]
    fam, nt = km.query_kernel(d, n, 8, None)
    kernel_name = {0: "kron_lds", 4: "kron_reg"}.get(fam, str(fam))
    stream = torch.cuda.current_stream()

    def timed(opt):
        def step():
            km.kronmult_xbatched(d, n, table, n, pX, pY, None, batch, stream=stream, opt=opt)

        for _ in range(args.warmup):
            step()
        torch.cuda.synchronize()
        ev0 = torch.cuda.Event(enable_timing=True)
        ev1 = torch.cuda.Event(enable_timing=True)
        t0 = time.perf_counter()
        ev0.record(stream)
        for _ in range(args.steps):
            step()
        ev1.record(stream)
        torch.cuda.synchronize()
        return ev0.elapsed_time(ev1) / args.steps, time.perf_counter() - t0

    results = {m: timed(modes[m]) for m in todo}
    best = min(results, key=lambda m: results[m][0])
    ms_per_step, wall = results[best]
    fl = flops_per_call(d, n, batch)
    tflops = fl / (ms_per_step * 1e-3) / 1e12

    # the same work in chunks, each with its own output-ordered plan,
    # captured once in a HIP graph and replayed.  Chunks are ranges of
    # outputs (the entries grouped by output, as an operator apply groups
    # its elements), so every output is read and written in one chunk only.
    graph = None
    if args.chunks > 0:
        k = args.chunks
        perm = np.argsort(yidx, kind="stable")
        ys = yidx[perm]
        bounds = [int(np.searchsorted(ys, c["ny"] * i // k)) for i in range(k)] + [batch]
        ctable = km.matrix_pointer_table(pools, aidx[perm])
        cpX = km.pointer_array(X, xidx[perm])
        cpY = km.pointer_array(Y, ys)
        cplans = [km.Plan(d, n, 8, cpY[bounds[i]:bounds[i + 1]], bounds[i + 1] - bounds[i])
                  for i in range(k)]

        def chunked():
            for i in range(k):
                b0, b1 = bounds[i], bounds[i + 1]
                km.kronmult_xbatched(d, n, ctable[d * b0:d * b1], n, cpX[b0:b1], cpY[b0:b1], None, b1 - b0,
                                     opt=km.options(plan=cplans[i], precision=prec))

        chunked()  # eager warm-up
        torch.cuda.synchronize()
        cg = torch.cuda.CUDAGraph()
        with torch.cuda.graph(cg):
            chunked()
        for _ in range(2):
            cg.replay()
        torch.cuda.synchronize()
        ev0 = torch.cuda.Event(enable_timing=True)
        ev1 = torch.cuda.Event(enable_timing=True)
        ev0.record(stream)
        for _ in range(args.steps):
            cg.replay()
        ev1.record(stream)
        torch.cuda.synchronize()
        gms = ev0.elapsed_time(ev1) / args.steps
        graph = {"chunks": k, "ms_per_replay": round(gms, 5),
                 "gflops": round(fl / (gms * 1e-3) / 1e9, 2),
                 "note": "chunks = output ranges, each: plan check + kron_reg launch; all captured "
                         "in one HIP graph"}
        del cg
        for cp in cplans:
            cp.close()
    by_compulsory = 8.0 * (c["nx"] * s + c["pool"] * d * n * n + 2 * c["ny"] * s)
    by_streamed = 8.0 * (batch * s + c["pool"] * d * n * n + 2 * c["ny"] * s)

    if args.copy_calibration:
        a = torch.empty(1 << 27, dtype=torch.float64, device=dev)  # 1 GiB
        b = torch.empty_like(a)
        a.fill_(1.0)
        for _ in range(3):
            b.copy_(a)
        torch.cuda.synchronize()

    cpu = None
    if not args.no_cpu:
        cpu = xpool_cpu_baseline([p.cpu().numpy() for p in pools], aidx, X, xidx, yidx)
    traffic = parse_traffic(os.path.join(ROOT, "profiles", f"{kernel_name}_xpool_d{d}n{n}_b{batch}.txt"))
    out = {
        "metric": "d=6 batched kronmult GFLOP/s + % HBM roofline; mixed vs FP64 at rel err <=1e-10",
        "value": round(fl / (ms_per_step * 1e-3) / 1e9, 2),
        "unit": "GFLOP/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 5),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (seeded U[-1,1) A pools and X on device, seeded indices)",
        "config": {
            "workload": f"kronmult{d}_xbatched FP64 pooled, d={d}, n={n}, batch={batch}, {c['pool']} A per "
                        f"mode, {c['nx']} X (uniform), {c['ny']} Y ({c['fan_in']} entries each)",
            "d": d, "n": n, "batch": batch,
            "reduction": {"plan": "plan (deterministic output-sorted, no atomics)",
                          "atomic": "atomic adds (drop-in)"}[best],
            "ms_per_step_by_reduction": {m: round(v[0], 5) for m, v in results.items()},
            "kernel": f"{kernel_name} ({nt} threads/workgroup)",
            "hbm_gbs_compulsory": round(by_compulsory / (ms_per_step * 1e-3) / 1e9, 1),
            "hbm_gbs_x_per_entry": round(by_streamed / (ms_per_step * 1e-3) / 1e9, 1),
            "graph_chunks": graph,
            "xpool_variant": args.xpool_variant,
            "parallelism": "single GPU",
        },
        "roofline": {
            "bound": "mfma",
            "achieved": round(tflops, 2),
            "peak": FP64_PEAK_TFLOPS,
            "unit": "TFLOP/s",
            "frac": round(tflops / FP64_PEAK_TFLOPS, 4),
            "traffic": traffic,
        },
        "cpu_baseline": cpu,
        "wall_s": round(wall, 4),
    }
    print(json.dumps(out))
    if plan is not None:
        plan.close()


def bench_grouped(args, dev):
    """Grouped variable-size batch through kronmult_grouped (one call per
    step: one on-chip launch per size bin).  Entry e has its own n_e drawn
    uniformly from GROUPED['sizes'], its own four n_e x n_e A ~ U[-1,1) and
    X ~ U[-1,1); the outputs are GROUPED['groups'] vectors per size and each
    entry adds into a random one of its size.  The same call with
    KM_FLAG_DETERMINISTIC and with KM_FLAG_PERSISTENT, and one
    kronmult4_vbatched call per size, are timed beside it."""
    import numpy as np
    import torch

    import kronmult_amd as km

    c = GROUPED
    d, batch, sizes, groups = c["d"], c["batch"], c["sizes"], c["groups"]
    rng = np.random.default_rng(SEED)
    ns = rng.choice(sizes, batch)
    ys = rng.integers(0, groups, batch)
    g = torch.Generator(device=dev)
    g.manual_seed(SEED)
    idx = {n: np.flatnonzero(ns == n) for n in sizes}
    X = {n: torch.rand((len(idx[n]), n ** d), generator=g, device=dev, dtype=torch.float64) * 2 - 1 for n in sizes}
    A = {n: torch.rand((len(idx[n]), d, n, n), generator=g, device=dev, dtype=torch.float64) * 2 - 1 for n in sizes}
    Y = {n: torch.zeros((groups, n ** d), device=dev, dtype=torch.float64) for n in sizes}
    # entry e is the slot-th entry of its size
    slot = np.empty(batch, dtype=np.int64)
    for n in sizes:
        slot[idx[n]] = np.arange(len(idx[n]))
    xaddr = np.array([X[n].data_ptr() + slot[e] * X[n].stride(0) * 8 for e, n in enumerate(ns)], dtype=np.int64)
    yaddr = np.array([Y[n].data_ptr() + ys[e] * Y[n].stride(0) * 8 for e, n in enumerate(ns)], dtype=np.int64)
    aaddr = np.array([[A[n].data_ptr() + (slot[e] * A[n].stride(0) + k * A[n].stride(1)) * 8 for k in range(d)]
                      for e, n in enumerate(ns)], dtype=np.int64).reshape(-1)
    pX = torch.tensor(xaddr, device=dev)
    pY = torch.tensor(yaddr, device=dev)
    table = torch.tensor(aaddr, device=dev)
    plan = km.GroupPlan(d, [int(v) for v in ns], pY)
    # the same entries as one kronmult4_vbatched call per size
    per_n = {n: (torch.tensor(xaddr[idx[n]], device=dev), torch.tensor(yaddr[idx[n]], device=dev),
                 torch.tensor(aaddr.reshape(batch, d)[idx[n]].reshape(-1), device=dev)) for n in sizes}
    stream = torch.cuda.current_stream()
    calls = {
        "grouped": lambda: km.kronmult_grouped(plan, table, pX, pY, stream=stream),
        "grouped_deterministic": lambda: km.kronmult_grouped(plan, table, pX, pY, stream=stream,
                                                             opt=km.options(deterministic=True)),
        "grouped_persistent": lambda: km.kronmult_grouped(plan, table, pX, pY, stream=stream,
                                                          opt=km.options(persistent=True)),
        "vbatched_per_size": lambda: [km.kronmult_vbatched([n] * d, [n] * d, per_n[n][2], per_n[n][0], per_n[n][1],
                                                           None, 0, len(idx[n]), stream=stream) for n in sizes],
    }

    def timed(fn):
        for _ in range(args.warmup):
            fn()
        torch.cuda.synchronize()
        ev0 = torch.cuda.Event(enable_timing=True)
        ev1 = torch.cuda.Event(enable_timing=True)
        t0 = time.perf_counter()
        ev0.record(stream)
        for _ in range(args.steps):
            fn()
        ev1.record(stream)
        torch.cuda.synchronize()
        return ev0.elapsed_time(ev1) / args.steps, time.perf_counter() - t0

    # --mode atomic / plan time the default or the deterministic call alone
    # (for profiling); the default mode times all four
    only = {"atomic": "grouped", "plan": "grouped_deterministic",
            "persistent": "grouped_persistent"}.get(args.mode)
    results = {name: timed(fn) for name, fn in calls.items() if only in (None, name)}
    ms_per_step, wall = results[only or "grouped"]

    if args.copy_calibration:
        a = torch.empty(1 << 27, dtype=torch.float64, device=dev)  # 1 GiB
        b = torch.empty_like(a)
        a.fill_(1.0)
        for _ in range(3):
            b.copy_(a)
        torch.cuda.synchronize()
    fl = float(sum(2.0 * d * int(n) ** (d + 1) for n in ns))
    by = 8.0 * (sum(int(n) ** d + d * int(n) ** 2 for n in ns) + 2 * groups * sum(n ** d for n in sizes))
    gbs = by / (ms_per_step * 1e-3) / 1e9

    cpu = None
    if not args.no_cpu:
        # the reference's CPU kronmult4_pbatched, one call per size, on the
        # first GROUPED_CPU_SAMPLE entries
        from oracle import oracle as O
        use_ref = "recursive" if O.ref_available("recursive") else None
        m = GROUPED_CPU_SAMPLE
        times = []
        for n in sizes:
            sel = idx[n][idx[n] < m]
            if len(sel) == 0:
                continue
            s_ = slot[sel]
            An = A[n][torch.as_tensor(s_, device=dev)].cpu().numpy()
            Xn = X[n][torch.as_tensor(s_, device=dev)].cpu().numpy()
            Yn = np.zeros((groups, n ** d))
            t = []
            for _ in range(3):
                if use_ref:
                    O.pbatched(An, Xn, Yn, ys[sel], use_ref=use_ref, timings=t)
                else:
                    O.pbatched(An, Xn, Yn, ys[sel], variant=O.RECURSIVE, timings=t)
            times.append(statistics.median(t))
        fl_s = float(sum(2.0 * d * int(n) ** (d + 1) for n in ns[:m]))
        cpu = {
            "value": round(fl_s / sum(times) / 1e9, 3),
            "unit": "GFLOP/s",
            "cores": O.num_threads(),
            "kind": "reference" if use_ref else "port",
            "sample": f"first {m} of {batch} entries, one kronmult{d}_pbatched call per size "
                      f"(default recursive build), median of 3 per size, OpenMP threads={O.num_threads()}",
        }
    kind = "kron_runs_" if only == "grouped_persistent" else ""
    traffic = parse_traffic(os.path.join(ROOT, "profiles", f"grouped_{kind}d{d}_b{batch}.txt"))
    out = {
        "metric": "d=6 batched kronmult GFLOP/s + % HBM roofline; mixed vs FP64 at rel err <=1e-10",
        "value": round(fl / (ms_per_step * 1e-3) / 1e9, 2),
        "unit": "GFLOP/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 5),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (seeded sizes, groups, U[-1,1) A and X on device)",
        "config": {
            "workload": f"kronmult_grouped FP64, d={d}, {batch} entries, n in {list(sizes)}, {groups} output "
                        f"groups per n, each entry into a random group of its n",
            "d": d, "batch": batch,
            "reduction": {"grouped_deterministic": "owned outputs (KM_FLAG_DETERMINISTIC)",
                          "grouped_persistent": "owned outputs, one launch (KM_FLAG_PERSISTENT)"}.get(
                              only, "atomic adds (per-size launches, split evenly)"),
            "ms_per_step_by_call": {k: round(v[0], 5) for k, v in results.items()},
            "kernel": ("kron_runs (one launch; a second short one combines split runs)"
                       if only == "grouped_persistent" else "one kron_reg / kron_lds launch per size bin"),
            "parallelism": "single GPU",
        },
        "roofline": {
            "bound": "hbm",
            "achieved": round(gbs, 1),
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": round(gbs / HBM_PEAK_GBS, 4),
            "traffic": traffic,
        },
        "cpu_baseline": cpu,
        "wall_s": round(wall, 4),
    }
    print(json.dumps(out))
    plan.close()


if __name__ == "__main__":
    main()
