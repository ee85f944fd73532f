"""Write the golden vectors under tests/golden/ from the reference's OWN CPU
path (oracle/_ref/libkronref_{recursive,alg993}.so, compiled from the
reference headers by oracle/Makefile; see oracle/ref_driver.cpp).

Run where the reference sources are present:
    make -C oracle all ref && python tests/golden/make_golden.py

Each .npz holds inputs and the reference's outputs for one entry point; the
CPU test suite checks the oracle restatement against them bitwise.  OpenMP is
set to one thread so that overlapping outputs receive their atomic adds in
entry order, which makes the reference's result deterministic.

Inputs: the reference tests' closed-form values
(test_kronmult6_xbatched.cpp:189-225, lda = n + 3 as at :65), seeded
U[-1,1) values, and (for the float cases) the same in single precision.
"""
from __future__ import annotations

import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.dirname(HERE))

from oracle import oracle as O  # noqa: E402
from _helpers import reference_inputs  # noqa: E402

BUILDS = ("recursive", "alg993")


def pbatched_cases():
    out = {}
    rng = np.random.default_rng(2026)
    for d in range(1, 7):
        for n in (1, 2, 3, 4, 5):
            if n ** d > 1024 and not (d == 6 and n == 4):
                continue
            batch = 5 if n ** d <= 256 else 2
            A = rng.uniform(-1, 1, (batch, d, n, n))
            X = rng.uniform(-1, 1, (batch, n ** d))
            nY = 2
            Y0 = rng.uniform(-1, 1, (nY, n ** d))
            yidx = np.arange(batch) % nY
            key = f"d{d}_n{n}"
            out[key + "_A"] = A
            out[key + "_X"] = X
            out[key + "_Y0"] = Y0
            out[key + "_yidx"] = yidx
            for b in BUILDS:
                out[key + f"_Y_{b}"] = O.pbatched(A, X, Y0, yidx, use_ref=b)
    return out


def batched_cases():
    out = {}
    rng = np.random.default_rng(7)
    for d, n in ((1, 6), (2, 5), (3, 4), (4, 3), (5, 2), (6, 3)):
        batch = 3
        A = rng.uniform(-1, 1, (batch, d, n, n))
        X = rng.uniform(-1, 1, (batch, n ** d))
        Y0 = rng.uniform(-1, 1, (batch, n ** d))
        key = f"d{d}_n{n}"
        out[key + "_A"] = A
        out[key + "_X"] = X
        out[key + "_Y0"] = Y0
        for b in BUILDS:
            out[key + f"_Y_{b}"] = O.batched(A, X, Y0, use_ref=b)
    return out


def xbatched_cases():
    """The reference test's own setup: closed-form inputs, lda = n + 3, every
    output pointer equal (use_overlap_in_Y, test_kronmult6_xbatched.cpp:
    255-262)."""
    out = {}
    for d in range(1, 7):
        for n in (2, 3, 4):
            if n ** d > 1024:
                continue
            batch = 6
            lda = n + 3
            A, X = reference_inputs(d, n, batch, lda)
            aidx = np.stack([np.arange(batch)] * d, axis=1)
            Y0 = np.zeros((1, n ** d))
            yidx = np.zeros(batch, dtype=np.int64)
            key = f"d{d}_n{n}"
            out[key + "_A"] = A
            out[key + "_X"] = X
            for b in BUILDS:
                out[key + f"_Y_{b}"] = O.xbatched([A[:, k] for k in range(d)], aidx, X, np.arange(batch), Y0,
                                                  yidx, use_ref=b)
    return out


def vbatched_cases():
    """Rectangular shapes.  Shapes are chosen so that the Alg. 993 build's
    stage sizes fit the reference's own workspace formula (the forward order
    is chosen, or backward needs no more room than forward), since the
    reference sizes its slots for the forward order only
    (kronmult_vbatched.hpp:114-130)."""
    out = {}
    rng = np.random.default_rng(31)
    shapes = [((3,), (5,)), ((2, 3), (4, 2)), ((1, 2, 3), (3, 2, 1)), ((2, 2, 2, 2), (3, 1, 2, 2)),
              ((2, 3, 2, 1, 2), (1, 2, 2, 3, 2)), ((2,) * 6, (3,) * 6), ((3,) * 6, (2,) * 6)]
    for si, (m, n) in enumerate(shapes):
        d = len(m)
        batch = 3
        mats = [[rng.uniform(-1, 1, (n[k], m[k])) for k in range(d)] for _ in range(batch)]
        X = rng.uniform(-1, 1, (batch, int(np.prod(n))))
        Y0 = rng.uniform(-1, 1, (batch, int(np.prod(m))))
        yidx = np.arange(batch)
        key = f"s{si}"
        out[key + "_m"] = np.array(m)
        out[key + "_n"] = np.array(n)
        for k in range(d):
            out[key + f"_A{k}"] = np.stack([mats[b][k] for b in range(batch)])
        out[key + "_X"] = X
        out[key + "_Y0"] = Y0
        for b in BUILDS:
            # generous capacity: one slot pair per entry
            out[key + f"_Y_{b}"] = O.vbatched(m, n, mats, X, Y0, yidx, use_ref=b)
    return out


def float_cases():
    out = {}
    rng = np.random.default_rng(5)
    for d, n in ((2, 4), (3, 3), (4, 4), (6, 2), (6, 3)):
        batch = 4
        A = rng.uniform(-1, 1, (batch, d, n, n)).astype(np.float32)
        X = rng.uniform(-1, 1, (batch, n ** d)).astype(np.float32)
        Y0 = rng.uniform(-1, 1, (2, n ** d)).astype(np.float32)
        yidx = np.arange(batch) % 2
        key = f"d{d}_n{n}"
        out[key + "_A"] = A
        out[key + "_X"] = X
        out[key + "_Y0"] = Y0
        out[key + "_yidx"] = yidx
        for b in BUILDS:
            out[key + f"_Y_{b}"] = O.pbatched(A, X, Y0, yidx, use_ref=b)
    return out


def flops_cases():
    shapes = [((6,) * 6, (6,) * 6), ((4,) * 6, (4,) * 6), ((8,) * 3, (8,) * 3), ((2, 3), (4, 5)),
              ((1, 2, 3, 4), (4, 3, 2, 1))]
    vals = []
    for m, n in shapes:
        L = O.ref_lib("recursive")
        mm = np.array(m, dtype=np.int32)
        nn = np.array(n, dtype=np.int32)
        vals.append(L.ref_kron_flops(len(m), mm.ctypes.data, nn.ctypes.data))
    return {"shapes_m": np.array([list(m) + [0] * (6 - len(m)) for m, _ in shapes]),
            "shapes_n": np.array([list(n) + [0] * (6 - len(n)) for _, n in shapes]),
            "d": np.array([len(m) for m, _ in shapes]),
            "flops": np.array(vals)}


def kgemm_cases():
    """kgemm_{nn,nt,tt}_batched: shapes from the reference tests' grid
    (test_kgemm_nt_batched.cpp:490-500) and slender ones, padded leading
    dimensions, beta in {0, 1, 1.2} (1: every entry adds into one shared C),
    double and float."""
    from _helpers import kgemm_inputs
    out = {}
    shapes = [(1, 1, 1), (8, 8, 8), (15, 1, 22), (29, 12, 8), (64, 8, 8), (1, 13, 22)]
    case = 0
    for op in ("nn", "nt", "tt"):
        for dt in (np.float64, np.float32):
            for si, (mm, nn, kk) in enumerate(shapes):
                for beta in (0.0, 1.0, 1.2):
                    kind = "reference" if si % 2 == 0 else "random"
                    shared = beta == 1.0
                    As, Bs, Cs = kgemm_inputs(op, mm, nn, kk, 2, kind=kind, seed=case, dtype=dt, pad=si % 3,
                                              shared_c=shared)
                    C = O.kgemm_batched(O.KGEMM_NN + ("nn", "nt", "tt").index(op), mm, nn, kk, 1.3, As, Bs,
                                        beta, Cs, use_ref=True)
                    key = f"c{case}"
                    out[key + "_meta"] = np.array([("nn", "nt", "tt").index(op), mm, nn, kk, int(shared)])
                    out[key + "_alpha"] = np.array([1.3], dtype=dt)
                    out[key + "_beta"] = np.array([beta], dtype=dt)
                    for name, bufs in (("A", As), ("B", Bs)):
                        for b, buf in enumerate(bufs):
                            out[f"{key}_{name}{b}"] = buf
                    cs = Cs[:1] if shared else Cs
                    outs = C[:1] if shared else C
                    for b, (c0, c1) in enumerate(zip(cs, outs)):
                        out[f"{key}_C0{b}"] = c0
                        out[f"{key}_C{b}"] = c1
                    case += 1
    out["ncases"] = np.array([case])
    return out


def main():
    for b in BUILDS:
        if not O.ref_available(b):
            raise SystemExit(f"reference CPU build {b} missing: run `make -C oracle ref`")
    O.set_num_threads(1)
    for name, fn in (("pbatched", pbatched_cases), ("batched", batched_cases), ("xbatched", xbatched_cases),
                     ("vbatched", vbatched_cases), ("pbatched_f32", float_cases), ("kron_flops", flops_cases),
                     ("kgemm", kgemm_cases)):
        data = fn()
        path = os.path.join(HERE, f"ref_{name}.npz")
        np.savez_compressed(path, **data)
        print(f"wrote {path}: {len(data)} arrays, {os.path.getsize(path)} bytes")


if __name__ == "__main__":
    main()
