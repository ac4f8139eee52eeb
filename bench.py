"""Benchmark of brute-force kNN on one MI355X (the headline workload):

  training set 200,000 x 128 and 20,000 queries drawn from a 64-component isotropic
  Gaussian mixture (means uniform in [-10,10]^128, sigma = 1), k = 10, euclidean,
  indices plus ascending distances.

A step is one da_nn_kneighbors_device_s call over the whole 20,000-query batch, with the
training set resident in GPU memory (da_nn_set_data_s) and the queries already in GPU
memory when the timed region starts. Prints one JSON line.
"""
import argparse
import ctypes as C
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "aocl-data-analytics_amd"))
sys.path.insert(0, ROOT)

METRIC = "kNN queries/s + p99 latency at recall@10=1.0; lowp vs FP32 speedup, 1 GPU"
HBM_PEAK = 8.0e12
PEAK = {"f32": 157.3e12, "f64": 78.6e12}


def mixture(rng, n, means, sigma=1.0, dtype=np.float32):
    comp = rng.integers(0, len(means), n)
    return (means[comp] + sigma * rng.standard_normal((n, means.shape[1]))).astype(dtype)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--n-train", type=int, default=200_000)
    ap.add_argument("--n-query", type=int, default=20_000)
    ap.add_argument("--dim", type=int, default=128)
    ap.add_argument("--k", type=int, default=10)
    ap.add_argument("--dtype", choices=["f32", "f64"], default="f32")
    ap.add_argument("--cpu-sample", type=int, default=64,
                    help="queries timed on the host with the oracle (0 disables)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; --gpus must be 1",
                          "n_gpus": args.gpus}))
        sys.exit(2)

    from aoclda import _lib as L
    import aoclda

    dtype = np.float32 if args.dtype == "f32" else np.float64
    rng = np.random.default_rng(0)
    means = rng.uniform(-10, 10, (64, args.dim))
    Xt = mixture(rng, args.n_train, means, dtype=dtype)
    Xq = mixture(rng, args.n_query, means, dtype=dtype)
    k, nq, d = args.k, args.n_query, args.dim

    nn = aoclda.neighbors.nearest_neighbors(n_neighbors=k).fit(Xt)
    h = nn.handle
    dq, di, dd = C.c_void_p(), C.c_void_p(), C.c_void_p()
    L.check(L.lib.da_amd_malloc(Xq.nbytes, C.byref(dq)))
    L.check(L.lib.da_amd_malloc(nq * k * 4, C.byref(di)))
    L.check(L.lib.da_amd_malloc(nq * k * Xq.itemsize, C.byref(dd)))
    L.check(L.lib.da_amd_memcpy_htod(dq, L.ptr(Xq), Xq.nbytes))

    def step():
        h.call("da_nn_kneighbors_device", nq, d, dq, d, di, dd, k, 1)

    for _ in range(args.warmup):
        step()
    L.check(L.lib.da_amd_synchronize(h.h))
    print(f"bench: warmup done ({args.warmup} steps)", file=sys.stderr, flush=True)

    # Timed region: K steps; per-step latency from host timestamps after a stream sync.
    lat = []
    t0 = time.perf_counter()
    for _ in range(args.steps):
        s0 = time.perf_counter()
        step()
        L.check(L.lib.da_amd_synchronize(h.h))
        lat.append(time.perf_counter() - s0)
    elapsed = time.perf_counter() - t0
    print(f"bench: timed {args.steps} steps in {elapsed:.3f}s", file=sys.stderr, flush=True)

    # Per-kernel durations (HIP events on the handle's stream) over a second pass.
    L.check(L.lib.da_amd_profile_enable(h.h, 1))
    L.check(L.lib.da_amd_profile_reset(h.h))
    for _ in range(args.steps):
        step()
    L.check(L.lib.da_amd_synchronize(h.h))
    kern = {}
    for name in ("knn_tile", "knn_merge", "pairwise_direct", "row_topk"):
        ms, cnt = C.c_double(), L.da_int()
        L.check(L.lib.da_amd_profile_get(h.h, name.encode(), C.byref(ms), C.byref(cnt)))
        if cnt.value:
            kern[name] = (ms.value, cnt.value)
    L.check(L.lib.da_amd_profile_enable(h.h, 0))

    ind = np.empty((nq, k), dtype=np.int32)
    dist = np.empty((nq, k), dtype=dtype)
    L.check(L.lib.da_amd_memcpy_dtoh(L.ptr(ind), di, ind.nbytes))
    L.check(L.lib.da_amd_memcpy_dtoh(L.ptr(dist), dd, dist.nbytes))
    for p in (dq, di, dd):
        L.lib.da_amd_free(p)

    flops = 2.0 * nq * args.n_train * d
    ms_step = elapsed / args.steps * 1e3
    roofline = None
    if kern:
        name = max(kern, key=lambda n: kern[n][0])
        tot_ms, cnt = kern[name]
        per_launch_ms = tot_ms / cnt
        launches_per_step = cnt / args.steps
        # algorithmic flops of one launch: the whole-batch contraction split evenly over the
        # launches of one step
        f_launch = flops / launches_per_step
        achieved = f_launch / (per_launch_ms * 1e-3)
        roofline = {"kernel": name, "bound": "mfma", "achieved": achieved / 1e12,
                    "peak": PEAK[args.dtype] / 1e12, "unit": "TFLOP/s",
                    "frac": achieved / PEAK[args.dtype], "traffic": None,
                    "avg_launch_ms": per_launch_ms, "launches_per_step": launches_per_step}

    cpu = None
    recall = None
    if args.cpu_sample > 0:
        from oracle import oracle
        s = min(args.cpu_sample, nq)
        t = time.perf_counter()
        ei, ed = oracle.knn(Xt.astype(np.float64), Xq[:s].astype(np.float64), k)
        ct = time.perf_counter() - t
        cpu = {"value": s / ct, "unit": "queries/s", "cores": 1, "kind": "port",
               "sample": f"{s} queries x {args.n_train} training rows, exact fp64 oracle, "
                         "1 thread"}
        hits = sum(len(set(ind[q]) & set(ei[q])) for q in range(s))
        recall = hits / (s * k)
        derr = float(np.max(np.abs(dist[:s].astype(np.float64) - ed) / np.maximum(ed, 1e-30)))
    out = {
        "metric": METRIC, "value": nq * args.steps / elapsed, "unit": "queries/s",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup, "ms_per_step": ms_step,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": args.dtype, "data": "synthetic: 64-component isotropic Gaussian mixture, seed 0",
        "config": {"workload": f"brute-force kNN, train {args.n_train}x{d}, {nq} queries, "
                               f"k={k}, euclidean", "n_train": args.n_train, "n_query": nq,
                   "dim": d, "k": k, "precision_form": "plain"},
        "p50_ms": float(np.percentile(np.array(lat) * 1e3, 50)),
        "p99_ms": float(np.percentile(np.array(lat) * 1e3, 99)),
        "recall_at_10": recall, "max_rel_dist_err": derr if cpu else None,
        "lowp_speedup": None,
        "roofline": roofline, "cpu_baseline": cpu,
        "kernels_ms": {n: {"total_ms": v[0], "launches": v[1]} for n, v in kern.items()},
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
