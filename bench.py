#!/usr/bin/env python3
"""Benchmark: full D-core decomposition on one MI355X.

Workload (the plain-GPU configuration the metric is quoted on): the seeded
directed R-MAT digraph of scale 20 -- V = 2^20, 16 * 2^20 samples with
(a, b, c, d) = (0.57, 0.19, 0.19, 0.05), self-loops and duplicates removed,
seed 1 -- resident in HBM.  One step = one full decomposition through the
C-ABI (dcx_decompose): the in-core (kmax) pass, the kmax-descending vertex
order, and every l-row k = 0..kmax.  Each dcx_decompose call returns after
its stream has finished, so host wall time brackets exactly K steps.

Prints one JSON line with the metric (GEdges/s = E / T_step), the roofline of
the dominant kernel (the batched rows kernel; algorithmic bytes
sum_k (8 E_k + 20 V_k) over the (k,0)-cores, divided by its HIP-event time on
the library's stream), and the CPU oracle timed on the host cores.
"""
import argparse
import json
import os
import platform
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "d-core-cuda_amd"))

HBM_PEAK_GBS = 8000.0


def core_sizes(kmaxes, src, dst, K):
    """V_k = |{v : kmax >= k}| and E_k = |{(u,v) : min(kmax) >= k}|, k=0..K."""
    vk = np.bincount(kmaxes, minlength=K + 1)[::-1].cumsum()[::-1]
    m = np.minimum(kmaxes[src], kmaxes[dst])
    ek = np.bincount(m, minlength=K + 1)[::-1].cumsum()[::-1]
    return vk.astype(np.int64), ek.astype(np.int64)


def read_traffic():
    """HBM bytes per rows-kernel launch from the committed rocprofv3 PMC
    summary (FETCH_SIZE + WRITE_SIZE, copy-kernel calibrated), or None."""
    p = os.path.join(ROOT, "profiles", "rows_batched_rmat20_pmc.txt")
    try:
        for line in open(p):
            if line.startswith("traffic_bytes_per_launch:"):
                return float(line.split(":")[1])
    except OSError:
        pass
    return None


def cpu_baseline(V, src, dst, threads):
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle

    g = oracle.csr_from_edges(V, src, dst)
    t0 = time.perf_counter()
    K, km, lmax, rows = oracle.decompose(g, nthreads=threads)
    t = time.perf_counter() - t0
    return t, K, lmax, km, rows


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--scale", type=int, default=20)
    ap.add_argument("--variant", default="batched", choices=["batched", "peel"])
    ap.add_argument("--no-plain", action="store_true", help="skip the plain-variant timing")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    ap.add_argument("--cpu-threads", type=int, default=0)
    args = ap.parse_args()

    world = int(os.environ.get("WORLD_SIZE", "1"))
    if args.gpus > 1 or world > 1:
        print("bench.py: this build runs on one GPU; the decomposition would shard by "
              "k-ranges across GPUs (see DESIGN.md), which is not built or measured here.",
              file=sys.stderr)
        sys.exit(2)

    import dcx

    scale = args.scale
    V = 1 << scale
    samples = 16 << scale
    t0 = time.perf_counter()
    src, dst = dcx.gen_rmat(scale, samples, seed=1)
    E = int(src.size)
    gen_s = time.perf_counter() - t0
    print(f"# rmat{scale}: V={V} E={E} (gen {gen_s:.1f}s)", file=sys.stderr, flush=True)

    variant = dcx.VARIANT_BATCHED if args.variant == "batched" else dcx.VARIANT_PEEL
    ctx = dcx.Context(0)
    g = ctx.upload_edges(V, src, dst, slack=0)
    for _ in range(args.warmup):
        g.decompose(variant)
    kms, rows_k, tot = [], [], []
    t0 = time.perf_counter()
    for _ in range(args.steps):
        K, lmax = g.decompose(variant)
        t = ctx.timing()
        kms.append(t.kmax_ms)
        rows_k.append(t.rows_kernel_ms)
        tot.append(t.total_ms)
    wall = time.perf_counter() - t0
    ms_per_step = wall / args.steps * 1e3
    timing = ctx.timing()
    print(f"# steps done: {ms_per_step:.2f} ms/step kmax={K} lmax={lmax}", file=sys.stderr,
          flush=True)

    kmaxes = g.kmaxes()
    vk, ek = core_sizes(kmaxes, src.astype(np.int64), dst.astype(np.int64), K)
    b_rows = float((8 * ek + 20 * vk).sum())
    b_kmax = float(8 * E + 20 * V)
    rows_ms = float(np.mean(rows_k))
    achieved = b_rows / (rows_ms * 1e-3) / 1e9
    roofline = {
        "bound": "hbm",
        "achieved": round(achieved, 2),
        "peak": HBM_PEAK_GBS,
        "unit": "GB/s",
        "frac": round(achieved / HBM_PEAK_GBS, 5),
        "traffic": read_traffic(),
        "kernel": "k_rows_batched" if variant == dcx.VARIANT_BATCHED else "k_pass_grid",
        "kernel_ms": round(rows_ms, 3),
        "alg_bytes": b_rows,
    }

    extra = {}
    if not args.no_plain and variant == dcx.VARIANT_BATCHED:
        g.decompose(dcx.VARIANT_PEEL)
        tp = ctx.timing()
        extra["plain_peel_ms"] = round(tp.total_ms, 2)
        extra["plain_rows_ms"] = round(tp.rows_ms, 2)
        extra["batched_speedup_vs_plain"] = round(tp.total_ms / float(np.mean(tot)), 2)

    cpu = None
    if not args.no_cpu:
        threads = args.cpu_threads or min(os.cpu_count() or 1, 16)
        t_cpu, Kc, Lc, kmc, rowsc = cpu_baseline(V, src, dst, threads)
        same = (Kc == K and Lc == lmax and np.array_equal(kmc, kmaxes)
                and np.array_equal(rowsc, g.rows()))
        cpu = {
            "value": round(E / t_cpu / 1e9, 6),
            "unit": "GEdges/s",
            "cores": threads,
            "kind": "port",
            "sample": (f"full decomposition of the same rmat{scale} graph by the C oracle "
                       f"(OpenMP over k-rows), {t_cpu:.2f} s on {platform.processor() or 'host'}"
                       f" {threads} threads; output equal to GPU: {bool(same)}"),
            "ms": round(t_cpu * 1e3, 1),
        }

    line = {
        "metric": "full D-core decomposition throughput (GEdges/s = E_dedup / T_full)",
        "value": round(E / (ms_per_step * 1e-3) / 1e9, 6),
        "unit": "GEdges/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "int32",
        "data": "synthetic",
        "config": {
            "workload": f"rmat{scale} full decomposition (kmax pass + all l-rows)",
            "V": V, "E": E, "samples": samples, "abcd": [0.57, 0.19, 0.19, 0.05], "seed": 1,
            "kmax": K, "lmax": lmax, "variant": args.variant,
            "rows_slots": timing.rows_slots,
        },
        "breakdown_ms": {
            "kmax_pass": round(float(np.mean(kms)), 3),
            "order": round(timing.order_ms, 3),
            "rows_kernel": round(rows_ms, 3),
            "rows_total": round(timing.rows_ms, 3),
        },
        "kmax_pass_roofline_frac": round(b_kmax / (np.mean(kms) * 1e-3) / 1e9 / HBM_PEAK_GBS, 5),
        "roofline": roofline,
        "cpu_baseline": cpu,
    }
    line.update(extra)
    g.free()
    ctx.close()
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
