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
the dominant kernel (the batched rows kernel k_rows_teams; algorithmic bytes
sum_k (8 E_k + 20 V_k) over the (k,0)-cores, divided by its HIP-event time on
the library's stream), and the CPU oracle timed on the host cores.  The
launch shape of the rows kernel is autotuned on the graph before timing
(dcx_autotune, untimed; every shape gives the same bits).

``--workload chunglu256`` instead times the batched workload: 256 seeded
Chung-Lu digraphs of log-uniform size decomposed by one union launch
sequence, against the per-graph, per-k launch loop (plain variant), with
every graph's output compared bit for bit.
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
    p = os.path.join(ROOT, "profiles", "k_rows_teams_rmat20_pmc.txt")
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
    ap.add_argument("--workload", default="headline",
                    choices=["headline", "rmat", "chunglu256", "serve", "learned", "ingest"])
    ap.add_argument("--schedule", default="learned", choices=["learned", "workshare"],
                    help="headline: rows planned by the cost model, or the fixed rule")
    ap.add_argument("--full-train", action="store_true",
                    help="headline: also train the cost model on passes logged from smaller "
                         "graphs at several team caps (default: the graph's own passes only)")
    ap.add_argument("--quick-train", action="store_true",
                    help="headline: the default (the graph's own passes only); kept for "
                         "older command lines")
    ap.add_argument("--plain", action="store_true",
                    help="headline: also time the plain one-k-per-launch peel")
    ap.add_argument("--no-rmat20", action="store_true",
                    help="headline: skip the R-MAT 20 seed 1 timing")
    ap.add_argument("--host-graph", action="store_true",
                    help="headline: sample and deduplicate the graph on the host and upload "
                         "it (the same graph) instead of sampling it on the device")
    ap.add_argument("--ingest-scale", type=int, default=22,
                    help="ingest: R-MAT scale of the edge list")
    ap.add_argument("--learned-scale", type=int, default=24,
                    help="learned: log2 of the vertex count of the two graphs")
    ap.add_argument("--no-plain-learned", action="store_true",
                    help="learned: skip the plain one-k-per-launch baseline")
    ap.add_argument("--serve-scale", type=int, default=22)
    ap.add_argument("--serve-batches", type=int, default=10000)
    ap.add_argument("--serve-loads", default="0.5,0.9",
                    help="serve: Poisson loads as fractions of the measured capacity")
    ap.add_argument("--serve-schedule", default="learned", choices=["learned", "workshare"],
                    help="serve: the recomputations' rows planned by a cost model trained on "
                         "the base graph (untimed), or the fixed rule")
    ap.add_argument("--scale", type=int, default=None,
                    help="R-MAT scale (headline: 24, rmat: 20)")
    ap.add_argument("--variant", default="batched", choices=["batched", "peel", "hindex"])
    ap.add_argument("--no-plain", action="store_true", help="skip the plain-variant timing")
    ap.add_argument("--no-hindex", action="store_true", help="skip the h-index timing")
    ap.add_argument("--no-tune", action="store_true", help="skip the launch-shape autotuner")
    ap.add_argument("--shape", default="",
                    help="rows launch shape SLOTS,TEAM_MAX,SHRINK_AT[,LDS_MEMBERS] instead of the "
                         "autotuner "
                         "(e.g. the shape an earlier run picked, for counter runs)")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    ap.add_argument("--cpu-threads", type=int, default=0)
    ap.add_argument("--baseline-graphs", type=int, default=256,
                    help="chunglu256: graphs timed in the per-graph loop")
    args = ap.parse_args()

    world = int(os.environ.get("WORLD_SIZE", "1"))
    if args.gpus > 1 or world > 1:
        print("bench.py: runs on one GPU only. Across GPUs the decomposition would shard "
              "by k-ranges (each GPU a replica of the core-ordered graph and its own kmax "
              "pass), which is not built or measured.",
              file=sys.stderr)
        sys.exit(2)
    if args.workload == "chunglu256":
        return run_chunglu(args)
    if args.workload == "serve":
        return run_serve(args)
    if args.workload == "learned":
        return run_learned(args)
    if args.workload == "ingest":
        return run_ingest(args)
    if args.workload == "rmat":
        args.scale = args.scale or 20
        return run_rmat(args)
    args.scale = args.scale or 24
    return run_headline(args)


def run_ingest(args):
    """Edge list -> CSR with slack on the device (dcx_graph_ingest_edges),
    the pairs already in HBM, against the host build plus upload
    (dcx_graph_upload_edges) and the device build from host arrays.  The
    device graph must equal the host-built one slot for slot."""
    import dcx

    scale, slack = args.ingest_scale, 1
    V = 1 << scale
    src, dst = dcx.gen_rmat(scale, 16 << scale, 1)
    E = int(src.size)
    print(f"# ingest: rmat{scale} V={V} E={E}", file=sys.stderr, flush=True)
    ctx = dcx.Context(0)

    def best(fn, reps):
        ts = []
        for _ in range(reps):
            t0 = time.perf_counter()
            g = fn()
            ts.append((time.perf_counter() - t0) * 1e3)
            g.free()
        return min(ts)

    ds, dd = ctx.to_device(src), ctx.to_device(dst)
    for _ in range(args.warmup):
        ctx.ingest_device(V, ds, dd, slack).free()
    dev_ms = best(lambda: ctx.ingest_device(V, ds, dd, slack), args.steps)
    print(f"# device ingest from HBM: {dev_ms:.2f} ms", file=sys.stderr, flush=True)
    host_arrays_ms = best(lambda: ctx.ingest_edges(V, src, dst, slack), 3)
    host_ms = best(lambda: ctx.upload_edges(V, src, dst, slack), 3)
    print(f"# device ingest from host arrays: {host_arrays_ms:.2f} ms; host build + upload: "
          f"{host_ms:.2f} ms", file=sys.stderr, flush=True)
    a = ctx.upload_edges(V, src, dst, slack)
    b = ctx.ingest_device(V, ds, dd, slack)
    ca, cb = a.download_csr(), b.download_csr()
    equal = all(np.array_equal(ca[k], cb[k]) for k in ca)
    b.free()
    ds.free()
    dd.free()
    # raw samples drawn on the device, self-loops dropped and deduplicated
    # there: the same graph as the host generator's output
    rs, rd = ctx.gen_rmat_device(scale, 16 << scale, 1)
    ctx.ingest(V, rs, rd, slack, drop_self_loops=True, dedup=True).free()
    raw_ms = best(lambda: ctx.ingest(V, rs, rd, slack, drop_self_loops=True, dedup=True),
                  args.steps)
    b = ctx.ingest(V, rs, rd, slack, drop_self_loops=True, dedup=True)
    cb = b.download_csr()
    raw_equal = all(np.array_equal(ca[k], cb[k]) for k in ca)
    b.free()
    a.free()
    rs.free()
    rd.free()
    ctx.close()
    print(f"# device ingest of raw samples with dedup: {raw_ms:.2f} ms, equal {raw_equal}",
          file=sys.stderr, flush=True)
    # algorithmic bytes: read the pairs (8 E), write both neighbour arrays
    # (8 E) and both sides' offsets, degrees and capacities (24 V)
    alg = 16.0 * E + 24.0 * V
    achieved = alg / (dev_ms * 1e-3) / 1e9
    line = {
        "metric": "device graph ingest throughput (GEdges/s = E / T, pairs in HBM)",
        "value": round(E / (dev_ms * 1e-3) / 1e9, 4),
        "unit": "GEdges/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(dev_ms, 3),
        "higher_is_better": True,
        "scaling": "none (one GPU)",
        "vs_baseline": None,
        "dtype": "uint32",
        "data": "synthetic",
        "config": {"workload": f"ingest: rmat{scale} edge list (16 * 2^{scale} samples, seed 1, "
                               f"deduplicated) to in/out CSR with slack {slack}",
                   "V": V, "E": E},
        "host_build_and_upload_ms": round(host_ms, 3),
        "host_build_threads": 1,
        "raw_samples_dedup_ingest_ms": round(raw_ms, 3),
        "raw_samples_dedup_equal_to_host_build": bool(raw_equal),
        "device_ingest_from_host_arrays_ms": round(host_arrays_ms, 3),
        "speedup_vs_host_build": round(host_ms / dev_ms, 2),
        "equal_to_host_build": bool(equal),
        "roofline": {"bound": "hbm", "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 5), "traffic": None,
                     "kernel": "whole ingest (count, scan, 2 radix sorts, placement)",
                     "alg_bytes": alg},
        "cpu_baseline": {"value": round(E / (host_ms * 1e-3) / 1e9, 4), "unit": "GEdges/s",
                         "cores": 1, "kind": "port",
                         "sample": "dcx_csr_build (single-thread host CSR build, the layout of "
                                   "Graph::readFile) plus its upload, same edge list"},
    }
    print(json.dumps(line), flush=True)


def read_pmc(name, key="traffic_bytes_per_launch"):
    """A figure of a committed rocprofv3 PMC summary under profiles/ (a
    '<key>: value' line, e.g. calibrated HBM bytes per launch), or None."""
    try:
        for line in open(os.path.join(ROOT, "profiles", name)):
            if line.startswith(key + ":"):
                return float(line.split(":")[1])
    except OSError:
        pass
    return None


def cpu_info():
    model = platform.processor() or platform.machine()
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    try:
        avail = len(os.sched_getaffinity(0))
    except AttributeError:
        avail = os.cpu_count() or 1
    return model, os.cpu_count() or 1, avail


def oracle_sample(g, V, threads, log):
    """The CPU oracle on a bounded sample of the same decomposition, on the
    graph downloaded from the device: the kmax pass, the relabelling, and
    every stride-th row on `threads` threads, then three rows on one thread;
    each scaled to the whole decomposition by the rows' algorithmic bytes."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle

    t0 = time.perf_counter()
    c = g.download_csr()
    csr = oracle.CSR(V, c["out_off"].astype(np.uint64), c["out_nbr"],
                     c["in_off"].astype(np.uint64), c["in_nbr"])
    del c
    log(f"cpu baseline: graph downloaded in {time.perf_counter() - t0:.1f}s")
    K, t, w_s, w_all = oracle.time_sample(csr, threads, 16)
    est = t["kmax_s"] + t["order_s"] + t["rows_s"] * w_all / w_s
    log(f"cpu baseline ({threads} threads): {t}, estimate {est:.1f}s")
    K1, t1, w1, _ = oracle.time_sample(csr, 1, max(K // 3, 1))
    est1 = t["kmax_s"] + t["order_s"] * threads + t1["rows_s"] * w_all / w1
    log(f"cpu baseline (1 thread): rows {t1['rows_s']:.1f}s, estimate {est1:.1f}s")
    return {"threads": threads, "parts_s": t, "rows_sampled_frac": round(w_s / w_all, 4),
            "estimate_s": est, "one_thread": {"rows_s": t1["rows_s"],
                                              "rows_sampled_frac": round(w1 / w_all, 4),
                                              "estimate_s": est1}}


# Settling runs at the end of the training protocol: each costs one
# decomposition (about 0.2 s at R-MAT 24), and more of them give the plan
# memory more measured plans to keep the fastest of.
SETTLING_RUNS = 8


def train_learned_plan(dcx, ctx, g, log, Xs=None, Ts=None):
    """The cost model's training protocol on graph g (untimed): the pass log
    of the fixed rule's last run (already decomposed by the caller), four
    exploration runs (every 8th team row an anchor: peeled with team caps
    32, 4 and 64, so the samples cover narrow and whole-XCD teams, and
    derived from row 0 with team cap 64),
    then two refinement rounds in which the model plans a run and is
    retrained with that run's log, and SETTLING_RUNS settling runs of the
    final model (each run's measured row times calibrate the next plan's
    predictions; the graph keeps the fastest plan measured so far and tries
    another only when it is predicted at least 3 % faster).  Returns the attached model, the sample count and the
    refinement and settling runs' times."""
    Xs = list(Xs or [])
    Ts = list(Ts or [])

    def collect():
        for wl in (True, False):
            X, T = g.pass_features(with_log=wl)
            Xs.append(X)
            Ts.append(T)

    collect()
    for spacing, tmax, far in ((8, 0, 0), (8, 4, 0), (8, 64, 0), (8, 64, 1)):
        ctx.set_rows_rule(spacing, tmax, far)
        g.decompose(dcx.VARIANT_BATCHED)
        kind = "derived from row 0" if far else "peeled"
        log(f"exploration run (every {spacing}th row an anchor, {kind}, team cap {tmax or 32}): "
            f"{ctx.timing().total_ms:.1f} ms")
        collect()
    ctx.set_rows_rule(0, 0, 0)
    if sum(len(t) for t in Ts) == 0:
        log("no team rows to learn from (every row ran whole-graph or LDS-resident): "
            "the fixed rule stays")
        return None, 0, []
    model = dcx.MLP(ctx, seed=7)
    refine_ms = []
    n = 0
    for rnd in range(3):
        X, T = np.concatenate(Xs), np.concatenate(Ts)
        loss = model.train(X, T, epochs=3000 if rnd == 0 else 1500, lr=3e-3)
        n = len(T)
        log(f"cost model round {rnd}: {n} samples, loss {loss:.4f}")
        ctx.set_scheduler(model)
        if rnd < 2:
            g.decompose(dcx.VARIANT_BATCHED)
            refine_ms.append(round(ctx.timing().total_ms, 2))
            collect()
    for _ in range(SETTLING_RUNS):
        g.decompose(dcx.VARIANT_BATCHED)
        refine_ms.append(round(ctx.timing().total_ms, 2))
    log(f"refinement and settling runs: {refine_ms} ms")
    return model, n, refine_ms


def run_headline(args):
    """R-MAT scale 24 (2^28 samples, seed 4), learned schedule: the
    learned-scheduling workload's R-MAT graph, the largest single-GPU
    decomposition the targets are stated on."""
    import dcx

    def log(msg):
        print(f"# {msg}", file=sys.stderr, flush=True)

    scale, seed = args.scale, (4 if args.scale == 24 else 1)
    V, samples = 1 << scale, 16 << scale
    ctx = dcx.Context(0)
    tg0 = time.perf_counter()
    if args.host_graph:
        hs, hd = dcx.gen_rmat(scale, samples, seed)
        tg1 = time.perf_counter()
        g = ctx.upload_edges(V, hs, hd)
        tg2 = time.perf_counter()
        del hs, hd
    else:
        s, d = ctx.gen_rmat_device(scale, samples, seed)  # returns after its stream ends
        tg1 = time.perf_counter()
        g = ctx.ingest(V, s, d, 0, drop_self_loops=True, dedup=True)
        tg2 = time.perf_counter()
        s.free()
        d.free()
    E = int(g.info()[1])
    log(f"rmat{scale}: V={V} E={E}; {'host' if args.host_graph else 'device'} sampling "
        f"{1e3 * (tg1 - tg0):.0f} ms, {'upload' if args.host_graph else 'dedup ingest'} "
        f"{1e3 * (tg2 - tg1):.0f} ms")

    # the fixed rule (every 32nd team row peeled, the rest derived from the
    # row above; widths by work share), then the cost model trained on the
    # GPU from what the decompositions logged: the fixed rule's runs, four
    # exploration runs (every 8th row an anchor: peeled on teams of up to
    # 32, 4 and 64 workgroups, and derived from row 0), and two refinement
    # rounds in which the model plans a run and is retrained with that run's
    # log as well
    ctx.set_scheduler(None)
    g.decompose(dcx.VARIANT_BATCHED)
    heur = []
    for _ in range(2):
        g.decompose(dcx.VARIANT_BATCHED)
        heur.append(ctx.timing().total_ms)
    log(f"fixed rule: {min(heur):.1f} ms")
    h_hash = g.state_hash()
    model = None
    n_samples = 0
    refine_ms = []
    if args.schedule == "learned":
        Xs, Ts = learned_training_set(dcx, log) if args.full_train else ([], [])
        model, n_samples, refine_ms = train_learned_plan(dcx, ctx, g, log, Xs, Ts)
    for _ in range(args.warmup):
        g.decompose(dcx.VARIANT_BATCHED)
    tims = []
    t0 = time.perf_counter()
    for _ in range(args.steps):
        K, lmax = g.decompose(dcx.VARIANT_BATCHED)
        tims.append(ctx.timing())
    wall = time.perf_counter() - t0
    ms_per_step = wall / args.steps * 1e3
    state_hash = g.state_hash()
    used_model = bool(tims[-1].rows_scheduler)
    log(f"steps done: {ms_per_step:.1f} ms/step kmax={K} lmax={lmax} (fixed rule "
        f"{min(heur):.1f} ms)")

    mean = lambda f: float(np.mean([getattr(t, f) for t in tims]))  # noqa: E731
    # algorithmic bytes per row: a peeled row reads its core's lists and
    # vertex words (8 E_k + 20 V_k); a derived row the list entries its
    # walks read (8 bytes each: the entry and its neighbour's value, counted
    # on the device) and 20 V_k of vertex words
    log_rows = g.pass_log()
    split = int(tims[-1].rows_split)
    first = 1 if mean("rows_first_ms") > 0 else 0
    team_b = lds_b = peel_equiv = 0.0
    n_peeled = n_derived = n_far = 0
    read_entries = core_entries = 0
    for p in log_rows[1:]:
        k, vk_, ek_ = p["k"], float(p["members"]), float(p["edges"])
        full = 8 * ek_ + 20 * vk_
        peel_equiv += full
        if p["variant"] == dcx.VARIANT_DERIVED:
            b = 8.0 * p["alive"] + 20 * vk_
            if p["levels"] == 2:
                n_far += 1  # an anchor derived from row 0
            else:
                n_derived += 1
            read_entries += p["alive"]
        else:
            b = full
            n_peeled += 1 if first <= k < split else 0
            read_entries += 2 * p["edges"] if first <= k < split else 0
        if first <= k < split:
            team_b += b
            core_entries += 2 * p["edges"]
        elif k >= split:
            lds_b += b
    b_kmax = 8.0 * E + 20.0 * V
    teams_ms, lds_ms = mean("rows_teams_ms"), mean("rows_lds_ms")
    kernels = {
        "k_rows_teams": (team_b, teams_ms, f"{first}..{split - 1} ({n_peeled} peeled, "
                                           f"{n_far} derived from row 0, "
                                           f"{n_derived} derived from the row above)"),
        "k_rows_lds": (lds_b, lds_ms, f"{split}..{K}"),
        "k_pass_grid (kmax pass)": (b_kmax, mean("kmax_ms"), "kmax pass"),
    }
    dom = max(kernels, key=lambda k: kernels[k][1])
    b_dom, ms_dom, rows_dom = kernels[dom]
    achieved = b_dom / (ms_dom * 1e-3) / 1e9
    roofline = {"bound": "hbm", "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS,
                "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 5),
                "traffic": read_pmc(f"k_rows_teams_rmat{scale}_pmc.txt")
                if dom == "k_rows_teams" else None,
                "kernel": dom.split(" ")[0], "kernel_ms": round(ms_dom, 3), "alg_bytes": b_dom,
                "rows": rows_dom}
    pmc_name = f"k_rows_teams_rmat{scale}_pmc.txt"
    fr_traffic = read_pmc(pmc_name, "fixed_rule_traffic_bytes_per_launch")
    fr_alg = read_pmc(pmc_name, "fixed_rule_alg_bytes_per_launch")
    b_all = team_b + lds_b + b_kmax + 8.0 * E + 20.0 * V  # + row 0 beside the kmax pass
    extra = {
        "kernels": {k: {"alg_bytes": v[0], "ms": round(v[1], 3), "rows": v[2],
                        "GBs": round(v[0] / (v[1] * 1e-3) / 1e9, 2) if v[1] > 0 else None}
                    for k, v in kernels.items()},
        "full_decomposition": {
            "alg_bytes": b_all, "ms": round(ms_per_step, 3),
            "roofline_frac": round(b_all / (ms_per_step * 1e-3) / 1e9 / HBM_PEAK_GBS, 5),
            "peel_equivalent_bytes": peel_equiv + b_kmax,
            "peel_equivalent_frac": round((peel_equiv + b_kmax) / (ms_per_step * 1e-3) / 1e9 /
                                          HBM_PEAK_GBS, 5),
            "team_rows_list_entries_read": int(read_entries),
            "team_rows_core_list_entries": int(core_entries)},
        "breakdown_ms": {"kmax_pass": round(mean("kmax_ms"), 3),
                         "order": round(mean("order_ms"), 3),
                         "rows_kernel": round(mean("rows_kernel_ms"), 3),
                         "rows_total": round(mean("rows_ms"), 3)},
        # counters of the team launch exist for the fixed rule's plan only
        # (profiles/k_rows_teams_rmat24_pmc.txt)
        "team_launch_traffic_fixed_rule": (
            {"traffic": fr_traffic, "alg_bytes": fr_alg,
             "traffic_over_alg": round(fr_traffic / fr_alg, 3)}
            if fr_traffic and fr_alg else None),
        "schedule": {"kind": args.schedule, "model_used": used_model,
                     "training": f"{n_samples} samples: this graph's logged team rows under the "
                                 "fixed rule, four exploration runs (every 8th row an anchor: "
                                 "peeled with team caps 32, 4 and 64, derived from row 0 with "
                                 "cap 64), two refinement rounds planned by the model and "
                                 f"{SETTLING_RUNS} settling runs (measured row times calibrate "
                                 "the predictions)" + (
                                     ", and smaller graphs' at several team caps"
                                     if args.full_train else ""),
                     "refinement_and_settling_runs_ms": refine_ms,
                     "fixed_rule_ms": round(min(heur), 2),
                     "learned_speedup_vs_fixed_rule": round(min(heur) / ms_per_step, 3),
                     "bit_exact_vs_fixed_rule": state_hash == h_hash},
        "ingest_ms": ({"host_sampling": round(1e3 * (tg1 - tg0), 1),
                       "upload": round(1e3 * (tg2 - tg1), 1)} if args.host_graph else
                      {"device_sampling": round(1e3 * (tg1 - tg0), 1),
                       "dedup_ingest": round(1e3 * (tg2 - tg1), 1)}),
    }
    gold = os.path.join(ROOT, "tests", "golden", "hashes.json")
    if os.path.exists(gold):
        want = json.load(open(gold)).get(f"rmat{scale}_s{seed}")
        if want:
            extra["oracle_hash_equal"] = want["hash"] == f"{state_hash:016x}"
    if model is not None:
        ctx.set_scheduler(None)
        model.free()
    if args.plain:
        g.decompose(dcx.VARIANT_PEEL)
        tp = ctx.timing()
        extra["plain_peel_ms"] = round(tp.total_ms, 1)
        extra["speedup_vs_plain"] = round(tp.total_ms / ms_per_step, 2)
        extra["plain_bit_exact"] = g.state_hash() == state_hash
        log(f"plain one-k-per-launch peel: {tp.total_ms:.0f} ms")
    cpu = None
    if not args.no_cpu:
        model_name, nproc, avail = cpu_info()
        threads = args.cpu_threads or min(avail, 16)
        smp = oracle_sample(g, V, threads, log)
        cpu = {"value": round(E / smp["estimate_s"] / 1e9, 6), "unit": "GEdges/s",
               "cores": threads, "kind": "port",
               "sample": (f"C oracle on the same rmat{scale} graph (downloaded from the device): "
                          f"kmax pass {smp['parts_s']['kmax_s']:.1f}s, kmax-rank relabelling "
                          f"{smp['parts_s']['order_s']:.1f}s, every 16th row "
                          f"({100 * smp['rows_sampled_frac']:.1f}% of the rows' bytes) "
                          f"{smp['parts_s']['rows_s']:.1f}s on {threads} threads, scaled by "
                          f"bytes to all rows: {smp['estimate_s']:.1f}s per decomposition"),
               "cpu_model": model_name, "nproc": nproc, "cpus_available": avail,
               "estimate_s": round(smp["estimate_s"], 2),
               "one_thread_estimate_s": round(smp["one_thread"]["estimate_s"], 1),
               "one_thread_value": round(E / smp["one_thread"]["estimate_s"] / 1e9, 6)}
    if not args.no_rmat20:
        g.free()
        extra["rmat20"] = rmat20_quick(ctx, dcx, log)
    else:
        g.free()
    ctx.close()
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
        "config": {"workload": f"rmat{scale} full decomposition (kmax pass + all l-rows), "
                               f"{args.schedule} schedule",
                   "V": V, "E": E, "samples": samples, "abcd": [0.57, 0.19, 0.19, 0.05],
                   "seed": seed, "kmax": K, "lmax": lmax, "state_hash": f"{state_hash:016x}"},
        "roofline": roofline,
        "cpu_baseline": cpu,
    }
    line.update(extra)
    print(json.dumps(line), flush=True)


def rmat20_quick(ctx, dcx, log):
    """R-MAT 20 seed 1 (the plain-GPU workload) in the same run: default
    launch shape, mean of 10 decompositions."""
    s, d = ctx.gen_rmat_device(20, 16 << 20, 1)
    g = ctx.ingest(1 << 20, s, d, 0, drop_self_loops=True, dedup=True)
    s.free()
    d.free()
    for _ in range(3):
        g.decompose(dcx.VARIANT_BATCHED)
    t0 = time.perf_counter()
    for _ in range(10):
        K, L = g.decompose(dcx.VARIANT_BATCHED)
    ms = (time.perf_counter() - t0) / 10 * 1e3
    out = {"ms_per_step": round(ms, 3), "E": int(g.info()[1]), "kmax": K, "lmax": L,
           "state_hash": f"{g.state_hash():016x}"}
    g.free()
    log(f"rmat20: {ms:.2f} ms per decomposition")
    return out


def run_rmat(args):
    import dcx

    scale = args.scale
    V = 1 << scale
    samples = 16 << scale
    t0 = time.perf_counter()
    src, dst = dcx.gen_rmat(scale, samples, seed=1)
    E = int(src.size)
    gen_s = time.perf_counter() - t0
    print(f"# rmat{scale}: V={V} E={E} (gen {gen_s:.1f}s)", file=sys.stderr, flush=True)

    variant = {"batched": dcx.VARIANT_BATCHED, "peel": dcx.VARIANT_PEEL,
               "hindex": dcx.VARIANT_HINDEX}[args.variant]
    shape = [int(x) for x in args.shape.split(",")] if args.shape else [0, 0, 0]
    shape += [0] * (4 - len(shape))
    ctx = dcx.Context(0, rows_slots=shape[0], team_max=shape[1], shrink_at=shape[2],
                      lds_members=shape[3])
    g = ctx.upload_edges(V, src, dst, slack=0)
    tune = None
    if variant == dcx.VARIANT_BATCHED and not args.no_tune and not args.shape:
        g.kmax()
        r = g.autotune(reps=2)
        tune = {"rows_slots": r.rows_slots, "team_max": r.team_max, "shrink_at": r.shrink_at,
                "lds_members": r.lds_members,
                "candidates": r.candidates, "best_rows_ms": round(r.best_ms, 3),
                "default_rows_ms": round(r.default_ms, 3)}
        print(f"# autotune: {tune}", file=sys.stderr, flush=True)
    for _ in range(args.warmup):
        g.decompose(variant)
    kms, rows_k, tot, teams_k, lds_k, first_k = [], [], [], [], [], []
    t0 = time.perf_counter()
    for _ in range(args.steps):
        K, lmax = g.decompose(variant)
        t = ctx.timing()
        kms.append(t.kmax_ms)
        rows_k.append(t.rows_kernel_ms)
        tot.append(t.total_ms)
        teams_k.append(t.rows_teams_ms)
        lds_k.append(t.rows_lds_ms)
        first_k.append(t.rows_first_ms)
    wall = time.perf_counter() - t0
    ms_per_step = wall / args.steps * 1e3
    timing = ctx.timing()
    print(f"# steps done: {ms_per_step:.2f} ms/step kmax={K} lmax={lmax}", file=sys.stderr,
          flush=True)

    # Fingerprint of the result (kmaxes and all rows, hashed on the device:
    # one streaming read of (K + 2) * V int32, also the known-size read that
    # calibrates the FETCH_SIZE counter).
    state_hash = g.state_hash()
    kmaxes = g.kmaxes()
    vk, ek = core_sizes(kmaxes, src.astype(np.int64), dst.astype(np.int64), K)
    per_row = 8 * ek + 20 * vk
    b_rows = float(per_row.sum())
    b_kmax = float(8 * E + 20 * V)
    rows_ms = float(np.mean(rows_k))
    lds_kernel = None
    # row 0 outside the rows launch: computed beside the kmax pass (or by a
    # grid pass of its own) because it bounds the other rows
    first = 1 if variant == dcx.VARIANT_BATCHED and float(np.mean(first_k)) > 0 else 0
    if variant == dcx.VARIANT_BATCHED and timing.rows_split <= K and np.mean(teams_k) > 0:
        # two launches side by side: the team launch (rows first <= k < split)
        # is the roofline kernel, the LDS-resident one (rows k >= split) is
        # reported beside it
        split = int(timing.rows_split)
        b_kernel = float(per_row[first:split].sum())
        kernel_ms = float(np.mean(teams_k))
        b_lds = float(per_row[split:].sum())
        lds_ms = float(np.mean(lds_k))
        lds_kernel = {"kernel": "k_rows_lds", "rows": f"{split}..{K}",
                      "kernel_ms": round(lds_ms, 3), "alg_bytes": b_lds,
                      "achieved_GBs": round(b_lds / (lds_ms * 1e-3) / 1e9, 2)}
    else:
        b_kernel, kernel_ms = b_rows, rows_ms
    achieved = b_kernel / (kernel_ms * 1e-3) / 1e9
    roofline = {
        "bound": "hbm",
        "achieved": round(achieved, 2),
        "peak": HBM_PEAK_GBS,
        "unit": "GB/s",
        "frac": round(achieved / HBM_PEAK_GBS, 5),
        "traffic": read_traffic() if variant == dcx.VARIANT_BATCHED else None,
        "kernel": {dcx.VARIANT_BATCHED: "k_rows_teams", dcx.VARIANT_PEEL: "k_pass_grid",
                   dcx.VARIANT_HINDEX: "k_hindex_row"}[variant],
        "kernel_ms": round(kernel_ms, 3),
        "alg_bytes": b_kernel,
        "rows": f"{first}..{int(timing.rows_split) - 1}" if lds_kernel else f"0..{K}",
    }

    extra = {}
    if not args.no_plain and variant == dcx.VARIANT_BATCHED:
        g.decompose(dcx.VARIANT_PEEL)
        tp = ctx.timing()
        extra["plain_peel_ms"] = round(tp.total_ms, 2)
        extra["plain_rows_ms"] = round(tp.rows_ms, 2)
        extra["batched_speedup_vs_plain"] = round(tp.total_ms / float(np.mean(tot)), 2)
    if not args.no_hindex and variant != dcx.VARIANT_HINDEX:
        g.decompose(dcx.VARIANT_HINDEX)
        th = ctx.timing()
        extra["hindex_ms"] = round(th.total_ms, 2)
        extra["hindex_rows_ms"] = round(th.rows_ms, 2)
        extra["hindex_kmax_ms"] = round(th.kmax_ms, 2)
    if tune:
        extra["autotune"] = tune
    extra["full_decomposition_roofline_frac"] = round(
        (b_rows + b_kmax) / (float(np.mean(tot)) * 1e-3) / 1e9 / HBM_PEAK_GBS, 5)

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
            "rows_slots": timing.rows_slots, "shape": args.shape or "autotuned",
            "state_hash": f"{state_hash:016x}",
        },
        # with row 0 on a second stream it runs beside the kmax pass and
        # the order kernels; "order" ends when both have
        "breakdown_ms": {
            "kmax_pass": round(float(np.mean(kms)), 3),
            "order": round(timing.order_ms, 3),
            "rows_kernel": round(rows_ms, 3),
            "rows_total": round(timing.rows_ms, 3),
        },
        "row0": ("second stream, beside the kmax pass and the order kernels"
                 if first and float(np.mean(first_k)) < 1.0 else
                 "grid pass before the rows launch" if first else "in the rows launch"),
        "kmax_pass_roofline_frac": round(
            b_kmax / (np.mean(kms) * 1e-3) / 1e9 / HBM_PEAK_GBS, 5),
        "roofline": roofline,
        "lds_kernel": lds_kernel,
        "cpu_baseline": cpu,
    }
    line.update(extra)
    g.free()
    ctx.close()
    print(json.dumps(line), flush=True)


def run_chunglu(args):
    import dcx
    import workloads

    t0 = time.perf_counter()
    graphs = workloads.chunglu_batch()
    E = int(sum(s.size for _, s, _ in graphs))
    Vt = int(sum(V for V, _, _ in graphs))
    print(f"# chunglu256: {len(graphs)} graphs V={Vt} E={E} (gen {time.perf_counter() - t0:.1f}s)",
          file=sys.stderr, flush=True)
    ctx = dcx.Context(0)
    g, vbase = ctx.upload_batch(graphs)
    g.kmax()
    tune = None
    if not args.no_tune:
        r = g.autotune(reps=2)
        tune = {"rows_slots": r.rows_slots, "team_max": r.team_max, "shrink_at": r.shrink_at,
                "lds_members": r.lds_members,
                "candidates": r.candidates, "best_rows_ms": round(r.best_ms, 3),
                "default_rows_ms": round(r.default_ms, 3)}
        print(f"# autotune: {tune}", file=sys.stderr, flush=True)
    for _ in range(args.warmup):
        g.decompose(dcx.VARIANT_BATCHED)
    rows_k, kms, first_k = [], [], []
    t0 = time.perf_counter()
    for _ in range(args.steps):
        K, lmax = g.decompose(dcx.VARIANT_BATCHED)
        rows_k.append(ctx.timing().rows_kernel_ms)
        kms.append(ctx.timing().kmax_ms)
        first_k.append(ctx.timing().rows_first_ms)
    ms_per_step = (time.perf_counter() - t0) / args.steps * 1e3
    timing = ctx.timing()
    print(f"# batched: {ms_per_step:.2f} ms/step kmax={K} lmax={lmax}", file=sys.stderr,
          flush=True)
    kmaxes = g.kmaxes()
    rp_u, vals_u = g.ragged()
    kk, ll = g.batch_summary(vbase)
    # algorithmic bytes of the union's passes (kmax pass + rows)
    src_u = np.concatenate([s.astype(np.int64) + int(vbase[i]) for i, (_, s, _) in
                            enumerate(graphs)])
    dst_u = np.concatenate([d.astype(np.int64) + int(vbase[i]) for i, (_, _, d) in
                            enumerate(graphs)])
    vk, ek = core_sizes(kmaxes, src_u, dst_u, K)
    del src_u, dst_u
    per_row = 8 * ek + 20 * vk
    # row 0 computed beside the kmax pass is not part of the rows kernels
    b_rows = float(per_row[1:].sum() if float(np.mean(first_k)) > 0 else per_row.sum())
    rows_ms = float(np.mean(rows_k))
    g.free()

    # baseline: per-graph, per-k launches (plain variant), bit-exact compare
    nb = min(args.baseline_graphs, len(graphs))
    base_ms, base_e, equal = 0.0, 0, True
    for i in range(nb):
        V, s, d = graphs[i]
        gi = ctx.upload_edges(V, s, d)
        t1 = time.perf_counter()
        Ki, Li = gi.decompose(dcx.VARIANT_PEEL)
        base_ms += (time.perf_counter() - t1) * 1e3
        base_e += s.size
        rp_i, vals_i = gi.ragged()
        a, b = int(vbase[i]), int(vbase[i + 1])
        ok = (Ki == kk[i] and Li == ll[i]
              and np.array_equal(gi.kmaxes(), kmaxes[a:b])
              and np.array_equal(vals_i, vals_u[rp_u[a]:rp_u[b]]))
        equal = equal and bool(ok)
        gi.free()
    print(f"# per-graph loop: {nb} graphs {base_ms:.1f} ms, equal={equal}", file=sys.stderr,
          flush=True)
    base_rate = base_e / (base_ms * 1e-3) / 1e9
    value = E / (ms_per_step * 1e-3) / 1e9
    achieved = b_rows / (rows_ms * 1e-3) / 1e9
    line = {
        "metric": "batched D-core decomposition throughput over 256 Chung-Lu digraphs "
                  "(GEdges/s = sum E / T)",
        "value": round(value, 6),
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
        "config": {"workload": "chunglu256: 256 Chung-Lu digraphs, samples log-uniform "
                               "1e4..2e6, avg degree 8, gamma 2.1, seeds 1000+i",
                   "graphs": len(graphs), "V": Vt, "E": E, "kmax": K, "lmax": lmax,
                   "rows_slots": timing.rows_slots},
        "breakdown_ms": {"kmax_pass": round(float(np.mean(kms)), 3),
                         "order": round(timing.order_ms, 3), "rows_kernel": round(rows_ms, 3)},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 5), "traffic": None,
                     # few rows (kmax <= 4) run as paired whole-graph passes
                     "kernel": "k_rows_teams" if timing.rows_teams_ms > 0 else "k_pass_pair",
                     "kernel_ms": round(rows_ms, 3),
                     "alg_bytes": b_rows},
        "per_graph_baseline": {"graphs": nb, "edges": base_e, "ms": round(base_ms, 2),
                               "GEdges_s": round(base_rate, 6),
                               "variant": "plain (one launch per k per graph)",
                               "output_equal_to_batched": equal},
        "speedup_vs_per_graph": round(value / base_rate, 2),
        "autotune": tune,
        "cpu_baseline": None,
    }
    ctx.close()
    print(json.dumps(line), flush=True)


def run_serve(args):
    """Continuous-batching maintenance of an R-MAT base graph under a seeded
    stream of update batches (sizes log-uniform in [1, 4096], 70% inserts of
    R-MAT-sampled new edges, 30% deletes of existing edges).  Capacity: the
    stream submitted as fast as possible (the server groups what is queued),
    run in four segments with the maintained state checked against a full
    GPU recomputation (dcx_state_hash) after each.  Then Poisson arrivals at
    the given fractions of that capacity on a fresh copy of the base graph;
    latency = completion - scheduled arrival.  With the learned schedule a
    cost model is trained on the base graph first (untimed, the headline's
    protocol) and stays attached: the recomputations a group falls back to
    plan their rows with it."""
    import dcx

    scale, nb = args.serve_scale, args.serve_batches
    V = 1 << scale
    t0 = time.perf_counter()
    src, dst = dcx.gen_rmat(scale, 16 << scale, seed=2)
    io, ip, do, dp = dcx.gen_update_stream(V, src, dst, scale, nb, max_batch=4096,
                                           p_insert=0.7, seed=3)
    reqs = [(ip[io[b]:io[b + 1]], dp[do[b]:do[b + 1]]) for b in range(nb)]
    edits = int(io[-1] + do[-1])
    print(f"# serve: rmat{scale} E={src.size}, {nb} batches, {edits} edits "
          f"(gen {time.perf_counter() - t0:.1f}s)", file=sys.stderr, flush=True)
    ctx = dcx.Context(0)

    def fresh():
        g = ctx.upload_edges(V, src, dst, slack=4)
        g.decompose()
        return g

    model = None
    if args.serve_schedule == "learned":
        def log(msg):
            print(f"# {msg}", file=sys.stderr, flush=True)
        g = fresh()
        model, n_samples, _ = train_learned_plan(dcx, ctx, g, log)
        g.free()

    def collect(sv, res, want, label):
        last = time.perf_counter()
        while len(res) < want:
            for r in sv.poll(wait_ms=500):
                res[r["id"]] = r
            if time.perf_counter() - last > 20:
                last = time.perf_counter()
                print(f"# {label}: {len(res)}/{want} done", file=sys.stderr, flush=True)

    # ---- capacity, with verification between segments
    g = fresh()
    seg = max(1, nb // 4)
    checks = []
    t_busy = 0.0
    done_total = 0
    for s0 in range(0, nb, seg):
        with g.serve() as sv:
            tseg = time.perf_counter()
            ids = [sv.submit(ins, dels) for ins, dels in reqs[s0:s0 + seg]]
            res = {}
            collect(sv, res, len(ids), f"capacity segment {s0 // seg}")
            t_busy += time.perf_counter() - tseg
            st = sv.stop()
        done_total += len(ids)
        bad = sum(1 for r in res.values() if r["status"] != 0)
        h_maint = g.state_hash()
        g.decompose()
        h_full = g.state_hash()
        checks.append({"after_batches": s0 + len(ids), "maintained_equals_recomputed":
                       h_maint == h_full, "failed_requests": bad, "groups": st["groups"]})
        print(f"# capacity segment {s0 // seg}: {len(ids)} requests in {st['groups']} groups, "
              f"state ok={h_maint == h_full}", file=sys.stderr, flush=True)
    capacity = done_total / t_busy
    g.free()

    # ---- Poisson loads
    loads = []
    for frac in [float(x) for x in args.serve_loads.split(",") if x]:
        lam = frac * capacity
        rng = np.random.default_rng(3)
        arrivals = np.cumsum(rng.exponential(1.0 / lam, size=nb))
        g = fresh()
        with g.serve() as sv:
            res, sub_lag = {}, {}
            base = time.perf_counter()
            last = base
            for i, (ins, dels) in enumerate(reqs):
                while True:
                    now = time.perf_counter() - base
                    if now >= arrivals[i]:
                        break
                    for r in sv.poll(wait_ms=0):
                        res[r["id"]] = r
                    time.sleep(min(0.0005, arrivals[i] - now))
                rid = sv.submit(ins, dels)
                sub_lag[rid] = (time.perf_counter() - base) - arrivals[i]
                if time.perf_counter() - last > 20:
                    last = time.perf_counter()
                    print(f"# load {frac}: submitted {i + 1}/{nb}", file=sys.stderr, flush=True)
            collect(sv, res, nb, f"load {frac}")
            wall = time.perf_counter() - base
            st = sv.stop()
        h_maint = g.state_hash()
        g.decompose()
        ok = h_maint == g.state_hash()
        g.free()
        lat = np.array([r["done_ms"] - r["submit_ms"] + 1e3 * sub_lag[r["id"]]
                        for r in res.values()])
        svc = np.array([r["done_ms"] - r["start_ms"] for r in res.values()])
        loads.append({"load": frac, "rate_req_s": round(lam, 2),
                      "p50_ms": round(float(np.percentile(lat, 50)), 3),
                      "p99_ms": round(float(np.percentile(lat, 99)), 3),
                      "service_p50_ms": round(float(np.percentile(svc, 50)), 3),
                      "service_p99_ms": round(float(np.percentile(svc, 99)), 3),
                      "throughput_req_s": round(nb / wall, 2),
                      "throughput_edits_s": round(edits / wall, 1),
                      "groups": st["groups"], "max_group": st["max_group"],
                      "final_state_equals_recomputed": bool(ok),
                      "failed_requests": sum(1 for r in res.values() if r["status"] != 0)})
        print(f"# load {frac}: {loads[-1]}", file=sys.stderr, flush=True)
    ctx.set_scheduler(None)
    if model is not None:
        model.free()
    ctx.close()
    head = loads[-1] if loads else {}
    line = {
        "metric": "p99 edge-batch update latency under continuous batching (ms)",
        "value": head.get("p99_ms"),
        "unit": "ms",
        "n_gpus": 1,
        "steps": nb,
        "warmup": 0,
        "ms_per_step": round(1e3 / capacity, 3),
        "higher_is_better": False,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "int32",
        "data": "synthetic",
        "config": {"workload": f"serve: rmat{scale} seed 2 base, {nb} update batches "
                               "(log-uniform 1..4096 edits, 70% insert / 30% delete, seed 3), "
                               "Poisson arrivals",
                   "V": V, "E": int(src.size), "edits": edits,
                   "load_of_headline": head.get("load"),
                   "schedule": ("learned (cost model trained on the base graph)"
                                if model is not None else "fixed rule")},
        "capacity_req_s": round(capacity, 2),
        "capacity_edits_s": round(edits / t_busy, 1),
        "loads": loads,
        "checks": checks,
        "cpu_baseline": None,
    }
    print(json.dumps(line), flush=True)



def learned_training_set(dcx, log):
    """Per-pass features and measured times the decomposition logs, from
    R-MAT 20 and Chung-Lu 2^20 decomposed at several team caps (so that the
    samples cover the widths the scheduler chooses between)."""
    Xs, Ts = [], []
    small = [("rmat20", 1 << 20) + tuple(dcx.gen_rmat(20, 16 << 20, 11)),
             ("chunglu20", 1 << 20) + tuple(dcx.gen_chunglu(1 << 20, 1 << 24, 12))]
    for name, V, s, d in small:
        for tm in (2, 4, 8, 16, 32, 64):
            with dcx.Context(0, team_max=tm) as c:
                g = c.upload_edges(V, s, d)
                for _ in range(2):
                    g.decompose(dcx.VARIANT_BATCHED)
                for wl in (True, False):
                    X, T = g.pass_features(with_log=wl)
                    Xs.append(X)
                    Ts.append(T)
                g.free()
        log(f"training data: {name} done, {sum(len(t) for t in Ts)} samples")
    return Xs, Ts


def run_learned(args):
    """R-MAT scale 24 and a Chung-Lu digraph of the same size, each decomposed
    with the fixed rule, with the plan of the GPU-trained cost model (trained
    as for the headline), and with the plain one-k-per-launch peel;
    bit-exactness between the three is checked by the device state hash."""
    import dcx

    def log(msg):
        print(f"# {msg}", file=sys.stderr, flush=True)

    sc = args.learned_scale
    V = 1 << sc
    t0 = time.perf_counter()
    ctx = dcx.Context(0)
    results = []
    for name, gen in ((f"rmat{sc}", lambda: ctx.gen_rmat_device(sc, 16 << sc, 4)),
                      (f"chunglu{sc}", lambda: dcx.gen_chunglu(V, 16 << sc, 5))):
        t1 = time.perf_counter()
        s, d = gen()
        if isinstance(s, dcx.DeviceArray):
            g = ctx.ingest(V, s, d, 0, drop_self_loops=True, dedup=True)
            s.free()
            d.free()
        else:
            g = ctx.ingest(V, s, d, 0)
        del s, d
        E = int(g.info()[1])
        log(f"{name}: V={V} E={E} generated and built in {time.perf_counter() - t1:.1f}s")
        # the fixed rule (every 32nd team row peeled, widths by work share)
        ctx.set_scheduler(None)
        heur = []
        for _ in range(3):
            K, lmax = g.decompose(dcx.VARIANT_BATCHED)
            t = ctx.timing()
            heur.append((t.total_ms, t.rows_kernel_ms, t.kmax_ms, t.order_ms))
        h_hash = g.state_hash()
        log(f"{name}: kmax={K} lmax={lmax} fixed rule {[round(x[0], 1) for x in heur]} ms")
        model, n_samples, refine_ms = train_learned_plan(dcx, ctx, g, log)
        learned = []
        for _ in range(3):
            g.decompose(dcx.VARIANT_BATCHED)
            t = ctx.timing()
            learned.append((t.total_ms, t.rows_kernel_ms, t.kmax_ms, t.order_ms,
                            t.rows_scheduler))
        m_hash = g.state_hash()
        ctx.set_scheduler(None)
        X, T = g.pass_features(with_log=True)
        corr = (float(np.corrcoef(model.predict(X), T)[0, 1])
                if model is not None and len(T) > 2 else None)
        if model is not None:
            model.free()
        log(f"{name}: learned {[round(x[0], 1) for x in learned]} ms, "
            f"hash equal {m_hash == h_hash}")
        plain = None
        if not args.no_plain_learned:
            g.decompose(dcx.VARIANT_PEEL)
            tp = ctx.timing()
            plain = {"total_ms": round(tp.total_ms, 2), "rows_ms": round(tp.rows_ms, 2),
                     "launches": tp.rows_launches,
                     "output_equal": g.state_hash() == h_hash}
            log(f"{name}: plain {tp.total_ms:.1f} ms, equal {plain['output_equal']}")
        hb = min(heur)
        lb = min(learned)
        results.append({
            "graph": name, "V": V, "E": E, "kmax": K, "lmax": lmax,
            "fixed_rule_ms": round(hb[0], 2), "fixed_rule_rows_kernel_ms": round(hb[1], 2),
            "learned_ms": round(lb[0], 2), "learned_rows_kernel_ms": round(lb[1], 2),
            "learned_scheduler_used": bool(lb[4]),
            "kmax_pass_ms": round(lb[2], 2), "order_ms": round(lb[3], 2),
            "learned_speedup_vs_fixed_rule": round(hb[0] / lb[0], 3),
            "plain": plain,
            "learned_speedup_vs_plain": round(plain["total_ms"] / lb[0], 2) if plain else None,
            "state_hash": f"{h_hash:016x}",
            "bit_exact_learned_vs_fixed_rule": m_hash == h_hash,
            "model_fit_corr": round(corr, 3) if corr is not None else None,
            "training_samples": n_samples,
        })
        g.free()
    ctx.close()
    E_tot = sum(r["E"] for r in results)
    T_tot = sum(r["learned_ms"] for r in results)
    line = {
        "metric": "learned-scheduled D-core decomposition throughput over R-MAT and Chung-Lu "
                  f"2^{sc} (GEdges/s = sum E / sum T)",
        "value": round(E_tot / (T_tot * 1e-3) / 1e9, 6),
        "unit": "GEdges/s",
        "n_gpus": 1,
        "steps": 3,
        "warmup": 0,
        "ms_per_step": round(T_tot, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "int32",
        "data": "synthetic",
        "config": {"workload": f"learned: R-MAT {sc} (16 * 2^{sc} samples, seed 4) and Chung-Lu "
                               f"V = 2^{sc}, 16 * 2^{sc} samples, gamma 2.1, seed 5; MLP 8-32-32-1 "
                               "trained on the GPU from each graph's logged team rows",
                   "graphs": [r["graph"] for r in results]},
        "graphs": results,
        "wall_s": round(time.perf_counter() - t0, 1),
        "cpu_baseline": None,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
