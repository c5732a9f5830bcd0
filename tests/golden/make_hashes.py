"""Writes tests/golden/hashes.json: oracle state hashes of the full-size
workloads, the data the GPU parity tests at those sizes compare against.

Each entry is the dcx_state_hash formula (include/dcx.h) evaluated by the
CPU oracle's streaming mode (oracle/dcore_oracle.c, dco_state_hash), which
never holds the dense rows, so the scale-24 graphs fit this machine.  Graphs
come from the library's own seeded host generators (no GPU needed):

* rmat24_s4:    R-MAT scale 24, 2^28 samples, seed 4 (learned-scheduling
                workload, R-MAT part);
* chunglu24_s5: Chung-Lu, V = 2^24, 2^28 samples, exponent 2.1, seed 5
                (learned-scheduling workload, Chung-Lu part);
* chunglu256:   the disjoint union of the 256 Chung-Lu graphs of
                workloads.chunglu_batch(), graph i at ids [vbase[i],
                vbase[i+1]) (the batched workload), with each graph's own
                kmax and lmax;
* serve_rmat22: R-MAT scale 22 seed 2 under dcx_gen_update_stream(10,000
                batches, max 4,096 edits, 70 % inserts, seed 3), the state
                after requests 0, 9, 99, 999, 4999 and 9999 (the serving
                workload);
* rmat20_s1:    R-MAT scale 20 seed 1 (the plain GPU workload).

Run from the repository root (about 30 minutes on 8 cores, ~40 GB of RAM):
    python tests/golden/make_hashes.py [names...]
"""
import json
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
for sub in ("d-core-cuda_amd", "oracle"):
    sys.path.insert(0, os.path.join(ROOT, sub))

import dcx  # noqa: E402  (host generators only)
import oracle  # noqa: E402
import workloads  # noqa: E402

OUT = os.path.join(HERE, "hashes.json")
SERVE_CHECKPOINTS = [0, 9, 99, 999, 4999, 9999]


def log(msg):
    print(f"[{time.strftime('%H:%M:%S')}] {msg}", flush=True)


def hash_graph(V, src, dst):
    t = time.time()
    g = oracle.csr_from_edges(V, src, dst)
    h, K, L = oracle.state_hash(g)
    log(f"  V={V} E={src.size} kmax={K} lmax={L} hash={h:016x} ({time.time() - t:.0f} s)")
    return {"V": int(V), "E": int(src.size), "kmax": K, "lmax": L, "hash": f"{h:016x}"}


def rmat(scale, seed):
    V, s, d = workloads.rmat(scale, seed)
    return dict(hash_graph(V, s, d), scale=scale, samples=16 << scale, seed=seed)


def chunglu24():
    V = 1 << 24
    s, d = dcx.gen_chunglu(V, 1 << 28, 5)
    return dict(hash_graph(V, s, d), samples=1 << 28, seed=5, gamma=2.1)


def chunglu256():
    graphs = workloads.chunglu_batch()
    vbase = np.zeros(len(graphs) + 1, np.int64)
    for i, (V, _, _) in enumerate(graphs):
        vbase[i + 1] = vbase[i] + V
    src = np.concatenate([s.astype(np.int64) + vbase[i] for i, (_, s, _) in enumerate(graphs)])
    dst = np.concatenate([d.astype(np.int64) + vbase[i] for i, (_, _, d) in enumerate(graphs)])
    out = hash_graph(int(vbase[-1]), src, dst)
    per = []
    for V, s, d in graphs:
        K, km, L, _ = oracle.decompose(oracle.csr_from_edges(V, s, d))
        per.append([K, L])
    out["per_graph_kmax_lmax"] = per
    out["n"] = len(graphs)
    return out


def serve22():
    V, s, d = workloads.rmat(22, 2)
    nb = 10000
    io, ip, do, dp = dcx.gen_update_stream(V, s, d, 22, nb, max_batch=4096, p_insert=0.7,
                                           seed=3)
    log(f"  stream: {int(io[-1])} inserts, {int(do[-1])} deletes")
    # edge set as sorted uint64 keys; a span of requests is applied at once
    # by keeping the last operation on each key (delete, then insert, per
    # request, in request order)
    cur = np.unique((s.astype(np.uint64) << np.uint64(32)) | d.astype(np.uint64))
    out = {"V": V, "base_E": int(s.size), "nbatches": nb, "max_batch": 4096, "p_insert": 0.7,
           "stream_seed": 3, "base_seed": 2, "checkpoints": {}}
    done = -1
    for cp in SERVE_CHECKPOINTS:
        keys, ops, seq = [], [], []
        for b in range(done + 1, cp + 1):
            dl = dp[do[b]:do[b + 1]]
            il = ip[io[b]:io[b + 1]]
            for arr, op, sub in ((dl, 0, 0), (il, 1, 1)):
                if arr.size:
                    keys.append((arr[:, 0].astype(np.uint64) << np.uint64(32)) |
                                arr[:, 1].astype(np.uint64))
                    ops.append(np.full(arr.shape[0], op, np.int8))
                    seq.append(np.full(arr.shape[0], 2 * b + sub, np.int64))
        done = cp
        if keys:
            k = np.concatenate(keys)
            o = np.concatenate(ops)
            q = np.concatenate(seq)
            order = np.lexsort((q, k))
            k, o = k[order], o[order]
            last = np.r_[k[1:] != k[:-1], True]
            k, o = k[last], o[last]
            cur = np.union1d(np.setdiff1d(cur, k[o == 0], assume_unique=True), k[o == 1])
        log(f"serve_rmat22 after request {cp}: E={cur.size}")
        e = hash_graph(V, (cur >> np.uint64(32)).astype(np.uint32),
                       (cur & np.uint64(0xffffffff)).astype(np.uint32))
        out["checkpoints"][str(cp)] = e
    return out


JOBS = {
    "rmat20_s1": lambda: rmat(20, 1),
    "chunglu256": chunglu256,
    "serve_rmat22": serve22,
    "rmat24_s4": lambda: rmat(24, 4),
    "chunglu24_s5": chunglu24,
}


def main(names):
    data = json.load(open(OUT)) if os.path.exists(OUT) else {}
    for name in names or list(JOBS):
        log(f"{name} ...")
        data[name] = JOBS[name]()
        with open(OUT + ".tmp", "w") as f:
            json.dump(data, f, indent=1, sort_keys=True)
        os.replace(OUT + ".tmp", OUT)
    log("done")


if __name__ == "__main__":
    main(sys.argv[1:])
