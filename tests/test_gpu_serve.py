"""Continuous-batching maintenance server on the GPU.

* every request's status is what it would get alone;
* the maintained decomposition after each group equals the oracle's full
  recomputation of the graph built from the same edits (state hashes,
  restated on the CPU from the oracle's arrays);
* a request's result is the same alone (one request per group) and in the
  mix (many requests per group, submitted in a burst): for every request
  that closes a group in the mix, its state hash equals the hash after the
  same request when served alone;
* admission by HBM budget: with a pool far smaller than the stream, requests
  wait for room and still all complete, in order, with the same results.
"""
import sys

import numpy as np
import pytest

import dcx
import graphs
import oracle

pytestmark = pytest.mark.gpu


def say(msg):
    print(f"[case] {msg}", flush=True)
    sys.stdout.flush()


def stream(scale=11, nb=40, max_batch=48, seed=5):
    V = 1 << scale
    src, dst = dcx.gen_rmat(scale, 16 << scale, 2)
    io, ip, do, dp = dcx.gen_update_stream(V, src, dst, scale, nb, max_batch=max_batch,
                                           seed=seed)
    reqs = [(ip[io[b]:io[b + 1]], dp[do[b]:do[b + 1]]) for b in range(nb)]
    return V, src, dst, reqs


def oracle_hashes(V, src, dst, reqs):
    """Hash after each request, by full recomputation on the CPU."""
    cur = set(zip(src.tolist(), dst.tolist()))
    out = []
    for ins, dels in reqs:
        cur.difference_update(map(tuple, dels.tolist()))
        cur.update(map(tuple, ins.tolist()))
        e = np.array(sorted(cur), np.uint32).reshape(-1, 2)
        K, km, lmax, rows = oracle.decompose(oracle.csr_from_edges(V, e[:, 0], e[:, 1]))
        out.append(graphs.state_hash(km, rows))
    return out


def serve(V, src, dst, reqs, burst, **kw):
    with dcx.Context(0) as c:
        g = c.upload_edges(V, src, dst, slack=2)
        g.decompose()
        with g.serve(hash_states=True, **kw) as sv:
            ids, res = [], {}
            for ins, dels in reqs:
                ids.append(sv.submit(ins, dels))
                if not burst:  # wait for it: one request per group
                    while ids[-1] not in res:
                        for r in sv.poll(wait_ms=200):
                            res[r["id"]] = r
            while len(res) < len(ids):
                for r in sv.poll(wait_ms=200):
                    res[r["id"]] = r
            stats = sv.stop()
        final = g.state_hash()
        g.free()
    return [res[i] for i in ids], stats, final


def test_state_hash_matches_restatement():
    V, src, dst, _ = stream(nb=1)
    ref = oracle.decompose(oracle.csr_from_edges(V, src, dst))
    say("state hash vs CPU restatement")
    with dcx.Context(0) as c:
        g = c.upload_edges(V, src, dst)
        g.decompose()
        h = g.state_hash()
        g.free()
    assert h == graphs.state_hash(ref[1], ref[3])


def test_serve_alone_and_in_the_mix_equal_recomputation():
    V, src, dst, reqs = stream()
    want = oracle_hashes(V, src, dst, reqs)
    say("serve alone (max_group=1)")
    alone, st_a, fin_a = serve(V, src, dst, reqs, burst=False, max_group=1)
    assert all(r["status"] == 0 for r in alone)
    assert [r["state_hash"] for r in alone] == want
    assert st_a["groups"] == len(reqs)
    say("serve in the mix (burst, max_group=16)")
    mix, st_m, fin_m = serve(V, src, dst, reqs, burst=True, max_group=16)
    assert [r["status"] for r in mix] == [r["status"] for r in alone]
    assert st_m["groups"] < len(reqs)  # requests really were grouped
    for i, r in enumerate(mix):
        last_of_group = i + 1 == len(mix) or mix[i + 1]["group"] != r["group"]
        if last_of_group:
            assert r["state_hash"] == alone[i]["state_hash"], i
        assert r["done_ms"] >= r["start_ms"] >= r["admit_ms"] >= r["submit_ms"]
    assert fin_m == fin_a == want[-1]


def test_serve_statuses_and_budget_admission():
    V, src, dst, reqs = stream(nb=24, seed=9)
    # plant requests that must fail alone: a delete of a missing edge, an
    # insert of an existing edge, and a self-loop (rejected at submit)
    present = set(zip(src.tolist(), dst.tolist()))
    for ins, _ in reqs:
        present.update(map(tuple, ins.tolist()))
    missing = next((0, v) for v in range(V - 1, 0, -1) if (0, v) not in present)
    none = np.zeros((0, 2), np.uint32)
    bad_missing = (none, np.array([missing], np.uint32))
    e0 = np.array([[src[0], dst[0]]], np.uint32)
    bad_dup = (e0, none)
    bad_loop = (np.array([[3, 3]], np.uint32), none)
    # the duplicate's edge may have been deleted by an earlier request (then
    # its insert is valid), so statuses are checked against the alone run
    mixed = (reqs[:5] + [bad_missing] + reqs[5:9] + [bad_loop] + reqs[9:12] + [bad_dup]
             + reqs[12:])
    say("statuses alone")
    alone, _, fin_a = serve(V, src, dst, mixed, burst=False, max_group=1)
    st = [r["status"] for r in alone]
    assert st[5] == 8  # DCX_ERR_NOT_FOUND
    assert st[10] == 1 and alone[10]["group"] == 0  # DCX_ERR_INVALID_ARG at submit
    say("statuses in the mix with a 6 KiB HBM budget (the largest request needs ~4.2 KiB)")
    mix, stats, fin_m = serve(V, src, dst, mixed, burst=True, max_group=8, hbm_budget=6 << 10)
    assert [r["status"] for r in mix] == st
    assert stats["max_waiting"] > 0  # the budget held requests back
    assert fin_m == fin_a
    gids = [r["group"] for r in mix if r["group"]]
    assert gids == sorted(gids)  # groups complete in arrival order


def test_serve_requests_that_depend_on_each_other_in_one_group():
    """Requests of one group that touch the same edges: a delete, a
    re-insert, a second delete, an insert of a new edge, a second insert of
    it (must fail: the edge exists after the first), a delete of an edge that
    was never there (must fail), a delete plus re-insert in one request.
    Statuses and the state after each request equal the one-request-per-
    group run, and the final state equals the oracle's recomputation."""
    V, src, dst, reqs = stream(nb=6, seed=11)
    # pairs the base graph or any request of the stream holds
    seen = set(zip(src.tolist(), dst.tolist()))
    for ins, dels in reqs:
        seen.update(map(tuple, dels.tolist()))
        seen.update(map(tuple, ins.tolist()))
    e = (int(src[1]), int(dst[1]))
    new = next((u, V - 1 - u) for u in range(V) if u != V - 1 - u and (u, V - 1 - u) not in seen)
    gone = next((V - 1, v) for v in range(V - 1) if (V - 1, v) not in seen and (V - 1, v) != new)
    none = np.zeros((0, 2), np.uint32)
    one = lambda p: np.array([p], np.uint32)
    dep = [(none, one(e)), (one(e), none), (none, one(e)), (one(new), none), (one(new), none),
           (none, one(gone)), (one(new), one(new))]
    # the first request must find e present: keep it out of the stream's edits
    base = [(ins, dels) for ins, dels in reqs
            if e not in set(map(tuple, dels.tolist())) | set(map(tuple, ins.tolist()))]
    mixed = base[:3] + dep + base[3:]
    say("dependent requests alone")
    alone, _, fin_a = serve(V, src, dst, mixed, burst=False, max_group=1)
    st = [r["status"] for r in alone]
    k = len(base[:3])
    assert st[k:k + 7] == [0, 0, 0, 0, 1, 8, 0], st
    say("dependent requests in one group")
    mix, stats, fin_m = serve(V, src, dst, mixed, burst=True, max_group=64)
    assert [r["status"] for r in mix] == st
    assert len({r["group"] for r in mix[k:k + 7]}) < 7  # grouped
    for i, r in enumerate(mix):
        last_of_group = i + 1 == len(mix) or mix[i + 1]["group"] != r["group"]
        if last_of_group:
            assert r["state_hash"] == alone[i]["state_hash"], i
    assert fin_m == fin_a
    accepted = [q for q, s in zip(mixed, st) if s == 0]
    assert fin_a == oracle_hashes(V, src, dst, accepted)[-1]
