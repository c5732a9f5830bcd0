"""Localized maintenance (the default path of dcx_update and the server):
after every batch the kmaxes and every l-row must equal a full oracle
recomputation of the current edge set -- the reference's criterion,
maintained == recomputed (main.cu:1905-1924, 2194-2211).  The batches mix
deletions and insertions, hit hubs, insert cliques among isolated vertices
(kmaxes that rise by more than one), delete whole cores, and grow and
shrink kmax.  These streams run without the work budget
(dcx_config.maintenance = 2) and every batch must take the local path (no
fallback) unless stated otherwise.  The recompute path (maintenance = 1) and
the budgeted default (maintenance = 0, with a budget small enough that some
batches are given to recomputation part way) give the same states.
"""
import sys

import numpy as np
import pytest

import dcx
import graphs
import oracle
from test_gpu_update import EdgeSet, check_against_oracle

pytestmark = pytest.mark.gpu


def say(msg):
    print(f"[case] {msg}", flush=True)
    sys.stdout.flush()


def run_stream(ctx, name, V, s, d, steps, make_batch, slack=2, seed=0, expect_local=True):
    es = EdgeSet(V, s, d)
    g = ctx.upload_edges(V, np.asarray(s), np.asarray(d), slack)
    g.decompose()
    rng = np.random.default_rng(seed)
    paths = []
    for step in range(steps):
        ins, dels = make_batch(rng, es, step)
        say(f"{name} batch {step}: {len(ins)} inserts, {len(dels)} deletes")
        st = g.update(ins=np.array(ins, np.uint32).reshape(-1, 2),
                      dels=np.array(dels, np.uint32).reshape(-1, 2))
        paths.append(st.path)
        es.apply(ins, dels)
        check_against_oracle(g, es, f"{name} batch {step}")
    g.free()
    if expect_local:
        assert all(p == 1 for p in paths), paths
    return paths


def mixed(n_ins, n_del, hub_bias=False):
    def make(rng, es, step):
        return es.random_inserts(rng, n_ins, hub_bias), es.random_deletes(rng, n_del)
    return make


@pytest.fixture(scope="module")
def ctx():
    c = dcx.Context(0, maintenance=2)
    yield c
    c.close()


@pytest.mark.parametrize("seed", range(3))
def test_local_random_mixed(ctx, seed):
    V, s, d = graphs.random_gnm(400, 5000, 100 + seed)
    run_stream(ctx, f"gnm400 seed {seed}", V, s, d, 12, mixed(40, 25), seed=seed)


def test_local_single_edges(ctx):
    """One edge per batch, as the reference's SINGLE_INSERT/DELETE modes."""
    V, s, d = graphs.random_gnm(200, 2200, 7)

    def make(rng, es, step):
        if step % 2 == 0:
            return es.random_inserts(rng, 1, hub_bias=True), []
        return [], es.random_deletes(rng, 1)
    run_stream(ctx, "single edges", V, s, d, 30, make, seed=3)


def test_local_rmat_hubs(ctx):
    s, d = dcx.gen_rmat(12, 16 << 12, 5)
    V = 1 << 12
    run_stream(ctx, "rmat12 hub-biased", V, s, d, 8, mixed(300, 150, hub_bias=True), seed=5)


def test_local_clique_among_isolated(ctx):
    """A clique of 12 vertices inserted among isolated vertices: their kmaxes
    rise from 0 to 11 in one batch (many rounds of +1), rows appear above
    the old kmax; then the clique is deleted again (kmax falls back)."""
    V, s, d = graphs.random_gnm(300, 900, 9)
    iso = list(range(300, 312))
    V2 = 312
    clique = [(a, b) for a in iso for b in iso if a != b]

    def make(rng, es, step):
        if step == 0:
            return clique, []
        if step == 1:
            return es.random_inserts(rng, 20), es.random_deletes(rng, 10)
        return [], [e for e in clique if e in es.edges]
    run_stream(ctx, "clique", V2, s, d, 3, make, seed=2)


def test_local_deletes_whole_core(ctx):
    """Deleting every edge of the densest core: kmax falls, rows above it
    empty out."""
    V, s, d = graphs.complete(15)
    s2, d2 = graphs.random_gnm(200, 1500, 4)[1:]
    s = np.concatenate([np.asarray(s), np.asarray(s2) + 15])
    d = np.concatenate([np.asarray(d), np.asarray(d2) + 15])
    V = 215

    def make(rng, es, step):
        if step == 0:
            return [], [e for e in es.edges if e[0] < 15 and e[1] < 15][:100]
        return [], [e for e in es.edges if e[0] < 15 and e[1] < 15]
    run_stream(ctx, "delete core", V, s, d, 2, make, seed=1)


def test_local_insert_and_delete_same_edge(ctx):
    """An edge deleted and inserted again in one batch is no net change;
    the batch still maintains the rest."""
    V, s, d = graphs.random_gnm(150, 1500, 11)

    def make(rng, es, step):
        dels = es.random_deletes(rng, 8)
        return dels[:3] + es.random_inserts(rng, 10), dels
    run_stream(ctx, "reinsert", V, s, d, 4, make, seed=8)


def test_local_equals_recompute_path():
    """maintenance = 2 (local, no budget), 1 (recompute) and 0 with a budget
    of one list entry per edge (local, abandoned for recomputation once a
    batch's tasks have read E list entries) give the same state hash after
    every batch of an R-MAT 13 update stream; the budgeted run keeps some
    batches local and gives some up (and sends some straight to
    recomputation while it backs off)."""
    V = 1 << 13
    s, d = dcx.gen_rmat(13, 16 << 13, 2)
    io, ip, do, dp = dcx.gen_update_stream(V, s, d, 13, 30, max_batch=256, seed=4)
    hashes, paths = {}, {}
    for mode in (2, 1, 0):
        with dcx.Context(0, maintenance=mode, local_budget=1 if mode == 0 else 0) as c:
            g = c.upload_edges(V, s, d, slack=2)
            g.decompose()
            out, ps = [], []
            for b in range(30):
                say(f"rmat13 stream mode {mode} batch {b}")
                st = g.update(ins=ip[io[b]:io[b + 1]], dels=dp[do[b]:do[b + 1]])
                ps.append(st.path)
                out.append(g.state_hash())
            hashes[mode], paths[mode] = out, ps
            g.free()
    assert all(p == 1 for p in paths[2]), paths[2]
    assert all(p == 0 for p in paths[1]), paths[1]
    assert {1, 2} <= set(paths[0]) <= {0, 1, 2}, paths[0]
    assert hashes[2] == hashes[1]
    assert hashes[0] == hashes[1]
    es_s, es_d = s, d
    # and the last state equals the oracle's
    cur = set(zip(es_s.tolist(), es_d.tolist()))
    for b in range(30):
        cur.difference_update(map(tuple, dp[do[b]:do[b + 1]].tolist()))
        cur.update(map(tuple, ip[io[b]:io[b + 1]].tolist()))
    e = np.array(sorted(cur), np.uint32)
    ref = oracle.decompose(oracle.csr_from_edges(V, e[:, 0], e[:, 1]))
    assert hashes[2][-1] == graphs.state_hash(ref[1], ref[3])


@pytest.mark.parametrize("mode", [0, 1])
def test_device_edge_batches_equal_host(ctx, mode):
    """getKmaxEdgeBatches on the device (from the resident kmaxes) equals the
    host restatement -- itself pinned by hand-derived batches -- edge for
    edge: R-MAT 12 with random pairs, a list dominated by one hub root,
    pairs among the densest (lowest-id) vertices, and an empty list."""
    s, d = dcx.gen_rmat(12, 16 << 12, 9)
    V = 1 << 12
    g = ctx.upload_edges(V, s, d)
    g.decompose()
    km = g.kmaxes()
    rng = np.random.default_rng(mode)
    cases = {
        "random": rng.integers(0, V, (3000, 2)).astype(np.uint32),
        "hub": np.stack([np.full(500, 0), rng.integers(1, V, 500)], 1).astype(np.uint32),
        "low ids": rng.integers(0, 64, (2000, 2)).astype(np.uint32),
        "empty": np.zeros((0, 2), np.uint32),
    }
    for name, p in cases.items():
        say(f"device edge batches mode {mode}: {name} ({len(p)} pairs)")
        want, nb_want = dcx.kmax_edge_batches(V, km, p, mode=mode)
        got, nb_got = g.kmax_edge_batches(p, mode=mode)
        assert nb_got == nb_want, name
        assert np.array_equal(got, want), name
    g.free()
