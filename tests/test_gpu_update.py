"""GPU edge-update maintenance: after every batch the maintained kmaxes and
l-rows must equal a full oracle recomputation on the current edge set.

This is the reference's STEP_TEST / SINGLE_INSERT_CHECK criterion
(main.cu:2176-2214, 1905-1924: maintained == recomputed), extended to
deletions and mixed batches.
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


@pytest.fixture(scope="module")
def ctx():
    c = dcx.Context(0)
    yield c
    c.close()


class EdgeSet:
    """Host mirror of the current edge set (the test's bookkeeping only)."""

    def __init__(self, V, s, d):
        self.V = V
        self.edges = set(zip(np.asarray(s).tolist(), np.asarray(d).tolist()))

    def arrays(self):
        e = sorted(self.edges)
        s = np.array([a for a, _ in e], np.int64)
        d = np.array([b for _, b in e], np.int64)
        return s, d

    def random_inserts(self, rng, n, hub_bias=False):
        out = set()
        while len(out) < n:
            if hub_bias:
                u = int(min(rng.integers(self.V), rng.integers(self.V)))
                v = int(min(rng.integers(self.V), rng.integers(self.V)))
            else:
                u, v = int(rng.integers(self.V)), int(rng.integers(self.V))
            if u != v and (u, v) not in self.edges and (u, v) not in out:
                out.add((u, v))
        return sorted(out)

    def random_deletes(self, rng, n):
        e = sorted(self.edges)
        idx = rng.choice(len(e), size=min(n, len(e)), replace=False)
        return [e[i] for i in sorted(idx)]

    def apply(self, ins, dels):
        for p in dels:
            self.edges.remove(tuple(p))
        for p in ins:
            self.edges.add(tuple(p))


def check_against_oracle(g, es, tag):
    s, d = es.arrays()
    ref = oracle.decompose(oracle.csr_from_edges(es.V, s, d))
    V, E, K, lmax = g.info()
    assert E == len(es.edges), tag
    assert K == ref[0], (tag, K, ref[0])
    assert lmax == ref[2], (tag, lmax, ref[2])
    assert np.array_equal(g.kmaxes(), ref[1]), tag
    assert np.array_equal(g.rows(), ref[3]), tag


@pytest.mark.parametrize("slack", [0, 2])
def test_mixed_batches_rmat(ctx, slack):
    rng = np.random.default_rng(11 + slack)
    s, d = dcx.gen_rmat(11, 12 << 11, seed=21)
    V = 1 << 11
    es = EdgeSet(V, s, d)
    g = ctx.upload_edges(V, s, d, slack=slack)
    say(f"update base rmat11 slack={slack}")
    g.decompose()
    check_against_oracle(g, es, "base")
    for step, (ni, nd) in enumerate([(1, 0), (0, 1), (7, 3), (64, 0), (0, 40), (300, 120),
                                     (1, 1), (900, 0), (0, 700)]):
        ins = es.random_inserts(rng, ni, hub_bias=step % 2 == 0)
        dels = es.random_deletes(rng, nd)
        say(f"update step {step}: +{ni} -{nd}")
        st = g.update(ins, dels)
        es.apply(ins, dels)
        check_against_oracle(g, es, f"step{step}")
        assert st.inserted == ni and st.deleted == nd
    g.free()


def test_step_build_from_empty(ctx):
    """STEP_TEST: start from V vertices without edges, insert the edges of a
    graph batch by batch, compare after every batch."""
    V, s, d = graphs.random_gnm(60, 400, 3)
    order = np.random.default_rng(4).permutation(len(s))
    es = EdgeSet(V, [], [])
    g = ctx.upload_edges(V, np.zeros(0, np.uint32), np.zeros(0, np.uint32), slack=0)
    g.decompose()
    check_against_oracle(g, es, "empty")
    for b in range(0, len(order), 25):
        idx = order[b:b + 25]
        ins = [(int(s[i]), int(d[i])) for i in idx]
        say(f"step-build batch at {b}")
        g.update(ins, None)
        es.apply(ins, [])
        check_against_oracle(g, es, f"build{b}")
    # and tear it down again with deletions
    for b in range(0, len(order), 40):
        idx = order[b:b + 40]
        dels = [(int(s[i]), int(d[i])) for i in idx]
        say(f"step-delete batch at {b}")
        g.update(None, dels)
        es.apply([], dels)
        check_against_oracle(g, es, f"teardown{b}")
    g.free()


def test_clique_among_isolated(ctx):
    """A batch can raise kmax by more than one: a 6-clique inserted among
    isolated vertices takes kmax from 0 to 5."""
    V, s, d = graphs.random_gnm(40, 60, 8)
    es = EdgeSet(V + 6, s, d)
    g = ctx.upload_edges(V + 6, s, d)
    g.decompose()
    clique = [(V + i, V + j) for i in range(6) for j in range(6) if i != j]
    say("clique insert")
    st = g.update(clique, None)
    es.apply(clique, [])
    check_against_oracle(g, es, "clique")
    assert st.kmax_after >= 5
    g.free()


def test_update_errors_leave_graph_unchanged(ctx):
    V, s, d = graphs.random_gnm(50, 300, 9)
    es = EdgeSet(V, s, d)
    g = ctx.upload_edges(V, s, d)
    g.decompose()
    e0 = next(iter(sorted(es.edges)))
    missing = next((u, v) for u in range(V) for v in range(V)
                   if u != v and (u, v) not in es.edges)
    cases = [
        (dict(ins=None, dels=[missing]), 8),   # absent edge
        (dict(ins=[e0], dels=None), 1),        # already present
        (dict(ins=[(3, 3)], dels=None), 1),    # self-loop
        (dict(ins=[(0, V)], dels=None), 1),    # id out of range
        (dict(ins=[missing, missing], dels=None), 1),  # listed twice
    ]
    for kw, code in cases:
        say(f"update error case {kw}")
        with pytest.raises(dcx.DcxError) as e:
            g.update(**kw)
        assert e.value.status == code
        check_against_oracle(g, es, "after error")
    # delete and re-insert the same edge in one batch is allowed
    st = g.update([e0], [e0])
    check_against_oracle(g, es, "delete+reinsert")
    g.free()


def test_structural_edits_and_download(ctx):
    V, s, d = graphs.random_gnm(200, 1500, 10)
    es = EdgeSet(V, s, d)
    g = ctx.upload_edges(V, s, d, slack=1)
    rng = np.random.default_rng(5)
    ins = es.random_inserts(rng, 150, hub_bias=True)
    dels = es.random_deletes(rng, 100)
    say("structural delete+insert")
    g.delete_edges(dels)
    g.insert_edges(ins)
    es.apply(ins, dels)
    csr = g.download_csr()
    got = set()
    for u in range(V):
        a, n = int(csr["out_off"][u]), int(csr["out_deg"][u])
        got.update((u, int(x)) for x in csr["out_nbr"][a:a + n])
    assert got == es.edges
    got_in = set()
    for v in range(V):
        a, n = int(csr["in_off"][v]), int(csr["in_deg"][v])
        got_in.update((int(x), v) for x in csr["in_nbr"][a:a + n])
    assert got_in == es.edges
    assert g.info()[2] == -1  # decomposition dropped by structural edits
    g.decompose()
    check_against_oracle(g, es, "after structural")
    g.free()
