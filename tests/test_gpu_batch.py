"""Many graphs in one decomposition, and the autotuner (GPU, bit-exact).

A batch of uneven graphs is uploaded as one disjoint union and decomposed
by the same launches as one graph; each graph's slice of the union's
kmaxes and rows must equal that graph's own oracle decomposition (rows
above its own kmax are zero), and the per-graph summary must equal the
oracle's (kmax, lmax).  The autotuner may pick any launch shape but the
bits must not move.
"""
import os
import sys

import numpy as np
import pytest

import dcx
import oracle
import workloads
from conftest import GOLDEN

pytestmark = pytest.mark.gpu


def say(msg):
    print(f"[case] {msg}", flush=True)
    sys.stdout.flush()


def small_batch():
    """Uneven graphs, with the edge cases a batch can hold: an empty graph,
    an edgeless one, a single vertex, the reference fixture."""
    graphs = [(5, np.zeros(0, np.uint32), np.zeros(0, np.uint32)),
              (1, np.zeros(0, np.uint32), np.zeros(0, np.uint32))]
    V, s, d = dcx.text_read(os.path.join(GOLDEN, "digraph"), prune=False)
    graphs.append((V, s, d))
    graphs += workloads.chunglu_batch(n=10, seed0=77, m_lo=2e3, m_hi=6e4)
    src, dst = dcx.gen_rmat(10, 16 << 10, 11)
    graphs.append((1 << 10, src, dst))
    graphs.append((0, np.zeros(0, np.uint32), np.zeros(0, np.uint32)))
    return graphs


@pytest.mark.parametrize("variant", [dcx.VARIANT_BATCHED, dcx.VARIANT_PEEL, dcx.VARIANT_HINDEX])
def test_union_batch_equals_each_graph(variant):
    graphs = small_batch()
    refs = [oracle.decompose(oracle.csr_from_edges(V, s, d)) for V, s, d in graphs]
    say(f"batch of {len(graphs)} graphs variant={variant}")
    with dcx.Context(0) as c:
        g, vbase = c.upload_batch(graphs)
        K, lmax = g.decompose(variant)
        km = g.kmaxes()
        rows = g.rows()
        kk, ll = g.batch_summary(vbase)
        g.free()
    assert K == max(r[0] for r in refs)
    assert lmax == max(r[2] for r in refs)
    for i, ((V, _, _), ref) in enumerate(zip(graphs, refs)):
        a, b = int(vbase[i]), int(vbase[i + 1])
        assert b - a == V
        assert np.array_equal(km[a:b], ref[1]), i
        Ki = ref[0]
        assert np.array_equal(rows[:Ki + 1, a:b], ref[3]), i
        assert not rows[Ki + 1:, a:b].any(), i
        assert kk[i] == Ki and ll[i] == ref[2], i


def test_autotune_keeps_bits():
    src, dst = dcx.gen_rmat(14, 16 << 14, 21)
    V = 1 << 14
    ref = oracle.decompose(oracle.csr_from_edges(V, src, dst))
    say("rmat14 autotune")
    with dcx.Context(0) as c:
        g = c.upload_edges(V, src, dst)
        g.kmax()
        r = g.autotune(reps=1)
        assert r.candidates >= 10
        assert 0 < r.best_ms <= r.default_ms
        assert r.team_max in (4, 8, 16, 32, 64, 128, 256)
        assert r.shrink_at in (1024, 2048, 4096, 8192, 16384)
        # the tuned rows are resident, and a new decomposition uses the shape
        assert np.array_equal(g.rows(), ref[3])
        K, lmax = g.decompose(dcx.VARIANT_BATCHED)
        assert K == ref[0] and lmax == ref[2]
        assert np.array_equal(g.kmaxes(), ref[1]) and np.array_equal(g.rows(), ref[3])
        assert c.timing().rows_slots <= max(r.rows_slots, 1)
        g.free()
