"""Device ingest (dcx_graph_ingest_edges): the CSR-with-slack graph built on
the GPU must equal, slot for slot, the one the host builds for
dcx_graph_upload_edges -- the layout of Graph::readFile
(src/graph.cpp:195-243): off[v+1]-off[v] = deg + slack, neighbours in input
order, free slots 0xFFFFFFFF, duplicates kept.  Input order within a list is
what a sort could lose, so the cases put many edges of one owner far apart
in the input (hubs, stars, shuffled pairs, repeated pairs).
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


def cases():
    out = []
    s, d = dcx.gen_rmat(14, 16 << 14, 5)
    out.append(("rmat14", 1 << 14, s, d))
    rng = np.random.default_rng(7)
    p = rng.permutation(s.size)
    out.append(("rmat14_shuffled", 1 << 14, s[p], d[p]))
    # repeated pairs, far apart in the input
    rep = rng.choice(s.size, 3000)
    out.append(("rmat14_duplicates", 1 << 14, np.concatenate([s, s[rep]]),
                np.concatenate([d, d[rep]])))
    s, d = dcx.gen_er(1 << 12, 5 << 12, 8)
    out.append(("er_isolated", (1 << 13) + 37, s, d))
    out.append(("out_star", *graphs.out_star(5000)))
    out.append(("in_star", *graphs.in_star(5000)))
    out.append(("complete40", *graphs.complete(40)))
    out.append(("one_edge", 2, np.array([1], np.uint32), np.array([0], np.uint32)))
    out.append(("two_edges", 3, np.array([2, 0], np.uint32), np.array([0, 2], np.uint32)))
    out.append(("no_edges", 5, np.zeros(0, np.uint32), np.zeros(0, np.uint32)))
    out.append(("single_vertex", 1, np.zeros(0, np.uint32), np.zeros(0, np.uint32)))
    return out


@pytest.mark.parametrize("slack", [0, 3])
def test_ingest_equals_host_build(ctx, slack):
    for name, V, s, d in cases():
        say(f"ingest {name} V={V} E={len(s)} slack={slack}")
        a = ctx.upload_edges(V, s, d, slack)
        b = ctx.ingest_edges(V, s, d, slack)
        ca, cb = a.download_csr(), b.download_csr()
        for key in ca:
            assert np.array_equal(ca[key], cb[key]), (name, key)
        assert a.info() == b.info(), name
        a.free()
        b.free()


def test_ingested_graph_decomposes_and_updates(ctx):
    """The ingested graph behaves as an uploaded one: decomposition equal
    to the oracle's, then an edit batch that relocates lists."""
    s, d = dcx.gen_rmat(13, 16 << 13, 9)
    V = 1 << 13
    say("ingest rmat13, decompose")
    g = ctx.ingest_edges(V, s, d, 1)
    K, lmax = g.decompose(dcx.VARIANT_BATCHED)
    ref = oracle.decompose(oracle.csr_from_edges(V, s, d))
    assert (K, lmax) == (ref[0], ref[2])
    assert np.array_equal(g.kmaxes(), ref[1])
    assert np.array_equal(g.rows(), ref[3])
    say("ingest rmat13, edit batch")
    have = set(zip(s.tolist(), d.tolist()))
    rng = np.random.default_rng(2)
    ins = []
    while len(ins) < 400:
        a, b = (int(x) for x in rng.integers(0, 64, 2))
        if a != b and (a, b) not in have:
            have.add((a, b))
            ins.append((a, b))
    ins = np.array(ins, np.uint32)
    g.update(ins=ins, dels=None)
    s2 = np.concatenate([s, ins[:, 0]])
    d2 = np.concatenate([d, ins[:, 1]])
    ref = oracle.decompose(oracle.csr_from_edges(V, s2, d2))
    assert np.array_equal(g.kmaxes(), ref[1])
    assert np.array_equal(g.rows(), ref[3])
    g.free()


@pytest.mark.parametrize("bad", ["self_loop", "out_of_range"])
def test_ingest_rejects_bad_pairs(ctx, bad):
    s = np.array([0, 1, 2, 3, 4], np.uint32)
    d = np.array([1, 2, 3, 4, 0], np.uint32)
    if bad == "self_loop":
        d[3] = 3
    else:
        s[2] = 99
    say(f"ingest rejects {bad}")
    with pytest.raises(dcx.DcxError) as e:
        ctx.ingest_edges(5, s, d)
    assert e.value.status == 1
    assert ("pair 3" if bad == "self_loop" else "pair 2") in str(e.value)
    # the context still works
    g = ctx.ingest_edges(5, np.array([0, 1], np.uint32), np.array([1, 0], np.uint32))
    assert g.decompose(dcx.VARIANT_BATCHED) == (1, 1)
    g.free()


def test_ingest_from_device_arrays(ctx):
    """on_device = 1: the pairs are read where they lie in HBM."""
    s, d = dcx.gen_rmat(14, 16 << 14, 11)
    V = 1 << 14
    say("ingest rmat14 from device arrays")
    ds, dd = ctx.to_device(s), ctx.to_device(d)
    a = ctx.upload_edges(V, s, d, 2)
    b = ctx.ingest_device(V, s.size, ds.ptr.value, dd.ptr.value, 2)
    ca, cb = a.download_csr(), b.download_csr()
    for key in ca:
        assert np.array_equal(ca[key], cb[key]), key
    a.free()
    b.free()
    ds.free()
    dd.free()
