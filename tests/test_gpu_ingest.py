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
    b = ctx.ingest_device(V, ds, dd, 2)
    same_csr(a, b, "rmat14 device")
    a.free()
    b.free()
    # a prefix of the device arrays
    say("ingest rmat14 prefix from device arrays")
    a = ctx.upload_edges(V, s[:1000], d[:1000], 0)
    b = ctx.ingest_device(V, ds, dd, 0, E=1000)
    same_csr(a, b, "rmat14 device prefix")
    a.free()
    b.free()
    with pytest.raises(ValueError):
        ctx.ingest_device(V, ds, dd, 0, E=s.size + 1)
    ds.free()
    dd.free()


def same_csr(a, b, name):
    ca, cb = a.download_csr(), b.download_csr()
    for key in ca:
        assert np.array_equal(ca[key], cb[key]), (name, key)
    assert a.info() == b.info(), name


def test_ingest_rmat18_slot_for_slot(ctx):
    """R-MAT 18 (more than 2^22 pairs: three digit passes of the owner sort,
    the owner bits starting off a byte boundary), shuffled so that every
    hub's edges lie far apart in the input."""
    s, d = dcx.gen_rmat(18, 16 << 18, 21)
    V = 1 << 18
    p = np.random.default_rng(3).permutation(s.size)
    s, d = s[p], d[p]
    for slack in (0, 1):
        say(f"ingest rmat18 shuffled E={s.size} slack={slack}")
        a = ctx.upload_edges(V, s, d, slack)
        b = ctx.ingest(V, s, d, slack)
        same_csr(a, b, f"rmat18 slack={slack}")
        a.free()
        b.free()


@pytest.mark.parametrize("dedup", [False, True])
def test_ingest_drops_self_loops(ctx, dedup):
    """With self-loops dropped (Graph::readFile skips them,
    src/graph.cpp:167) the graph is the host build of the list without
    them; with dedup as well, of the sorted list of distinct pairs."""
    rng = np.random.default_rng(5)
    V = 3000
    s = rng.integers(0, V, 60000).astype(np.uint32)
    d = rng.integers(0, V, 60000).astype(np.uint32)
    d[::7] = s[::7]  # many self-loops
    s = np.concatenate([s, s[:5000]])  # and repeated pairs
    d = np.concatenate([d, d[:5000]])
    keep = s != d
    if dedup:
        key = np.unique((s[keep].astype(np.uint64) << 32) | d[keep])
        hs, hd = (key >> 32).astype(np.uint32), (key & 0xffffffff).astype(np.uint32)
    else:
        hs, hd = s[keep], d[keep]
    for slack in (0, 2):
        say(f"ingest drop self-loops dedup={dedup} slack={slack}")
        a = ctx.upload_edges(V, hs, hd, slack)
        b = ctx.ingest(V, s, d, slack, drop_self_loops=True, dedup=dedup)
        same_csr(a, b, f"drop dedup={dedup}")
        a.free()
        b.free()
    # the text loader drops self-loops the same way
    if not dedup:
        say("ingest drop self-loops against dcx_text_read")
        import tempfile, os
        with tempfile.TemporaryDirectory() as tdir:
            path = os.path.join(tdir, "loops.txt")
            dcx.text_write(path, V, s, d)
            V2, ts, td = dcx.text_read(path, prune=False)
        a = ctx.upload_edges(V2, ts, td, 1)
        b = ctx.ingest(V, s, d, 1, drop_self_loops=True)
        same_csr(a, b, "text_read")
        a.free()
        b.free()


@pytest.mark.parametrize("scale,seed", [(12, 3), (16, 8)])
def test_device_rmat_samples_and_dedup(ctx, scale, seed):
    """Raw R-MAT samples drawn on the device equal the host generator's
    draws (their sorted distinct non-loop pairs are dcx_gen_rmat's output),
    and ingesting them with self-loop drop and dedup gives the graph of
    dcx_gen_rmat uploaded from the host, slot for slot."""
    V = 1 << scale
    samples = 16 << scale
    say(f"device rmat {scale} samples={samples}")
    ds, dd = ctx.gen_rmat_device(scale, samples, seed)
    rs, rd = ds.numpy(), dd.numpy()
    s, d = dcx.gen_rmat(scale, samples, seed)
    keep = rs != rd
    key = np.unique((rs[keep].astype(np.uint64) << 32) | rd[keep])
    assert np.array_equal((key >> 32).astype(np.uint32), s)
    assert np.array_equal((key & 0xffffffff).astype(np.uint32), d)
    for slack in (0, 1):
        say(f"device rmat {scale} ingest dedup slack={slack}")
        a = ctx.upload_edges(V, s, d, slack)
        b = ctx.ingest(V, ds, dd, slack, drop_self_loops=True, dedup=True)
        same_csr(a, b, f"device rmat {scale}")
        K, lmax = b.decompose(dcx.VARIANT_BATCHED)
        assert (K, lmax) == a.decompose(dcx.VARIANT_BATCHED)
        assert b.state_hash() == a.state_hash()
        a.free()
        b.free()
    ds.free()
    dd.free()


def test_ingest_device_rejects_bad_pairs(ctx):
    """A bad pair in device arrays fails with its index and leaves the
    context usable."""
    s = np.array([0, 1, 2, 3, 4], np.uint32)
    d = np.array([1, 2, 9, 4, 0], np.uint32)
    ds, dd = ctx.to_device(s), ctx.to_device(d)
    say("ingest device arrays with an id out of range")
    with pytest.raises(dcx.DcxError) as e:
        ctx.ingest_device(5, ds, dd)
    assert e.value.status == 1 and "pair 2" in str(e.value)
    ds.free()
    dd.free()
    g = ctx.ingest(5, np.array([0, 1], np.uint32), np.array([1, 0], np.uint32), 0,
                   drop_self_loops=True, dedup=True)
    assert g.decompose(dcx.VARIANT_BATCHED) == (1, 1)
    g.free()
