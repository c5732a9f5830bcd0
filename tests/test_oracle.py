"""Pins the CPU oracle before it is trusted as the reference of every GPU test (CPU only).

Checks, each independent of the oracle's bucket-queue peel:
  * the reference's golden vector dataset/digraph(.decomp);
  * brute force from the definition on random and exhaustive tiny graphs;
  * textbook reductions (Batagelj-Zaversnik in/out core numbers; networkx
    undirected cores on symmetric digraphs);
  * closed forms; reverse duality; invariants;
  * equality of the second exact path (h-index fixpoint) with the peel;
  * planted mistakes in the oracle are caught by these checks.
"""
import ctypes
import os
import subprocess

import networkx as nx
import numpy as np
import pytest

import graphs
import oracle
from conftest import GOLDEN, ROOT


def decomp(V, s, d):
    return oracle.decompose(oracle.csr_from_edges(V, s, d))


def test_fixture_decomp():
    V, s, d = oracle.load_text_graphcpu(os.path.join(GOLDEN, "digraph"))
    assert V == 8 and len(s) == 26
    K, km, lmax, rows = decomp(V, s, d)
    ref = np.fromfile(os.path.join(GOLDEN, "digraph.decomp"), dtype=np.int32)
    assert (K, lmax) == (ref[0], ref[1]) == (3, 3)
    assert np.array_equal(km, ref[2:2 + V])
    assert np.array_equal(rows.ravel(), ref[2 + V:])
    # decoded values of the fixture, for the record
    assert km.tolist() == [2, 3, 3, 2, 3, 3, 3, 3]
    assert rows[3].tolist() == [0, 1, 3, 0, 3, 3, 1, 3]


@pytest.mark.parametrize("seed", range(40))
def test_brute_force_random(seed):
    rng = np.random.default_rng(seed)
    V = int(rng.integers(1, 13))
    V, s, d = graphs.random_digraph(V, float(rng.uniform(0.05, 0.9)), seed)
    assert [np.asarray(x).tolist() if hasattr(x, "tolist") else x for x in decomp(V, s, d)] == \
        [np.asarray(x).tolist() if hasattr(x, "tolist") else x
         for x in graphs.brute_decomposition(V, s, d)]


def test_brute_force_exhaustive_3():
    n = 0
    for V, s, d in graphs.all_tiny_digraphs(3):
        a = decomp(V, s, d)
        b = graphs.brute_decomposition(V, s, d)
        assert a[0] == b[0] and a[2] == b[2]
        assert np.array_equal(a[1], b[1]) and np.array_equal(a[3], b[3])
        n += 1
    assert n == 64


@pytest.mark.parametrize("seed", range(6))
def test_bz_reductions(seed):
    V, s, d = graphs.random_gnm(300 + 200 * seed, 3000 + 1500 * seed, seed)
    K, km, lmax, rows = decomp(V, s, d)
    assert np.array_equal(km, graphs.in_core_numbers(V, s, d))
    assert np.array_equal(rows[0], graphs.out_core_numbers(V, s, d))


@pytest.mark.parametrize("seed", range(4))
def test_symmetric_networkx(seed):
    V, s, d = graphs.random_gnm(400, 2500, 100 + seed)
    V, s, d = graphs.symmetric(V, s, d)
    K, km, lmax, rows = decomp(V, s, d)
    G = nx.Graph()
    G.add_nodes_from(range(V))
    G.add_edges_from(zip(s.tolist(), d.tolist()))
    core = np.array([nx.core_number(G)[v] for v in range(V)], np.int32)
    assert np.array_equal(km, core)
    assert np.array_equal(rows[0], core)
    for k in range(K + 1):
        # the (k,k)-core of a symmetric digraph is the undirected k-core
        assert np.array_equal(rows[k] >= k, core >= k) or k == 0


def test_closed_forms():
    for n in (2, 5, 9):
        K, km, lmax, rows = decomp(*graphs.complete(n))
        assert K == n - 1 and lmax == n - 1 and (km == n - 1).all() and (rows == n - 1).all()
    K, km, lmax, rows = decomp(*graphs.cycle(11))
    assert K == 1 and lmax == 1 and (rows == 1).all()
    for V, s, d in (graphs.dag(25, 0.4, 3), graphs.out_star(9), graphs.in_star(9)):
        K, km, lmax, rows = decomp(V, s, d)
        assert K == 0 and lmax == 0 and (rows == 0).all()
    for a, b in ((2, 3), (3, 3), (4, 6)):
        K, km, lmax, rows = decomp(*graphs.bi_kab(a, b))
        m = min(a, b)
        assert K == m and (km == m).all() and (rows == m).all()
    for n, dd in ((10, 3), (17, 5)):
        K, km, lmax, rows = decomp(*graphs.circulant(n, dd))
        assert K == dd and (rows == dd).all()


@pytest.mark.parametrize("seed", range(5))
def test_reverse_duality(seed):
    V, s, d = graphs.random_gnm(250, 2500, 7 + seed)
    K, km, lmax, rows = decomp(V, s, d)
    KR, kmR, lmaxR, rowsR = decomp(*graphs.reverse(V, s, d))
    # kmaxes(G) = lmaxes(G^R)[0]
    assert np.array_equal(km, rowsR[0])
    # lmaxes_G[k][v] = max{l : v in D_{l,k}(G^R)}
    for k in range(K + 1):
        expect = np.zeros(V, np.int32)
        for l in range(KR + 1):
            inside = (kmR >= l) & (rowsR[l] >= k)
            expect[inside] = l
        expect[km < k] = 0
        assert np.array_equal(rows[k], expect)


@pytest.mark.parametrize("seed", range(3))
def test_invariants(seed):
    V, s, d = graphs.random_gnm(600, 9000, 40 + seed)
    K, km, lmax, rows = decomp(V, s, d)
    graphs.check_invariants(V, s, d, km, rows)


@pytest.mark.parametrize("seed", range(5))
def test_hindex_equals_peel(seed):
    V, s, d = graphs.random_gnm(500, 6000 + 2000 * seed, 60 + seed)
    g = oracle.csr_from_edges(V, s, d)
    K, km, lmax, rows = oracle.decompose(g)
    K2, km2, it = oracle.kmax_hindex(g)
    assert K2 == K and np.array_equal(km2, km)
    for k in range(K + 1):
        lk, l, it = oracle.klist_hindex(g, km, k)
        assert np.array_equal(l, rows[k])
        assert lk == int(rows[k].max())


def test_lrows_thread_count_invariant():
    V, s, d = graphs.random_gnm(800, 12000, 5)
    g = oracle.csr_from_edges(V, s, d)
    K, km = oracle.kmax(g)
    a = np.empty((K + 1, V), np.int32)
    b = np.empty((K + 1, V), np.int32)
    la = oracle.lib().dco_lrows(*g.args(), K, a, 1)
    lb = oracle.lib().dco_lrows(*g.args(), K, b, 4)
    assert la == lb and np.array_equal(a, b)


def _stream_cases():
    out = [oracle.load_text_graphcpu(os.path.join(GOLDEN, "digraph"))]
    out += [graphs.random_gnm(700, 9000, 80 + sd) for sd in range(3)]
    out += [graphs.random_digraph(12, 0.35, sd) for sd in range(5)]
    out += [graphs.complete(7), graphs.cycle(11), graphs.out_star(40), graphs.in_star(40),
            graphs.bi_kab(4, 9), graphs.circulant(30, 5), graphs.dag(40, 0.2, 3)]
    # isolated vertices, hubs, and a graph with no edges
    V, s, d = graphs.random_gnm(300, 2500, 7)
    out.append((V + 50, s, d))
    out.append((6, np.array([], np.int64), np.array([], np.int64)))
    return out


def test_streaming_hash_equals_dense():
    """dco_state_hash (each row peeled on its (k,0)-core in kmax-rank order,
    added to the hash and dropped) equals the dcx_state_hash formula applied
    to the dense decomposition of the pinned whole-graph peel, part by part
    (kmaxes and every row), on the reference's fixture, random graphs and
    closed forms; kmax and lmax agree too."""
    for V, s, d in _stream_cases():
        g = oracle.csr_from_edges(V, s, d)
        K, km, lmax, rows = oracle.decompose(g)
        h, K2, L2, parts = oracle.state_hash(g, rows=True)
        assert (K2, L2) == (K, lmax)
        assert h == graphs.state_hash(km, rows)
        assert int(parts[0]) == graphs.state_hash(km, [])
        # the sum of the parts is the hash (mod 2^64)
        assert int(np.sum(parts, dtype=np.uint64)) == h


def test_streaming_hash_catches_a_changed_cell():
    """The hash moves when one cell of one row changes (the check a wrong
    row value would fail)."""
    V, s, d = graphs.random_gnm(500, 7000, 3)
    g = oracle.csr_from_edges(V, s, d)
    K, km, lmax, rows = oracle.decompose(g)
    h, _, _ = oracle.state_hash(g)
    bad = rows.copy()
    bad[K // 2][int(np.argmax(rows[K // 2]))] -= 1
    assert graphs.state_hash(km, bad) != h


PLANTS = {
    1: "constraint threshold off by one (<= k instead of < k)",
    2: "key decrement applied to out-neighbours instead of in-neighbours",
    3: "level assigned one too high",
}


@pytest.mark.parametrize("plant", sorted(PLANTS))
def test_planted_mistakes_are_caught(tmp_path, plant):
    """Builds the oracle with a planted mistake and checks that the pinning
    tests above would fail on it (fixture or brute force)."""
    so = str(tmp_path / f"planted{plant}.so")
    subprocess.run(["gcc", "-O1", "-fopenmp", "-fPIC", "-shared", f"-DDCO_PLANT={plant}",
                    "-o", so, os.path.join(ROOT, "oracle", "dcore_oracle.c")], check=True)
    L = ctypes.CDLL(so)
    u64p = np.ctypeslib.ndpointer(np.uint64, flags="C_CONTIGUOUS")
    u32p = np.ctypeslib.ndpointer(np.uint32, flags="C_CONTIGUOUS")
    i32p = np.ctypeslib.ndpointer(np.int32, flags="C_CONTIGUOUS")
    L.dco_klist.argtypes = [ctypes.c_uint32, u64p, u32p, u64p, u32p, ctypes.c_int32, i32p]
    L.dco_klist.restype = ctypes.c_int32
    caught = False
    cases = [oracle.load_text_graphcpu(os.path.join(GOLDEN, "digraph"))]
    cases += [graphs.random_digraph(9, 0.4, sd) for sd in range(20)]
    for V, s, d in cases:
        g = oracle.csr_from_edges(V, s, d)
        BK, Bkm, Blmax, Brows = graphs.brute_decomposition(V, s, d)
        for k in range(BK + 1):
            l = np.empty(V, np.int32)
            L.dco_klist(*g.args(), k, l)
            if not np.array_equal(l, Brows[k]):
                caught = True
    assert caught, PLANTS[plant]
