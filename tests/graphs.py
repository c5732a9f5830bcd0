"""Small digraph constructors and independent checkers for the tests.

Edge lists are returned as (V, src, dst) with no self-loops and no
duplicates.  The checkers here are deliberately independent of the oracle's
algorithm: brute force from the definition of the (k,l)-core, and the
Batagelj-Zaversnik bucket algorithm for in/out core numbers.
"""
import itertools

import numpy as np


def dedup(V, pairs):
    seen = set()
    src, dst = [], []
    for s, t in pairs:
        if s == t or (s, t) in seen:
            continue
        seen.add((s, t))
        src.append(s)
        dst.append(t)
    return V, np.array(src, np.int64), np.array(dst, np.int64)


def complete(n):
    return dedup(n, [(i, j) for i in range(n) for j in range(n)])


def cycle(n):
    return dedup(n, [(i, (i + 1) % n) for i in range(n)])


def dag(n, p=0.5, seed=0):
    rng = np.random.default_rng(seed)
    return dedup(n, [(i, j) for i in range(n) for j in range(i + 1, n) if rng.random() < p])


def out_star(n):
    return dedup(n, [(0, j) for j in range(1, n)])


def in_star(n):
    return dedup(n, [(j, 0) for j in range(1, n)])


def bi_kab(a, b):
    pairs = []
    for i in range(a):
        for j in range(b):
            pairs += [(i, a + j), (a + j, i)]
    return dedup(a + b, pairs)


def circulant(n, d):
    return dedup(n, [(i, (i + o) % n) for i in range(n) for o in range(1, d + 1)])


def random_digraph(V, p, seed):
    rng = np.random.default_rng(seed)
    m = rng.random((V, V)) < p
    np.fill_diagonal(m, False)
    s, t = np.nonzero(m)
    return V, s.astype(np.int64), t.astype(np.int64)


def random_gnm(V, E, seed):
    rng = np.random.default_rng(seed)
    seen = set()
    pairs = []
    E = min(E, V * (V - 1))
    while len(pairs) < E:
        s, t = int(rng.integers(V)), int(rng.integers(V))
        if s != t and (s, t) not in seen:
            seen.add((s, t))
            pairs.append((s, t))
    return dedup(V, pairs)


def symmetric(V, src, dst):
    pairs = list(zip(src.tolist(), dst.tolist())) + list(zip(dst.tolist(), src.tolist()))
    return dedup(V, pairs)


def reverse(V, src, dst):
    return V, dst.copy(), src.copy()


# ---- brute force from the definition --------------------------------------

def dkl_core(V, src, dst, k, l):
    """D_{k,l}: iteratively delete vertices with in-degree < k or
    out-degree < l inside the remaining induced subgraph."""
    alive = np.ones(V, bool)
    while True:
        a = alive[src] & alive[dst]
        ind = np.bincount(dst[a], minlength=V)
        outd = np.bincount(src[a], minlength=V)
        bad = alive & ((ind < k) | (outd < l))
        if not bad.any():
            return alive
        alive &= ~bad


def brute_decomposition(V, src, dst):
    """(kmax, kmaxes, lmax, rows) straight from the definition."""
    kmaxes = np.zeros(V, np.int32)
    k = 0
    while True:
        core = dkl_core(V, src, dst, k, 0)
        if not core.any():
            break
        kmaxes[core] = k
        k += 1
    K = int(kmaxes.max()) if V else -1
    rows = np.zeros((K + 1, V), np.int32)
    for k in range(K + 1):
        l = 0
        while True:
            core = dkl_core(V, src, dst, k, l)
            if not core.any():
                break
            rows[k][core] = l
            l += 1
    lmax = int(rows.max()) if rows.size else 0
    return K, kmaxes, lmax, rows


def bz_core_numbers(V, adj_in_count_lists):
    """Batagelj-Zaversnik O(m) core numbers where the degree of v is the
    number of lists entries pointing at v: deg[v] = len(lists_of[v]) and
    removing u decrements deg[w] for every w in nbrs_affected[u]."""
    deg, affected = adj_in_count_lists
    deg = deg.copy()
    md = int(deg.max()) if V else 0
    bins = [0] * (md + 2)
    for d in deg:
        bins[d] += 1
    start = 0
    for d in range(md + 1):
        bins[d], start = start, start + bins[d]
    pos = [0] * V
    vert = [0] * V
    for v in range(V):
        pos[v] = bins[deg[v]]
        vert[pos[v]] = v
        bins[deg[v]] += 1
    for d in range(md, 0, -1):
        bins[d] = bins[d - 1]
    bins[0] = 0
    core = [0] * V
    for i in range(V):
        v = vert[i]
        core[v] = deg[v]
        for u in affected[v]:
            if deg[u] > deg[v]:
                du, pu = deg[u], pos[u]
                pw = bins[du]
                w = vert[pw]
                if u != w:
                    pos[u], pos[w] = pw, pu
                    vert[pu], vert[pw] = w, u
                bins[du] += 1
                deg[u] -= 1
    return np.array(core, np.int32)


def in_core_numbers(V, src, dst):
    """kmaxes by BZ: degree = in-degree; removing v lowers the in-degree of
    its out-neighbours."""
    outs = [[] for _ in range(V)]
    for s, t in zip(src.tolist(), dst.tolist()):
        outs[s].append(t)
    deg = np.bincount(dst, minlength=V).astype(np.int64)
    return bz_core_numbers(V, (deg, outs))


def out_core_numbers(V, src, dst):
    """lmaxes[0] by BZ: degree = out-degree; removing v lowers the
    out-degree of its in-neighbours."""
    ins = [[] for _ in range(V)]
    for s, t in zip(src.tolist(), dst.tolist()):
        ins[t].append(s)
    deg = np.bincount(src, minlength=V).astype(np.int64)
    return bz_core_numbers(V, (deg, ins))


def all_tiny_digraphs(n):
    """Every digraph on n labelled vertices (n <= 3 keeps this small)."""
    pairs = [(i, j) for i in range(n) for j in range(n) if i != j]
    for mask in range(1 << len(pairs)):
        sel = [pairs[b] for b in range(len(pairs)) if mask >> b & 1]
        yield dedup(n, sel)


def check_invariants(V, src, dst, kmaxes, rows):
    """Definition-level properties of a decomposition; raises AssertionError."""
    K = rows.shape[0] - 1
    ind = np.bincount(dst, minlength=V)
    outd = np.bincount(src, minlength=V)
    assert (kmaxes <= ind).all()
    for k in range(K + 1):
        assert (rows[k] <= outd).all()
        assert (rows[k][kmaxes < k] == 0).all()
        if k:
            assert (rows[k] <= rows[k - 1]).all()
    # core property and maximality of each level set, on a sample of (k,l)
    for k in sorted({0, K // 2, K}):
        for l in sorted({0, 1, int(rows[k].max()) // 2, int(rows[k].max())}):
            S = (rows[k] >= l) & (kmaxes >= k)
            a = S[src] & S[dst]
            if S.any():
                assert (np.bincount(dst[a], minlength=V)[S] >= k).all()
                assert (np.bincount(src[a], minlength=V)[S] >= l).all()
            # maximality: the brute-force core equals S (cheap at test sizes)
            if V <= 4000:
                assert (dkl_core(V, src, dst, k, l) == S).all()


def product_pairs(it):
    return list(itertools.product(*it))


# ---- state hash (restates dcx_state_hash, include/dcx.h) ----------------

_G1 = np.uint64(0x9E3779B97F4A7C15)
_G2 = np.uint64(0xC2B2AE3D27D4EB4F)


def _mix64(z):
    z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
    z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
    return z ^ (z >> np.uint64(31))


def state_hash(kmaxes, rows):
    """Sum over v and k = -1..kmax of mix64(v*G1 + (k+1)*G2 + value) mod
    2^64, value = kmaxes[v] for k = -1, rows[k][v] otherwise."""
    V = len(kmaxes)
    v = np.arange(V, dtype=np.uint64)
    total = np.uint64(0)
    with np.errstate(over="ignore"):
        layers = [np.asarray(kmaxes)] + [np.asarray(r) for r in rows]
        for kk, vals in enumerate(layers):
            x = np.asarray(vals, dtype=np.int64).astype(np.uint32).astype(np.uint64)
            t = _mix64(v * _G1 + np.uint64(kk) * _G2 + x)
            total = total + np.sum(t, dtype=np.uint64)
    return int(total)
