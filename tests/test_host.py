"""Host-side logic of libdcx.so (CPU only, no GPU calls).

  * the library loads and exports every function include/dcx.h declares;
  * the seeded generators: contract (distinct, no self-loops, sorted),
    determinism across thread counts, pinned output hashes;
  * the CSR-with-slack builder against the oracle's compact CSR;
  * the text loader against the restatement of src/graph_cpu.cpp:38-102;
  * byte compatibility with the reference's binary CSR cache and .decomp
    golden files (dataset/digraph-binary-g=1, dataset/digraph.decomp).
"""
import ctypes
import hashlib
import os
import re
import subprocess
import sys

import numpy as np
import pytest

import dcx
import oracle
from conftest import GOLDEN, ROOT


def declared_functions():
    text = open(os.path.join(ROOT, "include", "dcx.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(dcx_[a-z0-9_]+)\s*\(", text)))


def test_exports_every_declared_symbol():
    L = dcx.lib()
    names = declared_functions()
    assert len(names) >= 30
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing, missing


def test_status_strings():
    L = dcx.lib()
    for code, name in dcx.STATUS.items():
        assert L.dcx_status_str(code).decode() == name


def digest(*arrays):
    h = hashlib.sha256()
    for a in arrays:
        h.update(np.ascontiguousarray(a).tobytes())
    return h.hexdigest()[:16]


def check_edge_contract(V, s, d):
    assert s.dtype == np.uint32 and d.dtype == np.uint32
    assert (s < V).all() and (d < V).all()
    assert (s != d).all()
    key = s.astype(np.uint64) << np.uint64(32) | d.astype(np.uint64)
    assert (np.diff(key.astype(np.int64)) > 0).all()  # sorted, distinct


def test_gen_er():
    s, d = dcx.gen_er(16384, 131072, seed=1)
    assert s.size == 131072
    check_edge_contract(16384, s, d)
    s2, d2 = dcx.gen_er(16384, 131072, seed=1)
    assert np.array_equal(s, s2) and np.array_equal(d, d2)
    s3, _ = dcx.gen_er(16384, 131072, seed=2)
    assert not np.array_equal(s, s3)
    # uniform endpoints: mean id close to V/2
    assert abs(s.mean() / 16384 - 0.5) < 0.01 and abs(d.mean() / 16384 - 0.5) < 0.01
    assert digest(s, d) == ER16K_SEED1


def test_gen_rmat():
    s, d = dcx.gen_rmat(12, 16 << 12, seed=3)
    V = 1 << 12
    check_edge_contract(V, s, d)
    assert 0.8 * (16 << 12) < s.size <= 16 << 12
    # quadrant of the top bit: P(src top bit 0) = a + b = 0.76 (duplicates
    # removal skews it a little; the drawn samples are exact)
    top = (s >> 11) & 1
    assert abs((top == 0).mean() - 0.76) < 0.05
    assert digest(s, d) == RMAT12_SEED3


def test_gen_rmat_thread_invariant():
    code = ("import sys; sys.path.insert(0, %r); import dcx, hashlib, numpy as np; "
            "s, d = dcx.gen_rmat(14, 16 << 14, seed=9); "
            "print(hashlib.sha256(s.tobytes() + d.tobytes()).hexdigest())"
            % os.path.join(ROOT, "d-core-cuda_amd"))
    outs = set()
    for nt in ("1", "3", "8"):
        env = dict(os.environ, OMP_NUM_THREADS=nt)
        outs.add(subprocess.run([sys.executable, "-c", code], env=env, check=True,
                                capture_output=True, text=True).stdout.strip())
    assert len(outs) == 1


def test_gen_chunglu():
    s, d = dcx.gen_chunglu(5000, 40000, seed=1001)
    check_edge_contract(5000, s, d)
    outdeg = np.bincount(s, minlength=5000)
    # power law: the top vertex carries far more than the mean degree
    assert outdeg.max() > 20 * outdeg.mean()
    assert digest(s, d) == CL5K_SEED1001


def test_gen_invalid_args():
    L = dcx.lib()
    a = np.empty(4, np.uint32)
    assert L.dcx_gen_er(ctypes.c_uint32(1), ctypes.c_uint64(1), ctypes.c_uint64(0),
                        dcx.P(a), dcx.P(a)) == 1
    assert L.dcx_gen_er(ctypes.c_uint32(3), ctypes.c_uint64(7), ctypes.c_uint64(0),
                        dcx.P(a), dcx.P(a)) == 1  # more than V(V-1) pairs


@pytest.mark.parametrize("slack", [0, 1, 3])
def test_csr_build_matches_oracle(slack):
    s, d = dcx.gen_rmat(10, 8 << 10, seed=5)
    V = 1 << 10
    g = dcx.csr_build(V, s, d, slack)
    ref = oracle.csr_from_edges(V, s, d)
    oo, on, io, inn = dcx.compact_csr(g)
    assert np.array_equal(oo, ref.out_off) and np.array_equal(on, ref.out_nbr)
    assert np.array_equal(io, ref.in_off) and np.array_equal(inn, ref.in_nbr)
    assert (np.diff(g["out_off"].astype(np.int64)) == g["out_deg"] + slack).all()
    free = np.ones(g["out_nbr"].size, bool)
    for v in range(V):
        free[g["out_off"][v]:g["out_off"][v] + g["out_deg"][v]] = False
    assert (g["out_nbr"][free] == 0xFFFFFFFF).all()


def test_csr_build_rejects_bad_ids():
    L = dcx.lib()
    s = np.array([0, 5], np.uint32)
    d = np.array([1, 1], np.uint32)
    bufs = [np.empty(8, np.uint32) for _ in range(6)]
    st = L.dcx_csr_build(ctypes.c_uint32(3), ctypes.c_uint64(2), dcx.P(s), dcx.P(d),
                         ctypes.c_uint32(0), *[dcx.P(b) for b in bufs])
    assert st == 1


def test_text_loader_matches_graphcpu():
    path = os.path.join(GOLDEN, "digraph")
    V, s, d = dcx.text_read(path, prune=False)
    V2, s2, d2 = oracle.load_text_graphcpu(path)
    assert V == V2 == 8
    assert s.tolist() == s2.tolist() and d.tolist() == d2.tolist()


def test_text_loader_prune_and_selfloops(tmp_path):
    p = tmp_path / "g.txt"
    p.write_text("# 10\n0 3\n3 3\n3 7\n7 0\n9 9\n")
    V, s, d = dcx.text_read(str(p), prune=False)
    assert V == 10 and list(zip(s.tolist(), d.tolist())) == [(0, 3), (3, 7), (7, 0)]
    V, s, d = dcx.text_read(str(p), prune=True)
    # vertices 0,3,7 keep their order: 0->0, 3->1, 7->2 (9 only had a loop)
    assert V == 3 and list(zip(s.tolist(), d.tolist())) == [(0, 1), (1, 2), (2, 0)]


def test_text_roundtrip(tmp_path):
    s, d = dcx.gen_er(300, 2000, seed=4)
    p = str(tmp_path / "er.txt")
    dcx.text_write(p, 300, s, d)
    V, s2, d2 = dcx.text_read(p)
    assert V == 300 and np.array_equal(s, s2) and np.array_equal(d, d2)


def test_text_loader_errors(tmp_path):
    L = dcx.lib()
    V = ctypes.c_uint32()
    n = ctypes.c_uint64()
    assert L.dcx_text_count(b"/nonexistent/file", ctypes.byref(V), ctypes.byref(n)) == 6
    p = tmp_path / "bad.txt"
    p.write_text("# 3\n0 5\n")
    with pytest.raises(dcx.DcxError) as e:
        dcx.text_read(str(p))
    assert e.value.status == 7


def test_binary_golden_matches_text_build():
    """dataset/digraph-binary-g=1 equals the CSR built from dataset/digraph
    with one slack slot per vertex, except the unwritten gap slots."""
    V, E, gb = dcx.binary_read(os.path.join(GOLDEN, "digraph-binary-g=1"), gap=1)
    assert (V, E) == (8, 26)
    assert gb["in_off"].tolist() == [0, 3, 8, 14, 17, 21, 26, 30, 34]
    assert gb["out_off"].tolist() == [0, 3, 7, 13, 16, 22, 27, 30, 34]
    Vt, s, d = dcx.text_read(os.path.join(GOLDEN, "digraph"))
    gt = dcx.csr_build(Vt, s, d, slack=1)
    for side in ("out", "in"):
        assert np.array_equal(gb[side + "_off"], gt[side + "_off"])
        assert np.array_equal(gb[side + "_deg"], gt[side + "_deg"])
        for v in range(V):
            a, n = gb[side + "_off"][v], gb[side + "_deg"][v]
            assert gb[side + "_nbr"][a:a + n].tolist() == gt[side + "_nbr"][a:a + n].tolist()


def test_binary_write_layout(tmp_path):
    Vt, s, d = dcx.text_read(os.path.join(GOLDEN, "digraph"))
    gt = dcx.csr_build(Vt, s, d, slack=1)
    p = str(tmp_path / "digraph-binary-g=1")
    dcx.binary_write(p, Vt, len(s), 1, gt)
    ours = open(p, "rb").read()
    ref = open(os.path.join(GOLDEN, "digraph-binary-g=1"), "rb").read()
    assert len(ours) == len(ref)
    a = np.frombuffer(ours, np.uint32)
    b = np.frombuffer(ref, np.uint32)
    diff = np.nonzero(a != b)[0]
    # only gap slots differ: the reference's are uninitialised (0xCDCDCDCD)
    assert (b[diff] == 0xCDCDCDCD).all() and (a[diff] == 0xFFFFFFFF).all()
    assert diff.size == 16  # one gap slot per vertex, both directions


def test_decomp_golden_roundtrip(tmp_path):
    K, Lm, km, rows = dcx.decomp_read(os.path.join(GOLDEN, "digraph.decomp"), 8)
    assert (K, Lm) == (3, 3)
    p = str(tmp_path / "x.decomp")
    dcx.decomp_write(p, 8, K, Lm, km, rows)
    assert open(p, "rb").read() == open(os.path.join(GOLDEN, "digraph.decomp"), "rb").read()


def test_decomp_read_size_mismatch():
    with pytest.raises(dcx.DcxError) as e:
        dcx.decomp_read(os.path.join(GOLDEN, "digraph.decomp"), 9)
    assert e.value.status == 7


def test_decompcpu_format(tmp_path):
    K, Lm, km, rows = dcx.decomp_read(os.path.join(GOLDEN, "digraph.decomp"), 8)
    rp, vals = oracle.ragged(km, rows)
    p = str(tmp_path / "x.decompCPU")
    dcx.decompcpu_write(p, 8, rp, vals)
    lines = open(p).read().split("\n")
    assert lines[0] == "8" and len(lines) == 9
    assert lines[1] == "2 2 2"          # vertex 0: kmax 2 -> lmaxes[0..2][0]
    assert lines[3] == "3 3 3 3"        # vertex 2


def test_no_device_is_reported():
    """Without a GPU, creating a context fails with a status, not a crash."""
    if dcx.device_count() > 0:
        pytest.skip("a HIP device is present")
    with pytest.raises(dcx.DcxError) as e:
        dcx.Context(0)
    assert e.value.status == 2


# pinned generator outputs (splitmix64 streams are platform independent)
ER16K_SEED1 = "317d929d8c944663"
RMAT12_SEED3 = "b48e861089c2e588"
CL5K_SEED1001 = "22a21e4350670d46"


def test_chunglu_batch_sizes_pinned():
    """The batched workload's sizes come from SplitMix64(1000+i) draws:
    log-uniform in [1e4, 2e6], V = ceil(m/8)."""
    import math
    import workloads
    s = workloads.chunglu_sizes()
    assert len(s) == 256
    assert s[:4] == [(4338, 34704), (7001, 56003), (175766, 1406123), (30654, 245228)]
    assert all(1e4 <= m <= 2e6 and V == math.ceil(m / 8) for V, m in s)
    assert sum(m for _, m in s) == 87328167
    # the textbook SplitMix64 first output for state 0
    assert workloads.splitmix64(0)[1] == 0xE220A8397B1DCDAF


def test_update_stream_is_valid_and_seeded():
    """Every batch of the serving stream is applicable to the graph the
    previous batches produced: deletes name current edges, inserts new
    ones, no self-loops or repeats, sizes in [1, max_batch]."""
    scale = 10
    V, src, dst = 1 << scale, *dcx.gen_rmat(scale, 16 << scale, 2)
    io, ip, do, dp = dcx.gen_update_stream(V, src, dst, scale, 300, max_batch=64, seed=3)
    io2, ip2, do2, dp2 = dcx.gen_update_stream(V, src, dst, scale, 300, max_batch=64, seed=3)
    assert np.array_equal(ip, ip2) and np.array_equal(dp, dp2)
    cur = set(zip(src.tolist(), dst.tolist()))
    n_ins = n_del = 0
    for b in range(300):
        ins = [tuple(x) for x in ip[io[b]:io[b + 1]].tolist()]
        dels = [tuple(x) for x in dp[do[b]:do[b + 1]].tolist()]
        assert 1 <= len(ins) + len(dels) <= 64
        assert len(set(ins)) == len(ins) and len(set(dels)) == len(dels)
        assert all(d in cur for d in dels)
        assert all(i not in cur and i[0] != i[1] and max(i) < V for i in ins)
        cur.difference_update(dels)
        cur.update(ins)
        n_ins += len(ins)
        n_del += len(dels)
    frac = n_ins / (n_ins + n_del)
    assert 0.64 < frac < 0.76, frac
    # sizing call alone gives the same offsets
    io3, _, do3, _ = dcx.gen_update_stream(V, src, dst, scale, 300, max_batch=64, seed=3)
    assert np.array_equal(io3, io) and np.array_equal(do3, do)


def test_state_hash_restatement_on_fixture():
    """The hash restatement is order independent and sees every cell."""
    import graphs
    V, s, d = dcx.text_read(os.path.join(GOLDEN, "digraph"), prune=False)
    K0, L0, km0, rows0 = dcx.decomp_read(os.path.join(GOLDEN, "digraph.decomp"), V)
    h = graphs.state_hash(km0, rows0)
    rows1 = rows0.copy()
    rows1[3, 0] += 1
    assert graphs.state_hash(km0, rows1) != h
    km1 = km0.copy()
    km1[[0, 1]] = km1[[1, 0]]
    assert graphs.state_hash(km1, rows0) != h


def test_mlp_reference_gradient_matches_finite_differences():
    """The CPU restatement of the cost model is pinned by central finite
    differences (so it can judge the device kernels)."""
    import mlp_ref
    rng = np.random.default_rng(3)
    p = rng.normal(0, 0.3, mlp_ref.N_PARAMS)
    X = rng.uniform(0, 1, (20, 8))
    T = rng.normal(0, 1, 20)
    g, loss = mlp_ref.gradient(p, X, T)
    y, _ = mlp_ref.forward(p, X)
    assert abs(loss - 0.5 * ((y - T) ** 2).sum()) < 1e-12
    idx = rng.choice(p.size, 60, replace=False)
    h = 1e-6
    for i in idx:
        q = p.copy(); q[i] += h
        lp = 0.5 * ((mlp_ref.forward(q, X)[0] - T) ** 2).sum()
        q[i] -= 2 * h
        lm = 0.5 * ((mlp_ref.forward(q, X)[0] - T) ** 2).sum()
        assert abs((lp - lm) / (2 * h) - g[i]) < 1e-5 * max(1.0, abs(g[i])), i


def test_kmax_edge_batches_hand_derived():
    """getKmaxEdgeBatches (main.cu:1441-1525) on a case worked out by hand.

    Root of an edge: the endpoint with the larger kmax, the target on ties
    (main.cu:1449-1451); its kmax is the edge's bucket.
      kmaxes: 0:3 1:3 2:1 3:5 4:1 5:3
      e0 0->1 root 1 (tie: target) K3   e1 1->0 root 0 K3
      e2 2->3 root 3 K5                 e3 3->2 root 3 (larger: source) K5
      e4 4->1 root 1 K3                 e5 5->1 root 1 (tie) K3
      e6 2->4 root 4 (tie) K1           e7 4->2 root 2 (tie) K1
    KEDGE (mode 0): per bucket, ascending, repeated passes taking edges
    whose root is not yet in the pass's batch:
      K1: {e6,e7};  K3: {e0,e1}, {e4}, {e5};  K5: {e2}, {e3}
    KMAX (mode 1): per round, the first remaining edge of each bucket whose
    kmax differs by more than one from the round's first edge:
      {e6,e0,e2}, {e7,e1,e3}, {e4}, {e5}
    A source-on-ties root would put e5 in the first K3 batch; a
    smaller-kmax root would move e2 to bucket 1."""
    import dcx
    km = np.array([3, 3, 1, 5, 1, 3], np.int32)
    pairs = np.array([[0, 1], [1, 0], [2, 3], [3, 2], [4, 1], [5, 1], [2, 4], [4, 2]],
                     np.uint32)
    b, nb = dcx.kmax_edge_batches(6, km, pairs, mode=0)
    assert nb == 6
    assert b.tolist() == [1, 1, 4, 5, 2, 3, 0, 0]
    b, nb = dcx.kmax_edge_batches(6, km, pairs, mode=1)
    assert nb == 4
    assert b.tolist() == [0, 1, 0, 1, 2, 3, 0, 1]
    # KMAX compares each bucket with the round's FIRST edge only: buckets
    # 1, 3, 4 all join one round (4 - 1 > 1 even though 4 - 3 = 1)
    km2 = np.array([1, 1, 3, 3, 4, 4], np.int32)
    p2 = np.array([[0, 1], [2, 3], [4, 5]], np.uint32)
    b, nb = dcx.kmax_edge_batches(6, km2, p2, mode=1)
    assert nb == 1 and b.tolist() == [0, 0, 0]
    # and a bucket one away from the first waits for the next round
    km3 = np.array([1, 1, 2, 2, 5, 5], np.int32)
    b, nb = dcx.kmax_edge_batches(6, km3, p2, mode=1)
    assert nb == 2 and b.tolist() == [0, 1, 0]
