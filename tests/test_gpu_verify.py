"""Device-side verification (no download of the decomposition):

* dcx_compare_i32, the verified counterpart of the reference's compareArray
  kernel (src/cuda_utils.cuh:79-93): mismatch count and first index equal
  numpy's on arrays with planted differences, sizes that are not multiples
  of a workgroup, empty input;
* dcx_compare_decompositions: the reference's "maintained == recomputed"
  check (main.cu:1905-1924) on the device -- a graph maintained through
  mixed update batches against the same edge set uploaded and decomposed
  afresh gives 0 mismatches after every batch; two different graphs give
  exactly the mismatches a host comparison of their downloaded arrays
  counts, with the same first cell.
"""
import sys

import numpy as np
import pytest

import dcx

pytestmark = pytest.mark.gpu


def say(msg):
    print(f"[case] {msg}", flush=True)
    sys.stdout.flush()


@pytest.mark.parametrize("n", [0, 1, 255, 1000, 1 << 20])
def test_compare_arrays(n):
    rng = np.random.default_rng(n + 3)
    a = rng.integers(0, 1 << 31, n, dtype=np.int64).astype(np.uint32)
    b = a.copy()
    if n:
        idx = rng.choice(n, size=min(n, 1 + n // 97), replace=False)
        b[idx] ^= 1
    say(f"compare_i32 n={n}")
    with dcx.Context(0) as c:
        da, db = dcx.DeviceArray(a), dcx.DeviceArray(b)
        got = c.compare_arrays(da, db)
        same = c.compare_arrays(da, da)
        da.free()
        db.free()
    diff = np.nonzero(a != b)[0]
    assert got == (len(diff), int(diff[0]) if len(diff) else None)
    assert same == (0, None)


def _host_cells(g):
    km = g.kmaxes()
    rows = g.rows()
    return np.concatenate([km.astype(np.int64), rows.reshape(-1).astype(np.int64)])


def test_compare_decompositions_maintained_and_different():
    V = 1 << 13
    s, d = dcx.gen_rmat(13, 12 << 13, 21)
    rng = np.random.default_rng(5)
    edges = set(zip(s.tolist(), d.tolist()))
    with dcx.Context(0) as c:
        g = c.upload_edges(V, s, d, slack=2)
        g.decompose()
        for b in range(4):
            e = sorted(edges)
            dels = [e[i] for i in sorted(rng.choice(len(e), size=40, replace=False))]
            ins = set()
            while len(ins) < 60:
                u, v = int(rng.integers(V)), int(rng.integers(V))
                if u != v and (u, v) not in edges:
                    ins.add((u, v))
            ins = sorted(ins)
            say(f"verify: batch {b} +{len(ins)} -{len(dels)}")
            g.update(ins=np.array(ins, np.uint32), dels=np.array(dels, np.uint32))
            edges -= set(dels)
            edges |= set(ins)
            es = sorted(edges)
            h = c.upload_edges(V, np.array([x for x, _ in es], np.uint32),
                               np.array([y for _, y in es], np.uint32))
            h.decompose()
            assert g.compare(h) == (0, None)
            assert h.compare(g) == (0, None)
            h.free()
        # a different graph: one more edge between hubs changes cells
        say("verify: different graph")
        es = sorted(edges)
        extra = next((u, v) for u in range(64) for v in range(64)
                     if u != v and (u, v) not in edges)
        es2 = sorted(es + [extra])
        h = c.upload_edges(V, np.array([x for x, _ in es2], np.uint32),
                           np.array([y for _, y in es2], np.uint32))
        h.decompose()
        got = g.compare(h)
        a, b = _host_cells(g), _host_cells(h)
        n = max(len(a), len(b))
        a = np.pad(a, (0, n - len(a)))
        b = np.pad(b, (0, n - len(b)))
        diff = np.nonzero(a != b)[0]
        assert got == (len(diff), int(diff[0]) if len(diff) else None)
        h.free()
        g.free()
