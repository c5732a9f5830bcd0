"""ctypes front end of the CPU oracle (test infrastructure only).

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg import
this module: it is the checker, never the thing measured or shipped.

Also holds two pure-Python restatements used by the tests:
  * ``load_text_graphcpu`` -- the text loader conventions of the reference's
    CPU graph (src/graph_cpu.cpp:38-102): V from the ``# V`` header line,
    self-loops dropped, input order kept, no pruning, no dedup.
  * ``csr_from_edges`` -- compact CSR in/out with the reference's neighbour
    order (input order, src/graph.cpp:185-190, 224-236).
"""
import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "libdcore_oracle.so")

_u64p = np.ctypeslib.ndpointer(np.uint64, flags="C_CONTIGUOUS")
_u32p = np.ctypeslib.ndpointer(np.uint32, flags="C_CONTIGUOUS")
_i32p = np.ctypeslib.ndpointer(np.int32, flags="C_CONTIGUOUS")
_lib = None


def build():
    subprocess.run(["make", "-s", "-C", HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            build()
        L = ctypes.CDLL(LIB)
        csr = [ctypes.c_uint32, _u64p, _u32p, _u64p, _u32p]
        L.dco_klist.argtypes = csr + [ctypes.c_int32, _i32p]
        L.dco_klist.restype = ctypes.c_int32
        L.dco_kmax.argtypes = csr + [_i32p]
        L.dco_kmax.restype = ctypes.c_int32
        L.dco_lrows.argtypes = csr + [ctypes.c_int32, _i32p, ctypes.c_int]
        L.dco_lrows.restype = ctypes.c_int32
        L.dco_klist_hindex.argtypes = csr + [_i32p, ctypes.c_int32, _i32p,
                                             ctypes.POINTER(ctypes.c_int32)]
        L.dco_klist_hindex.restype = ctypes.c_int32
        L.dco_kmax_hindex.argtypes = [ctypes.c_uint32, _u64p, _u32p, _i32p,
                                      ctypes.POINTER(ctypes.c_int32)]
        L.dco_kmax_hindex.restype = ctypes.c_int32
        L.dco_state_hash.argtypes = csr + [ctypes.c_int, ctypes.POINTER(ctypes.c_int32),
                                           ctypes.POINTER(ctypes.c_int32), ctypes.c_void_p]
        L.dco_state_hash.restype = ctypes.c_uint64
        L.dco_time_sample.argtypes = csr + [ctypes.c_int, ctypes.c_int32, ctypes.c_void_p,
                                            ctypes.c_void_p]
        L.dco_time_sample.restype = ctypes.c_int32
        _lib = L
    return _lib


class CSR:
    """Compact in/out CSR of a digraph (uint64 offsets, uint32 ids)."""

    def __init__(self, V, out_off, out_nbr, in_off, in_nbr):
        self.V = int(V)
        self.out_off = np.ascontiguousarray(out_off, dtype=np.uint64)
        self.out_nbr = np.ascontiguousarray(out_nbr, dtype=np.uint32)
        self.in_off = np.ascontiguousarray(in_off, dtype=np.uint64)
        self.in_nbr = np.ascontiguousarray(in_nbr, dtype=np.uint32)

    @property
    def E(self):
        return int(self.out_off[-1])

    def args(self):
        return (self.V, self.out_off, self.out_nbr, self.in_off, self.in_nbr)

    def reverse(self):
        return CSR(self.V, self.in_off, self.in_nbr, self.out_off, self.out_nbr)


def csr_from_edges(V, src, dst):
    """Compact CSR with neighbours in input order (stable grouping)."""
    src = np.asarray(src, dtype=np.int64)
    dst = np.asarray(dst, dtype=np.int64)
    o = np.argsort(src, kind="stable")
    out_off = np.zeros(V + 1, dtype=np.uint64)
    np.cumsum(np.bincount(src, minlength=V), out=out_off[1:])
    i = np.argsort(dst, kind="stable")
    in_off = np.zeros(V + 1, dtype=np.uint64)
    np.cumsum(np.bincount(dst, minlength=V), out=in_off[1:])
    return CSR(V, out_off, dst[o], in_off, src[i])


def load_text_graphcpu(path):
    """Text edge list per src/graph_cpu.cpp:38-102 (returns V, src, dst)."""
    with open(path) as f:
        V = int(f.readline()[2:])
        src, dst = [], []
        for line in f:
            p = line.split()
            if len(p) < 2:
                continue
            s, t = int(p[0]), int(p[1])
            if s == t:
                continue
            src.append(s)
            dst.append(t)
    return V, np.array(src, dtype=np.int64), np.array(dst, dtype=np.int64)


def kmax(g):
    km = np.empty(g.V, dtype=np.int32)
    k = lib().dco_kmax(*g.args(), km)
    return int(k), km


def klist(g, k):
    l = np.empty(g.V, dtype=np.int32)
    lk = lib().dco_klist(*g.args(), int(k), l)
    return int(lk), l


def decompose(g, nthreads=0):
    """Returns (kmax, kmaxes[V], lmax, lmaxes[(kmax+1), V])."""
    K, km = kmax(g)
    rows = np.empty((max(K + 1, 0), g.V), dtype=np.int32)
    lmax = int(lib().dco_lrows(*g.args(), K, rows, int(nthreads))) if K >= 0 else 0
    return K, km, lmax, rows


def state_hash(g, nthreads=0, rows=False):
    """dcx_state_hash of the decomposition of g, computed row by row on each
    (k,0)-core without storing the rows.  Returns (hash, kmax, lmax), plus
    the per-part hashes (kmaxes, then row 0..kmax) when rows=True."""
    K = ctypes.c_int32(0)
    L = ctypes.c_int32(0)
    part = None
    if rows:
        kk, _ = kmax(g)
        part = np.zeros(max(kk + 2, 1), np.uint64)
    h = lib().dco_state_hash(*g.args(), int(nthreads), ctypes.byref(K), ctypes.byref(L),
                             None if part is None else part.ctypes.data_as(ctypes.c_void_p))
    if rows:
        return int(h), K.value, L.value, part
    return int(h), K.value, L.value


def time_sample(g, nthreads=0, stride=1):
    """Times the oracle on a sample of one decomposition (see dco_time_sample):
    returns (kmax, {kmax_s, order_s, rows_s}, sampled row bytes, all row bytes)."""
    t = np.zeros(3, np.float64)
    w = np.zeros(2, np.float64)
    K = lib().dco_time_sample(*g.args(), int(nthreads), int(stride),
                              t.ctypes.data_as(ctypes.c_void_p),
                              w.ctypes.data_as(ctypes.c_void_p))
    return int(K), {"kmax_s": float(t[0]), "order_s": float(t[1]), "rows_s": float(t[2])}, \
        float(w[0]), float(w[1])


def klist_hindex(g, kmaxes, k):
    l = np.empty(g.V, dtype=np.int32)
    it = ctypes.c_int32(0)
    lk = lib().dco_klist_hindex(*g.args(), np.ascontiguousarray(kmaxes, np.int32),
                                int(k), l, ctypes.byref(it))
    return int(lk), l, it.value


def kmax_hindex(g):
    km = np.empty(g.V, dtype=np.int32)
    it = ctypes.c_int32(0)
    K = lib().dco_kmax_hindex(g.V, g.in_off, g.in_nbr, km, ctypes.byref(it))
    return int(K), km, it.value


def ragged(kmaxes, rows):
    """Per-vertex ragged layout (the .decompCPU order, src/utils.h:110-124):
    row_ptr[V+1] (int64) and values = lmaxes[0..kmaxes[v]][v] for each v."""
    V = kmaxes.size
    cnt = kmaxes.astype(np.int64) + 1
    row_ptr = np.zeros(V + 1, dtype=np.int64)
    np.cumsum(cnt, out=row_ptr[1:])
    vals = np.empty(int(row_ptr[-1]), dtype=np.int32)
    for v in range(V):
        vals[row_ptr[v]:row_ptr[v + 1]] = rows[: kmaxes[v] + 1, v]
    return row_ptr, vals
