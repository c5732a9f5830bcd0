"""ctypes binding of libdcx.so (the C-ABI in include/dcx.h).

This is the host-side mirror used by the tests and bench.py.  It adds no
compute of its own: every result comes from the HIP library.  Loading fails
loudly when libdcx.so is missing -- there is no CPU fallback.
"""
import ctypes
import weakref
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libdcx.so")

VARIANT_AUTO, VARIANT_PEEL, VARIANT_BATCHED, VARIANT_HINDEX = 0, 1, 2, 3
# pass log only: a batched team row derived from the row above
VARIANT_DERIVED = 4
INGEST_ON_DEVICE, INGEST_DROP_SELF_LOOPS, INGEST_DEDUP = 1, 2, 4
STATUS = {
    0: "DCX_OK", 1: "DCX_ERR_INVALID_ARG", 2: "DCX_ERR_NO_DEVICE", 3: "DCX_ERR_HIP",
    4: "DCX_ERR_OUT_OF_MEMORY", 5: "DCX_ERR_CAPACITY", 6: "DCX_ERR_IO",
    7: "DCX_ERR_FORMAT", 8: "DCX_ERR_NOT_FOUND", 9: "DCX_ERR_STATE",
    10: "DCX_ERR_BUDGET", 11: "DCX_ERR_TIMEOUT", 12: "DCX_ERR_INTERNAL",
}


class DcxError(RuntimeError):
    def __init__(self, status, where, detail=""):
        self.status = status
        super().__init__(f"{where}: {STATUS.get(status, status)} {detail}".strip())


class Config(ctypes.Structure):
    _fields_ = [("device", ctypes.c_int32), ("variant", ctypes.c_int32),
                ("rows_slots", ctypes.c_int32), ("team_max", ctypes.c_int32),
                ("shrink_at", ctypes.c_int32), ("trace_phases", ctypes.c_int32),
                ("scratch_budget", ctypes.c_uint64), ("pull_ratio", ctypes.c_int32),
                ("lds_members", ctypes.c_int32), ("maintenance", ctypes.c_int32),
                ("local_budget", ctypes.c_int32), ("chain_rows", ctypes.c_int32),
                ("anchors", ctypes.c_int32), ("reserved", ctypes.c_uint64 * 1)]


class Timing(ctypes.Structure):
    _fields_ = [("kmax_ms", ctypes.c_double), ("order_ms", ctypes.c_double),
                ("rows_ms", ctypes.c_double), ("rows_kernel_ms", ctypes.c_double),
                ("total_ms", ctypes.c_double), ("rows_launches", ctypes.c_int32),
                ("rows_slots", ctypes.c_int32), ("rows_scheduler", ctypes.c_int32),
                ("rows_split", ctypes.c_int32), ("rows_teams_ms", ctypes.c_float),
                ("rows_lds_ms", ctypes.c_float), ("rows_first_ms", ctypes.c_float),
                ("rows_chains_ms", ctypes.c_float)]


class PassLog(ctypes.Structure):
    _fields_ = [("k", ctypes.c_int32), ("variant", ctypes.c_int32),
                ("members", ctypes.c_uint32), ("phases", ctypes.c_uint32),
                ("levels", ctypes.c_uint32), ("team", ctypes.c_uint32),
                ("edges", ctypes.c_uint64), ("ms", ctypes.c_double),
                ("failed", ctypes.c_uint32), ("removed", ctypes.c_uint32),
                ("alive", ctypes.c_uint32), ("level", ctypes.c_int32)]


class PhaseTrace(ctypes.Structure):
    _fields_ = [("level", ctypes.c_int32), ("seed", ctypes.c_int32),
                ("removed", ctypes.c_uint32), ("alive", ctypes.c_uint32),
                ("us", ctypes.c_double)]


class TuneResult(ctypes.Structure):
    _fields_ = [("rows_slots", ctypes.c_int32), ("team_max", ctypes.c_int32),
                ("shrink_at", ctypes.c_int32), ("candidates", ctypes.c_int32),
                ("best_ms", ctypes.c_double), ("default_ms", ctypes.c_double),
                ("lds_members", ctypes.c_int32), ("reserved32", ctypes.c_int32),
                ("reserved", ctypes.c_double * 3)]


class ServeConfig(ctypes.Structure):
    _fields_ = [("hbm_budget", ctypes.c_uint64), ("max_inflight", ctypes.c_uint32),
                ("max_group", ctypes.c_uint32), ("max_group_edges", ctypes.c_uint32),
                ("hash_states", ctypes.c_int32), ("reserved", ctypes.c_uint64 * 4)]


class ServeResult(ctypes.Structure):
    _fields_ = [("id", ctypes.c_uint64), ("status", ctypes.c_int32), ("group", ctypes.c_uint32),
                ("inserted", ctypes.c_uint32), ("deleted", ctypes.c_uint32),
                ("kmax_after", ctypes.c_int32), ("rows_recomputed", ctypes.c_int32),
                ("submit_ms", ctypes.c_double), ("admit_ms", ctypes.c_double),
                ("start_ms", ctypes.c_double), ("done_ms", ctypes.c_double),
                ("state_hash", ctypes.c_uint64)]


class ServeStats(ctypes.Structure):
    _fields_ = [("requests", ctypes.c_uint64), ("groups", ctypes.c_uint64),
                ("max_group", ctypes.c_uint32), ("max_waiting", ctypes.c_uint32),
                ("busy_ms", ctypes.c_double), ("elapsed_ms", ctypes.c_double),
                ("reserved", ctypes.c_uint64 * 4)]


class UpdateStats(ctypes.Structure):
    _fields_ = [("inserted", ctypes.c_uint32), ("deleted", ctypes.c_uint32),
                ("relocated", ctypes.c_uint32), ("m_del", ctypes.c_int32),
                ("m_ins", ctypes.c_int32), ("rows_recomputed", ctypes.c_int32),
                ("kmax_before", ctypes.c_int32), ("kmax_after", ctypes.c_int32),
                ("insert_ms", ctypes.c_double), ("kmax_ms", ctypes.c_double),
                ("rows_ms", ctypes.c_double), ("total_ms", ctypes.c_double),
                ("path", ctypes.c_int32), ("net_deleted", ctypes.c_int32),
                ("net_inserted", ctypes.c_int32), ("reserved", ctypes.c_int32 * 5)]


_lib = None


def lib():
    """Loads libdcx.so (raises if it was not built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is missing: build it first "
                               "(python -c 'import __graft_entry__ as g; g.build()')")
        _lib = ctypes.CDLL(LIB_PATH)
        _lib.dcx_status_str.restype = ctypes.c_char_p
        _lib.dcx_ctx_last_error.restype = ctypes.c_char_p
        _lib.dcx_ctx_last_error.argtypes = [ctypes.c_void_p]
    return _lib


def P(a):
    """Pointer to a numpy array's data (None passes NULL)."""
    return None if a is None else a.ctypes.data_as(ctypes.c_void_p)


def check(status, where, ctx=None):
    if status != 0:
        detail = ""
        if ctx is not None:
            detail = lib().dcx_ctx_last_error(ctx).decode(errors="replace")
        raise DcxError(status, where, detail)


u32, u64, i32, f64 = ctypes.c_uint32, ctypes.c_uint64, ctypes.c_int32, ctypes.c_double


# ---- host-side graph building ------------------------------------------

def gen_er(V, E, seed):
    src = np.empty(E, np.uint32)
    dst = np.empty(E, np.uint32)
    check(lib().dcx_gen_er(u32(V), u64(E), u64(seed), P(src), P(dst)), "dcx_gen_er")
    return src, dst


def gen_rmat(scale, samples, seed, a=0.57, b=0.19, c=0.19):
    src = np.empty(samples, np.uint32)
    dst = np.empty(samples, np.uint32)
    E = ctypes.c_uint64()
    check(lib().dcx_gen_rmat(u32(scale), u64(samples), f64(a), f64(b), f64(c), u64(seed),
                             P(src), P(dst), u64(samples), ctypes.byref(E)), "dcx_gen_rmat")
    return src[:E.value].copy(), dst[:E.value].copy()


def gen_chunglu(V, samples, seed, gamma=2.1):
    src = np.empty(samples, np.uint32)
    dst = np.empty(samples, np.uint32)
    E = ctypes.c_uint64()
    check(lib().dcx_gen_chunglu(u32(V), u64(samples), f64(gamma), u64(seed), P(src), P(dst),
                                u64(samples), ctypes.byref(E)), "dcx_gen_chunglu")
    return src[:E.value].copy(), dst[:E.value].copy()


def gen_update_stream(V, src, dst, scale, nbatches, max_batch=4096, p_insert=0.7, seed=3,
                      a=0.57, b=0.19, c=0.19):
    """Update batches against the R-MAT graph (V=2^scale, src, dst):
    returns (ins_off, ins_pairs, del_off, del_pairs), batch b's inserts being
    ins_pairs[ins_off[b]:ins_off[b+1]] as (n, 2) rows."""
    src = np.ascontiguousarray(src, np.uint32)
    dst = np.ascontiguousarray(dst, np.uint32)
    io = np.zeros(nbatches + 1, np.uint64)
    do = np.zeros(nbatches + 1, np.uint64)
    args = (u32(V), u64(src.size), P(src), P(dst), u32(scale), f64(a), f64(b), f64(c),
            u32(nbatches), u32(max_batch), f64(p_insert), u64(seed), P(io), P(do))
    check(lib().dcx_gen_update_stream(*args, None, u64(0), None, u64(0)), "dcx_gen_update_stream")
    ip = np.empty((int(io[-1]), 2), np.uint32)
    dp = np.empty((int(do[-1]), 2), np.uint32)
    check(lib().dcx_gen_update_stream(*args, P(ip), u64(ip.size), P(dp), u64(dp.size)),
          "dcx_gen_update_stream")
    return io.astype(np.int64), ip, do.astype(np.int64), dp


def csr_build(V, src, dst, slack=0):
    src = np.ascontiguousarray(src, np.uint32)
    dst = np.ascontiguousarray(dst, np.uint32)
    E = src.size
    cap = E + slack * V
    oo = np.empty(V + 1, np.uint32)
    io = np.empty(V + 1, np.uint32)
    od = np.empty(V, np.int32)
    idg = np.empty(V, np.int32)
    on = np.empty(max(cap, 1), np.uint32)
    inn = np.empty(max(cap, 1), np.uint32)
    check(lib().dcx_csr_build(u32(V), u64(E), P(src), P(dst), u32(slack), P(oo), P(od), P(on),
                              P(io), P(idg), P(inn)), "dcx_csr_build")
    return dict(out_off=oo, out_deg=od, out_nbr=on[:cap], in_off=io, in_deg=idg,
                in_nbr=inn[:cap])


def text_read(path, prune=False):
    Vh = ctypes.c_uint32()
    n = ctypes.c_uint64()
    check(lib().dcx_text_count(path.encode(), ctypes.byref(Vh), ctypes.byref(n)),
          "dcx_text_count")
    src = np.empty(max(n.value, 1), np.uint32)
    dst = np.empty(max(n.value, 1), np.uint32)
    V = ctypes.c_uint32()
    E = ctypes.c_uint64()
    check(lib().dcx_text_read(path.encode(), i32(1 if prune else 0), P(src), P(dst),
                              u64(n.value), ctypes.byref(V), ctypes.byref(E)), "dcx_text_read")
    return V.value, src[:E.value].copy(), dst[:E.value].copy()


def text_write(path, V, src, dst):
    src = np.ascontiguousarray(src, np.uint32)
    dst = np.ascontiguousarray(dst, np.uint32)
    check(lib().dcx_text_write(path.encode(), u32(V), u64(src.size), P(src), P(dst)),
          "dcx_text_write")


def binary_read(path, gap):
    V = ctypes.c_uint32()
    E = ctypes.c_uint32()
    check(lib().dcx_binary_header(path.encode(), ctypes.byref(V), ctypes.byref(E)),
          "dcx_binary_header")
    V, E = V.value, E.value
    cap = E + gap * V
    g = dict(out_off=np.empty(V + 1, np.uint32), out_deg=np.empty(V, np.int32),
             out_nbr=np.empty(max(cap, 1), np.uint32), in_off=np.empty(V + 1, np.uint32),
             in_deg=np.empty(V, np.int32), in_nbr=np.empty(max(cap, 1), np.uint32))
    check(lib().dcx_binary_read(path.encode(), u32(gap), P(g["out_off"]), P(g["out_deg"]),
                                P(g["out_nbr"]), P(g["in_off"]), P(g["in_deg"]),
                                P(g["in_nbr"])), "dcx_binary_read")
    g["out_nbr"] = g["out_nbr"][:cap]
    g["in_nbr"] = g["in_nbr"][:cap]
    return V, E, g


def binary_write(path, V, E, gap, g):
    check(lib().dcx_binary_write(path.encode(), u32(V), u32(E), u32(gap), P(g["out_off"]),
                                 P(g["out_deg"]), P(g["out_nbr"]), P(g["in_off"]),
                                 P(g["in_deg"]), P(g["in_nbr"])), "dcx_binary_write")


def decomp_write(path, V, kmax, lmax, kmaxes, rows):
    kmaxes = np.ascontiguousarray(kmaxes, np.int32)
    rows = np.ascontiguousarray(rows, np.int32)
    check(lib().dcx_decomp_write(path.encode(), u32(V), i32(kmax), i32(lmax), P(kmaxes),
                                 P(rows)), "dcx_decomp_write")


def decomp_read(path, V):
    K = ctypes.c_int32()
    Lm = ctypes.c_int32()
    check(lib().dcx_decomp_header(path.encode(), ctypes.byref(K), ctypes.byref(Lm)),
          "dcx_decomp_header")
    km = np.empty(V, np.int32)
    rows = np.empty((max(K.value + 1, 0), V), np.int32)
    check(lib().dcx_decomp_read(path.encode(), u32(V), P(km), P(rows)), "dcx_decomp_read")
    return K.value, Lm.value, km, rows


def decompcpu_write(path, V, row_ptr, values):
    row_ptr = np.ascontiguousarray(row_ptr, np.int64)
    values = np.ascontiguousarray(values, np.int32)
    check(lib().dcx_decompcpu_write(path.encode(), u32(V), P(row_ptr), P(values)),
          "dcx_decompcpu_write")


# ---- device ---------------------------------------------------------------

def device_count():
    n = ctypes.c_int32()
    check(lib().dcx_device_count(ctypes.byref(n)), "dcx_device_count")
    return n.value


class Context:
    """A dcx_ctx on one device (library-owned stream and scratch)."""

    def __init__(self, device=0, variant=VARIANT_AUTO, rows_slots=0, scratch_budget=0,
                 team_max=0, shrink_at=0, trace_phases=0, pull_ratio=0, lds_members=0,
                 maintenance=0, local_budget=0, chain_rows=0, anchors=0):
        cfg = Config()
        check(lib().dcx_config_init(ctypes.byref(cfg)), "dcx_config_init")
        cfg.device = device
        cfg.variant = variant
        cfg.rows_slots = rows_slots
        cfg.team_max = team_max
        cfg.shrink_at = shrink_at
        cfg.trace_phases = trace_phases
        cfg.scratch_budget = scratch_budget
        cfg.pull_ratio = pull_ratio
        cfg.lds_members = lds_members
        cfg.maintenance = maintenance
        cfg.local_budget = local_budget
        cfg.chain_rows = chain_rows
        cfg.anchors = anchors
        h = ctypes.c_void_p()
        check(lib().dcx_ctx_create(ctypes.byref(cfg), ctypes.byref(h)), "dcx_ctx_create")
        self.h = h
        # graphs and models of this context: freed before it (the C-ABI
        # requires it), also when an exception leaves them behind
        self._children = weakref.WeakSet()

    def _adopt(self, obj):
        self._children.add(obj)

    def close(self):
        if self.h:
            for ch in list(self._children):
                ch.free()
            lib().dcx_ctx_destroy(self.h)
            self.h = None

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def timing(self):
        t = Timing()
        check(lib().dcx_last_timing(self.h, ctypes.byref(t)), "dcx_last_timing", self.h)
        return t

    def set_scheduler(self, mlp):
        """Batched team rows are planned by the cost model -- peeled or
        derived from the row above, and team widths (None: the fixed rule)."""
        check(lib().dcx_set_scheduler(self.h, mlp.h if mlp is not None else None),
              "dcx_set_scheduler", self.h)

    def compare_arrays(self, a, b):
        """dcx_compare_i32 on two device arrays (DeviceArray) of equal size:
        (mismatching entries, first such index or None)."""
        if a.size != b.size:
            raise ValueError("arrays of different sizes")
        n, f = ctypes.c_uint64(), ctypes.c_uint64()
        check(lib().dcx_compare_i32(self.h, a.ptr, b.ptr, ctypes.c_uint64(a.size),
                                    ctypes.byref(n), ctypes.byref(f)), "dcx_compare_i32", self.h)
        return n.value, (None if n.value == 0 else f.value)

    def set_rows_rule(self, chain_rows=0, team_max=0, anchors=0):
        """cfg.chain_rows, cfg.team_max and cfg.anchors for the next
        decompositions (chain_rows 0: the default rule, or the cost model's
        plan when one is attached; 1: every team row an anchor; n > 1: every
        n-th row from row 0 an anchor; team_max 0: the default widest team;
        anchors 0: peeled, 1: derived from row 0)."""
        check(lib().dcx_ctx_set_rows_rule(self.h, i32(chain_rows), i32(team_max), i32(anchors)),
              "dcx_ctx_set_rows_rule", self.h)

    def last_trace(self):
        """Phase trace of the last grid peel pass (needs trace_phases=1):
        structured numpy array (level, seed, removed, alive, us)."""
        n = ctypes.c_uint32()
        check(lib().dcx_last_trace(self.h, None, u32(0), ctypes.byref(n)), "dcx_last_trace", self.h)
        arr = (PhaseTrace * max(n.value, 1))()
        check(lib().dcx_last_trace(self.h, arr, u32(n.value), ctypes.byref(n)), "dcx_last_trace",
              self.h)
        out = np.zeros(n.value, [("level", np.int32), ("seed", np.int32), ("removed", np.uint32),
                                 ("alive", np.uint32), ("us", np.float64)])
        for i in range(n.value):
            out[i] = (arr[i].level, arr[i].seed, arr[i].removed, arr[i].alive, arr[i].us)
        return out

    def upload_edges(self, V, src, dst, slack=0):
        return Graph(self, V, src, dst, slack)

    def ingest_edges(self, V, src, dst, slack=0):
        """The graph of upload_edges, built on the GPU (dcx_graph_ingest_edges)."""
        src = np.ascontiguousarray(src, np.uint32)
        dst = np.ascontiguousarray(dst, np.uint32)
        h = ctypes.c_void_p()
        check(lib().dcx_graph_ingest_edges(self.h, u32(V), u64(src.size), P(src), P(dst),
                                           ctypes.c_int32(0), u32(slack), ctypes.byref(h)),
              "dcx_graph_ingest_edges", self.h)
        g = Graph.__new__(Graph)
        g.ctx = self
        g.h = h
        self._adopt(g)
        g.V = V
        return g

    def to_device(self, a):
        """A device copy of a uint32 array, allocated by the HIP runtime the
        library uses (libamdhip64.so.7; another runtime in the same process,
        such as torch's, cannot share the device with it)."""
        return DeviceArray(np.ascontiguousarray(a, np.uint32))

    def ingest_device(self, V, src, dst, slack=0, E=None):
        """dcx_graph_ingest_edges from two DeviceArray objects (the first E
        entries of each, default all of them)."""
        return self.ingest(V, src, dst, slack=slack, E=E)

    def ingest(self, V, src, dst, slack=0, drop_self_loops=False, dedup=False, E=None):
        """dcx_graph_ingest: src/dst are numpy arrays (copied to the device)
        or DeviceArray objects; E (device arrays only) takes a prefix."""
        flags = (INGEST_DROP_SELF_LOOPS if drop_self_loops else 0) | (INGEST_DEDUP if dedup else 0)
        if isinstance(src, DeviceArray) != isinstance(dst, DeviceArray):
            raise TypeError("src and dst must both be host arrays or both DeviceArray")
        if isinstance(src, DeviceArray):
            n = min(src.size, dst.size)
            E = n if E is None else int(E)
            if E < 0 or E > n:
                raise ValueError(f"E = {E} exceeds the device arrays ({src.size}, {dst.size})")
            flags |= INGEST_ON_DEVICE
            ps, pd = src.ptr, dst.ptr
        else:
            if E is not None:
                raise ValueError("E applies to device arrays only")
            src = np.ascontiguousarray(src, np.uint32)
            dst = np.ascontiguousarray(dst, np.uint32)
            if src.size != dst.size:
                raise ValueError("src and dst differ in length")
            E = src.size
            ps, pd = P(src), P(dst)
        h = ctypes.c_void_p()
        check(lib().dcx_graph_ingest(self.h, u32(V), u64(E), ps, pd, u32(flags), u32(slack),
                                     ctypes.byref(h)),
              "dcx_graph_ingest", self.h)
        g = Graph.__new__(Graph)
        g.ctx = self
        g.h = h
        self._adopt(g)
        g.V = V
        return g

    def gen_rmat_device(self, scale, samples, seed, a=0.57, b=0.19, c=0.19):
        """Raw R-MAT samples (those of gen_rmat before its dedup) drawn on
        the device: two DeviceArray of `samples` uint32."""
        src = DeviceArray.empty(samples)
        dst = DeviceArray.empty(samples)
        check(lib().dcx_gen_rmat_device(self.h, u32(scale), u64(samples), ctypes.c_double(a),
                                        ctypes.c_double(b), ctypes.c_double(c), u64(seed),
                                        src.ptr, dst.ptr),
              "dcx_gen_rmat_device", self.h)
        return src, dst

    def upload_csr(self, V, E, csr):
        return Graph(self, V, None, None, 0, csr=(E, csr))

    def upload_batch(self, graphs, slack=0):
        """graphs: list of (V, src, dst).  Returns (Graph of the disjoint
        union, vbase[n+1])."""
        n = len(graphs)
        Vs = np.array([g[0] for g in graphs], np.uint32)
        srcs = [np.ascontiguousarray(g[1], np.uint32) for g in graphs]
        dsts = [np.ascontiguousarray(g[2], np.uint32) for g in graphs]
        Es = np.array([s.size for s in srcs], np.uint64)
        SP = ctypes.c_void_p * max(n, 1)
        sp = SP(*[P(s) for s in srcs])
        dp = SP(*[P(d) for d in dsts])
        vbase = np.empty(n + 1, np.uint32)
        h = ctypes.c_void_p()
        check(lib().dcx_graph_upload_batch(self.h, u32(n), P(Vs), P(Es), sp, dp, u32(slack),
                                           P(vbase), ctypes.byref(h)),
              "dcx_graph_upload_batch", self.h)
        g = Graph.__new__(Graph)
        g.ctx = self
        g.h = h
        self._adopt(g)
        g.V = int(vbase[-1])
        return g, vbase


class DeviceArray:
    """Device memory holding a copy of a host uint32 array (hipMalloc and a
    synchronous copy through the library's own HIP runtime)."""

    def __init__(self, a=None, size=None):
        lib()
        hip = ctypes.CDLL("libamdhip64.so.7")
        self._hip = hip
        self.ptr = ctypes.c_void_p()
        self.size = a.size if a is not None else int(size)
        nbytes = 4 * self.size
        if hip.hipMalloc(ctypes.byref(self.ptr), ctypes.c_size_t(max(nbytes, 4))) != 0:
            raise DcxError(3, "hipMalloc")
        if a is not None and nbytes and hip.hipMemcpy(self.ptr, P(a), ctypes.c_size_t(nbytes),
                                                      1) != 0:
            self.free()
            raise DcxError(3, "hipMemcpy")

    @classmethod
    def empty(cls, n):
        """Uninitialised device memory for n uint32."""
        return cls(None, size=n)

    def numpy(self):
        out = np.empty(self.size, np.uint32)
        if self.size and self._hip.hipMemcpy(P(out), self.ptr, ctypes.c_size_t(4 * self.size),
                                             2) != 0:
            raise DcxError(3, "hipMemcpy")
        return out

    def free(self):
        if self.ptr:
            self._hip.hipFree(self.ptr)
            self.ptr = ctypes.c_void_p()


class Graph:
    """A device-resident graph (dcx_graph) with its resident decomposition."""

    def __init__(self, ctx, V, src, dst, slack=0, csr=None):
        self.ctx = ctx
        h = ctypes.c_void_p()
        if csr is None:
            src = np.ascontiguousarray(src, np.uint32)
            dst = np.ascontiguousarray(dst, np.uint32)
            check(lib().dcx_graph_upload_edges(ctx.h, u32(V), u64(src.size), P(src), P(dst),
                                               u32(slack), ctypes.byref(h)),
                  "dcx_graph_upload_edges", ctx.h)
        else:
            E, g = csr
            check(lib().dcx_graph_upload_csr(ctx.h, u32(V), u64(E), P(g["out_off"]),
                                             P(g["out_deg"]), P(g["out_nbr"]), P(g["in_off"]),
                                             P(g["in_deg"]), P(g["in_nbr"]), ctypes.byref(h)),
                  "dcx_graph_upload_csr", ctx.h)
        self.h = h
        self.V = V
        ctx._adopt(self)

    def free(self):
        if self.h:
            if self.ctx.h:
                lib().dcx_graph_free(self.h)
            self.h = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass

    def info(self):
        V = ctypes.c_uint32()
        E = ctypes.c_uint64()
        K = ctypes.c_int32()
        Lm = ctypes.c_int32()
        check(lib().dcx_graph_info(self.h, ctypes.byref(V), ctypes.byref(E), ctypes.byref(K),
                                   ctypes.byref(Lm)), "dcx_graph_info")
        return V.value, E.value, K.value, Lm.value

    def kmax(self):
        km = np.empty(self.V, np.int32)
        K = ctypes.c_int32()
        check(lib().dcx_kmax(self.ctx.h, self.h, P(km), ctypes.byref(K)), "dcx_kmax", self.ctx.h)
        return K.value, km

    def klist(self, k, variant=VARIANT_AUTO):
        l = np.empty(self.V, np.int32)
        lk = ctypes.c_int32()
        check(lib().dcx_klist(self.ctx.h, self.h, i32(k), i32(variant), P(l), ctypes.byref(lk)),
              "dcx_klist", self.ctx.h)
        return lk.value, l

    def klist_noalloc(self, k, out):
        """dcx_klist_noalloc: row k on the whole graph into `out`, a
        DeviceArray of V uint32 (read back as int32 by the caller); returns
        lmax_k."""
        if out.size < self.V:
            raise ValueError("output array smaller than V")
        lk = ctypes.c_int32()
        check(lib().dcx_klist_noalloc(self.ctx.h, self.h, i32(k), out.ptr, ctypes.byref(lk)),
              "dcx_klist_noalloc", self.ctx.h)
        return lk.value

    def decompose(self, variant=VARIANT_AUTO):
        K = ctypes.c_int32()
        Lm = ctypes.c_int32()
        check(lib().dcx_decompose(self.ctx.h, self.h, i32(variant), ctypes.byref(K),
                                  ctypes.byref(Lm)), "dcx_decompose", self.ctx.h)
        return K.value, Lm.value

    def kmaxes(self):
        km = np.empty(self.V, np.int32)
        check(lib().dcx_get_kmaxes(self.ctx.h, self.h, P(km)), "dcx_get_kmaxes", self.ctx.h)
        return km

    def rows(self, k0=0, nrows=None):
        K = self.info()[2]
        if nrows is None:
            nrows = K + 1 - k0
        out = np.empty((nrows, self.V), np.int32)
        check(lib().dcx_get_rows(self.ctx.h, self.h, i32(k0), i32(nrows), P(out)),
              "dcx_get_rows", self.ctx.h)
        return out

    def ragged(self):
        n = ctypes.c_uint64()
        check(lib().dcx_ragged_size(self.ctx.h, self.h, ctypes.byref(n)), "dcx_ragged_size",
              self.ctx.h)
        rp = np.empty(self.V + 1, np.int64)
        vals = np.empty(max(n.value, 1), np.int32)
        check(lib().dcx_get_ragged(self.ctx.h, self.h, P(rp), P(vals)), "dcx_get_ragged",
              self.ctx.h)
        return rp, vals[:n.value]

    def autotune(self, reps=2):
        """Tunes the batched rows launch on this graph (dcx_autotune); the
        context keeps the winning shape.  Returns TuneResult."""
        r = TuneResult()
        check(lib().dcx_autotune(self.ctx.h, self.h, i32(reps), ctypes.byref(r)), "dcx_autotune",
              self.ctx.h)
        return r

    def pass_log(self):
        """Device log of the resident decomposition's passes (dcx_graph_pass_log):
        a list of dicts, entry 0 the kmax pass, entry 1+k row k."""
        n = ctypes.c_uint32()
        check(lib().dcx_graph_pass_log(self.ctx.h, self.h, None, u32(0), ctypes.byref(n)),
              "dcx_graph_pass_log", self.ctx.h)
        arr = (PassLog * max(n.value, 1))()
        check(lib().dcx_graph_pass_log(self.ctx.h, self.h, arr, u32(n.value), ctypes.byref(n)),
              "dcx_graph_pass_log", self.ctx.h)
        return [{f: getattr(arr[i], f) for f, _ in PassLog._fields_} for i in range(n.value)]

    def pass_features(self, with_log=True):
        """(X, T) training samples from the last batched rows launch."""
        n = ctypes.c_uint32()
        check(lib().dcx_graph_pass_features(self.ctx.h, self.h, i32(1 if with_log else 0),
                                            u32(0), None, None, ctypes.byref(n)),
              "dcx_graph_pass_features", self.ctx.h)
        X = np.empty((max(n.value, 1), MLP.N_IN), np.float32)
        T = np.empty(max(n.value, 1), np.float32)
        check(lib().dcx_graph_pass_features(self.ctx.h, self.h, i32(1 if with_log else 0),
                                            u32(n.value), P(X), P(T), ctypes.byref(n)),
              "dcx_graph_pass_features", self.ctx.h)
        return X[:n.value], T[:n.value]

    def state_hash(self):
        """dcx_state_hash of the resident decomposition."""
        h = ctypes.c_uint64()
        check(lib().dcx_state_hash(self.ctx.h, self.h, ctypes.byref(h)), "dcx_state_hash",
              self.ctx.h)
        return h.value

    def compare(self, other):
        """dcx_compare_decompositions against another resident graph of the
        same V: (mismatching cells, first such cell or None); cell v is
        kmaxes[v], cell V * (1 + k) + v is row k of v."""
        n, f = ctypes.c_uint64(), ctypes.c_uint64()
        check(lib().dcx_compare_decompositions(self.ctx.h, self.h, other.h, ctypes.byref(n),
                                               ctypes.byref(f)),
              "dcx_compare_decompositions", self.ctx.h)
        return n.value, (None if n.value == 0 else f.value)

    def serve(self, **kw):
        """Starts a maintenance server on this graph (dcx_serve_start)."""
        return Server(self, **kw)

    def batch_summary(self, vbase):
        """Per-part (kmax, lmax) arrays of a decomposed batch."""
        vbase = np.ascontiguousarray(vbase, np.uint32)
        n = vbase.size - 1
        km = np.empty(n, np.int32)
        lm = np.empty(n, np.int32)
        check(lib().dcx_batch_summary(self.ctx.h, self.h, u32(n), P(vbase), P(km), P(lm)),
              "dcx_batch_summary", self.ctx.h)
        return km, lm

    def kmax_edge_batches(self, pairs, mode=0):
        """getKmaxEdgeBatches on the device from the resident kmaxes
        (dcx_kmax_edge_batches_device): (batch of each edge, batch count)."""
        pairs = _pairs(pairs)
        n = pairs.size // 2
        out = np.empty(max(n, 1), np.uint32)
        nb = ctypes.c_uint32()
        check(lib().dcx_kmax_edge_batches_device(self.ctx.h, self.h, P(pairs), u32(n), i32(mode),
                                                 P(out), ctypes.byref(nb)),
              "dcx_kmax_edge_batches_device", self.ctx.h)
        return out[:n], nb.value

    def update(self, ins=None, dels=None):
        """Maintained batch (dcx_update): deletions first, then insertions.
        ins/dels: arrays of shape (n, 2).  Returns UpdateStats."""
        ins = _pairs(ins)
        dels = _pairs(dels)
        st = UpdateStats()
        check(lib().dcx_update(self.ctx.h, self.h, P(ins), u32(ins.size // 2), P(dels),
                               u32(dels.size // 2), ctypes.byref(st)), "dcx_update", self.ctx.h)
        return st

    def insert_edges(self, pairs):
        pairs = _pairs(pairs)
        check(lib().dcx_graph_insert_edges(self.ctx.h, self.h, P(pairs), u32(pairs.size // 2)),
              "dcx_graph_insert_edges", self.ctx.h)

    def delete_edges(self, pairs):
        pairs = _pairs(pairs)
        check(lib().dcx_graph_delete_edges(self.ctx.h, self.h, P(pairs), u32(pairs.size // 2)),
              "dcx_graph_delete_edges", self.ctx.h)

    def download_csr(self):
        oc = ctypes.c_uint64()
        ic = ctypes.c_uint64()
        check(lib().dcx_graph_capacity(self.h, ctypes.byref(oc), ctypes.byref(ic)),
              "dcx_graph_capacity")
        V = self.V
        g = dict(out_off=np.empty(V + 1, np.uint32), out_deg=np.empty(V, np.int32),
                 out_nbr=np.empty(max(oc.value, 1), np.uint32),
                 in_off=np.empty(V + 1, np.uint32), in_deg=np.empty(V, np.int32),
                 in_nbr=np.empty(max(ic.value, 1), np.uint32))
        check(lib().dcx_graph_download_csr(self.ctx.h, self.h, P(g["out_off"]), P(g["out_deg"]),
                                           P(g["out_nbr"]), P(g["in_off"]), P(g["in_deg"]),
                                           P(g["in_nbr"])), "dcx_graph_download_csr",
              self.ctx.h)
        g["out_nbr"] = g["out_nbr"][:oc.value]
        g["in_nbr"] = g["in_nbr"][:ic.value]
        return g


def _pairs(p):
    if p is None:
        return np.zeros(0, np.uint32)
    return np.ascontiguousarray(np.asarray(p, dtype=np.uint32).reshape(-1))


def kmax_edge_batches(V, kmaxes, pairs, mode=0):
    """getKmaxEdgeBatches restated in C (host only): batch id per edge."""
    kmaxes = np.ascontiguousarray(kmaxes, np.int32)
    pairs = _pairs(pairs)
    n = pairs.size // 2
    out = np.empty(max(n, 1), np.uint32)
    nb = ctypes.c_uint32()
    check(lib().dcx_kmax_edge_batches(u32(V), P(kmaxes), P(pairs), u32(n), i32(mode), P(out),
                                      ctypes.byref(nb)), "dcx_kmax_edge_batches")
    return out[:n], nb.value


def compact_csr(g):
    """Compact (no slack) out/in lists from a CSR-with-slack dict, as
    (out_off, out_nbr, in_off, in_nbr) with uint64 offsets."""
    res = []
    for side in ("out", "in"):
        off = g[side + "_off"].astype(np.int64)
        deg = g[side + "_deg"].astype(np.int64)
        V = deg.size
        coff = np.zeros(V + 1, np.uint64)
        np.cumsum(deg, out=coff[1:])
        idx = np.repeat(off[:-1], deg) + (np.arange(int(coff[-1])) - np.repeat(coff[:-1].astype(np.int64), deg))
        res += [coff, g[side + "_nbr"][idx].astype(np.uint32)]
    return res


class MLP:
    """Cost model of the batched-rows scheduler (dcx_mlp_*): 8 inputs, two
    hidden layers of 32 ReLU units, one output, fp32, on the device."""

    N_IN = 8

    def __init__(self, ctx, seed=1):
        self.ctx = ctx
        self.h = ctypes.c_void_p()
        check(lib().dcx_mlp_create(ctx.h, u64(seed), ctypes.byref(self.h)), "dcx_mlp_create",
              ctx.h)
        self.n_params = int(lib().dcx_mlp_param_count())
        ctx._adopt(self)

    def free(self):
        if self.h:
            if self.ctx.h:
                check(lib().dcx_mlp_free(self.h), "dcx_mlp_free")
            self.h = None

    @property
    def params(self):
        out = np.empty(self.n_params, np.float32)
        check(lib().dcx_mlp_get_params(self.h, P(out)), "dcx_mlp_get_params", self.ctx.h)
        return out

    @params.setter
    def params(self, p):
        p = np.ascontiguousarray(p, np.float32)
        assert p.size == self.n_params
        check(lib().dcx_mlp_set_params(self.h, P(p)), "dcx_mlp_set_params", self.ctx.h)

    def predict(self, X):
        X = np.ascontiguousarray(X, np.float32).reshape(-1, self.N_IN)
        y = np.empty(X.shape[0], np.float32)
        check(lib().dcx_mlp_predict(self.h, P(X), u32(X.shape[0]), P(y)), "dcx_mlp_predict",
              self.ctx.h)
        return y

    def gradient(self, X, T):
        X = np.ascontiguousarray(X, np.float32).reshape(-1, self.N_IN)
        T = np.ascontiguousarray(T, np.float32)
        g = np.empty(self.n_params, np.float32)
        loss = ctypes.c_float()
        check(lib().dcx_mlp_gradient(self.h, P(X), P(T), u32(X.shape[0]), P(g),
                                     ctypes.byref(loss)), "dcx_mlp_gradient", self.ctx.h)
        return g, loss.value

    def train(self, X, T, epochs=500, lr=3e-3):
        X = np.ascontiguousarray(X, np.float32).reshape(-1, self.N_IN)
        T = np.ascontiguousarray(T, np.float32)
        loss = ctypes.c_float()
        check(lib().dcx_mlp_train(self.h, P(X), P(T), u32(X.shape[0]), u32(epochs),
                                  ctypes.c_float(lr), ctypes.byref(loss)), "dcx_mlp_train",
              self.ctx.h)
        return loss.value


class Server:
    """Continuous-batching maintenance server (dcx_serve_*).  While it runs,
    the graph refuses other calls; stop() drains every submitted request."""

    def __init__(self, graph, hbm_budget=0, max_inflight=0, max_group=0, max_group_edges=0,
                 hash_states=False):
        cfg = ServeConfig()
        check(lib().dcx_serve_config_init(ctypes.byref(cfg)), "dcx_serve_config_init")
        if hbm_budget:
            cfg.hbm_budget = hbm_budget
        if max_inflight:
            cfg.max_inflight = max_inflight
        if max_group:
            cfg.max_group = max_group
        if max_group_edges:
            cfg.max_group_edges = max_group_edges
        cfg.hash_states = 1 if hash_states else 0
        self.graph = graph
        self.h = ctypes.c_void_p()
        check(lib().dcx_serve_start(graph.ctx.h, graph.h, ctypes.byref(cfg), ctypes.byref(self.h)),
              "dcx_serve_start", graph.ctx.h)
        lib().dcx_serve_last_error.restype = ctypes.c_char_p

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        if self.h:
            try:
                self.stop()
            except DcxError:
                if exc[0] is None:
                    raise
        return False

    def __del__(self):
        # never leave the worker thread running into interpreter shutdown
        if getattr(self, "h", None):
            try:
                self.stop()
            except Exception:
                pass

    def _check(self, status, where):
        if status != 0:
            detail = lib().dcx_serve_last_error(self.h).decode(errors="replace")
            raise DcxError(status, where, detail)

    def submit(self, ins=None, dels=None):
        """Queues one request; returns its id."""
        ins = _pairs(ins)
        dels = _pairs(dels)
        rid = ctypes.c_uint64()
        self._check(lib().dcx_serve_submit(self.h, P(ins), u32(ins.size // 2), P(dels),
                                           u32(dels.size // 2), ctypes.byref(rid)),
                    "dcx_serve_submit")
        return rid.value

    def poll(self, cap=4096, wait_ms=0):
        """Finished results (list of dicts), waiting up to wait_ms for one."""
        arr = (ServeResult * cap)()
        n = ctypes.c_uint32()
        self._check(lib().dcx_serve_poll(self.h, arr, u32(cap), ctypes.byref(n), i32(wait_ms)),
                    "dcx_serve_poll")
        return [{f: getattr(arr[i], f) for f, _ in ServeResult._fields_} for i in range(n.value)]

    def stop(self):
        """Drains, stops and frees the server; returns its stats as a dict."""
        st = ServeStats()
        detail = lib().dcx_serve_last_error(self.h).decode(errors="replace")
        h, self.h = self.h, None
        status = lib().dcx_serve_stop(h, ctypes.byref(st))
        if status != 0:
            raise DcxError(status, "dcx_serve_stop", detail)
        return {f: getattr(st, f) for f, _ in ServeStats._fields_ if f != "reserved"}
