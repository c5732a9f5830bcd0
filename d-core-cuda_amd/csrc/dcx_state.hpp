// Context and graph state behind the dcx C-ABI, and the helpers every entry
// point uses (error capture, device allocation, events).  Included once, by
// dcx_api.hip.
#pragma once
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <memory>
#include <new>
#include <vector>

#include "dcx_internal.hpp"

using namespace dcx;

struct dcx_ctx {
    dcx_config cfg{};
    int device = 0;
    int num_cus = 0;
    int grid_blocks = 0;    // workgroups of a whole-grid pass / of the rows launch
    int hindex_grid = 0;
    int rows_per_cu = 1;    // team rows workgroups that fit one CU (occupancy)
    uint64_t timeout_ticks = 0;  // wall-clock ticks one team barrier may take
    double wall_khz = 100000.0;  // device wall clock rate
    uint32_t* trace = nullptr;   // phase trace of the last grid pass (cfg.trace_phases)
    uint32_t* fail_rec = nullptr;  // [kFailWords] first failing pass member's view
    dcx_mlp* sched_mlp = nullptr;  // cost model of the rows scheduler (nullptr: heuristic)
    uint32_t lds_cap = 0;      // most members of a row run LDS-resident (0: none)
    uint32_t lds_cap_max = 0;  // what the device's LDS allows (lds_cap <= this)
    hipStream_t stream = nullptr;
    // second stream (default priority) for the LDS-resident rows, which run
    // beside the team launch; fork/join order it against `stream`
    hipStream_t side = nullptr;
    hipEvent_t fork = nullptr, join = nullptr;
    // third stream (default priority): row 0's pass beside the kmax pass
    // and the order kernels; row0_done marks its end (timed), order_done
    // the end of the order kernels on `stream`
    hipStream_t side2 = nullptr;
    hipEvent_t row0_done = nullptr, order_done = nullptr;
    hipEvent_t clear_done = nullptr;  // the rows' clear queued on side2 after row 0
    // [0], [1]: brackets of the LDS-resident launch on `side`; [3], [2]:
    // start and end of the team launch on `stream` (ev[5] opens the rows
    // kernels, row 0's grid pass first when the other rows use its bound)
    hipEvent_t sev[4] = {};
    int32_t* zeros = nullptr;  // V zero words (dcx_compare_decompositions)
    size_t zeros_n = 0;
    std::mutex mu;
    std::string err;
    dcx_timing timing{};
    // this device's copy of the static scratch array of dcx_klist_noalloc
    // (from hipGetSymbolAddress at context creation)
    void* noalloc_scratch = nullptr;
    // grown-on-demand device scratch
    void* scratch = nullptr;
    size_t scratch_bytes = 0;
    void* tmp = nullptr;  // core-graph build temporaries
    size_t tmp_bytes = 0;
    // small device words: [0,12) grid-pass phase slots, [16,20) h-index sweep
    // flags, [24] batched error word, [64, 1024) grid barrier words
    uint32_t* ctrs = nullptr;
    uint32_t* ctrs2 = nullptr;  // slots and barrier words of the second pass of a pair
    int32_t* d_scalars = nullptr;  // 64 words of kernel results
    int32_t* h_scalars = nullptr;  // pinned mirror (words 32.. hold error words)
    hipEvent_t ev[8] = {};
    // team schedule of the rows launch (device copies, grown on demand)
    void* sched = nullptr;
    size_t sched_bytes = 0;
    uint32_t* pools = nullptr;  // team sync words + team slots
    size_t pools_words = 0;
    // localized maintenance: per-slot worklists, task list, small words
    // ([0] task counter, [1] recorded kmax changes, [2] failure bits,
    // [3] largest kmax, [4] largest row-0 value, [5] edits not found), and
    // the per-edge old kmax minimum of the deletions
    uint32_t* local_scratch = nullptr;
    size_t local_scratch_bytes = 0;
    LocalTaskDesc* local_tasks = nullptr;
    LocalTaskDesc* local_tasks_host = nullptr;  // pinned
    size_t local_tasks_cap = 0;
    uint32_t* local_words = nullptr;
    int32_t* local_dm = nullptr;
    size_t local_dm_cap = 0;
    unsigned long long local_counters[20] = {};  // of the last local maintenance
    // staging for update batches (device + pinned host), grown on demand
    uint32_t* upd_dev = nullptr;
    size_t upd_dev_words = 0;
    uint32_t* upd_host = nullptr;
    size_t upd_host_words = 0;
};

namespace {
// ctrs layout (32-bit words): [0, kSlotWords) phase slots of grid passes;
// three h-index change flags kSlotStride apart from kCtrFlags; the error
// word of single-workgroup passes; the row queue of the LDS-resident rows
// kernel; the grid barrier's words from kCtrSync.  Every group starts its
// own 128-byte line.
constexpr int kCtrFlags = 96, kCtrErr = 192, kCtrQueue = 224, kCtrSync = 256, kCtrWords = 1024;
constexpr uint32_t kMaxGridBlocks = kCtrWords - kCtrSync - kSyncFlags;
// Workgroups per CU for grid-wide passes: far below the occupancy limit, so
// residency of the whole grid is never in question.
constexpr int kGridBlocksPerCu = 2;
constexpr double kBarrierTimeoutS = 30.0;
// Vertex ids must stay below 2^31 (the peel combines 2 * id + side keys).
constexpr uint32_t kMaxVertices = 0x7fffffffu;
}  // namespace

// Cost model of the variant scheduler (mlp_kernels.hip): parameters,
// Adam moments, gradient and a staging buffer on the device.
struct dcx_mlp {
    dcx_ctx* ctx = nullptr;
    float* p = nullptr;     // [kMlpParams]
    float* m = nullptr;     // Adam first moment
    float* v = nullptr;     // Adam second moment
    float* grad = nullptr;  // [kMlpParams]
    float* loss = nullptr;  // [1]
    int step = 0;           // Adam steps taken
    float* buf = nullptr;   // staging for samples and predictions
    size_t buf_floats = 0;
};

struct dcx_graph {
    dcx_ctx* ctx = nullptr;
    bool served = false;  // owned by a running dcx_server: API calls refused
    bool broken = false;  // an edit failed half-way: the adjacency is not trusted
    bool audited = false; // the adjacency passed the audit and was not edited since
    uint64_t edits = 0;   // edit batches applied (plans measured before one are stale)
    uint32_t V = 0;
    uint64_t E = 0;
    // Adjacency arrays: lists start at off[v] with room for capv[v] entries;
    // *_cap is the used prefix (a bump pointer for relocated lists), *_alloc
    // the allocated length.  Free slots hold 0xFFFFFFFF.
    uint64_t out_cap = 0, in_cap = 0;
    uint64_t out_alloc = 0, in_alloc = 0;
    uint32_t *out_off = nullptr, *in_off = nullptr, *out_nbr = nullptr, *in_nbr = nullptr;
    int32_t *out_deg = nullptr, *in_deg = nullptr;
    uint32_t *out_capv = nullptr, *in_capv = nullptr;
    uint64_t relocations = 0;
    // resident results
    int32_t* kmaxes = nullptr;
    int32_t kmax = -1;
    bool kmax_valid = false;
    uint32_t* order = nullptr;   // [V] vertices by kmax, descending (rank -> id)
    uint32_t* vk = nullptr;      // [kmax+2] members per k
    uint32_t* hist = nullptr;    // [kmax+2]
    uint64_t* ekh = nullptr;     // [kmax+3] edge histogram by min kmax, queue length
    uint64_t* ekq = nullptr;     // its queue of long-list slices
    size_t ekq_cap = 0;
    size_t vk_cap = 0;
    std::vector<uint32_t> h_vk;  // host copy of vk
    std::vector<uint64_t> h_ek;  // E_k, k = 0..kmax
    // core-ordered graph (rebuilt with every kmax computation)
    uint32_t* rank = nullptr;
    uint32_t *cout_off = nullptr, *cin_off = nullptr, *cout_nbr = nullptr, *cin_nbr = nullptr;
    int32_t *cout_deg = nullptr, *cin_deg = nullptr;
    size_t core_e_cap = 0;
    bool core_valid = false;
    int32_t* kmax_old = nullptr; // [V] previous kmaxes during an update
    int32_t* row_bound = nullptr; // [V] row 0 by rank: bound of the batched rows' values
    int32_t rows_first_grid = 0;  // batched: leading rows not run by the rows launch (row 0)
    // Row 0 computed beside the kmax pass (it needs no kmaxes): values by
    // original id, its pass-log words and lmax; valid until the edges change.
    int32_t* row0 = nullptr;
    uint32_t* row0_stats = nullptr;
    bool row0_ready = false;
    bool row0_pending = false;  // row 0's pass queued on side2, not yet joined
    // Sixteenths of the grid row 0's pass takes beside the kmax pass (2..8),
    // adapted after every such run: both passes are chains of dependent
    // phases on power-law graphs (few workgroups suffice) but
    // throughput-bound on graphs with few levels (half each).
    int32_t row0_16ths = 4;
    size_t rows_cleared = 0;  // cells of `rows` zeroed on side2 for the next rows launch
    uint32_t* pstats = nullptr;  // [(kmax+2) * kStatWords] device pass log: kmax pass, rows
    size_t pstats_cap = 0;
    std::vector<uint32_t> h_pstats;      // host copy of pstats after the last rows launch
    // Work statistics of the team rows over the batched launches so far,
    // value + 1 (0: row k never ran that way): phases of its last peel;
    // rounds and list entries read of its last derivation from the row
    // above, and of its last derivation from row 0 (a chain's supports
    // set-up excluded).  Cost-model features.
    std::vector<uint32_t> h_peel_phases, h_der_rounds, h_far_rounds;
    std::vector<uint64_t> h_der_reads, h_far_reads;
    // Time (ms, 0: none) and team width of row k's last run of each kind of
    // team row (peeled, first derived, derived, derived from row 0): the
    // cost model's predictions for row k and that kind are scaled to it.
    std::vector<float> h_meas_ms[4];
    std::vector<uint16_t> h_meas_team[4];
    // A plan of the cost model for the team rows 1..ksplit-1 (anchors,
    // their kinds and chains, widths and weights) with the team launch's
    // measured time: the fastest one measured since the last edit, and the
    // one of the last launch until its time is known.
    struct RowsPlan {
        bool valid = false;
        int32_t ksplit = -1;
        uint64_t edits = 0;
        std::vector<char> peel, far;
        std::vector<int32_t> chain_last;
        std::vector<uint32_t> want;
        std::vector<double> wts;
        float teams_ms = 0.0f;
    };
    RowsPlan best_plan, last_plan;
    std::vector<int8_t> h_pass_variant;  // [kmax+2] variant that last ran each pass
    std::vector<uint16_t> h_pass_team;   // [kmax+2] workgroups that ran it
    int32_t* rows = nullptr;     // (kmax+1) x V
    size_t rows_cap = 0;         // rows allocated
    int32_t rows_k = 0;          // rows currently held (kmax+1 when valid)
    int32_t* lmax_rows = nullptr;
    size_t lmax_rows_cap = 0;
    int32_t lmax = -1;
    bool rows_valid = false;
    // the kmax order, V_k and E_k follow the resident kmaxes (false after
    // local maintenance changed kmaxes; recomputed when a pass needs them)
    bool order_valid = false;
    // host time of the last full decomposition or recomputation of this
    // graph (ms; 0 = none yet): the time budget of local maintenance
    double full_ms = 0.0;
    // budgeted local maintenance: batches still to send to recomputation,
    // and the current backoff (0 after a batch kept local)
    uint32_t local_skip = 0, local_backoff = 0;
    // localized maintenance: the batch's entries at the end of each list
    // ([V], zero outside a maintenance), and the vertices whose kmax fell
    int32_t* ins_cnt_in = nullptr;
    int32_t* ins_cnt_out = nullptr;
    uint32_t* changed = nullptr;  // [2 V]
};

namespace {

struct Fail {
    dcx_status st;
};

void set_err(dcx_ctx* c, const char* fmt, ...) {
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    std::vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    c->err = buf;
}

#define HIPX(ctx, call)                                                                      \
    do {                                                                                     \
        hipError_t e_ = (call);                                                              \
        if (e_ != hipSuccess) {                                                              \
            set_err((ctx), "%s failed: %s (%s:%d)", #call, hipGetErrorString(e_), __FILE__, \
                    __LINE__);                                                               \
            throw Fail{e_ == hipErrorOutOfMemory ? DCX_ERR_OUT_OF_MEMORY : DCX_ERR_HIP};     \
        }                                                                                    \
    } while (0)

template <class F>
dcx_status guarded(dcx_ctx* c, F&& f) {
    if (!c) return DCX_ERR_INVALID_ARG;
    std::lock_guard<std::mutex> lk(c->mu);
    c->err.clear();
    try {
        if (hipSetDevice(c->device) != hipSuccess) {
            set_err(c, "hipSetDevice(%d) failed", c->device);
            return DCX_ERR_NO_DEVICE;
        }
        return f();
    } catch (const Fail& x) {
        // a pass queued on the second side stream may still run on the
        // scratch arrays: let it end before the next call reuses them
        if (c->side2) (void)hipStreamSynchronize(c->side2);
        return x.st;
    } catch (const std::bad_alloc&) {
        set_err(c, "host allocation failed");
        return DCX_ERR_OUT_OF_MEMORY;
    }
}

template <class T>
void dev_alloc(dcx_ctx* c, T** p, size_t n) {
    *p = nullptr;
    if (n == 0) n = 1;
    HIPX(c, hipMalloc(reinterpret_cast<void**>(p), n * sizeof(T)));
}

template <class T>
void dev_free(T*& p) {
    if (p) (void)hipFree(p);
    p = nullptr;
}

void ensure_bytes(dcx_ctx* c, void** p, size_t* have, size_t bytes) {
    if (*have >= bytes) return;
    if (*p) (void)hipFree(*p);
    *p = nullptr;
    *have = 0;
    HIPX(c, hipMalloc(p, bytes ? bytes : 1));
    *have = bytes;
}

void ensure_scratch(dcx_ctx* c, size_t bytes) {
    if (c->cfg.scratch_budget && bytes > c->cfg.scratch_budget) {
        set_err(c, "pass scratch of %zu bytes exceeds the budget of %llu", bytes,
                (unsigned long long)c->cfg.scratch_budget);
        throw Fail{DCX_ERR_BUDGET};
    }
    ensure_bytes(c, &c->scratch, &c->scratch_bytes, bytes);
}

void sync(dcx_ctx* c) { HIPX(c, hipStreamSynchronize(c->stream)); }

float elapsed(dcx_ctx* c, int a, int b) {
    float ms = 0;
    HIPX(c, hipEventElapsedTime(&ms, c->ev[a], c->ev[b]));
    return ms;
}

GridSync gsync(dcx_ctx* c, int grid) {
    GridSync g;
    g.w = c->ctrs + kCtrSync;
    g.nblocks = (uint32_t)grid;
    g.rank = 0;
    g.timeout_ticks = c->timeout_ticks;
    g.epoch = 0;
    return g;
}

// Clears the kernel error words before a series of launches.
void clear_kernel_errors(dcx_ctx* c) {
    HIPX(c, hipMemsetAsync(c->ctrs + kCtrSync + kSyncError, 0, sizeof(uint32_t), c->stream));
    HIPX(c, hipMemsetAsync(c->ctrs + kCtrErr, 0, sizeof(uint32_t), c->stream));
}

// Reads the kernel error words (synchronises) and turns them into a status.
// `team_err` is an extra device error word (team pools) or nullptr.
void check_kernel_errors(dcx_ctx* c, const char* what, const uint32_t* team_err = nullptr,
                         uint32_t nteam = 0, uint32_t stride = 0) {
    HIPX(c, hipMemcpyAsync(c->h_scalars + 32, c->ctrs + kCtrSync + kSyncError, 4,
                           hipMemcpyDeviceToHost, c->stream));
    HIPX(c, hipMemcpyAsync(c->h_scalars + 33, c->ctrs + kCtrErr, 4, hipMemcpyDeviceToHost,
                           c->stream));
    std::vector<uint32_t> te(nteam ? (size_t)nteam * stride : 0);
    if (team_err && nteam)
        HIPX(c, hipMemcpyAsync(te.data(), team_err, 4ull * nteam * stride, hipMemcpyDeviceToHost,
                               c->stream));
    sync(c);
    int32_t e = std::max(c->h_scalars[32], c->h_scalars[33]);
    int32_t first_team = -1;
    for (uint32_t t = 0; t < nteam; ++t) {
        const int32_t x = (int32_t)te[(size_t)t * stride];
        if (x && first_team < 0) first_team = (int32_t)t;
        e = std::max(e, x);
    }
    if (e == (int32_t)kErrTimeout) {
        set_err(c, "%s: a team barrier timed out (workgroups not co-resident?)", what);
        throw Fail{DCX_ERR_TIMEOUT};
    }
    if (e != 0) {
        set_err(c, "%s: kernel counters inconsistent (code %d; grid word %d, workgroup word %d, "
                   "first team %d)",
                what, e, c->h_scalars[32], c->h_scalars[33], first_team);
        throw Fail{DCX_ERR_INTERNAL};
    }
}

dcx_status check_graph(dcx_ctx* ctx, const dcx_graph* g) {
    if (!ctx || !g) return DCX_ERR_INVALID_ARG;
    if (g->ctx != ctx) {
        set_err(ctx, "graph belongs to another context");
        return DCX_ERR_INVALID_ARG;
    }
    if (g->served) {
        set_err(ctx, "graph is being served (dcx_serve_stop first)");
        return DCX_ERR_STATE;
    }
    if (g->broken) {
        set_err(ctx, "graph is inconsistent after a failed edit; upload it again");
        return DCX_ERR_STATE;
    }
    return DCX_OK;
}

}  // namespace
