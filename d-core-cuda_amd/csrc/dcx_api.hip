// C-ABI entry points of the dcx library: contexts, device graphs and the
// decomposition drivers.  Every entry point returns a dcx_status; HIP
// failures are caught at each call and reported, never ignored.
#include <chrono>

#include "dcx_state.hpp"
#include "rng.hpp"

namespace {

constexpr int32_t kChainRows = 32;  // batched: one team row in kChainRows is peeled
constexpr int32_t kFewRows = 4;  // batched: at most this many rows above row 0 run whole-graph

// kmax pass: the peel with in and out swapped and k = 0 (swapInOut,
// src/common.h:114-120), whole graph, output straight into kmaxes.
PassSpec kmax_spec(const dcx_graph* g) {
    PassSpec p{};
    p.V = g->V;
    p.K = Csr{g->in_off, g->in_deg, g->in_nbr};
    p.C = Csr{g->out_off, g->out_deg, g->out_nbr};
    p.core = false;
    p.nmembers = g->V;
    p.k = 0;
    p.lval = g->kmaxes;
    p.stats = g->pstats;
    return p;
}

// One row on the whole graph (KList semantics, no kmaxes needed).
PassSpec full_row_spec(const dcx_graph* g, int32_t k, int32_t* lval) {
    PassSpec p{};
    p.V = g->V;
    p.K = Csr{g->out_off, g->out_deg, g->out_nbr};
    p.C = Csr{g->in_off, g->in_deg, g->in_nbr};
    p.core = false;
    p.nmembers = g->V;
    p.k = k;
    p.lval = lval;
    return p;
}

// One row on the core-ordered graph (needs the resident kmaxes).
PassSpec core_row_spec(const dcx_graph* g, int32_t k, int32_t* out_row) {
    PassSpec p{};
    p.V = g->V;
    p.K = Csr{g->cout_off, g->cout_deg, g->cout_nbr};
    p.C = Csr{g->cin_off, g->cin_deg, g->cin_nbr};
    p.core = true;
    p.order = g->order;
    p.nmembers = k <= g->kmax ? g->h_vk[k] : 0u;
    p.k = k;
    p.out_row = out_row;
    p.stats = g->pstats + (size_t)kStatWords * (k + 1);
    return p;
}

// Runs one pass on the whole grid; returns its last level.  Events ev[a],
// ev[b] (if a >= 0) bracket the kernel.
constexpr uint32_t kTraceCap = 1u << 16;

// Adds the phase trace to a pass when the context asks for it:
// trace_phases 1 traces grid passes, 2 + k the team pass of row k in the
// batched launch.
PassSpec traced(dcx_ctx* c, PassSpec p, bool team = false) {
    p.trace_k = -1;
    p.pull_ratio = c->cfg.pull_ratio < 0 ? 0u
                   : c->cfg.pull_ratio == 0 ? kDefaultPullRatio
                                            : (uint32_t)c->cfg.pull_ratio;
    if (c->cfg.trace_phases <= 0 || (c->cfg.trace_phases == 1) == team) return p;
    if (team) p.trace_k = c->cfg.trace_phases - 2;
    if (!c->trace) {
        dev_alloc(c, &c->trace, (size_t)kTraceWords * (kTraceCap + 1));
        HIPX(c, hipMemsetAsync(c->trace, 0, 4, c->stream));
    }
    p.trace = c->trace;
    p.trace_cap = kTraceCap;
    return p;
}

int32_t run_grid_pass(dcx_ctx* c, const PassSpec& p0, int a = -1, int b = -1) {
    const PassSpec p = traced(c, p0);
    ensure_scratch(c, (size_t)kScratchArrays * std::max<uint32_t>(p.V, 1) * 4);
    PassScratch s = scratch_at(c->scratch, std::max<uint32_t>(p.V, 1));
    clear_kernel_errors(c);
    if (a >= 0) HIPX(c, hipEventRecord(c->ev[a], c->stream));
    HIPX(c, launch_pass_grid(p, s, c->ctrs, c->d_scalars, gsync(c, c->grid_blocks), c->stream));
    if (b >= 0) HIPX(c, hipEventRecord(c->ev[b], c->stream));
    HIPX(c, hipMemcpyAsync(c->h_scalars, c->d_scalars, sizeof(int32_t), hipMemcpyDeviceToHost,
                           c->stream));
    check_kernel_errors(c, "grid peel pass");
    return c->h_scalars[0];
}

// Checks that both adjacency sides hold exactly E valid entries before a
// rebuild sizes its buffers from E (one read of the lists; a graph that
// fails it is marked broken instead of being read out of bounds).
void audit_adjacency(dcx_ctx* c, dcx_graph* g) {
    unsigned long long* d = reinterpret_cast<unsigned long long*>(c->d_scalars + 24);
    HIPX(c, hipMemsetAsync(d, 0, 32, c->stream));
    // (the queue of long-list slices is sized for a valid side of E
    // entries; slices that do not fit count as bad entries)
    ensure_bytes(c, &c->tmp, &c->tmp_bytes, audit_tmp_bytes(g->E));
    HIPX(c, launch_audit(g->V, Csr{g->out_off, g->out_deg, g->out_nbr}, g->out_alloc, g->E, d,
                         c->tmp, c->stream));
    HIPX(c, launch_audit(g->V, Csr{g->in_off, g->in_deg, g->in_nbr}, g->in_alloc, g->E, d + 2,
                         c->tmp, c->stream));
    HIPX(c, hipMemcpyAsync(c->h_scalars + 24, d, 32, hipMemcpyDeviceToHost, c->stream));
    sync(c);
    uint64_t w[4];
    std::memcpy(w, c->h_scalars + 24, 32);
    if (w[0] != g->E || w[2] != g->E || w[1] || w[3]) {
        g->broken = true;
        g->kmax_valid = g->rows_valid = false;
        set_err(c, "adjacency audit: degree sums out %llu in %llu, E %llu, bad ids %llu/%llu",
                (unsigned long long)w[0], (unsigned long long)w[2], (unsigned long long)g->E,
                (unsigned long long)w[1], (unsigned long long)w[3]);
        throw Fail{DCX_ERR_INTERNAL};
    }
    g->audited = true;
}

// Relabels by kmax and sorts every list (see core_graph.hip); with_ek:
// the E_k histogram (g->ekh) comes out of the same pass, and the in side is
// built on the side stream beside the out side.
void build_core(dcx_ctx* c, dcx_graph* g, bool with_ek = false) {
    const uint32_t V = g->V;
    if (!g->audited) audit_adjacency(c, g);  // once per edit, not per rebuild
    if (!g->rank) {
        dev_alloc(c, &g->rank, V);
        dev_alloc(c, &g->cout_off, (size_t)V + 1);
        dev_alloc(c, &g->cin_off, (size_t)V + 1);
        dev_alloc(c, &g->cout_deg, V);
        dev_alloc(c, &g->cin_deg, V);
    }
    if (g->core_e_cap < g->E || !g->cout_nbr) {
        dev_free(g->cout_nbr);
        dev_free(g->cin_nbr);
        const size_t cap = g->E + g->E / 8 + 1024;
        dev_alloc(c, &g->cout_nbr, cap);
        dev_alloc(c, &g->cin_nbr, cap);
        g->core_e_cap = cap;
    }
    ensure_bytes(c, &c->tmp, &c->tmp_bytes, core_graph_tmp_bytes(V, g->E, g->kmax));
    HIPX(c, build_core_graph(V, g->E, g->order, g->kmaxes, g->kmax,
                             Csr{g->out_off, g->out_deg, g->out_nbr},
                             Csr{g->in_off, g->in_deg, g->in_nbr}, g->rank, g->cout_off,
                             g->cout_deg, g->cout_nbr, g->cin_off, g->cin_deg, g->cin_nbr, c->tmp,
                             g->vk, with_ek ? reinterpret_cast<unsigned long long*>(g->ekh) : nullptr,
                             c->stream, with_ek ? c->side : nullptr, c->fork, c->join));
    g->core_valid = true;
}

// Builds the core-ordered graph for the resident kmaxes unless it is
// current; the time goes to timing.order_ms.
void ensure_order(dcx_ctx* c, dcx_graph* g);

void ensure_core(dcx_ctx* c, dcx_graph* g) {
    ensure_order(c, g);
    if (g->core_valid) return;
    HIPX(c, hipEventRecord(c->ev[7], c->stream));
    build_core(c, g);
    HIPX(c, hipEventRecord(c->ev[2], c->stream));
    sync(c);
    c->timing.order_ms += elapsed(c, 7, 2);
}

// Grows a device array to n elements keeping its first `keep` elements.
template <class T>
void grow_keep(dcx_ctx* c, T** p, size_t* cap, size_t n, size_t keep) {
    if (*cap >= n) return;
    T* q = nullptr;
    dev_alloc(c, &q, n);
    if (*p && keep)
        HIPX(c, hipMemcpyAsync(q, *p, keep * sizeof(T), hipMemcpyDeviceToDevice, c->stream));
    sync(c);
    dev_free(*p);
    *p = q;
    *cap = n;
}

// In-core numbers, then the kmax order, V_k, E_k and the core-ordered
// graph.  The plain variant peels (the reference's kmax pass); the others
// run the h-index fixpoint over in-neighbours, from the in-degree or --
// `warm`, valid only when no kmax can have grown (a deletion-only update) --
// from the resident kmaxes.
// The kmax pass (in-degree peel) and row 0 (out-degree peel, which needs no
// kmaxes) in one launch, each on half of the grid: both are chains of
// ~600 barrier-separated phases, so running them side by side costs little
// more than one.  Row 0 goes to g->row0 (original ids) with its log words
// and lmax in g->row0_stats; returns kmax.
int32_t run_kmax_with_row0(dcx_ctx* c, dcx_graph* g) {
    const uint32_t V = g->V;
    const int G1 = std::max(1, c->grid_blocks * g->row0_16ths / 16), G0 = c->grid_blocks - G1;
    ensure_scratch(c, 2ull * kScratchArrays * V * 4);
    if (!c->ctrs2) {
        dev_alloc(c, &c->ctrs2, kCtrWords);
        HIPX(c, hipMemsetAsync(c->ctrs2, 0, kCtrWords * 4, c->stream));
    }
    if (!g->row0) dev_alloc(c, &g->row0, V);
    if (!g->row0_stats) dev_alloc(c, &g->row0_stats, kStatWords + 1);
    const PassSpec p0 = traced(c, kmax_spec(g));
    PassSpec p1 = full_row_spec(g, 0, g->row0);
    p1.pull_ratio = p0.pull_ratio;
    p1.stats = g->row0_stats;
    GridSync g0 = gsync(c, G0), g1 = gsync(c, G1);
    g1.w = c->ctrs2 + kCtrSync;
    clear_kernel_errors(c);
    HIPX(c, hipMemsetAsync(g1.w + kSyncError, 0, 4, c->stream));
    // row 0 on a stream of its own: it goes on beside the order kernels
    // once the kmax pass is done (joined in compute_kmax)
    HIPX(c, hipEventRecord(c->fork, c->stream));
    HIPX(c, hipStreamWaitEvent(c->side2, c->fork, 0));
    HIPX(c, launch_pass_grid(p1,
                             scratch_at(static_cast<uint32_t*>(c->scratch) +
                                            (size_t)kScratchArrays * V,
                                        V),
                             c->ctrs2, reinterpret_cast<int32_t*>(g->row0_stats + kStatWords),
                             g1, c->side2));
    HIPX(c, hipEventRecord(c->row0_done, c->side2));
    g->row0_pending = true;
    HIPX(c, hipEventRecord(c->ev[0], c->stream));
    HIPX(c, launch_pass_grid(p0, scratch_at(c->scratch, V), c->ctrs, c->d_scalars, g0,
                             c->stream));
    HIPX(c, hipEventRecord(c->ev[1], c->stream));
    HIPX(c, hipMemcpyAsync(c->h_scalars, c->d_scalars, sizeof(int32_t), hipMemcpyDeviceToHost,
                           c->stream));
    check_kernel_errors(c, "kmax pass");
    return c->h_scalars[0];
}

// The E_k histogram's queue of long-list slices (E / 2048 + 64 entries:
// a list of n > 4096 entries gives at most n / 2048 slices).
uint64_t* ensure_ekq(dcx_ctx* c, dcx_graph* g) {
    const size_t need = (size_t)(g->E / 2048) + 64;
    if (g->ekq_cap < need) {
        dev_free(g->ekq);
        g->ekq_cap = 0;
        dev_alloc(c, &g->ekq, need);
        g->ekq_cap = need;
    }
    return g->ekq;
}

// clear_rows: every row follows (a full decomposition); rows_zero: the rows
// array still holds 0 in every cell (k, v) with k > kmax(v) for the kmaxes
// being replaced (it held a valid decomposition of them).
void compute_kmax(dcx_ctx* c, dcx_graph* g, int variant = DCX_VARIANT_BATCHED,
                  bool warm = false, bool clear_rows = false, bool rows_zero = false) {
    if (!g->kmaxes) dev_alloc(c, &g->kmaxes, g->V);
    g->core_valid = false;
    g->order_valid = false;
    if (g->V == 0) {
        g->kmax = -1;
        g->kmax_valid = true;
        return;
    }
    grow_keep(c, &g->pstats, &g->pstats_cap, (size_t)kStatWords, 0);
    // Level-synchronous peel by default; the h-index fixpoint when asked
    // for, and for deletion-only updates, where it starts from the old
    // values (an upper bound) and settles in a few sweeps.
    const bool use_peel = variant == DCX_VARIANT_PEEL || (variant != DCX_VARIANT_HINDEX && !warm);
    g->row0_ready = false;
    g->row0_pending = false;
    g->rows_cleared = 0;
    if (use_peel) {
        // the batched rows start from row 0 when it is a team row (it bounds
        // every other row): compute it beside the kmax pass
        const bool pair = variant == DCX_VARIANT_BATCHED && g->V > c->lds_cap &&
                          c->grid_blocks >= 2 * (int)kXcds;
        // Valid rows hold 0 in every cell (k, v) with k > kmax(v): a new
        // decomposition, which writes every cell of every core member, then
        // only has to zero the cells of vertices whose kmax fell (and rows
        // beyond the old row count), not the whole array.
        const bool keep = pair && clear_rows && rows_zero && g->kmax_valid && g->rows;
        const int32_t old_rows = keep ? g->rows_k : 0;
        if (keep) {
            if (!g->kmax_old) dev_alloc(c, &g->kmax_old, g->V);
            HIPX(c, hipMemcpyAsync(g->kmax_old, g->kmaxes, 4ull * g->V, hipMemcpyDeviceToDevice,
                                   c->stream));
        }
        g->kmax = pair ? run_kmax_with_row0(c, g) : run_grid_pass(c, kmax_spec(g), 0, 1);
        if (pair && clear_rows && g->kmax >= 0) {
            const size_t cells = ((size_t)g->kmax + 1) * g->V;
            g->rows_valid = false;
            if (keep && g->rows_cap >= cells) {
                HIPX(c, launch_clear_stale(g->kmaxes, g->kmax_old, g->V, old_rows, g->kmax + 1,
                                           g->rows, c->stream));
                if ((size_t)old_rows < (size_t)g->kmax + 1)
                    HIPX(c, hipMemsetAsync(g->rows + (size_t)old_rows * g->V, 0,
                                           (cells - (size_t)old_rows * g->V) * sizeof(int32_t),
                                           c->stream));
                HIPX(c, hipEventRecord(c->clear_done, c->stream));
            } else {
                // every row follows: clear them on the second stream after
                // row 0, beside the order kernels
                grow_keep(c, &g->rows, &g->rows_cap, cells, 0);
                HIPX(c, hipMemsetAsync(g->rows, 0, cells * sizeof(int32_t), c->side2));
                HIPX(c, hipEventRecord(c->clear_done, c->side2));
            }
            g->rows_cleared = cells;
        }
    } else {
        const int32_t* prev = nullptr;
        if (warm && g->kmax_valid) {
            if (!g->kmax_old) dev_alloc(c, &g->kmax_old, g->V);
            HIPX(c, hipMemcpyAsync(g->kmax_old, g->kmaxes, 4ull * g->V, hipMemcpyDeviceToDevice,
                                   c->stream));
            prev = g->kmax_old;
        }
        ensure_scratch(c, (size_t)kScratchArrays * g->V * 4);
        clear_kernel_errors(c);
        HIPX(c, hipEventRecord(c->ev[0], c->stream));
        HIPX(c, launch_hindex_row(kmax_spec(g), scratch_at(c->scratch, g->V), prev,
                                  c->ctrs + kCtrFlags, c->d_scalars + 1,
                                  gsync(c, c->hindex_grid), c->stream));
        HIPX(c, hipEventRecord(c->ev[1], c->stream));
        HIPX(c, hipMemcpyAsync(c->h_scalars + 1, c->d_scalars + 1, 4, hipMemcpyDeviceToHost,
                               c->stream));
        check_kernel_errors(c, "kmax h-index");
        g->kmax = c->h_scalars[1];
    }
    HIPX(c, hipEventRecord(c->ev[7], c->stream));
    const size_t need = (size_t)g->kmax + 2;
    grow_keep(c, &g->pstats, &g->pstats_cap, need * kStatWords, (size_t)kStatWords);
    g->h_pass_variant.resize(need, 0);
    g->h_pass_team.resize(need, 0);
    g->h_pass_variant[0] = (int8_t)(use_peel ? DCX_VARIANT_PEEL : DCX_VARIANT_HINDEX);
    g->h_pass_team[0] = (uint16_t)(use_peel ? c->grid_blocks : c->hindex_grid);
    if (g->vk_cap < need) {
        dev_free(g->vk);
        dev_free(g->hist);
        dev_free(g->ekh);
        dev_alloc(c, &g->vk, need);
        dev_alloc(c, &g->hist, need);
        dev_alloc(c, &g->ekh, need + 1);
        g->vk_cap = need;
    }
    if (!g->order) dev_alloc(c, &g->order, g->V);
    HIPX(c, launch_order_by_kmax(g->kmaxes, g->V, g->kmax, g->hist, g->vk, g->order, c->stream));
    // The core is built here unless the rows will be whole-graph passes
    // (few rows after row 0); otherwise when a pass first needs it
    // (ensure_core).  Its out-side list pass also counts the E_k histogram
    // (it reads every list entry's endpoints anyway); without a build, or
    // when the build cannot count it, the histogram runs on its own, on
    // the side stream beside the build.
    const bool few_rows = variant == DCX_VARIANT_BATCHED && g->row0_pending && g->kmax >= 1 &&
                          g->kmax <= kFewRows;
    const bool ek_in_build = !few_rows && core_graph_writes_ek(g->kmax);
    if (!ek_in_build) {
        HIPX(c, hipEventRecord(c->fork, c->stream));
        HIPX(c, hipStreamWaitEvent(c->side, c->fork, 0));
        HIPX(c, launch_core_edges(g->kmaxes, Csr{g->out_off, g->out_deg, g->out_nbr}, g->V,
                                  g->kmax, g->ekh, ensure_ekq(c, g), c->side));
        HIPX(c, hipEventRecord(c->join, c->side));
    }
    if (!few_rows) build_core(c, g, ek_in_build);
    if (!ek_in_build) HIPX(c, hipStreamWaitEvent(c->stream, c->join, 0));
    if (g->row0_pending) {
        HIPX(c, hipEventRecord(c->order_done, c->stream));
        HIPX(c, hipStreamWaitEvent(c->stream, c->row0_done, 0));
    }
    if (g->rows_cleared) HIPX(c, hipStreamWaitEvent(c->stream, c->clear_done, 0));
    g->h_vk.assign(need, 0);
    std::vector<uint64_t> eh(need, 0);
    HIPX(c, hipMemcpyAsync(g->h_vk.data(), g->vk, need * sizeof(uint32_t), hipMemcpyDeviceToHost,
                           c->stream));
    HIPX(c, hipMemcpyAsync(eh.data(), g->ekh, need * sizeof(uint64_t), hipMemcpyDeviceToHost,
                           c->stream));
    HIPX(c, hipEventRecord(c->ev[2], c->stream));
    sync(c);
    g->h_ek.assign((size_t)g->kmax + 1, 0);
    uint64_t acc = 0;
    for (int32_t k = g->kmax; k >= 0; --k) {
        acc += eh[k];
        g->h_ek[k] = acc;
    }
    c->timing.kmax_ms = elapsed(c, 0, 1);
    c->timing.order_ms = elapsed(c, 7, 2);
    if (g->row0_pending) {
        g->row0_pending = false;
        // row 0's pass has ended (the sync above waited for it)
        HIPX(c, hipMemcpy(c->h_scalars + 34, c->ctrs2 + kCtrSync + kSyncError, 4,
                          hipMemcpyDeviceToHost));
        if (c->h_scalars[34] != 0) {
            set_err(c, "row 0 pass: %s (code %d)",
                    c->h_scalars[34] == (int32_t)kErrTimeout ? "a barrier timed out"
                                                             : "kernel counters inconsistent",
                    c->h_scalars[34]);
            throw Fail{c->h_scalars[34] == (int32_t)kErrTimeout ? DCX_ERR_TIMEOUT
                                                                 : DCX_ERR_INTERNAL};
        }
        g->row0_ready = true;
        // next run's share: row 0 should end while the order kernels run
        float r_end = 0, o_end = 0;
        HIPX(c, hipEventElapsedTime(&r_end, c->ev[0], c->row0_done));
        HIPX(c, hipEventElapsedTime(&o_end, c->ev[0], c->order_done));
        if (r_end > o_end && g->row0_16ths < 8)
            ++g->row0_16ths;  // the library stream waited for row 0
        else if (r_end < c->timing.kmax_ms && g->row0_16ths > 2)
            --g->row0_16ths;  // row 0 ended before the kmax pass
    }
    g->kmax_valid = true;
    g->order_valid = true;
}

// The kmax order, V_k and E_k of the resident kmaxes when local maintenance
// changed the kmaxes since they were computed (the core-ordered graph is
// rebuilt from them when a pass needs it).
void ensure_order(dcx_ctx* c, dcx_graph* g) {
    if (g->order_valid || !g->kmax_valid) return;
    g->core_valid = false;
    if (g->V == 0 || g->kmax < 0) {
        g->order_valid = true;
        return;
    }
    const size_t need = (size_t)g->kmax + 2;
    grow_keep(c, &g->pstats, &g->pstats_cap, need * kStatWords, 0);
    g->h_pass_variant.assign(need, 0);
    g->h_pass_team.assign(need, 0);
    if (g->vk_cap < need) {
        dev_free(g->vk);
        dev_free(g->hist);
        dev_free(g->ekh);
        dev_alloc(c, &g->vk, need);
        dev_alloc(c, &g->hist, need);
        dev_alloc(c, &g->ekh, need + 1);
        g->vk_cap = need;
    }
    if (!g->order) dev_alloc(c, &g->order, g->V);
    HIPX(c, launch_order_by_kmax(g->kmaxes, g->V, g->kmax, g->hist, g->vk, g->order, c->stream));
    HIPX(c, launch_core_edges(g->kmaxes, Csr{g->out_off, g->out_deg, g->out_nbr}, g->V, g->kmax,
                              g->ekh, ensure_ekq(c, g), c->stream));
    g->h_vk.assign(need, 0);
    std::vector<uint64_t> eh(need, 0);
    HIPX(c, hipMemcpyAsync(g->h_vk.data(), g->vk, need * sizeof(uint32_t), hipMemcpyDeviceToHost,
                           c->stream));
    HIPX(c, hipMemcpyAsync(eh.data(), g->ekh, need * sizeof(uint64_t), hipMemcpyDeviceToHost,
                           c->stream));
    sync(c);
    g->h_ek.assign((size_t)g->kmax + 1, 0);
    uint64_t acc = 0;
    for (int32_t k = g->kmax; k >= 0; --k) {
        acc += eh[k];
        g->h_ek[k] = acc;
    }
    g->order_valid = true;
}

// Static schedule of the passes k = 0..kr onto teams of workgroups: a pass
// whose work (2 E_k + V_k) exceeds the per-workgroup share gets its own
// team sized to it (up to team_max and to one XCD's workgroups, using at
// most 3/4 of the grid), placed on the XCD with the most free workgroups;
// the remaining passes go to single-workgroup teams on the free
// workgroups, longest first, each to the least loaded team.  Team scratch
// regions start on 128-byte lines.
struct Sched {
    std::vector<TeamDesc> teams;
    std::vector<uint32_t> block_map;  // [G] team of each block (~0u idle), [G] rank
    std::vector<int32_t> ks;
    uint64_t scratch_words = 0;
    uint32_t sync_stride = 0;
};
constexpr uint32_t kTeamMax = 32;
constexpr uint32_t kShrinkAt = 4096;
constexpr int kRowsBlocksPerCu = 2;

// want/wts (optional): team width and weight of each pass chosen by the
// cost model; without them the width follows the pass's share of the work.
// take (optional): only the passes k with take[k] set are scheduled.
Sched make_schedule(const dcx_graph* g, int32_t k_first, int32_t kr, uint32_t G,
                    uint32_t team_max, const std::vector<uint32_t>* want = nullptr,
                    const std::vector<double>* wts = nullptr,
                    const std::vector<char>* take = nullptr) {
    const uint32_t np = (uint32_t)kr + 1, k0 = (uint32_t)std::max(k_first, 0);
    std::vector<double> w(np, 0.0);
    double W = 0;
    for (uint32_t k = k0; k < np; ++k) {
        if (take && !(*take)[k]) continue;
        w[k] = wts ? (*wts)[k] : 2.0 * (double)g->h_ek[k] + (double)g->h_vk[k] + 64.0;
        W += w[k];
    }
    // widths follow each pass's share of the work within the 3/4 of the
    // grid that wide teams may take; when no pass is left for single
    // workgroups, the teams take the whole grid instead
    const uint32_t tcap = std::min(team_max, G / kXcds);
    uint32_t big_budget = std::max<uint32_t>(G * 3 / 4, 1);
    double share = W / big_budget;
    bool any_small = false;
    for (uint32_t k = k0; k < np && !any_small; ++k)
        if (!take || (*take)[k])
            any_small = want ? std::min((*want)[k], tcap) < 2
                             : std::min<double>(tcap, w[k] / share) < 2.0;
    if (!any_small) {
        big_budget = std::max<uint32_t>(G, 1);
        share = W / big_budget;
    }
    std::vector<uint32_t> idx;
    for (uint32_t k = k0; k < np; ++k)
        if (!take || (*take)[k]) idx.push_back(k);
    std::stable_sort(idx.begin(), idx.end(), [&](uint32_t a, uint32_t b) { return w[a] > w[b]; });
    // free workgroups of each XCD, in increasing block order
    std::vector<std::vector<uint32_t>> xfree(kXcds);
    for (uint32_t b = G; b-- > 0;) xfree[b % kXcds].push_back(b);  // back() = lowest
    Sched s;
    s.sync_stride = (sync_words(team_max) + kSlotStride - 1) / kSlotStride * kSlotStride;
    s.block_map.assign(2ull * G, ~0u);
    uint32_t used = 0;
    std::vector<uint32_t> small;
    std::vector<std::vector<uint32_t>> team_blocks;
    for (uint32_t k : idx) {
        uint32_t t = want ? std::min((*want)[k], tcap)
                          : (uint32_t)std::min<double>(tcap, w[k] / share);
        uint32_t x = 0;
        for (uint32_t y = 1; y < kXcds; ++y)
            if (xfree[y].size() > xfree[x].size()) x = y;
        t = std::min<uint32_t>(t, (uint32_t)xfree[x].size());
        if (t >= 2 && used + t <= big_budget) {
            TeamDesc td{};
            td.xcd = x;
            td.size = t;
            td.pass_begin = (uint32_t)s.ks.size();
            s.ks.push_back((int32_t)k);
            td.pass_end = (uint32_t)s.ks.size();
            td.scratch_cap = std::max<uint32_t>(g->h_vk[k], 1);
            std::vector<uint32_t> blocks;
            for (uint32_t r = 0; r < t; ++r) {
                blocks.push_back(xfree[x].back());
                xfree[x].pop_back();
            }
            team_blocks.push_back(std::move(blocks));
            s.teams.push_back(td);
            used += t;
        } else {
            small.push_back(k);
        }
    }
    std::vector<uint32_t> rest;
    for (uint32_t x = 0; x < kXcds; ++x)
        for (uint32_t b : xfree[x]) rest.push_back(b);
    std::sort(rest.begin(), rest.end());
    const uint32_t nsingle = std::min<uint32_t>((uint32_t)rest.size(), (uint32_t)small.size());
    std::vector<std::vector<int32_t>> lists(nsingle);
    std::vector<double> load(nsingle, 0.0);
    for (uint32_t k : small) {  // already longest first
        const uint32_t t = (uint32_t)(std::min_element(load.begin(), load.end()) - load.begin());
        lists[t].push_back((int32_t)k);
        load[t] += w[k];
    }
    for (uint32_t t = 0; t < nsingle; ++t) {
        TeamDesc td{};
        td.xcd = rest[t] % kXcds;
        td.size = 1;
        td.pass_begin = (uint32_t)s.ks.size();
        uint32_t capv = 1;
        for (int32_t k : lists[t]) {
            s.ks.push_back(k);
            capv = std::max<uint32_t>(capv, g->h_vk[k]);
        }
        td.pass_end = (uint32_t)s.ks.size();
        td.scratch_cap = capv;
        s.teams.push_back(td);
        team_blocks.push_back({rest[t]});
    }
    uint64_t off = 0;
    for (uint32_t t = 0; t < s.teams.size(); ++t) {
        TeamDesc& td = s.teams[t];
        td.scratch_cap = (td.scratch_cap + kSlotStride - 1) / kSlotStride * kSlotStride;
        td.scratch_off = off;
        off += (uint64_t)kScratchArrays * td.scratch_cap;
        td.sync_off = t * s.sync_stride;
        td.slot_off = t * kSlotWords;
        for (uint32_t r = 0; r < team_blocks[t].size(); ++r) {
            s.block_map[team_blocks[t][r]] = t;
            s.block_map[G + team_blocks[t][r]] = r;
        }
    }
    s.scratch_words = off;
    return s;
}

// Kinds of team row: peeled; derived from the row above as the first row
// of a chain (its time includes the supports' set-up); derived later in a
// chain; derived from row 0 as the first row of a chain (an anchor that is
// not peeled).
enum : int { kPassPeel = 0, kPassFirst = 1, kPassDerived = 2, kPassFar = 3, kPassKinds = 4 };

// Kind of row k's last run in the batched launch (-1: none, or another
// variant), from the pass log.
int last_pass_kind(const dcx_graph* g, int32_t k) {
    if ((size_t)k + 1 >= g->h_pass_variant.size() || g->h_pass_team[k + 1] == 0) return -1;
    if (g->h_pass_variant[k + 1] == DCX_VARIANT_BATCHED) return k >= g->rows_first_grid ? 0 : -1;
    if (g->h_pass_variant[k + 1] != DCX_VARIANT_DERIVED) return -1;
    if (g->h_pstats.size() < ((size_t)k + 2) * kStatWords) return -1;
    const uint32_t lv = g->h_pstats[(size_t)(k + 1) * kStatWords + kStatLevels];
    return lv == 2 ? kPassFar : lv ? kPassFirst : kPassDerived;
}

// Folds the team rows of the last batched launch (rows 0..kr) into the
// per-row work statistics and measured times (wall_khz: device clock).
void record_row_stats(dcx_graph* g, int32_t kr, double wall_khz) {
    const size_t n = (size_t)g->kmax + 1;
    for (int kind = 0; kind < kPassKinds; ++kind) {
        g->h_meas_ms[kind].resize(n, 0.0f);
        g->h_meas_team[kind].resize(n, 0);
    }
    g->h_peel_phases.resize(n, 0);
    g->h_der_rounds.resize(n, 0);
    g->h_der_reads.resize(n, 0);
    g->h_far_rounds.resize(n, 0);
    g->h_far_reads.resize(n, 0);
    for (int32_t k = 0; k <= kr && (size_t)k < n; ++k) {
        const int kind = last_pass_kind(g, k);
        if (kind < 0) continue;
        const uint32_t* w = g->h_pstats.data() + (size_t)(k + 1) * kStatWords;
        const uint64_t ticks = (uint64_t)w[kStatTicksLo] | ((uint64_t)w[kStatTicksHi] << 32);
        g->h_meas_ms[kind][k] = (float)std::max(ticks / wall_khz, 1e-3);
        g->h_meas_team[kind][k] = g->h_pass_team[k + 1];
        if (kind == kPassPeel) {
            g->h_peel_phases[k] = w[kStatPhases] + 1;
        } else {
            // a first row's walks include the supports' set-up: one pass
            // over both member prefixes of every member, 2 E_k entries
            uint64_t reads = w[kStatAlive];
            if (kind != kPassDerived) reads -= std::min<uint64_t>(reads, 2 * g->h_ek[k]);
            if (kind == kPassFar) {
                g->h_far_rounds[k] = w[kStatPhases] + 1;
                g->h_far_reads[k] = reads + 1;
            } else {
                g->h_der_rounds[k] = w[kStatPhases] + 1;
                g->h_der_reads[k] = reads + 1;
            }
        }
    }
}

// Statistic of row k (value + 1 as stored), or of the nearest row that has
// one (up to 64 rows away); 0 if none.
template <typename T>
T row_stat(const std::vector<T>& v, int32_t k) {
    const int32_t n = (int32_t)v.size();
    for (int32_t d = 0; d <= 64; ++d) {
        if (k - d >= 0 && k - d < n && v[k - d]) return v[k - d];
        if (k + d < n && v[k + d]) return v[k + d];
    }
    return 0;
}

// Scheduler features of row k run as `kind` (kPassPeel: peeled; kPassFirst:
// derived from the row above as the first row of a chain, its time
// including the supports' set-up; kPassDerived: derived, later in a chain;
// kPassFar: derived from row 0 as the first row of a chain) on a team of
// `team` workgroups (kMlpIn floats, about [0, 1]): sizes of the (k,0)-core
// and the vertices that left it since row k-1, the kind, the width, and the
// work row k took when it last ran that way (phases of a peel; rounds and
// list entries read of a derivation, plus the set-up's 2 E_k entries for a
// first row), from this graph's batched launches (or the nearest row's;
// with_stats = false: those two as unknown).
void pass_features(const dcx_graph* g, int32_t k, uint32_t team, int kind, bool with_stats,
                   float* f) {
    const double Vk = g->h_vk[k], Ek = (double)g->h_ek[k];
    const double Vp = k > 0 ? g->h_vk[k - 1] : Vk;
    double steps = 0.0, reads = 0.0;
    if (with_stats) {
        if (kind == kPassPeel) {
            const uint32_t ph = row_stat(g->h_peel_phases, k);
            steps = ph ? ph - 1.0 : 0.0;
            reads = 2.0 * Ek;
        } else {
            const bool far = kind == kPassFar;
            const uint32_t r = row_stat(far ? g->h_far_rounds : g->h_der_rounds, k);
            const uint64_t e = row_stat(far ? g->h_far_reads : g->h_der_reads, k);
            steps = r ? r - 1.0 : 0.0;
            reads = e ? (double)(e - 1) : 0.0;
            if (kind != kPassDerived) reads += 2.0 * Ek;
        }
    }
    f[0] = (float)(std::log2(1.0 + Vk) / 24.0);
    f[1] = (float)(std::log2(1.0 + Ek) / 32.0);
    f[2] = (float)(std::log2(1.0 + Vp - Vk) / 24.0);
    f[3] = (float)kind / 3.0f;
    f[4] = (float)(std::log2((double)std::max(team, 1u)) / 8.0);
    f[5] = (float)(std::log2(1.0 + steps) / 12.0);
    f[6] = (float)(std::log2(1.0 + reads) / 36.0);
    f[7] = (float)(std::log2(1.0 + Ek / std::max(Vk, 1.0)) / 12.0);
}

// True once a batched launch has logged team rows of g (the cost model's
// plan needs their work statistics).
bool have_row_stats(const dcx_graph* g) {
    for (uint32_t x : g->h_peel_phases)
        if (x) return true;
    for (uint32_t x : g->h_der_rounds)
        if (x) return true;
    return false;
}

void mlp_stage(dcx_ctx* c, dcx_mlp* m, size_t floats) {
    if (m->buf_floats >= floats) return;
    dev_free(m->buf);
    dev_alloc(c, &m->buf, floats);
    m->buf_floats = floats;
}

// Predictions (log2 ms) for n samples, computed on the device.
std::vector<float> mlp_predict(dcx_ctx* c, dcx_mlp* m, const std::vector<float>& X, uint32_t n) {
    mlp_stage(c, m, (size_t)n * (kMlpIn + 1));
    std::vector<float> y(n);
    if (!n) return y;
    HIPX(c, hipMemcpyAsync(m->buf, X.data(), 4ull * n * kMlpIn, hipMemcpyHostToDevice, c->stream));
    HIPX(c, launch_mlp_forward(m->p, m->buf, n, m->buf + (size_t)n * kMlpIn, c->stream));
    HIPX(c, hipMemcpyAsync(y.data(), m->buf + (size_t)n * kMlpIn, 4ull * n, hipMemcpyDeviceToHost,
                           c->stream));
    sync(c);
    return y;
}

// Plan of the team rows 1..kr (row 0 is ready before the launch) from the
// cost model.  A unit is an anchor a with the chain a+1..b derived from it
// on the same team: row 0 (ready before the launch), a row the team peels,
// or a row the team derives from row 0 (far[a]).  Its predicted time at
// width t is, for a peeled anchor, peel(a, t) + first(a+1, t) + the sum
// over k = a+2..b of derived(k, t); for an anchor derived from row 0,
// far(a, t) + the sum over k = a+1..b of derived(k, t); for row 0 the
// derived rows alone (each prediction scaled by the ratio of measured to
// predicted time of the row's last run of that kind).  For a makespan T
// the rows are cut into units from
// the top, each extended while some width and anchor kind keep it within
// T, and each unit takes the narrowest width that does, with the cheaper
// kind there; T is the smallest value (binary search) for which the wide
// units fit 3/4 of the grid (all of it when no unit is left for a single
// workgroup) and the single-workgroup units, packed longest first, fit the
// rest.  With `fixed` set, the anchors, their kinds and chains are given
// and only the widths are chosen.  Outputs: peel[k] (row k is an anchor),
// far[k] (the anchor k is derived from row 0), chain_last[a] (last row of
// a's chain, -1: none), and per anchor the width want[a] and the weight
// wts[a] (predicted workgroup time, which places the units longest first).
// With `best` (a plan measured on this graph), the new plan is taken only
// if its predicted makespan beats best's, predicted the same way, by 3 %;
// else best's plan is returned (*used_best).  Returns false if no plan fits.
bool mlp_rows_plan(dcx_ctx* c, const dcx_graph* g, int32_t kr, uint32_t G, uint32_t team_max,
                   bool fixed, std::vector<char>& peel, std::vector<char>& far,
                   std::vector<int32_t>& chain_last, std::vector<uint32_t>& want,
                   std::vector<double>& wts, const dcx_graph::RowsPlan* best = nullptr,
                   bool* used_best = nullptr) {
    if (used_best) *used_best = false;
    const uint32_t np = (uint32_t)kr + 1;
    const uint32_t cap = std::min(team_max, G / kXcds);
    std::vector<uint32_t> cands{1};
    for (uint32_t t = 2; t <= cap; t *= 2) cands.push_back(t);
    const uint32_t nc = (uint32_t)cands.size();
    if (!have_row_stats(g)) return false;  // nothing logged yet: the fixed rule
    // predictions: [kind][k][width]
    std::vector<float> X((size_t)kPassKinds * np * nc * kMlpIn);
    for (int kind = 0; kind < kPassKinds; ++kind)
        for (uint32_t k = 0; k < np; ++k)
            for (uint32_t j = 0; j < nc; ++j)
                pass_features(g, (int32_t)k, cands[j], kind, true,
                              &X[(((size_t)kind * np + k) * nc + j) * kMlpIn]);
    const std::vector<float> y = mlp_predict(c, c->sched_mlp, X, kPassKinds * np * nc);
    // calibration: where row k last ran as `kind` on a candidate width, the
    // predictions for row k and that kind are scaled so that this width's
    // prediction equals the measured time (the model keeps the shape over
    // widths, the measurement fixes the level)
    std::vector<double> cal((size_t)kPassKinds * np, 1.0);
    for (int kind = 0; kind < kPassKinds; ++kind) {
        const auto& ms = g->h_meas_ms[kind];
        const auto& tm = g->h_meas_team[kind];
        for (uint32_t k = 0; k < np && k < ms.size(); ++k) {
            if (ms[k] <= 0.0f) continue;
            for (uint32_t j = 0; j < nc; ++j)
                if (cands[j] == tm[k]) {
                    const double p = std::exp2((double)y[((size_t)kind * np + k) * nc + j]);
                    cal[(size_t)kind * np + k] = (double)ms[k] / std::max(p, 1e-6);
                }
        }
    }
    auto d = [&](int kind, uint32_t k, uint32_t j) {
        return cal[(size_t)kind * np + k] * std::exp2((double)y[((size_t)kind * np + k) * nc + j]);
    };
    // kind of row k in the unit anchored at a (alt 1: a derived from row 0)
    auto kind_of = [&](uint32_t a, int alt, uint32_t k) {
        if (k == a) return alt ? kPassFar : kPassPeel;
        return k == a + 1 && a > 0 && !alt ? kPassFirst : kPassDerived;
    };
    auto unit_cost = [&](uint32_t a, int alt, uint32_t b, uint32_t j) {
        double t = 0.0;
        for (uint32_t k = a > 0 ? a : 1; k <= b; ++k) t += d(kind_of(a, alt, k), k, j);
        return t;
    };
    struct Unit {
        uint32_t a, b, j;
        int alt;
        double cost;
    };
    auto plan = [&](double T, std::vector<Unit>& units) -> bool {
        units.clear();
        if (fixed) {
            for (uint32_t a = 0; a < np; ++a) {
                if (!(a == 0 ? chain_last[0] >= 0 : peel[a] != 0)) continue;
                const uint32_t b = chain_last[a] >= 0 ? (uint32_t)chain_last[a] : a;
                const int alt = a > 0 && far[a] ? 1 : 0;
                uint32_t j = 0;
                while (j + 1 < nc && unit_cost(a, alt, b, j) > T) ++j;
                if (unit_cost(a, alt, b, j) > T) return false;
                units.push_back({a, b, j, alt, unit_cost(a, alt, b, j)});
            }
        } else {
            uint32_t a = 0;
            // cost of a..k at each width, per anchor kind (row 0: both 0)
            std::vector<double> run[2] = {std::vector<double>(nc, 0.0),
                                          std::vector<double>(nc, 0.0)};
            for (uint32_t k = 1; k <= np; ++k) {
                bool extend = false;
                if (k < np) {
                    for (int alt = 0; alt < 2; ++alt)
                        for (uint32_t j = 0; j < nc; ++j)
                            extend = extend || run[alt][j] + d(kind_of(a, alt, k), k, j) <= T;
                    if (extend)
                        for (int alt = 0; alt < 2; ++alt)
                            for (uint32_t j = 0; j < nc; ++j)
                                run[alt][j] += d(kind_of(a, alt, k), k, j);
                }
                if (extend) continue;
                // close the unit a..k-1 (row 0 alone is no unit)
                if (a > 0 || k > 1) {
                    uint32_t j = 0;
                    while (j + 1 < nc && std::min(run[0][j], run[1][j]) > T) ++j;
                    const int alt = a > 0 && run[1][j] < run[0][j] ? 1 : 0;
                    if (run[alt][j] > T) return false;
                    units.push_back({a, k - 1, j, alt, run[alt][j]});
                }
                if (k == np) break;
                a = k;  // k becomes an anchor
                for (uint32_t j = 0; j < nc; ++j) {
                    run[0][j] = d(kPassPeel, k, j);
                    run[1][j] = d(kPassFar, k, j);
                }
            }
        }
        uint32_t big = 0;
        std::vector<double> singles;
        for (const Unit& u : units) {
            if (cands[u.j] >= 2)
                big += cands[u.j];
            else
                singles.push_back(u.cost);
        }
        if (big > (singles.empty() ? G : G * 3 / 4)) return false;
        if (singles.empty()) return true;
        const uint32_t slots = G - big;
        if (!slots) return false;
        std::sort(singles.begin(), singles.end(), std::greater<double>());
        std::vector<double> load(std::min<size_t>(slots, singles.size()), 0.0);
        for (double x : singles) {
            auto it = std::min_element(load.begin(), load.end());
            *it += x;
            if (*it > T) return false;
        }
        return true;
    };
    // one workgroup doing every row, each of every kind, is within this bound
    double hi = 0.0;
    for (uint32_t k = 1; k < np; ++k)
        for (int kind = 0; kind < kPassKinds; ++kind) hi += d(kind, k, 0);
    hi = 2.0 * hi + 1.0;
    std::vector<Unit> units;
    if (!plan(hi, units)) return false;
    double lo = 0.0;
    // to 0.1 % of the makespan (each step evaluates a whole plan, and this
    // runs on the host between the order kernels and the rows launch)
    for (int it = 0; it < 60 && hi - lo > 1e-3 * hi; ++it) {
        const double mid = 0.5 * (lo + hi);
        if (plan(mid, units))
            hi = mid;
        else
            lo = mid;
    }
    plan(hi, units);
    if (!fixed && best && best->valid && best->want.size() == np &&
        best->peel.size() == peel.size()) {
        // best's predicted makespan: its units at their widths
        double tb = 0.0;
        bool ok = true;
        for (uint32_t a = 0; a < np && ok; ++a) {
            if (!(a == 0 ? best->chain_last[0] >= 0 : best->peel[a] != 0)) continue;
            const uint32_t b = best->chain_last[a] >= 0 ? (uint32_t)best->chain_last[a] : a;
            const int alt = a > 0 && best->far[a] ? 1 : 0;
            uint32_t j = 0;
            while (j < nc && cands[j] != best->want[a]) ++j;
            if (j == nc) {
                ok = false;
                break;
            }
            tb = std::max(tb, unit_cost(a, alt, b, j));
        }
        if (ok && tb <= 1.03 * hi) {
            peel = best->peel;
            far = best->far;
            chain_last = best->chain_last;
            want = best->want;
            wts = best->wts;
            if (used_best) *used_best = true;
            return true;
        }
    }
    if (!fixed) {
        peel.assign(peel.size(), 1);
        far.assign(far.size(), 0);
        std::fill(chain_last.begin(), chain_last.end(), -1);
    }
    want.assign(np, 1);
    wts.assign(np, 0.0);
    for (const Unit& u : units) {
        if (!fixed) {
            for (uint32_t k = u.a + 1; k <= u.b; ++k) peel[k] = 0;
            chain_last[u.a] = u.b > u.a ? (int32_t)u.b : -1;
            far[u.a] = (char)u.alt;
        }
        want[u.a] = cands[u.j];
        wts[u.a] = u.cost * cands[u.j];
    }
    return true;
}

// Appends the device log of the failed row passes (k <= kr) to the error
// text: which pass, on how many workgroups, and where its counters stood.
void describe_failed_passes(dcx_ctx* c, const dcx_graph* g, int32_t kr) {
    std::vector<uint32_t> st((size_t)(kr + 2) * kStatWords);
    if (hipMemcpy(st.data(), g->pstats, st.size() * 4, hipMemcpyDeviceToHost) != hipSuccess)
        return;
    std::string d = c->err;
    int shown = 0;
    for (int32_t k = 0; k <= kr && shown < 4; ++k) {
        const uint32_t* w = st.data() + (size_t)(k + 1) * kStatWords;
        if (!w[kStatFailed]) continue;
        char buf[256];
        std::snprintf(buf, sizeof buf,
                      "; row %d failed: team %u, members %u, removed %u, alive %u, level %u, "
                      "levels %u, phases %u",
                      k, (unsigned)g->h_pass_team[k + 1], g->h_vk[k], w[kStatRemoved],
                      w[kStatAlive], w[kStatLevel], w[kStatLevels], w[kStatPhases]);
        d += buf;
        ++shown;
    }
    c->err = d;
}

// Appends the record the first failing pass member wrote.
void describe_fail_record(dcx_ctx* c) {
    uint32_t r[kFailWords] = {};
    if (!c->fail_rec ||
        hipMemcpy(r, c->fail_rec, sizeof r, hipMemcpyDeviceToHost) != hipSuccess || !r[0])
        return;
    char buf[400];
    std::snprintf(buf, sizeof buf,
                  "; first failure: k %u member %u/%u (rank %u) phase %u levels %u removed %u "
                  "of %u alive %u min %u level %d epoch %u leader-phase %u release %u nblocks %u",
                  r[10], r[1], r[2], r[11], r[3], r[4], r[5], r[6], r[7], r[8], (int32_t)r[9],
                  r[12], r[13], r[14], r[15]);
    c->err += buf;
}

// Computes rows k = 0..kr (kr < 0: all rows 0..kmax) with the given
// variant; rows above kr keep their values (they must be valid already).
void compute_rows(dcx_ctx* c, dcx_graph* g, int variant, int32_t kr = -1) {
    ensure_order(c, g);
    const uint32_t V = g->V;
    const int32_t K = g->kmax;
    const bool had_rows = g->rows_valid;
    const int32_t old_rows = had_rows ? g->rows_k : 0;
    g->rows_valid = false;
    g->rows_first_grid = 0;  // set below by the batched paths
    if (K < 0) {
        g->lmax = 0;
        g->rows_k = 0;
        g->rows_valid = true;
        return;
    }
    if (kr < 0 || kr > K || !had_rows) kr = K;
    const size_t nrows = (size_t)K + 1;
    const size_t keep_rows = std::min<size_t>((size_t)old_rows, nrows);
    grow_keep(c, &g->rows, &g->rows_cap, nrows * V, keep_rows * V);
    grow_keep(c, &g->lmax_rows, &g->lmax_rows_cap, nrows, keep_rows);
    const size_t nre = (size_t)kr + 1;
    // Few rows on a big graph (a Chung-Lu 2^24 graph has kmax 2): with row
    // 0 already computed beside the kmax pass, rows 1..kr run as
    // whole-graph passes two per launch, bounded by row 0, and the
    // core-ordered graph is never built (its sort would cost more than the
    // rows themselves).
    const bool few = variant == DCX_VARIANT_BATCHED && g->row0_ready && kr >= 1 &&
                     kr <= kFewRows;
    if (!few) ensure_core(c, g);
    uint32_t nteam_err = 0, team_err_stride = 0;
    const uint32_t* team_err = nullptr;
    clear_kernel_errors(c);
    if (!c->fail_rec) dev_alloc(c, &c->fail_rec, kFailWords);
    HIPX(c, hipMemsetAsync(c->fail_rec, 0, kFailWords * 4, c->stream));
    HIPX(c, hipEventRecord(c->ev[3], c->stream));
    if (g->rows_cleared != nre * V)
        HIPX(c, hipMemsetAsync(g->rows, 0, nre * V * sizeof(int32_t), c->stream));
    g->rows_cleared = 0;
    int launches = 0, slots_used = 1;
    int32_t split = kr + 1;  // batched: first LDS-resident row
    bool teams_ran = false, lds_ran = false, first_ran = false;
    std::vector<std::pair<uint32_t, int32_t>> last_teams;  // (size, first k) for diagnosis
    if (few) {
        g->rows_first_grid = kr + 1;  // no row is a team pass
        ensure_scratch(c, 2ull * kScratchArrays * V * 4);
        HIPX(c, hipMemsetAsync(c->ctrs2 + kCtrSync + kSyncError, 0, 4, c->stream));
        HIPX(c, hipEventRecord(c->ev[5], c->stream));
        HIPX(c, hipMemcpyAsync(g->rows, g->row0, 4ull * V, hipMemcpyDeviceToDevice, c->stream));
        HIPX(c, hipMemcpyAsync(g->pstats + kStatWords, g->row0_stats, 4ull * kStatWords,
                               hipMemcpyDeviceToDevice, c->stream));
        HIPX(c, hipMemcpyAsync(g->lmax_rows, g->row0_stats + kStatWords, 4,
                               hipMemcpyDeviceToDevice, c->stream));
        g->h_pass_variant[1] = DCX_VARIANT_BATCHED;
        g->h_pass_team[1] = (uint16_t)c->grid_blocks;
        HIPX(c, hipEventRecord(c->sev[3], c->stream));
        const PassScratch s0 = scratch_at(c->scratch, V);
        const PassScratch s1 =
            scratch_at(static_cast<uint32_t*>(c->scratch) + (size_t)kScratchArrays * V, V);
        auto spec = [&](int32_t k) {
            PassSpec p = traced(c, full_row_spec(g, k, g->rows + (size_t)k * V));
            p.trace = nullptr;
            p.stats = g->pstats + (size_t)kStatWords * (k + 1);
            p.ub = g->row0;
            return p;
        };
        for (int32_t k = 1; k <= kr; k += 2) {
            if (k + 1 <= kr) {
                const int G1 = c->grid_blocks / 2, G0 = c->grid_blocks - G1;
                GridSync g0 = gsync(c, G0), g1 = gsync(c, G1);
                g1.w = c->ctrs2 + kCtrSync;
                HIPX(c, launch_pass_pair(spec(k), s0, c->ctrs, g->lmax_rows + k, g0, spec(k + 1),
                                         s1, c->ctrs2, g->lmax_rows + k + 1, g1, c->stream));
                g->h_pass_team[k + 1] = (uint16_t)G0;
                g->h_pass_team[k + 2] = (uint16_t)G1;
                g->h_pass_variant[k + 2] = DCX_VARIANT_BATCHED;
            } else {
                HIPX(c, launch_pass_grid(spec(k), s0, c->ctrs, g->lmax_rows + k,
                                         gsync(c, c->grid_blocks), c->stream));
                g->h_pass_team[k + 1] = (uint16_t)c->grid_blocks;
            }
            g->h_pass_variant[k + 1] = DCX_VARIANT_BATCHED;
            ++launches;
        }
        team_err = c->ctrs2 + kCtrSync + kSyncError;
        nteam_err = 1;
        team_err_stride = 1;
        first_ran = true;
    } else if (variant == DCX_VARIANT_HINDEX) {
        // rows in k order, each started from the row above (l_k <= l_{k-1})
        ensure_scratch(c, 2ull * kScratchArrays * V * 4);
        HIPX(c, hipEventRecord(c->ev[5], c->stream));
        for (int32_t k = 0; k <= kr; ++k) {
            PassScratch s = scratch_at(static_cast<uint32_t*>(c->scratch) +
                                           (size_t)(k & 1) * kScratchArrays * V,
                                       V);
            PassScratch prev = scratch_at(static_cast<uint32_t*>(c->scratch) +
                                              (size_t)((k + 1) & 1) * kScratchArrays * V,
                                          V);
            PassSpec p = core_row_spec(g, k, g->rows + (size_t)k * V);
            HIPX(c, launch_hindex_row(p, s, k ? prev.lv : nullptr, c->ctrs + kCtrFlags,
                                      g->lmax_rows + k, gsync(c, c->hindex_grid), c->stream));
            ++launches;
            g->h_pass_variant[k + 1] = DCX_VARIANT_HINDEX;
            g->h_pass_team[k + 1] = (uint16_t)c->hindex_grid;
        }
    } else if (variant == DCX_VARIANT_PEEL) {
        ensure_scratch(c, (size_t)kScratchArrays * V * 4);
        HIPX(c, hipEventRecord(c->ev[5], c->stream));
        for (int32_t k = 0; k <= kr; ++k) {
            PassSpec p = traced(c, core_row_spec(g, k, g->rows + (size_t)k * V));
            HIPX(c, launch_pass_grid(p, scratch_at(c->scratch, V), c->ctrs, g->lmax_rows + k,
                                     gsync(c, c->grid_blocks), c->stream));
            ++launches;
            g->h_pass_variant[k + 1] = DCX_VARIANT_PEEL;
            g->h_pass_team[k + 1] = (uint16_t)c->grid_blocks;
        }
    } else {
        // batched: rows whose core fits the LDS run LDS-resident, one
        // workgroup each, taken longest first from a queue; the larger rows
        // run on teams of workgroups in one persistent launch
        c->timing.rows_scheduler = 0;
        int32_t ksplit = 0;  // rows [0, ksplit) on teams, [ksplit, kr] in LDS
        while (ksplit <= kr && g->h_vk[ksplit] > c->lds_cap) ++ksplit;
        // every workgroup of the launch must be resident at once (a team's
        // members wait for each other): never more than the CUs hold
        const uint32_t G = (uint32_t)std::min<int64_t>(
            c->cfg.rows_slots > 0 ? c->cfg.rows_slots : c->num_cus * kRowsBlocksPerCu,
            (int64_t)c->num_cus * c->rows_per_cu);
        const uint32_t tmax = c->cfg.team_max > 0 ? (uint32_t)c->cfg.team_max : kTeamMax;
        const uint32_t shrink = c->cfg.shrink_at > 0 ? (uint32_t)c->cfg.shrink_at : kShrinkAt;
        // Row 0 first, on the whole grid, when other rows run on teams: a
        // (k,l)-core lies inside the (0,l)-core, so row 0 bounds every other
        // row, and the team and LDS passes seed with that bound
        // (PassSpec::ub); most vertices then leave in the seeding phase of
        // their level instead of at the end of a cascade.
        // (When the kmax pass computed row 0 beside it, only its copy and
        // the bound by rank are left to do here.)
        const bool have_row0 = g->row0_ready;
        const int32_t k_first = ksplit >= 2 || have_row0 ? 1 : 0;
        g->rows_first_grid = k_first;
        // Anchors and derived team rows: a row is either an anchor --
        // peeled, or derived from row 0 (far) -- or derived from the row
        // above; the rows a+1..b derived from anchor a run on the team
        // that computed a, right after it (row 0, ready before the launch,
        // anchors a chain of its own).  The plan comes from the cost model
        // when one is attached and cfg.chain_rows is 0, else every
        // spacing-th row from row 0 is an anchor, of the kind cfg.anchors
        // names (the fixed rule); the model then still picks the widths.
        const int32_t spacing =
            c->cfg.chain_rows == 0 ? kChainRows : std::max<int32_t>(c->cfg.chain_rows, 1);
        std::vector<char> peel(kr + 1, 1);
        std::vector<char> far(kr + 1, c->cfg.anchors == 1 ? 1 : 0);
        std::vector<int32_t> chain_last(kr + 1, -1);
        const bool chains_ok = k_first == 1 && ksplit >= 2;
        if (chains_ok && spacing > 1) {
            for (int32_t a = 0; a < ksplit; a += spacing) {
                const int32_t b = std::min(a + spacing - 1, ksplit - 1);
                chain_last[a] = b > a ? b : -1;
                for (int32_t k = a + 1; k <= b; ++k) peel[k] = 0;
            }
        }
        std::vector<uint32_t> want;
        std::vector<double> wts;
        bool by_model = false;
        g->last_plan.valid = false;
        if (c->sched_mlp && chains_ok) {
            const bool fixed = c->cfg.chain_rows != 0;
            std::vector<char> p2 = peel, f2 = far;
            std::vector<int32_t> c2 = chain_last;
            // without a configured cap the model may give a unit a whole
            // XCD (the fixed rule stops at kTeamMax)
            const uint32_t mcap = c->cfg.team_max > 0 ? tmax : G / kXcds;
            // the fastest plan measured since the last edit, for the same rows
            const dcx_graph::RowsPlan* best =
                !fixed && g->best_plan.valid && g->best_plan.edits == g->edits &&
                        g->best_plan.ksplit == ksplit
                    ? &g->best_plan
                    : nullptr;
            by_model = mlp_rows_plan(c, g, ksplit - 1, G, mcap, fixed, p2, f2, c2, want, wts,
                                     best);
            if (by_model) {
                peel = p2;
                far = f2;
                chain_last = c2;
                if (!fixed) {
                    dcx_graph::RowsPlan& lp = g->last_plan;
                    lp.valid = true;
                    lp.ksplit = ksplit;
                    lp.edits = g->edits;
                    lp.peel = peel;
                    lp.far = far;
                    lp.chain_last = chain_last;
                    lp.want = want;
                    lp.wts = wts;
                }
            }
        }
        // units: anchors with their chains, scheduled by unit (team) weight
        std::vector<char> unit(kr + 1, 0);
        std::vector<double> uw(kr + 1, 0.0);
        uint32_t nunits = 0;
        for (int32_t a = 0; a < ksplit; ++a) {
            const bool peeled = a >= k_first && peel[a];
            if (!peeled && chain_last[a] < 0) continue;
            unit[a] = 1;
            ++nunits;
            if (peeled) uw[a] += 2.0 * (double)g->h_ek[a] + (double)g->h_vk[a] + 64.0;
            if (chain_last[a] >= 0)
                uw[a] += 2.0 * (double)g->h_ek[a + 1] + (double)g->h_vk[a + 1] + 64.0;
        }
        Sched s;
        uint32_t busy = 0;
        std::vector<int32_t> pkb;  // per pass of s.ks: -1 peel, else last derived row
        if (nunits) {
            s = make_schedule(g, 0, ksplit - 1, G,
                              by_model && c->cfg.team_max <= 0 ? G / kXcds : tmax,
                              by_model ? &want : nullptr, by_model ? &wts : &uw, &unit);
            c->timing.rows_scheduler = by_model ? 1 : 0;
            busy = (uint32_t)std::count_if(s.block_map.begin(), s.block_map.begin() + G,
                                           [](uint32_t t) { return t != ~0u; });
            // expand every unit into its passes: the anchor's peel, then its chain
            // (an anchor derived from row 0 is one chain a-1 -> b that
            // starts from row 0's values: only when row 0 is ready before
            // the launch)
            std::vector<int32_t> pks;
            for (TeamDesc& td : s.teams) {
                const uint32_t b0 = (uint32_t)pks.size();
                for (uint32_t i = td.pass_begin; i < td.pass_end; ++i) {
                    const int32_t a = s.ks[i];
                    if (k_first == 1 && a >= 1 && peel[a] && far[a]) {
                        const int32_t b = chain_last[a] >= 0 ? chain_last[a] : a;
                        pks.push_back(a - 1);
                        pkb.push_back(b | kChainFromRow0);
                        for (int32_t k = a; k <= b; ++k) {
                            g->h_pass_variant[k + 1] = DCX_VARIANT_DERIVED;
                            g->h_pass_team[k + 1] = (uint16_t)td.size;
                        }
                        continue;
                    }
                    if (a >= k_first && peel[a]) {
                        pks.push_back(a);
                        pkb.push_back(-1);
                        g->h_pass_variant[a + 1] = DCX_VARIANT_BATCHED;
                        g->h_pass_team[a + 1] = (uint16_t)td.size;
                    }
                    if (chain_last[a] >= 0) {
                        pks.push_back(a);
                        pkb.push_back(chain_last[a]);
                        for (int32_t k = a + 1; k <= chain_last[a]; ++k) {
                            g->h_pass_variant[k + 1] = DCX_VARIANT_DERIVED;
                            g->h_pass_team[k + 1] = (uint16_t)td.size;
                        }
                    }
                }
                td.pass_begin = b0;
                td.pass_end = (uint32_t)pks.size();
                last_teams.emplace_back(td.size, td.pass_begin < td.pass_end ? pks[b0] : -1);
            }
            s.ks = pks;
        }
        std::vector<int32_t> lks;
        for (int32_t k = ksplit; k <= kr; ++k) {
            lks.push_back(k);
            g->h_pass_variant[k + 1] = DCX_VARIANT_BATCHED;
            g->h_pass_team[k + 1] = 0;  // LDS-resident
        }
        std::stable_sort(lks.begin(), lks.end(),
                         [&](int32_t a, int32_t b) { return g->h_ek[a] > g->h_ek[b]; });
        const uint32_t lcap =
            lks.empty() ? 0u : (std::max<uint32_t>(g->h_vk[ksplit], 1) + 31) / 32 * 32;
        const uint32_t lblocks = (uint32_t)std::min<size_t>(lks.size(), (size_t)c->num_cus);
        // device copies: team table, block map, team passes, LDS passes
        const size_t nt = s.teams.size(), nb = nunits ? G : 0, nk = s.ks.size();
        const size_t bytes = nt * sizeof(TeamDesc) + 8 * nb + 8 * nk + 4 * lks.size() + 64;
        ensure_bytes(c, &c->sched, &c->sched_bytes, bytes);
        char* base = static_cast<char*>(c->sched);
        TeamDesc* d_teams = reinterpret_cast<TeamDesc*>(base);
        uint32_t* d_bt = reinterpret_cast<uint32_t*>(base + nt * sizeof(TeamDesc));
        int32_t* d_ks = reinterpret_cast<int32_t*>(d_bt + 2 * nb);
        int32_t* d_kb = d_ks + nk;
        int32_t* d_lks = d_kb + nk;
        std::vector<char> host(bytes);
        std::memcpy(host.data(), s.teams.data(), nt * sizeof(TeamDesc));
        std::memcpy(host.data() + nt * sizeof(TeamDesc), s.block_map.data(), 8 * nb);
        std::memcpy(host.data() + nt * sizeof(TeamDesc) + 8 * nb, s.ks.data(), 4 * nk);
        std::memcpy(host.data() + nt * sizeof(TeamDesc) + 8 * nb + 4 * nk, pkb.data(), 4 * nk);
        std::memcpy(host.data() + nt * sizeof(TeamDesc) + 8 * nb + 8 * nk, lks.data(),
                    4 * lks.size());
        HIPX(c, hipMemcpyAsync(c->sched, host.data(), bytes, hipMemcpyHostToDevice, c->stream));
        const size_t pool_words = nt * ((size_t)s.sync_stride + kSlotWords);
        if (nt && c->pools_words < pool_words) {
            dev_free(c->pools);
            dev_alloc(c, &c->pools, pool_words);
            c->pools_words = pool_words;
        }
        if (nt) HIPX(c, hipMemsetAsync(c->pools, 0, 4 * pool_words, c->stream));
        // the two launches run side by side, each on its own scratch (row
        // 0's grid pass runs before them on the same memory)
        const uint64_t lds_off = (s.scratch_words + 31) / 32 * 32;
        ensure_scratch(c, 4 * std::max<uint64_t>(lds_off + (uint64_t)lblocks * kScratchArrays * lcap,
                                                 k_first ? (uint64_t)kScratchArrays * V : 0));
        if (!lks.empty()) HIPX(c, hipMemsetAsync(c->ctrs + kCtrQueue, 0, 4, c->stream));
        PassSpec basep = traced(c, core_row_spec(g, 0, g->rows), true);
        basep.fail_rec = c->fail_rec;
        if (k_first && !g->row_bound) dev_alloc(c, &g->row_bound, V);

        HIPX(c, hipEventRecord(c->ev[5], c->stream));
        if (k_first && have_row0) {
            HIPX(c, hipMemcpyAsync(g->rows, g->row0, 4ull * V, hipMemcpyDeviceToDevice,
                                   c->stream));
            HIPX(c, hipMemcpyAsync(g->pstats + kStatWords, g->row0_stats, 4ull * kStatWords,
                                   hipMemcpyDeviceToDevice, c->stream));
            HIPX(c, hipMemcpyAsync(g->lmax_rows, g->row0_stats + kStatWords, 4,
                                   hipMemcpyDeviceToDevice, c->stream));
            HIPX(c, launch_gather_rank(g->row0, g->order, V, g->row_bound, c->stream));
            basep.ub = g->row_bound;
            g->h_pass_variant[1] = DCX_VARIANT_BATCHED;
            g->h_pass_team[1] = (uint16_t)c->grid_blocks;
        } else if (k_first) {
            const PassScratch s0 = scratch_at(c->scratch, V);
            HIPX(c, launch_pass_grid(traced(c, core_row_spec(g, 0, g->rows)), s0, c->ctrs,
                                     g->lmax_rows, gsync(c, c->grid_blocks), c->stream));
            ++launches;
            // the pass leaves its levels by rank in the scratch
            HIPX(c, hipMemcpyAsync(g->row_bound, s0.lv, 4ull * V, hipMemcpyDeviceToDevice,
                                   c->stream));
            basep.ub = g->row_bound;
            g->h_pass_variant[1] = DCX_VARIANT_BATCHED;
            g->h_pass_team[1] = (uint16_t)c->grid_blocks;
        }
        HIPX(c, hipEventRecord(c->sev[3], c->stream));
        if (!lks.empty()) {
            // LDS rows on the side stream: each of its workgroups needs a
            // CU's LDS to itself and waits for nothing, so it starts where
            // team workgroups have finished (or have not started yet); it
            // takes rows from its queue until the queue drains, so a team
            // member that finds no free CU can wait until then (bounded by
            // the barrier timeout, as every wait)
            HIPX(c, hipEventRecord(c->fork, c->stream));
            HIPX(c, hipStreamWaitEvent(c->side, c->fork, 0));
        }
        if (nt) {
            // row 0's chain is set up on the whole grid first (its team
            // would spend the time of several derived rows on it)
            uint32_t chain0 = 0;
            for (const TeamDesc& td : s.teams)
                for (uint32_t i = td.pass_begin; i < td.pass_end && !chain0; ++i)
                    if (s.ks[i] == 0 && pkb[i] >= 0) {
                        HIPX(c, launch_chain_preset(
                                    basep, g->vk, g->rows,
                                    static_cast<uint32_t*>(c->scratch) + td.scratch_off,
                                    td.scratch_cap, c->pools + nt * s.sync_stride + td.slot_off,
                                    V, g->h_ek.size() > 1 ? g->h_ek[1] : 0, c->stream));
                        chain0 = 1;
                    }
            HIPX(c, launch_rows_teams(basep, g->vk, g->rows, g->lmax_rows, d_teams, d_bt, d_ks,
                                      d_kb, (uint32_t)nb, c->scratch, c->pools,
                                      c->pools + nt * s.sync_stride, c->timeout_ticks,
                                      c->ctrs + kCtrErr, shrink, chain0, c->stream));
            ++launches;
            team_err = c->pools + kSyncError;
            nteam_err = (uint32_t)nt;
            team_err_stride = s.sync_stride;
            HIPX(c, hipEventRecord(c->sev[2], c->stream));
        }
        if (!lks.empty()) {
            HIPX(c, hipEventRecord(c->sev[0], c->side));
            HIPX(c, launch_rows_lds(basep, g->vk, g->rows, g->lmax_rows, d_lks,
                                    (uint32_t)lks.size(), c->ctrs + kCtrQueue, lblocks,
                                    static_cast<uint32_t*>(c->scratch) + lds_off, lcap,
                                    c->ctrs + kCtrErr, c->side));
            ++launches;
            HIPX(c, hipEventRecord(c->sev[1], c->side));
            HIPX(c, hipEventRecord(c->join, c->side));
            HIPX(c, hipStreamWaitEvent(c->stream, c->join, 0));
        }
        slots_used = (int)std::max(busy, lblocks);
        split = ksplit;
        first_ran = k_first > 0;
        teams_ran = nt > 0;
        lds_ran = !lks.empty();
    }
    HIPX(c, hipEventRecord(c->ev[6], c->stream));
    try {
        check_kernel_errors(c, "l-rows", team_err, nteam_err, team_err_stride);
    } catch (const Fail&) {
        describe_failed_passes(c, g, (int32_t)nre - 1);
        describe_fail_record(c);
        if (!last_teams.empty()) {
            std::string d = c->err + "; teams:";
            for (size_t t = 0; t < last_teams.size() && t < 6; ++t) {
                char buf[96];
                std::snprintf(buf, sizeof buf, " [%zu: size %u, k %d]", t, last_teams[t].first,
                              last_teams[t].second);
                d += buf;
            }
            c->err = d;
        }
        throw;
    }
    std::vector<int32_t> lm(nrows);
    HIPX(c, hipMemcpyAsync(lm.data(), g->lmax_rows, nrows * sizeof(int32_t),
                           hipMemcpyDeviceToHost, c->stream));
    HIPX(c, hipEventRecord(c->ev[4], c->stream));
    sync(c);
    c->timing.rows_ms = elapsed(c, 3, 4);
    c->timing.rows_kernel_ms = elapsed(c, 5, 6);
    c->timing.rows_launches = launches;
    c->timing.rows_slots = slots_used;
    c->timing.rows_split = split;
    c->timing.rows_teams_ms = 0.0f;
    c->timing.rows_lds_ms = 0.0f;
    c->timing.rows_first_ms = 0.0f;
    c->timing.rows_chains_ms = 0.0f;
    if (teams_ran) {
        float ms = 0;
        HIPX(c, hipEventElapsedTime(&ms, c->sev[3], c->sev[2]));
        c->timing.rows_teams_ms = ms;
        // the cost model's plan, measured: it becomes the best plan if it
        // beat it (or is it, measured again)
        dcx_graph::RowsPlan& lp = g->last_plan;
        dcx_graph::RowsPlan& bp = g->best_plan;
        if (lp.valid) {
            lp.teams_ms = ms;
            const bool same = bp.valid && bp.edits == lp.edits && bp.ksplit == lp.ksplit &&
                              bp.peel == lp.peel && bp.far == lp.far &&
                              bp.chain_last == lp.chain_last && bp.want == lp.want;
            if (!bp.valid || bp.edits != lp.edits || bp.ksplit != lp.ksplit || same ||
                ms < bp.teams_ms)
                bp = lp;
            lp.valid = false;
        }
    }
    if (first_ran) {
        float ms = 0;
        HIPX(c, hipEventElapsedTime(&ms, c->ev[5], c->sev[3]));
        c->timing.rows_first_ms = ms;
    }
    if (lds_ran) {
        float ms = 0;
        HIPX(c, hipEventElapsedTime(&ms, c->sev[0], c->sev[1]));
        c->timing.rows_lds_ms = ms;
    }
    g->lmax = 0;
    for (int32_t x : lm) g->lmax = std::max(g->lmax, x);
    g->h_pstats.resize(((size_t)K + 2) * kStatWords);
    HIPX(c, hipMemcpy(g->h_pstats.data(), g->pstats, g->h_pstats.size() * 4,
                      hipMemcpyDeviceToHost));
    g->rows_k = K + 1;
    g->rows_valid = true;
    if (variant == DCX_VARIANT_BATCHED) record_row_stats(g, kr, c->wall_khz);
}

__global__ void k_ragged_gather(const int32_t* rows, const int32_t* kmaxes, const int64_t* row_ptr,
                                uint32_t V, int32_t* values) {
    for (uint32_t v = blockIdx.x * blockDim.x + threadIdx.x; v < V; v += gridDim.x * blockDim.x) {
        const int64_t b = row_ptr[v];
        const int32_t km = kmaxes[v];
        for (int32_t k = 0; k <= km; ++k) values[b + k] = rows[(size_t)k * V + v];
    }
}

// Per-vertex capacities of one direction, summed (the length of a repacked
// adjacency array).
uint64_t capacity_sum(dcx_ctx* c, const uint32_t* capv, uint32_t V) {
    std::vector<uint32_t> cv(V);
    if (V) HIPX(c, hipMemcpy(cv.data(), capv, 4ull * V, hipMemcpyDeviceToHost));
    uint64_t t = 0;
    for (uint32_t x : cv) t += x;
    return t;
}

// Device lists of one direction -> reference layout (off[v+1]-off[v] =
// capacity of v, lists in place, free slots 0xFFFFFFFF).
void download_side(dcx_ctx* c, uint32_t V, const uint32_t* d_off, const int32_t* d_deg,
                   const uint32_t* d_capv, const uint32_t* d_nbr, uint64_t used, uint32_t* off,
                   int32_t* deg, uint32_t* nbr) {
    std::vector<uint32_t> o(V + 1), cv(V), all(used ? used : 1);
    HIPX(c, hipMemcpy(o.data(), d_off, 4ull * (V + 1), hipMemcpyDeviceToHost));
    if (V) {
        HIPX(c, hipMemcpy(deg, d_deg, 4ull * V, hipMemcpyDeviceToHost));
        HIPX(c, hipMemcpy(cv.data(), d_capv, 4ull * V, hipMemcpyDeviceToHost));
    }
    if (used) HIPX(c, hipMemcpy(all.data(), d_nbr, 4ull * used, hipMemcpyDeviceToHost));
    uint64_t pos = 0;
    for (uint32_t v = 0; v < V; ++v) {
        off[v] = (uint32_t)pos;
        for (uint32_t j = 0; j < cv[v]; ++j)
            nbr[pos + j] = j < (uint32_t)deg[v] ? all[o[v] + j] : 0xffffffffu;
        pos += cv[v];
    }
    off[V] = (uint32_t)pos;
}

void free_graph_arrays(dcx_graph* g) {
    dev_free(g->out_off);
    dev_free(g->in_off);
    dev_free(g->out_deg);
    dev_free(g->in_deg);
    dev_free(g->out_nbr);
    dev_free(g->in_nbr);
    dev_free(g->out_capv);
    dev_free(g->in_capv);
    dev_free(g->kmax_old);
    dev_free(g->row_bound);
    dev_free(g->row0);
    dev_free(g->row0_stats);
    dev_free(g->pstats);
    dev_free(g->kmaxes);
    dev_free(g->order);
    dev_free(g->vk);
    dev_free(g->hist);
    dev_free(g->ekh);
    dev_free(g->ekq);
    dev_free(g->rank);
    dev_free(g->cout_off);
    dev_free(g->cin_off);
    dev_free(g->cout_deg);
    dev_free(g->cin_deg);
    dev_free(g->cout_nbr);
    dev_free(g->cin_nbr);
    dev_free(g->rows);
    dev_free(g->lmax_rows);
    dev_free(g->ins_cnt_in);
    dev_free(g->ins_cnt_out);
    dev_free(g->changed);
}

}  // namespace

extern "C" {

const char* dcx_status_str(dcx_status s) {
    switch (s) {
        case DCX_OK: return "DCX_OK";
        case DCX_ERR_INVALID_ARG: return "DCX_ERR_INVALID_ARG";
        case DCX_ERR_NO_DEVICE: return "DCX_ERR_NO_DEVICE";
        case DCX_ERR_HIP: return "DCX_ERR_HIP";
        case DCX_ERR_OUT_OF_MEMORY: return "DCX_ERR_OUT_OF_MEMORY";
        case DCX_ERR_CAPACITY: return "DCX_ERR_CAPACITY";
        case DCX_ERR_IO: return "DCX_ERR_IO";
        case DCX_ERR_FORMAT: return "DCX_ERR_FORMAT";
        case DCX_ERR_NOT_FOUND: return "DCX_ERR_NOT_FOUND";
        case DCX_ERR_STATE: return "DCX_ERR_STATE";
        case DCX_ERR_BUDGET: return "DCX_ERR_BUDGET";
        case DCX_ERR_TIMEOUT: return "DCX_ERR_TIMEOUT";
        case DCX_ERR_INTERNAL: return "DCX_ERR_INTERNAL";
    }
    return "DCX_ERR_UNKNOWN";
}

dcx_status dcx_config_init(dcx_config* cfg) {
    if (!cfg) return DCX_ERR_INVALID_ARG;
    std::memset(cfg, 0, sizeof *cfg);
    cfg->variant = DCX_VARIANT_AUTO;
    return DCX_OK;
}

dcx_status dcx_device_count(int32_t* n) {
    if (!n) return DCX_ERR_INVALID_ARG;
    int c = 0;
    if (hipGetDeviceCount(&c) != hipSuccess) c = 0;
    *n = c;
    return DCX_OK;
}

dcx_status dcx_ctx_set_rows_rule(dcx_ctx* ctx, int32_t chain_rows, int32_t team_max,
                                 int32_t anchors) {
    if (!ctx || chain_rows < -1 || team_max < 0 || team_max > 256 || anchors < 0 || anchors > 1)
        return DCX_ERR_INVALID_ARG;
    std::lock_guard<std::mutex> lk(ctx->mu);
    ctx->cfg.chain_rows = chain_rows;
    ctx->cfg.team_max = team_max;
    ctx->cfg.anchors = anchors;
    return DCX_OK;
}

dcx_status dcx_ctx_create(const dcx_config* cfg, dcx_ctx** out) {
    if (!out) return DCX_ERR_INVALID_ARG;
    *out = nullptr;
    dcx_config def;
    dcx_config_init(&def);
    if (!cfg) cfg = &def;
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n <= 0) return DCX_ERR_NO_DEVICE;
    if (cfg->device < 0 || cfg->device >= n) return DCX_ERR_NO_DEVICE;
    if (cfg->variant < DCX_VARIANT_AUTO || cfg->variant > DCX_VARIANT_HINDEX)
        return DCX_ERR_INVALID_ARG;
    if (cfg->rows_slots < 0 || cfg->rows_slots > (int32_t)kMaxGridBlocks || cfg->team_max < 0 ||
        cfg->team_max > 256 || cfg->shrink_at < 0)
        return DCX_ERR_INVALID_ARG;
    if (cfg->maintenance < 0 || cfg->maintenance > 2 || cfg->local_budget < 0 ||
        cfg->chain_rows < -1 || cfg->anchors < 0 || cfg->anchors > 1)
        return DCX_ERR_INVALID_ARG;
    std::unique_ptr<dcx_ctx> c(new (std::nothrow) dcx_ctx);
    if (!c) return DCX_ERR_OUT_OF_MEMORY;
    c->cfg = *cfg;
    c->device = cfg->device;
    if (hipSetDevice(c->device) != hipSuccess) return DCX_ERR_NO_DEVICE;
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, c->device) != hipSuccess) return DCX_ERR_HIP;
    c->num_cus = prop.multiProcessorCount;
    c->grid_blocks = (int)std::min<uint32_t>(
        kMaxGridBlocks, c->num_cus * std::min(kGridBlocksPerCu, pass_grid_blocks_per_cu()));
    c->rows_per_cu = std::max(1, rows_teams_blocks_per_cu());
    c->hindex_grid = (int)std::min<uint32_t>(
        kMaxGridBlocks, c->num_cus * std::min(kGridBlocksPerCu, hindex_blocks_per_cu()));
    int clock_khz = 0;
    if (hipDeviceGetAttribute(&clock_khz, hipDeviceAttributeWallClockRate, c->device) !=
            hipSuccess ||
        clock_khz <= 0)
        clock_khz = 100000;  // 100 MHz, the CDNA constant clock
    // LDS-resident rows: 4 bytes of counters and 1 removed-bit per member
    // in what the workgroup's LDS leaves beside the kernel's static part
    int max_lds = 0;
    if (hipDeviceGetAttribute(&max_lds, hipDeviceAttributeMaxSharedMemoryPerBlock, c->device) !=
            hipSuccess)
        max_lds = 0;
    const int64_t room = (int64_t)max_lds - (int64_t)rows_lds_static_bytes() - 256;
    uint32_t cap = room > 0 ? (uint32_t)std::min<int64_t>(kLdsMaxMembers, room * 32 / 132) : 0u;
    cap = cap / 32 * 32;
    c->lds_cap_max = cap;
    if (cfg->lds_members < 0)
        cap = 0;
    else if (cfg->lds_members > 0)
        cap = std::min<uint32_t>(cap, (uint32_t)cfg->lds_members);
    c->lds_cap = cap;
    c->timeout_ticks = (uint64_t)(kBarrierTimeoutS * 1000.0 * clock_khz);
    c->wall_khz = clock_khz;
    if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess)
        return DCX_ERR_HIP;
    for (auto& e : c->ev)
        if (hipEventCreate(&e) != hipSuccess) return DCX_ERR_HIP;
    if (hipStreamCreateWithFlags(&c->side2, hipStreamNonBlocking) != hipSuccess ||
        hipEventCreate(&c->row0_done) != hipSuccess || hipEventCreate(&c->order_done) != hipSuccess ||
        hipEventCreateWithFlags(&c->clear_done, hipEventDisableTiming) != hipSuccess)
        return DCX_ERR_HIP;
    if (hipStreamCreateWithFlags(&c->side, hipStreamNonBlocking) != hipSuccess ||
        hipEventCreateWithFlags(&c->fork, hipEventDisableTiming) != hipSuccess ||
        hipEventCreateWithFlags(&c->join, hipEventDisableTiming) != hipSuccess)
        return DCX_ERR_HIP;
    for (auto& e : c->sev)
        if (hipEventCreate(&e) != hipSuccess) return DCX_ERR_HIP;
    // the static scratch of dcx_klist_noalloc: its address on this device
    if (noalloc_scratch_address(&c->noalloc_scratch) != hipSuccess || !c->noalloc_scratch)
        return DCX_ERR_HIP;
    if (hipMalloc(&c->ctrs, kCtrWords * sizeof(uint32_t)) != hipSuccess)
        return DCX_ERR_OUT_OF_MEMORY;
    if (hipMemset(c->ctrs, 0, kCtrWords * sizeof(uint32_t)) != hipSuccess) return DCX_ERR_HIP;
    if (hipMalloc(&c->d_scalars, 64 * sizeof(int32_t)) != hipSuccess)
        return DCX_ERR_OUT_OF_MEMORY;
    if (hipHostMalloc(&c->h_scalars, 64 * sizeof(int32_t), hipHostMallocDefault) != hipSuccess)
        return DCX_ERR_OUT_OF_MEMORY;
    *out = c.release();
    return DCX_OK;
}

dcx_status dcx_ctx_destroy(dcx_ctx* c) {
    if (!c) return DCX_ERR_INVALID_ARG;
    (void)hipSetDevice(c->device);
    if (c->stream) (void)hipStreamSynchronize(c->stream);
    dev_free(c->scratch);
    dev_free(c->tmp);
    dev_free(c->sched);
    dev_free(c->pools);
    dev_free(c->trace);
    dev_free(c->fail_rec);
    dev_free(c->zeros);
    dev_free(c->ctrs);
    dev_free(c->ctrs2);
    dev_free(c->d_scalars);
    dev_free(c->upd_dev);
    dev_free(c->local_scratch);
    dev_free(c->local_tasks);
    if (c->local_tasks_host) (void)hipHostFree(c->local_tasks_host);
    dev_free(c->local_words);
    dev_free(c->local_dm);
    if (c->h_scalars) (void)hipHostFree(c->h_scalars);
    if (c->upd_host) (void)hipHostFree(c->upd_host);
    for (auto& e : c->ev)
        if (e) (void)hipEventDestroy(e);
    if (c->stream) (void)hipStreamDestroy(c->stream);
    if (c->side) (void)hipStreamDestroy(c->side);
    if (c->side2) {
        (void)hipStreamSynchronize(c->side2);
        (void)hipStreamDestroy(c->side2);
    }
    if (c->row0_done) (void)hipEventDestroy(c->row0_done);
    if (c->order_done) (void)hipEventDestroy(c->order_done);
    if (c->clear_done) (void)hipEventDestroy(c->clear_done);
    if (c->fork) (void)hipEventDestroy(c->fork);
    if (c->join) (void)hipEventDestroy(c->join);
    for (auto& e : c->sev)
        if (e) (void)hipEventDestroy(e);
    delete c;
    return DCX_OK;
}

const char* dcx_ctx_last_error(const dcx_ctx* c) { return c ? c->err.c_str() : ""; }

dcx_status dcx_last_timing(const dcx_ctx* c, dcx_timing* t) {
    if (!c || !t) return DCX_ERR_INVALID_ARG;
    *t = c->timing;
    return DCX_OK;
}

dcx_status dcx_graph_upload_csr(dcx_ctx* ctx, uint32_t V, uint64_t E, const uint32_t* out_off,
                                const int32_t* out_deg, const uint32_t* out_nbr,
                                const uint32_t* in_off, const int32_t* in_deg,
                                const uint32_t* in_nbr, dcx_graph** out) {
    if (!ctx || !out || !out_off || !in_off || (V && (!out_deg || !in_deg)))
        return DCX_ERR_INVALID_ARG;
    return guarded(ctx, [&]() -> dcx_status {
        *out = nullptr;
        if (V > kMaxVertices) {
            set_err(ctx, "V = %u exceeds the supported %u vertices", V, kMaxVertices);
            return DCX_ERR_CAPACITY;
        }
        // validate the layout on the host: monotone offsets, degrees within
        // capacity, ids in range, sum of degrees equal on both sides
        uint64_t so = 0, si = 0;
        if (out_off[0] != 0 || in_off[0] != 0) {
            set_err(ctx, "offsets must start at 0");
            return DCX_ERR_INVALID_ARG;
        }
        for (uint32_t v = 0; v < V; ++v) {
            if (out_off[v + 1] < out_off[v] || in_off[v + 1] < in_off[v] || out_deg[v] < 0 ||
                in_deg[v] < 0 || (uint32_t)out_deg[v] > out_off[v + 1] - out_off[v] ||
                (uint32_t)in_deg[v] > in_off[v + 1] - in_off[v]) {
                set_err(ctx, "bad CSR entry at vertex %u", v);
                return DCX_ERR_INVALID_ARG;
            }
            so += (uint32_t)out_deg[v];
            si += (uint32_t)in_deg[v];
            for (int32_t j = 0; j < out_deg[v]; ++j)
                if (out_nbr[out_off[v] + j] >= V || out_nbr[out_off[v] + j] == v) {
                    set_err(ctx, "out-neighbour out of range or self-loop at vertex %u", v);
                    return DCX_ERR_INVALID_ARG;
                }
            for (int32_t j = 0; j < in_deg[v]; ++j)
                if (in_nbr[in_off[v] + j] >= V || in_nbr[in_off[v] + j] == v) {
                    set_err(ctx, "in-neighbour out of range or self-loop at vertex %u", v);
                    return DCX_ERR_INVALID_ARG;
                }
        }
        if (so != E || si != E) {
            set_err(ctx, "degree sums (%llu out, %llu in) differ from E=%llu",
                    (unsigned long long)so, (unsigned long long)si, (unsigned long long)E);
            return DCX_ERR_INVALID_ARG;
        }
        std::unique_ptr<dcx_graph> g(new dcx_graph);
        g->ctx = ctx;
        g->V = V;
        g->E = E;
        g->out_cap = out_off[V];
        g->in_cap = in_off[V];
        // headroom for lists relocated by later inserts
        auto headroom = [](uint64_t n) { return n + n / 16 + 4096; };
        g->out_alloc = std::min<uint64_t>(headroom(g->out_cap), 0xffffffffull);
        g->in_alloc = std::min<uint64_t>(headroom(g->in_cap), 0xffffffffull);
        std::vector<uint32_t> capo(V), capi(V);
        for (uint32_t v = 0; v < V; ++v) {
            capo[v] = out_off[v + 1] - out_off[v];
            capi[v] = in_off[v + 1] - in_off[v];
        }
        try {
            dev_alloc(ctx, &g->out_off, (size_t)V + 1);
            dev_alloc(ctx, &g->in_off, (size_t)V + 1);
            dev_alloc(ctx, &g->out_deg, V);
            dev_alloc(ctx, &g->in_deg, V);
            dev_alloc(ctx, &g->out_capv, V);
            dev_alloc(ctx, &g->in_capv, V);
            dev_alloc(ctx, &g->out_nbr, g->out_alloc);
            dev_alloc(ctx, &g->in_nbr, g->in_alloc);
            hipStream_t st = ctx->stream;
            HIPX(ctx, hipMemcpyAsync(g->out_off, out_off, 4ull * (V + 1), hipMemcpyHostToDevice, st));
            HIPX(ctx, hipMemcpyAsync(g->in_off, in_off, 4ull * (V + 1), hipMemcpyHostToDevice, st));
            if (V) {
                HIPX(ctx, hipMemcpyAsync(g->out_deg, out_deg, 4ull * V, hipMemcpyHostToDevice, st));
                HIPX(ctx, hipMemcpyAsync(g->in_deg, in_deg, 4ull * V, hipMemcpyHostToDevice, st));
                HIPX(ctx, hipMemcpyAsync(g->out_capv, capo.data(), 4ull * V, hipMemcpyHostToDevice,
                                         st));
                HIPX(ctx, hipMemcpyAsync(g->in_capv, capi.data(), 4ull * V, hipMemcpyHostToDevice,
                                         st));
            }
            HIPX(ctx, hipMemsetAsync(g->out_nbr, 0xff, 4ull * g->out_alloc, st));
            HIPX(ctx, hipMemsetAsync(g->in_nbr, 0xff, 4ull * g->in_alloc, st));
            if (g->out_cap)
                HIPX(ctx, hipMemcpyAsync(g->out_nbr, out_nbr, 4ull * g->out_cap,
                                         hipMemcpyHostToDevice, st));
            if (g->in_cap)
                HIPX(ctx, hipMemcpyAsync(g->in_nbr, in_nbr, 4ull * g->in_cap,
                                         hipMemcpyHostToDevice, st));
            sync(ctx);
        } catch (...) {
            dcx_graph* raw = g.release();
            free_graph_arrays(raw);
            delete raw;
            throw;
        }
        *out = g.release();
        return DCX_OK;
    });
}

dcx_status dcx_graph_upload_edges(dcx_ctx* ctx, uint32_t V, uint64_t E, const uint32_t* src,
                                  const uint32_t* dst, uint32_t slack, dcx_graph** out) {
    if (!ctx || !out || (E && (!src || !dst))) return DCX_ERR_INVALID_ARG;
    std::vector<uint32_t> oo(V + 1), io(V + 1);
    std::vector<int32_t> od(V), id(V);
    uint64_t cap = E + (uint64_t)slack * V;
    std::vector<uint32_t> on(cap ? cap : 1), in(cap ? cap : 1);
    dcx_status st = dcx_csr_build(V, E, src, dst, slack, oo.data(), od.data(), on.data(),
                                  io.data(), id.data(), in.data());
    if (st != DCX_OK) return st;
    return dcx_graph_upload_csr(ctx, V, E, oo.data(), od.data(), on.data(), io.data(), id.data(),
                                in.data(), out);
}

dcx_status dcx_graph_free(dcx_graph* g) {
    if (!g) return DCX_ERR_INVALID_ARG;
    dcx_ctx* c = g->ctx;
    std::lock_guard<std::mutex> lk(c->mu);
    (void)hipSetDevice(c->device);
    (void)hipStreamSynchronize(c->stream);
    free_graph_arrays(g);
    delete g;
    return DCX_OK;
}

dcx_status dcx_graph_info(const dcx_graph* g, uint32_t* V, uint64_t* E, int32_t* kmax,
                          int32_t* lmax) {
    if (!g) return DCX_ERR_INVALID_ARG;
    if (V) *V = g->V;
    if (E) *E = g->E;
    if (kmax) *kmax = g->kmax_valid ? g->kmax : -1;
    if (lmax) *lmax = g->rows_valid ? g->lmax : -1;
    return DCX_OK;
}

dcx_status dcx_graph_pass_log(dcx_ctx* ctx, const dcx_graph* g, dcx_pass_log* out,
                              uint32_t cap, uint32_t* n) {
    dcx_status s = check_graph(ctx, g);
    if (s != DCX_OK) return s;
    if (!n || (cap && !out)) return DCX_ERR_INVALID_ARG;
    return guarded(ctx, [&]() -> dcx_status {
        *n = 0;
        if (!g->kmax_valid || g->V == 0) return DCX_OK;
        // sizes of the current cores (local maintenance may have moved them;
        // the passes' own words describe the last decomposition that ran)
        ensure_order(ctx, const_cast<dcx_graph*>(g));
        const uint32_t np = (uint32_t)g->kmax + 2;
        *n = np;
        if (!cap) return DCX_OK;
        const uint32_t m = std::min(cap, np);
        std::vector<uint32_t> st((size_t)m * kStatWords);
        sync(ctx);
        HIPX(ctx, hipMemcpy(st.data(), g->pstats, st.size() * 4, hipMemcpyDeviceToHost));
        for (uint32_t i = 0; i < m; ++i) {
            dcx_pass_log& e = out[i];
            std::memset(&e, 0, sizeof e);
            e.k = (int32_t)i - 1;
            // a row's entry is reported once it was ever computed (after a
            // failed launch too, for diagnosis)
            const bool ran = i == 0 || g->h_pass_variant[i] != 0;
            e.variant = ran ? g->h_pass_variant[i] : 0;
            e.team = ran ? g->h_pass_team[i] : 0;
            e.members = i == 0 ? g->V : g->h_vk[i - 1];
            e.edges = i == 0 ? g->E : g->h_ek[i - 1];
            if (!ran) continue;
            const uint32_t* w = st.data() + (size_t)i * kStatWords;
            e.phases = w[kStatPhases];
            e.levels = w[kStatLevels];
            const uint64_t ticks = (uint64_t)w[kStatTicksLo] | ((uint64_t)w[kStatTicksHi] << 32);
            e.ms = (double)ticks / ctx->wall_khz;
            e.failed = w[kStatFailed];
            e.removed = w[kStatRemoved];
            e.alive = w[kStatAlive];
            e.level = (int32_t)w[kStatLevel];
        }
        return DCX_OK;
    });
}

dcx_status dcx_last_trace(dcx_ctx* ctx, dcx_phase_trace* out, uint32_t cap, uint32_t* n) {
    if (!ctx || !n || (cap && !out)) return DCX_ERR_INVALID_ARG;
    return guarded(ctx, [&]() -> dcx_status {
        *n = 0;
        if (!ctx->trace) return DCX_OK;
        sync(ctx);
        uint32_t cnt = 0;
        HIPX(ctx, hipMemcpy(&cnt, ctx->trace, 4, hipMemcpyDeviceToHost));
        *n = cnt;
        const uint32_t m = std::min(cap, cnt);
        if (!m) return DCX_OK;
        std::vector<uint32_t> w((size_t)m * kTraceWords);
        HIPX(ctx, hipMemcpy(w.data(), ctx->trace + kTraceWords, w.size() * 4,
                            hipMemcpyDeviceToHost));
        for (uint32_t i = 0; i < m; ++i) {
            const uint32_t* e = w.data() + (size_t)i * kTraceWords;
            out[i].level = (int32_t)(e[kTraceLevel] & kTraceLevelMask);
            out[i].seed = (e[kTraceLevel] & kTraceSeedBit) ? 1 : (e[kTraceLevel] & kTracePullBit) ? 2 : 0;
            out[i].removed = e[kTraceFrontier];
            out[i].alive = e[kTraceAlive];
            out[i].us = 1000.0 * e[kTraceTicks] / ctx->wall_khz;
        }
        return DCX_OK;
    });
}

dcx_status dcx_graph_capacity(const dcx_graph* g, uint64_t* out_cap, uint64_t* in_cap) {
    if (!g) return DCX_ERR_INVALID_ARG;
    dcx_ctx* ctx = g->ctx;
    return guarded(ctx, [&]() -> dcx_status {
        if (out_cap) *out_cap = capacity_sum(ctx, g->out_capv, g->V);
        if (in_cap) *in_cap = capacity_sum(ctx, g->in_capv, g->V);
        return DCX_OK;
    });
}

dcx_status dcx_graph_download_csr(dcx_ctx* ctx, const dcx_graph* g, uint32_t* out_off,
                                  int32_t* out_deg, uint32_t* out_nbr, uint32_t* in_off,
                                  int32_t* in_deg, uint32_t* in_nbr) {
    dcx_status s = check_graph(ctx, g);
    if (s != DCX_OK) return s;
    if (!out_off || !out_deg || !out_nbr || !in_off || !in_deg || !in_nbr)
        return DCX_ERR_INVALID_ARG;
    return guarded(ctx, [&]() -> dcx_status {
        sync(ctx);
        download_side(ctx, g->V, g->out_off, g->out_deg, g->out_capv, g->out_nbr, g->out_cap,
                      out_off, out_deg, out_nbr);
        download_side(ctx, g->V, g->in_off, g->in_deg, g->in_capv, g->in_nbr, g->in_cap, in_off,
                      in_deg, in_nbr);
        return DCX_OK;
    });
}

dcx_status dcx_kmax(dcx_ctx* ctx, dcx_graph* g, int32_t* kmaxes, int32_t* kmax) {
    dcx_status s = check_graph(ctx, g);
    if (s != DCX_OK) return s;
    return guarded(ctx, [&]() -> dcx_status {
        compute_kmax(ctx, g);
        g->rows_valid = false;
        if (kmaxes && g->V)
            HIPX(ctx, hipMemcpy(kmaxes, g->kmaxes, 4ull * g->V, hipMemcpyDeviceToHost));
        if (kmax) *kmax = g->kmax;
        return DCX_OK;
    });
}

dcx_status dcx_klist(dcx_ctx* ctx, dcx_graph* g, int32_t k, int32_t variant, int32_t* lvals,
                     int32_t* lmax_k) {
    dcx_status s = check_graph(ctx, g);
    if (s != DCX_OK) return s;
    if (k < 0) return DCX_ERR_INVALID_ARG;
    if (variant != DCX_VARIANT_AUTO && variant != DCX_VARIANT_PEEL &&
        variant != DCX_VARIANT_HINDEX) {
        set_err(ctx, "dcx_klist: variant %d not available for a single row", variant);
        return DCX_ERR_INVALID_ARG;
    }
    return guarded(ctx, [&]() -> dcx_status {
        const uint32_t V = g->V;
        if (V == 0) {
            if (lmax_k) *lmax_k = -1;
            return DCX_OK;
        }
        int32_t* row = nullptr;
        dev_alloc(ctx, &row, V);
        int32_t lk = 0;
        try {
            // with resident kmaxes the pass runs on the core-ordered graph;
            // otherwise it runs on the whole graph exactly like KList.  The
            // h-index fixpoint needs the core, so it computes kmaxes first.
            if (variant == DCX_VARIANT_HINDEX && !g->kmax_valid) compute_kmax(ctx, g);
            if (g->kmax_valid) ensure_core(ctx, g);
            if (g->kmax_valid && g->core_valid) {
                HIPX(ctx, hipMemsetAsync(row, 0, 4ull * V, ctx->stream));
                PassSpec p = core_row_spec(g, k, row);
                if (p.nmembers == 0) {
                    lk = 0;
                } else if (variant == DCX_VARIANT_HINDEX) {
                    ensure_scratch(ctx, (size_t)kScratchArrays * V * 4);
                    clear_kernel_errors(ctx);
                    HIPX(ctx, launch_hindex_row(p, scratch_at(ctx->scratch, V), nullptr,
                                                ctx->ctrs + kCtrFlags, ctx->d_scalars,
                                                gsync(ctx, ctx->hindex_grid), ctx->stream));
                    HIPX(ctx, hipMemcpyAsync(ctx->h_scalars, ctx->d_scalars, 4,
                                             hipMemcpyDeviceToHost, ctx->stream));
                    check_kernel_errors(ctx, "h-index row");
                    lk = ctx->h_scalars[0];
                } else {
                    lk = run_grid_pass(ctx, p);
                }
            } else {
                lk = run_grid_pass(ctx, full_row_spec(g, k, row));
            }
            if (lvals) HIPX(ctx, hipMemcpy(lvals, row, 4ull * V, hipMemcpyDeviceToHost));
        } catch (...) {
            dev_free(row);
            throw;
        }
        dev_free(row);
        if (lmax_k) *lmax_k = lk;
        return DCX_OK;
    });
}

dcx_status dcx_klist_noalloc(dcx_ctx* ctx, dcx_graph* g, int32_t k, int32_t* lvals,
                             int32_t* lmax_k) {
    dcx_status s = check_graph(ctx, g);
    if (s != DCX_OK) return s;
    if (k < 0 || !lvals || !lmax_k) return DCX_ERR_INVALID_ARG;
    return guarded(ctx, [&]() -> dcx_status {
        const uint32_t V = g->V;
        if (V > kNoAllocMaxVertices) {
            set_err(ctx, "dcx_klist_noalloc: V = %u exceeds its %u vertices of static scratch", V,
                    kNoAllocMaxVertices);
            return DCX_ERR_CAPACITY;
        }
        if (ctx->cfg.trace_phases) {
            set_err(ctx, "dcx_klist_noalloc: the phase trace allocates; turn trace_phases off");
            return DCX_ERR_STATE;
        }
        if (V == 0) {
            *lmax_k = -1;
            return DCX_OK;
        }
        // the whole-graph pass of dcx_klist (KList semantics), with its
        // scratch in the static array and its output in the caller's array;
        // the pass initialises every scratch word it reads
        PassSpec p = full_row_spec(g, k, lvals);
        p.trace_k = -1;
        p.pull_ratio = ctx->cfg.pull_ratio < 0    ? 0u
                       : ctx->cfg.pull_ratio == 0 ? kDefaultPullRatio
                                                  : (uint32_t)ctx->cfg.pull_ratio;
        clear_kernel_errors(ctx);
        HIPX(ctx, launch_pass_grid(p, scratch_at(ctx->noalloc_scratch, V), ctx->ctrs,
                                   ctx->d_scalars, gsync(ctx, ctx->grid_blocks), ctx->stream));
        HIPX(ctx, hipMemcpyAsync(ctx->h_scalars, ctx->d_scalars, sizeof(int32_t),
                                 hipMemcpyDeviceToHost, ctx->stream));
        check_kernel_errors(ctx, "allocation-free row pass");
        *lmax_k = ctx->h_scalars[0];
        return DCX_OK;
    });
}

dcx_status dcx_decompose(dcx_ctx* ctx, dcx_graph* g, int32_t variant, int32_t* kmax,
                         int32_t* lmax) {
    dcx_status s = check_graph(ctx, g);
    if (s != DCX_OK) return s;
    if (variant == DCX_VARIANT_AUTO) variant = ctx->cfg.variant;
    if (variant == DCX_VARIANT_AUTO) variant = DCX_VARIANT_BATCHED;
    if (variant != DCX_VARIANT_PEEL && variant != DCX_VARIANT_BATCHED &&
        variant != DCX_VARIANT_HINDEX) {
        set_err(ctx, "dcx_decompose: unknown variant %d", variant);
        return DCX_ERR_INVALID_ARG;
    }
    return guarded(ctx, [&]() -> dcx_status {
        ctx->timing = dcx_timing{};
        const bool rows_zero = g->rows_valid && g->kmax_valid;
        g->rows_valid = false;
        const auto t0 = std::chrono::steady_clock::now();
        compute_kmax(ctx, g, variant, false, true, rows_zero);
        compute_rows(ctx, g, variant);
        sync(ctx);
        g->full_ms = std::chrono::duration<double, std::milli>(
                         std::chrono::steady_clock::now() - t0).count();
        ctx->timing.total_ms = ctx->timing.kmax_ms + ctx->timing.order_ms + ctx->timing.rows_ms;
        if (kmax) *kmax = g->kmax;
        if (lmax) *lmax = g->lmax;
        return DCX_OK;
    });
}

dcx_status dcx_get_kmaxes(dcx_ctx* ctx, const dcx_graph* g, int32_t* kmaxes) {
    dcx_status s = check_graph(ctx, g);
    if (s != DCX_OK) return s;
    if (!kmaxes && g->V) return DCX_ERR_INVALID_ARG;
    if (!g->kmax_valid) return DCX_ERR_STATE;
    return guarded(ctx, [&]() -> dcx_status {
        if (g->V) HIPX(ctx, hipMemcpy(kmaxes, g->kmaxes, 4ull * g->V, hipMemcpyDeviceToHost));
        return DCX_OK;
    });
}

dcx_status dcx_get_rows(dcx_ctx* ctx, const dcx_graph* g, int32_t k0, int32_t nrows,
                        int32_t* rows) {
    dcx_status s = check_graph(ctx, g);
    if (s != DCX_OK) return s;
    if (!g->rows_valid) return DCX_ERR_STATE;
    if (k0 < 0 || nrows < 0 || (int64_t)k0 + nrows > (int64_t)g->kmax + 1 || (nrows && !rows))
        return DCX_ERR_INVALID_ARG;
    return guarded(ctx, [&]() -> dcx_status {
        size_t n = (size_t)nrows * g->V;
        if (n)
            HIPX(ctx, hipMemcpy(rows, g->rows + (size_t)k0 * g->V, 4 * n, hipMemcpyDeviceToHost));
        return DCX_OK;
    });
}

dcx_status dcx_ragged_size(dcx_ctx* ctx, const dcx_graph* g, uint64_t* n) {
    dcx_status s = check_graph(ctx, g);
    if (s != DCX_OK) return s;
    if (!n) return DCX_ERR_INVALID_ARG;
    if (!g->kmax_valid) return DCX_ERR_STATE;
    uint64_t t = 0;
    for (int32_t k = 0; k <= g->kmax; ++k) t += g->h_vk[k];
    *n = t;
    return DCX_OK;
}

dcx_status dcx_get_ragged(dcx_ctx* ctx, const dcx_graph* g, int64_t* row_ptr, int32_t* values) {
    dcx_status s = check_graph(ctx, g);
    if (s != DCX_OK) return s;
    if (!row_ptr || (g->V && !values)) return DCX_ERR_INVALID_ARG;
    if (!g->rows_valid) return DCX_ERR_STATE;
    return guarded(ctx, [&]() -> dcx_status {
        const uint32_t V = g->V;
        std::vector<int32_t> km(V);
        if (V) HIPX(ctx, hipMemcpy(km.data(), g->kmaxes, 4ull * V, hipMemcpyDeviceToHost));
        row_ptr[0] = 0;
        for (uint32_t v = 0; v < V; ++v) row_ptr[v + 1] = row_ptr[v] + km[v] + 1;
        const uint64_t n = (uint64_t)row_ptr[V];
        if (n == 0) return DCX_OK;
        int64_t* d_rp = nullptr;
        int32_t* d_val = nullptr;
        dev_alloc(ctx, &d_rp, (size_t)V + 1);
        try {
            dev_alloc(ctx, &d_val, n);
            HIPX(ctx, hipMemcpyAsync(d_rp, row_ptr, 8ull * (V + 1), hipMemcpyHostToDevice,
                                     ctx->stream));
            hipLaunchKernelGGL(k_ragged_gather, dim3(std::min<uint32_t>((V + 255) / 256, 4096)),
                               dim3(256), 0, ctx->stream, g->rows, g->kmaxes, d_rp, V, d_val);
            HIPX(ctx, hipGetLastError());
            HIPX(ctx, hipMemcpyAsync(values, d_val, 4ull * n, hipMemcpyDeviceToHost, ctx->stream));
            sync(ctx);
        } catch (...) {
            dev_free(d_rp);
            dev_free(d_val);
            throw;
        }
        dev_free(d_rp);
        dev_free(d_val);
        return DCX_OK;
    });
}

}  // extern "C"

// Edge-update maintenance and graph batches (same translation unit: they
// use the context and graph internals above).
#include "dcx_update_impl.hpp"
#include "dcx_batch_impl.hpp"
#include "dcx_tune_impl.hpp"
#include "dcx_serve_impl.hpp"
#include "dcx_mlp_impl.hpp"
#include "dcx_ingest_impl.hpp"
