// Edge-update maintenance: included at the end of dcx_api.hip.
//
// Mirrors the reference's maintenance drivers (maintenance /
// maintenanceFast, main.cu:1536-1751): apply a batch to the device graph,
// then bring kmaxes and the l-rows back to the values a full recomputation
// would give.  The reference raises kmaxes by h-index refinement and then
// recomputes rows k = 0..M+1 (main.cu:1731-1741).  Here:
//   * the batch is validated, then applied in place on the device,
//     deletions first (capacity is checked before anything changes; lists
//     that would overflow are moved to the tail region);
//   * by default (dcx_config.maintenance = 0) maintenance is local
//     (local_maintain, tasks in local_kernels.hip): the batch's net edits
//     D and I are maintained as "D on G - D, then I": the insertions are
//     taken out of the lists again, kmaxes and rows are lowered by h-index
//     refinement from their old values (seeds: endpoints of D, vertices
//     leaving a core), the insertions are put back at the ends of the
//     lists, kmaxes and row 0 are raised in +1 rounds over candidate
//     regions, and rows k >= 1 over candidate regions bounded by the new
//     row 0.  A task whose worklist overflows makes the batch fall back to
//     recomputation;
//   * recomputation (maintenance = 1, and the fallback): kmaxes re-peeled
//     (a warm h-index fixpoint from the old kmaxes when the batch only
//     deletes), then rows k <= R recomputed in one batched launch, where
//     D_{k,l} can change only if a deleted edge had both endpoints in the
//     old core or an inserted edge has both in the new one: R =
//     max(M_del(old kmaxes), M_ins(new kmaxes)), the getMValue quantity
//     (main.cu:344-364).
// The bound uses the new kmaxes for insertions because a batch of inserts
// can raise kmaxes by more than one (e.g. a clique inserted among isolated
// vertices), which a bound from the old kmaxes would miss.
#pragma once

#include <chrono>
#include <unordered_map>
#include <unordered_set>

namespace {

struct Groups {
    std::vector<uint32_t> key, start, cnt, val;
};

// Sorts 64-bit keys: an LSD radix sort by bytes for large inputs, skipping
// the bytes every key shares (vertex ids use few of their 32 bits), so a
// batch of 40 K pairs sorts in a fraction of std::sort's ~3 ms on the
// serving path.
void sort_u64(std::vector<uint64_t>& v) {
    const size_t n = v.size();
    if (n < 4096) {
        std::sort(v.begin(), v.end());
        return;
    }
    std::vector<uint64_t> tmp(n);
    for (int shift = 0; shift < 64; shift += 8) {
        size_t cnt[257] = {};
        for (uint64_t x : v) ++cnt[((x >> shift) & 0xff) + 1];
        if (cnt[((v[0] >> shift) & 0xff) + 1] == n) continue;  // one bucket: nothing moves
        for (int d = 0; d < 256; ++d) cnt[d + 1] += cnt[d];
        for (uint64_t x : v) tmp[cnt[(x >> shift) & 0xff]++] = x;
        v.swap(tmp);
    }
}

// Groups pairs by source (out-lists) or by target (in-lists).
Groups group_pairs(const uint32_t* pairs, uint32_t n, bool by_src) {
    std::vector<uint64_t> kv(n);
    for (uint32_t i = 0; i < n; ++i) {
        const uint64_t a = pairs[2 * i], b = pairs[2 * i + 1];
        kv[i] = by_src ? (a << 32 | b) : (b << 32 | a);
    }
    sort_u64(kv);
    Groups g;
    g.val.resize(n);
    for (uint32_t i = 0; i < n; ++i) {
        const uint32_t key = (uint32_t)(kv[i] >> 32);
        if (g.key.empty() || g.key.back() != key) {
            g.key.push_back(key);
            g.start.push_back(i);
            g.cnt.push_back(0);
        }
        g.cnt.back()++;
        g.val[i] = (uint32_t)kv[i];
    }
    return g;
}

// Shape checks of a pair list (ids in range, no self-loop, no repeat); the
// detail goes to c's error text when c is given.
dcx_status validate_pairs(dcx_ctx* c, uint32_t V, const uint32_t* p, uint32_t n,
                          const char* what) {
    std::vector<uint64_t> keys(n);
    for (uint32_t i = 0; i < n; ++i) {
        const uint32_t u = p[2 * i], v = p[2 * i + 1];
        if (u >= V || v >= V) {
            if (c) set_err(c, "%s edge %u: vertex id out of range (V=%u)", what, i, V);
            return DCX_ERR_INVALID_ARG;
        }
        if (u == v) {
            if (c) set_err(c, "%s edge %u: self-loop %u->%u", what, i, u, v);
            return DCX_ERR_INVALID_ARG;
        }
        keys[i] = (uint64_t)u << 32 | v;
    }
    sort_u64(keys);
    for (uint32_t i = 1; i < n; ++i)
        if (keys[i] == keys[i - 1]) {
            if (c)
                set_err(c, "%s batch lists %u->%u twice", what, (uint32_t)(keys[i] >> 32),
                        (uint32_t)keys[i]);
            return DCX_ERR_INVALID_ARG;
        }
    return DCX_OK;
}

// Device + pinned host staging of at least `words` uint32 words.
void ensure_staging(dcx_ctx* c, size_t words) {
    if (c->upd_dev_words < words) {
        dev_free(c->upd_dev);
        c->upd_dev_words = 0;
        const size_t n = std::max<size_t>(words, 1 << 16);
        dev_alloc(c, &c->upd_dev, n);
        c->upd_dev_words = n;
    }
    if (c->upd_host_words < words) {
        if (c->upd_host) (void)hipHostFree(c->upd_host);
        c->upd_host = nullptr;
        c->upd_host_words = 0;
        const size_t n = std::max<size_t>(words, 1 << 16);
        HIPX(c, hipHostMalloc(reinterpret_cast<void**>(&c->upd_host), n * 4,
                              hipHostMallocDefault));
        c->upd_host_words = n;
    }
}

struct DevGroups {
    uint32_t *key, *start, *cnt, *val;
    uint32_t n;
};

// Grows one adjacency array to hold `need` words, keeping [0, used).
void grow_adjacency(dcx_ctx* c, uint32_t** nbr, uint64_t* alloc, uint64_t used, uint64_t need) {
    if (need <= *alloc) return;
    uint64_t n = std::max<uint64_t>(need + need / 4, *alloc + *alloc / 2);
    n = std::min<uint64_t>(n, 0xffffffffull);
    uint32_t* q = nullptr;
    dev_alloc(c, &q, n);
    HIPX(c, hipMemsetAsync(q, 0xff, 4ull * n, c->stream));
    if (used) HIPX(c, hipMemcpyAsync(q, *nbr, 4ull * used, hipMemcpyDeviceToDevice, c->stream));
    sync(c);
    dev_free(*nbr);
    *nbr = q;
    *alloc = n;
}

// Moves lists whose free room is below the batch's demand; `ovf` holds
// deg+1 for such groups (0 otherwise).  Returns the lists moved.
uint32_t relocate_side(dcx_ctx* c, dcx_graph* g, bool out, const Groups& G,
                       const uint32_t* ovf) {
    std::vector<uint32_t> rk, ro, rc;
    uint64_t& used = out ? g->out_cap : g->in_cap;
    uint64_t end = used;
    for (size_t i = 0; i < G.key.size(); ++i) {
        if (!ovf[i]) continue;
        const uint64_t need = (uint64_t)(ovf[i] - 1) + G.cnt[i];
        const uint64_t cap = std::max<uint64_t>(2 * need, need + 8);
        rk.push_back(G.key[i]);
        ro.push_back((uint32_t)end);
        rc.push_back((uint32_t)cap);
        end += cap;
    }
    if (rk.empty()) return 0;
    if (end > 0xffffffffull) {
        set_err(c, "adjacency array would exceed 2^32 slots");
        throw Fail{DCX_ERR_CAPACITY};
    }
    if (out)
        grow_adjacency(c, &g->out_nbr, &g->out_alloc, used, end);
    else
        grow_adjacency(c, &g->in_nbr, &g->in_alloc, used, end);
    const uint32_t nr = (uint32_t)rk.size();
    uint32_t* d = nullptr;
    dev_alloc(c, &d, 3ull * nr);
    try {
        HIPX(c, hipMemcpyAsync(d, rk.data(), 4ull * nr, hipMemcpyHostToDevice, c->stream));
        HIPX(c, hipMemcpyAsync(d + nr, ro.data(), 4ull * nr, hipMemcpyHostToDevice, c->stream));
        HIPX(c, hipMemcpyAsync(d + 2 * nr, rc.data(), 4ull * nr, hipMemcpyHostToDevice, c->stream));
        if (out)
            HIPX(c, launch_relocate(g->out_off, g->out_deg, g->out_capv, g->out_nbr, d, d + nr,
                                    d + 2 * nr, nr, c->stream));
        else
            HIPX(c, launch_relocate(g->in_off, g->in_deg, g->in_capv, g->in_nbr, d, d + nr,
                                    d + 2 * nr, nr, c->stream));
        sync(c);
    } catch (...) {
        dev_free(d);
        throw;
    }
    dev_free(d);
    used = end;
    return nr;
}

int32_t batch_mvalue(dcx_ctx* c, const int32_t* kmaxes, const uint32_t* d_pairs, uint32_t n) {
    if (!n) return -1;
    HIPX(c, launch_fill_i32(c->d_scalars + 2, -1, 1, c->stream));
    HIPX(c, launch_mvalue(kmaxes, d_pairs, n, c->d_scalars + 2, c->stream));
    HIPX(c, hipMemcpyAsync(c->h_scalars + 2, c->d_scalars + 2, 4, hipMemcpyDeviceToHost,
                           c->stream));
    sync(c);
    return c->h_scalars[2];
}

using Clock = std::chrono::steady_clock;
double ms_since(Clock::time_point t) {
    return std::chrono::duration<double, std::milli>(Clock::now() - t).count();
}

// Device layout of one update batch: the pair lists, their groupings by
// source and by target, and the result words the checks write.
struct Staged {
    Groups G[4];  // deletions by source / by target, insertions by source / by target
    std::vector<size_t> at;
    size_t in_words = 0, words = 0;
    size_t i_del = 0, i_ins = 0, i_grp[4][4] = {}, i_fdel = 0, i_fins = 0, i_ovo = 0, i_ovi = 0,
           i_miss = 0;
    uint32_t n_ins = 0, n_del = 0;
};

Staged layout_batch(const uint32_t* ins, uint32_t n_ins, const uint32_t* del, uint32_t n_del) {
    Staged L;
    L.n_ins = n_ins;
    L.n_del = n_del;
    L.G[0] = group_pairs(del, n_del, true);
    L.G[1] = group_pairs(del, n_del, false);
    L.G[2] = group_pairs(ins, n_ins, true);
    L.G[3] = group_pairs(ins, n_ins, false);
    auto reserve = [&](size_t n) {
        L.at.push_back(L.words);
        L.words += n;
        return L.at.size() - 1;
    };
    L.i_del = reserve(2ull * n_del);
    L.i_ins = reserve(2ull * n_ins);
    for (int q = 0; q < 4; ++q) {
        L.i_grp[q][0] = reserve(L.G[q].key.size());
        L.i_grp[q][1] = reserve(L.G[q].start.size());
        L.i_grp[q][2] = reserve(L.G[q].cnt.size());
        L.i_grp[q][3] = reserve(L.G[q].val.size());
    }
    L.in_words = L.words;
    L.i_fdel = reserve(n_del);
    L.i_fins = reserve(n_ins);
    L.i_ovo = reserve(L.G[2].key.size());
    L.i_ovi = reserve(L.G[3].key.size());
    L.i_miss = reserve(1);
    return L;
}

// Writes the batch's inputs (words [0, in_words)) into host memory H.
void fill_batch(const Staged& L, uint32_t* H, const uint32_t* ins, const uint32_t* del) {
    auto put = [&](size_t idx, const uint32_t* src, size_t n) {
        if (n) std::memcpy(H + L.at[idx], src, 4 * n);
    };
    put(L.i_del, del, 2ull * L.n_del);
    put(L.i_ins, ins, 2ull * L.n_ins);
    for (int q = 0; q < 4; ++q) {
        put(L.i_grp[q][0], L.G[q].key.data(), L.G[q].key.size());
        put(L.i_grp[q][1], L.G[q].start.data(), L.G[q].start.size());
        put(L.i_grp[q][2], L.G[q].cnt.data(), L.G[q].cnt.size());
        put(L.i_grp[q][3], L.G[q].val.data(), L.G[q].val.size());
    }
}

// Applies one validated-by-shape batch to the device graph: checks that
// deleted edges exist and inserted ones do not (unless deleted first in
// the same batch), moves lists that would overflow, then deletes and
// inserts.  H/D: host (pinned) and device staging of L.words words whose
// inputs are already on the device when `uploaded` (else uploaded here on
// the library stream).  On a state error (NOT_FOUND, duplicate insert) the
// graph is unchanged.  *m_del = M over the deletions with the kmaxes
// resident before the batch (-1 if none or no kmaxes).
dcx_status edit_batch(dcx_ctx* c, dcx_graph* g, const Staged& L, uint32_t* H, uint32_t* D,
                      bool uploaded, const uint32_t* ins, const uint32_t* del, int32_t* m_del,
                      uint32_t* relocated) {
    g->audited = false;  // the lists change below
    ++g->edits;
    g->row0_ready = false;
    const uint32_t n_ins = L.n_ins, n_del = L.n_del;
    auto dptr = [&](size_t idx) { return D + L.at[idx]; };
    if (!uploaded && L.in_words)
        HIPX(c, hipMemcpyAsync(D, H, 4 * L.in_words, hipMemcpyHostToDevice, c->stream));
    HIPX(c, hipMemsetAsync(D + L.in_words, 0, 4 * (L.words - L.in_words), c->stream));
    DevGroups dg[4];
    for (int q = 0; q < 4; ++q)
        dg[q] = DevGroups{dptr(L.i_grp[q][0]), dptr(L.i_grp[q][1]), dptr(L.i_grp[q][2]),
                          dptr(L.i_grp[q][3]), (uint32_t)L.G[q].key.size()};

    // ---- existence: deletions must exist, insertions must not (unless the
    // same edge is deleted first in this batch)
    const Csr so{g->out_off, g->out_deg, g->out_nbr}, si{g->in_off, g->in_deg, g->in_nbr};
    HIPX(c, launch_find_edges(so, si, dptr(L.i_del), n_del, dptr(L.i_fdel), c->stream));
    HIPX(c, launch_find_edges(so, si, dptr(L.i_ins), n_ins, dptr(L.i_fins), c->stream));
    HIPX(c, launch_check_room(g->out_deg, g->out_capv, dg[2].key, dg[2].cnt, dg[2].n,
                              dptr(L.i_ovo), c->stream));
    HIPX(c, launch_check_room(g->in_deg, g->in_capv, dg[3].key, dg[3].cnt, dg[3].n,
                              dptr(L.i_ovi), c->stream));
    if (L.words > L.in_words)
        HIPX(c, hipMemcpyAsync(H + L.in_words, D + L.in_words, 4 * (L.words - L.in_words),
                               hipMemcpyDeviceToHost, c->stream));
    sync(c);
    for (uint32_t i = 0; i < n_del; ++i)
        if (!H[L.at[L.i_fdel] + i]) {
            set_err(c, "delete: edge %u->%u is not in the graph", del[2 * i], del[2 * i + 1]);
            return DCX_ERR_NOT_FOUND;
        }
    if (n_ins) {
        std::unordered_set<uint64_t> deleted;
        for (uint32_t i = 0; i < n_del; ++i)
            deleted.insert((uint64_t)del[2 * i] << 32 | del[2 * i + 1]);
        for (uint32_t i = 0; i < n_ins; ++i)
            if (H[L.at[L.i_fins] + i] &&
                !deleted.count((uint64_t)ins[2 * i] << 32 | ins[2 * i + 1])) {
                set_err(c, "insert: edge %u->%u is already in the graph", ins[2 * i],
                        ins[2 * i + 1]);
                return DCX_ERR_INVALID_ARG;
            }
    }

    // ---- M over deletions with the old kmaxes (before anything changes)
    *m_del = g->kmax_valid ? batch_mvalue(c, g->kmaxes, dptr(L.i_del), n_del) : -1;

    // ---- structural update
    std::vector<uint32_t> ovo(H + L.at[L.i_ovo], H + L.at[L.i_ovo] + L.G[2].key.size());
    std::vector<uint32_t> ovi(H + L.at[L.i_ovi], H + L.at[L.i_ovi] + L.G[3].key.size());
    uint32_t rel = relocate_side(c, g, true, L.G[2], ovo.data());
    rel += relocate_side(c, g, false, L.G[3], ovi.data());
    g->relocations += rel;
    if (relocated) *relocated = rel;
    HIPX(c, launch_apply_deletes(g->out_off, g->out_deg, g->out_nbr, dg[0].key, dg[0].start,
                                 dg[0].cnt, dg[0].val, dg[0].n, dptr(L.i_miss), c->stream));
    HIPX(c, launch_apply_deletes(g->in_off, g->in_deg, g->in_nbr, dg[1].key, dg[1].start,
                                 dg[1].cnt, dg[1].val, dg[1].n, dptr(L.i_miss), c->stream));
    HIPX(c, launch_apply_inserts(g->out_off, g->out_deg, g->out_nbr, dg[2].key, dg[2].start,
                                 dg[2].cnt, dg[2].val, dg[2].n, c->stream));
    HIPX(c, launch_apply_inserts(g->in_off, g->in_deg, g->in_nbr, dg[3].key, dg[3].start,
                                 dg[3].cnt, dg[3].val, dg[3].n, c->stream));
    HIPX(c, hipMemcpyAsync(H + L.at[L.i_miss], dptr(L.i_miss), 4, hipMemcpyDeviceToHost,
                           c->stream));
    sync(c);
    if (H[L.at[L.i_miss]]) {
        g->broken = true;
        g->kmax_valid = g->rows_valid = false;
        set_err(c, "delete: %u edges vanished between check and apply", H[L.at[L.i_miss]]);
        throw Fail{DCX_ERR_INTERNAL};
    }
    g->E = g->E + n_ins - n_del;
    return DCX_OK;
}

// Brings kmaxes and the rows back after edits: `ins` (host, n_ins pairs)
// are all edges inserted since the decomposition was last valid, m_del the
// largest M of the deletions (old kmaxes), deletion_only whether nothing
// was inserted.  Fills the maintenance fields of *S.
void maintain(dcx_ctx* c, dcx_graph* g, const uint32_t* ins, uint32_t n_ins, int32_t m_del,
              bool changed, bool deletion_only, dcx_update_stats* S) {
    const auto t1 = Clock::now();
    const bool rows = g->rows_valid;
    const int var = c->cfg.variant == DCX_VARIANT_AUTO ? DCX_VARIANT_BATCHED : c->cfg.variant;
    // deletions alone can only lower kmaxes: the old ones are an upper
    // bound and the h-index fixpoint starts from them
    if (changed) compute_kmax(c, g, var, deletion_only);
    S->m_ins = -1;
    if (n_ins) {
        ensure_staging(c, 2ull * n_ins);
        std::memcpy(c->upd_host, ins, 8ull * n_ins);
        HIPX(c, hipMemcpyAsync(c->upd_dev, c->upd_host, 8ull * n_ins, hipMemcpyHostToDevice,
                               c->stream));
        S->m_ins = batch_mvalue(c, g->kmaxes, c->upd_dev, n_ins);
    }
    S->kmax_ms = ms_since(t1);
    const int32_t R = std::max(m_del, S->m_ins);
    S->rows_recomputed = 0;
    if (rows) {
        const auto t2 = Clock::now();
        if (R >= 0) {
            const int32_t kr = std::min(R, g->kmax);
            compute_rows(c, g, var, kr);
            S->rows_recomputed = kr + 1;
        }
        S->rows_ms = ms_since(t2);
    }
}

// ---- localized maintenance ---------------------------------------------

// Net effect of validated edits applied in order (per request: deletions,
// then insertions): edges present before and absent after (del), and
// absent before and present after (ins).  An edge removed and put back, or
// put in and removed again, is in neither.
struct NetEdits {
    std::vector<uint32_t> del, ins;
};

struct EditLog {
    std::unordered_map<uint64_t, uint8_t> first_last;  // bit 0: first op inserts, bit 1: last op inserts
    void add(const uint32_t* p, uint32_t n, bool insert) {
        for (uint32_t i = 0; i < n; ++i) {
            const uint64_t key = (uint64_t)p[2 * i] << 32 | p[2 * i + 1];
            auto it = first_last.find(key);
            if (it == first_last.end())
                first_last.emplace(key, insert ? 3 : 0);
            else
                it->second = (uint8_t)((it->second & 1) | (insert ? 2 : 0));
        }
    }
    NetEdits net() const {
        std::vector<uint64_t> d, ins;
        for (const auto& [key, fl] : first_last) {
            if (fl == 0) d.push_back(key);        // deleted, not re-inserted
            else if (fl == 3) ins.push_back(key); // inserted, not deleted again
        }
        sort_u64(d);
        sort_u64(ins);
        NetEdits out;
        for (uint64_t k : d) {
            out.del.push_back((uint32_t)(k >> 32));
            out.del.push_back((uint32_t)k);
        }
        for (uint64_t k : ins) {
            out.ins.push_back((uint32_t)(k >> 32));
            out.ins.push_back((uint32_t)k);
        }
        return out;
    }
};

// Whether this batch goes to local maintenance.  In the budgeted mode
// (maintenance = 0) a batch given up sends the next local_backoff batches
// straight to recomputation, the backoff doubling (up to 64) with each
// further batch given up and cleared by one kept: a stream whose batches
// all exceed the budget then costs about 1 + 1.25/64 recomputations per
// batch, one whose batches are cheap stays local.
bool local_enabled(const dcx_ctx* c, dcx_graph* g) {
    if (c->cfg.maintenance == 1 || !g->kmax_valid || !g->rows_valid || g->kmax < 0 ||
        g->rows_k != g->kmax + 1 || g->V == 0)
        return false;
    if (c->cfg.maintenance == 0 && g->local_skip > 0) {
        --g->local_skip;
        return false;
    }
    return true;
}

// Updates the backoff after local maintenance returned `path`.
void local_result(const dcx_ctx* c, dcx_graph* g, int path) {
    if (c->cfg.maintenance != 0) return;
    if (path == 1) {
        g->local_backoff = 0;
    } else {
        g->local_backoff = g->local_backoff ? std::min<uint32_t>(2 * g->local_backoff, 64) : 1;
        g->local_skip = g->local_backoff;
    }
}

// Maintains kmaxes and rows after the edits N (already applied to the
// graph) by local tasks (local_kernels.hip).  Returns 1, or 2 when a task's
// worklist overflowed and everything was recomputed instead.
int local_maintain(dcx_ctx* c, dcx_graph* g, const NetEdits& N, dcx_update_stats* S) {
    const auto t0 = Clock::now();
    const uint32_t V = g->V;
    const uint32_t nD = (uint32_t)(N.del.size() / 2), nI = (uint32_t)(N.ins.size() / 2);
    const int32_t K = g->kmax;
    hipStream_t st = c->stream;
    const Staged L = layout_batch(N.ins.data(), nI, N.del.data(), nD);
    ensure_staging(c, L.words);
    fill_batch(L, c->upd_host, N.ins.data(), N.del.data());
    uint32_t* Dv = c->upd_dev;
    if (L.in_words)
        HIPX(c, hipMemcpyAsync(Dv, c->upd_host, 4 * L.in_words, hipMemcpyHostToDevice, st));
    auto dptr = [&](size_t idx) { return Dv + L.at[idx]; };
    DevGroups dg[4];
    for (int q = 0; q < 4; ++q)
        dg[q] = DevGroups{dptr(L.i_grp[q][0]), dptr(L.i_grp[q][1]), dptr(L.i_grp[q][2]),
                          dptr(L.i_grp[q][3]), (uint32_t)L.G[q].key.size()};
    // worklists: one slot per resident workgroup
    const uint32_t qcap = (uint32_t)std::min<uint64_t>(1u << 18, 2ull * V + 64);
    const uint32_t slots = (uint32_t)std::max(1, 2 * c->num_cus);
    ensure_bytes(c, reinterpret_cast<void**>(&c->local_scratch), &c->local_scratch_bytes,
                 (size_t)slots * local_slot_words(qcap) * 4);
    if (!c->local_words) dev_alloc(c, &c->local_words, 16 + 2 * 20);
    if (!g->ins_cnt_in) {
        dev_alloc(c, &g->ins_cnt_in, V);
        dev_alloc(c, &g->ins_cnt_out, V);
        HIPX(c, hipMemsetAsync(g->ins_cnt_in, 0, 4ull * V, st));
        HIPX(c, hipMemsetAsync(g->ins_cnt_out, 0, 4ull * V, st));
    }
    if (!g->changed) dev_alloc(c, &g->changed, 2ull * V);
    if (c->local_dm_cap < std::max<size_t>(nD, 1)) {
        dev_free(c->local_dm);
        c->local_dm_cap = 0;
        dev_alloc(c, &c->local_dm, std::max<size_t>(nD, 1024));
        c->local_dm_cap = std::max<size_t>(nD, 1024);
    }
    uint32_t* W = c->local_words;
    HIPX(c, hipMemsetAsync(W, 0, (16 + 2 * 20) * 4, st));
    HIPX(c, launch_edge_min_kmax(g->kmaxes, dptr(L.i_del), nD, c->local_dm, st));
    // the insertions come out again: deletions are maintained on G - D
    if (nI) {
        HIPX(c, launch_apply_deletes(g->out_off, g->out_deg, g->out_nbr, dg[2].key, dg[2].start,
                                     dg[2].cnt, dg[2].val, dg[2].n, W + 5, st));
        HIPX(c, launch_apply_deletes(g->in_off, g->in_deg, g->in_nbr, dg[3].key, dg[3].start,
                                     dg[3].cnt, dg[3].val, dg[3].n, W + 5, st));
    }
    LocalArgs A{};
    A.V = V;
    A.out = Csr{g->out_off, g->out_deg, g->out_nbr};
    A.in = Csr{g->in_off, g->in_deg, g->in_nbr};
    A.kmaxes = g->kmaxes;
    A.rows = g->rows;
    A.task_ctr = W;
    A.n_changed = W + 1;
    A.fail = W + 2;
    A.scratch = c->local_scratch;
    A.qcap = qcap;
    A.del_pairs = dptr(L.i_del);
    A.n_del = nD;
    A.del_m = c->local_dm;
    A.del_targets = dg[1].key;
    A.n_del_targets = dg[1].n;
    A.changed = g->changed;
    A.changed_cap = V;
    A.ins_pairs = dptr(L.i_ins);
    A.n_ins = nI;
    A.gin_key = dg[3].key;
    A.gin_cnt = dg[3].cnt;
    A.n_gin = dg[3].n;
    A.gout_key = dg[2].key;
    A.gout_cnt = dg[2].cnt;
    A.n_gout = dg[2].n;
    A.ins_cnt_in = g->ins_cnt_in;
    A.ins_cnt_out = g->ins_cnt_out;
    A.counters = reinterpret_cast<unsigned long long*>(W + 16);
    // budget: local_budget list entries per edge when set; else a quarter
    // of the last full recomputation's time (at least 0.2 ms), so that a
    // batch given up costs at most 1.25 recomputations; else (no time
    // known yet) 12 entries per edge, ~1/8 of a recomputation's list reads
    A.t_start = reinterpret_cast<unsigned long long*>(W + 8);
    if (c->cfg.maintenance == 2) {
    } else if (c->cfg.local_budget > 0) {
        A.budget = (unsigned long long)c->cfg.local_budget * g->E;
    } else if (g->full_ms > 0) {
        A.ticks = (unsigned long long)(std::max(0.25 * g->full_ms, 0.2) * c->wall_khz);
    } else {
        A.budget = std::max<unsigned long long>(12ull * g->E, 1ull << 20);
    }
    // task lists go through a pinned buffer, each launch's at its own offset
    // (no copy is overwritten before it ran)
    std::vector<LocalTaskDesc> tasks;
    size_t toff = 0;
    auto run = [&]() {
        if (tasks.empty()) return;
        const size_t need = toff + tasks.size();
        if (c->local_tasks_cap < need) {
            sync(c);
            dev_free(c->local_tasks);
            if (c->local_tasks_host) (void)hipHostFree(c->local_tasks_host);
            c->local_tasks_host = nullptr;
            c->local_tasks_cap = 0;
            const size_t cap = std::max<size_t>(2 * need, 4096);
            dev_alloc(c, &c->local_tasks, cap);
            HIPX(c, hipHostMalloc(reinterpret_cast<void**>(&c->local_tasks_host),
                                  cap * sizeof(LocalTaskDesc), hipHostMallocDefault));
            c->local_tasks_cap = cap;
        }
        std::memcpy(c->local_tasks_host + toff, tasks.data(), tasks.size() * sizeof(LocalTaskDesc));
        HIPX(c, hipMemcpyAsync(c->local_tasks + toff, c->local_tasks_host + toff,
                               tasks.size() * sizeof(LocalTaskDesc), hipMemcpyHostToDevice, st));
        HIPX(c, hipMemsetAsync(W, 0, 4, st));  // task counter
        A.tasks = c->local_tasks + toff;
        A.ntasks = (uint32_t)tasks.size();
        A.rows = g->rows;
        HIPX(c, launch_local_tasks(A, std::min<uint32_t>(slots, A.ntasks), st));
        toff = need;
        tasks.clear();
    };
    if (nD) {
        tasks.push_back({kLocalKmaxDel, 0});
        run();
        for (int32_t k = 0; k <= K; ++k) tasks.push_back({kLocalRowDel, k});
        run();
    }
    if (nI) {
        HIPX(c, launch_apply_inserts(g->out_off, g->out_deg, g->out_nbr, dg[2].key, dg[2].start,
                                     dg[2].cnt, dg[2].val, dg[2].n, st));
        HIPX(c, launch_apply_inserts(g->in_off, g->in_deg, g->in_nbr, dg[3].key, dg[3].start,
                                     dg[3].cnt, dg[3].val, dg[3].n, st));
        HIPX(c, launch_set_counts(g->ins_cnt_in, dg[3].key, dg[3].cnt, dg[3].n, true, st));
        HIPX(c, launch_set_counts(g->ins_cnt_out, dg[2].key, dg[2].cnt, dg[2].n, true, st));
        tasks.push_back({kLocalCoreIns, 0});
        tasks.push_back({kLocalCoreIns, 1});
        run();
        HIPX(c, launch_set_counts(g->ins_cnt_in, dg[3].key, dg[3].cnt, dg[3].n, false, st));
        HIPX(c, launch_set_counts(g->ins_cnt_out, dg[2].key, dg[2].cnt, dg[2].n, false, st));
    }
    HIPX(c, launch_max_masked(g->kmaxes, V, reinterpret_cast<int32_t*>(W + 3), st));
    uint32_t* hw = reinterpret_cast<uint32_t*>(c->h_scalars + 40);
    HIPX(c, hipMemcpyAsync(hw, W, 8 * 4, hipMemcpyDeviceToHost, st));
    sync(c);
    S->kmax_ms = ms_since(t0);
    auto fallback = [&]() {
        const int var = c->cfg.variant == DCX_VARIANT_AUTO ? DCX_VARIANT_BATCHED : c->cfg.variant;
        if (hw[5]) {
            g->broken = true;
            g->kmax_valid = g->rows_valid = false;
            set_err(c, "local maintenance: %u edges missing while taking insertions out", hw[5]);
            throw Fail{DCX_ERR_INTERNAL};
        }
        const auto tf = Clock::now();
        compute_kmax(c, g, var);
        compute_rows(c, g, var);
        sync(c);
        g->full_ms = ms_since(tf);
        S->rows_recomputed = g->kmax + 1;
        return 2;
    };
    if (hw[2] || hw[5]) return fallback();
    const int32_t K2 = (int32_t)hw[3];
    if (nI) {
        if (K2 > K) {
            const size_t nrows = (size_t)K2 + 1;
            grow_keep(c, &g->rows, &g->rows_cap, nrows * V, (size_t)(K + 1) * V);
            grow_keep(c, &g->lmax_rows, &g->lmax_rows_cap, nrows, (size_t)K + 1);
            HIPX(c, hipMemsetAsync(g->rows + (size_t)(K + 1) * V, 0,
                                   (size_t)(K2 - K) * V * sizeof(int32_t), st));
        }
        // rows 1..K2, one task each, bounded by the new row 0; the largest
        // rows (lowest k) start first
        for (int32_t k = 1; k <= K2; ++k) tasks.push_back({kLocalRowIns, k});
        run();
    }
    HIPX(c, launch_max_masked(g->rows, V, reinterpret_cast<int32_t*>(W + 4), st));
    HIPX(c, hipMemcpyAsync(hw, W, 8 * 4, hipMemcpyDeviceToHost, st));
    sync(c);
    if (hw[2]) return fallback();
    HIPX(c, hipMemcpy(c->local_counters, W + 16, sizeof c->local_counters, hipMemcpyDeviceToHost));
    g->kmax = K2;
    g->lmax = (int32_t)hw[4];
    g->rows_k = K2 + 1;
    g->order_valid = false;
    g->core_valid = false;
    g->row0_ready = false;
    S->rows_recomputed = 0;
    S->rows_ms = ms_since(t0) - S->kmax_ms;
    return 1;
}

dcx_status do_update(dcx_ctx* c, dcx_graph* g, const uint32_t* ins, uint32_t n_ins,
                     const uint32_t* del, uint32_t n_del, bool maintain_values,
                     dcx_update_stats* st) {
    const auto t0 = Clock::now();
    dcx_update_stats S{};
    const uint32_t V = g->V;
    dcx_status s = validate_pairs(c, V, del, n_del, "delete");
    if (s != DCX_OK) return s;
    s = validate_pairs(c, V, ins, n_ins, "insert");
    if (s != DCX_OK) return s;
    if ((uint64_t)n_ins > g->E + (uint64_t)n_del + 0xffffffffull) return DCX_ERR_CAPACITY;
    const Staged L = layout_batch(ins, n_ins, del, n_del);
    ensure_staging(c, L.words);
    fill_batch(L, c->upd_host, ins, del);
    S.kmax_before = g->kmax_valid ? g->kmax : -1;
    int32_t m_del = -1;
    s = edit_batch(c, g, L, c->upd_host, c->upd_dev, false, ins, del, &m_del, &S.relocated);
    if (s != DCX_OK) return s;
    S.inserted = n_ins;
    S.deleted = n_del;
    S.insert_ms = ms_since(t0);
    S.m_del = m_del;
    S.m_ins = -1;
    if (!(maintain_values && g->kmax_valid)) {
        g->kmax_valid = false;
        g->rows_valid = false;
    } else if (local_enabled(c, g)) {
        EditLog log;
        log.add(del, n_del, false);
        log.add(ins, n_ins, true);
        const NetEdits N = log.net();
        S.net_deleted = (int32_t)(N.del.size() / 2);
        S.net_inserted = (int32_t)(N.ins.size() / 2);
        S.path = local_maintain(c, g, N, &S);
        local_result(c, g, S.path);
    } else {
        maintain(c, g, ins, n_ins, m_del, n_ins + n_del > 0, n_ins == 0, &S);
    }
    S.kmax_after = g->kmax_valid ? g->kmax : -1;
    S.total_ms = ms_since(t0);
    if (st) *st = S;
    return DCX_OK;
}

}  // namespace

extern "C" {

dcx_status dcx_update(dcx_ctx* ctx, dcx_graph* g, const uint32_t* ins, uint32_t n_ins,
                      const uint32_t* del, uint32_t n_del, dcx_update_stats* stats) {
    dcx_status s = check_graph(ctx, g);
    if (s != DCX_OK) return s;
    if ((n_ins && !ins) || (n_del && !del)) return DCX_ERR_INVALID_ARG;
    return guarded(ctx, [&]() { return do_update(ctx, g, ins, n_ins, del, n_del, true, stats); });
}

dcx_status dcx_last_maintenance_counters(const dcx_ctx* ctx, uint64_t* out, uint32_t n) {
    if (!ctx || !out || n < 20) return DCX_ERR_INVALID_ARG;
    for (int i = 0; i < 20; ++i) out[i] = ctx->local_counters[i];
    return DCX_OK;
}

dcx_status dcx_kmax_edge_batches_device(dcx_ctx* ctx, dcx_graph* g, const uint32_t* pairs,
                                        uint32_t n, int32_t mode, uint32_t* batch_of_edge,
                                        uint32_t* n_batches) {
    dcx_status s = check_graph(ctx, g);
    if (s != DCX_OK) return s;
    if (!n_batches || (n && (!pairs || !batch_of_edge)) || (mode != 0 && mode != 1))
        return DCX_ERR_INVALID_ARG;
    return guarded(ctx, [&]() -> dcx_status {
        if (!g->kmax_valid) {
            set_err(ctx, "dcx_kmax_edge_batches_device: no resident kmaxes (decompose first)");
            return DCX_ERR_STATE;
        }
        for (uint32_t i = 0; i < 2 * n; ++i)
            if (pairs[i] >= g->V) {
                set_err(ctx, "edge %u: vertex id out of range (V=%u)", i / 2, g->V);
                return DCX_ERR_INVALID_ARG;
            }
        const uint32_t nbuckets = (uint32_t)std::max(g->kmax, 0) + 1;
        std::vector<void*> tmp;
        struct Free {
            std::vector<void*>& t;
            ~Free() {
                for (void* p : t) (void)hipFree(p);
            }
        } free_tmp{tmp};
        uint32_t *d_pairs = nullptr, *d_batch = nullptr, *d_nb = nullptr;
        void* work = nullptr;
        dev_alloc(ctx, &d_pairs, 2ull * std::max<uint32_t>(n, 1));
        tmp.push_back(d_pairs);
        dev_alloc(ctx, &d_batch, std::max<uint32_t>(n, 1));
        tmp.push_back(d_batch);
        dev_alloc(ctx, &d_nb, 1);
        tmp.push_back(d_nb);
        HIPX(ctx, hipMalloc(&work, edge_batch_tmp_bytes(n, nbuckets)));
        tmp.push_back(work);
        if (n)
            HIPX(ctx, hipMemcpyAsync(d_pairs, pairs, 8ull * n, hipMemcpyHostToDevice, ctx->stream));
        HIPX(ctx, launch_edge_batches(g->kmaxes, g->V, d_pairs, n, mode, nbuckets, d_batch, d_nb,
                                      work, ctx->stream));
        if (n)
            HIPX(ctx, hipMemcpyAsync(batch_of_edge, d_batch, 4ull * n, hipMemcpyDeviceToHost,
                                     ctx->stream));
        HIPX(ctx, hipMemcpyAsync(n_batches, d_nb, 4, hipMemcpyDeviceToHost, ctx->stream));
        sync(ctx);
        return DCX_OK;
    });
}

dcx_status dcx_graph_insert_edges(dcx_ctx* ctx, dcx_graph* g, const uint32_t* pairs,
                                  uint32_t n) {
    dcx_status s = check_graph(ctx, g);
    if (s != DCX_OK) return s;
    if (n && !pairs) return DCX_ERR_INVALID_ARG;
    return guarded(ctx,
                   [&]() { return do_update(ctx, g, pairs, n, nullptr, 0, false, nullptr); });
}

dcx_status dcx_graph_delete_edges(dcx_ctx* ctx, dcx_graph* g, const uint32_t* pairs,
                                  uint32_t n) {
    dcx_status s = check_graph(ctx, g);
    if (s != DCX_OK) return s;
    if (n && !pairs) return DCX_ERR_INVALID_ARG;
    return guarded(ctx,
                   [&]() { return do_update(ctx, g, nullptr, 0, pairs, n, false, nullptr); });
}

}  // extern "C"
