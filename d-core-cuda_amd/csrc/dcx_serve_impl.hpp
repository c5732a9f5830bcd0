// Continuous-batching maintenance server: included at the end of
// dcx_api.hip (after the update path it reuses).
//
// The reference applies one batch at a time from its host loop
// (maintenanceFast, main.cu:1665-1751, driven by the experiment drivers in
// main.cu:1905-1924).  Here a server owns one device graph and takes a
// stream of update requests:
//   * submit (any client thread) checks the request's shape, lays it out
//     (groupings by source and target) and admits it as soon as its staging
//     fits in a fixed HBM pool and an in-flight slot is free; admission
//     copies it into a pinned host block and uploads it on a copy stream, so
//     uploads of newly arrived requests overlap the maintenance running on
//     the compute stream;
//   * the worker takes every admitted request (up to max_group, in arrival
//     order), applies each one's edits after the edits of the ones before
//     it -- each is checked against the graph as it then is, so a request's
//     status is the one it would get alone -- and then runs ONE maintenance
//     for the whole group: kmaxes re-peeled, rows k <= R recomputed with
//     R = max(M over the group's deletions with the kmaxes before the group,
//     M over its insertions with the kmaxes after).  That bound is valid for
//     any mix of edits: rows above M_del survive the deletions and rows
//     above M_ins survive the insertions that follow them;
//   * results carry the server-clock times of submit / admission / start /
//     completion and, on request, a hash of the whole decomposition after
//     the request's group.
#pragma once

#include <condition_variable>
#include <deque>
#include <string>
#include <thread>

namespace {

// Order-independent hash of kmaxes and rows 0..kmax: the 64-bit sum over
// all cells of mix64(v * G1 + (k + 1) * G2 + value), k = -1 for kmaxes.
constexpr uint64_t kHashG1 = 0x9E3779B97F4A7C15ull, kHashG2 = 0xC2B2AE3D27D4EB4Full;

__device__ __forceinline__ uint64_t hmix(uint64_t z) {
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
    return z ^ (z >> 31);
}

__global__ void k_state_hash(const int32_t* kmaxes, const int32_t* rows, uint32_t V, int32_t K,
                             unsigned long long* out) {
    const uint64_t cells = (uint64_t)V * (uint64_t)(K + 2);
    uint64_t acc = 0;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < cells;
         i += (uint64_t)gridDim.x * blockDim.x) {
        const uint64_t kk = i / V;  // 0: kmaxes, 1 + k: row k
        const uint64_t v = i - kk * V;
        const int32_t x = kk == 0 ? kmaxes[v] : rows[(kk - 1) * V + v];
        acc += hmix(v * kHashG1 + kk * kHashG2 + (uint64_t)(uint32_t)x);
    }
    for (int d = 32; d > 0; d >>= 1) acc += __shfl_xor(acc, d, 64);
    if ((threadIdx.x & 63) == 0) atomicAdd(out, (unsigned long long)acc);
}

uint64_t state_hash(dcx_ctx* c, dcx_graph* g) {
    if (!g->kmax_valid || !g->rows_valid || g->V == 0) return 0;
    unsigned long long* d = reinterpret_cast<unsigned long long*>(c->d_scalars + 16);
    HIPX(c, hipMemsetAsync(d, 0, 8, c->stream));
    const uint64_t cells = (uint64_t)g->V * (uint64_t)(g->kmax + 2);
    const unsigned blocks = (unsigned)std::min<uint64_t>((cells + 255) / 256, 8192);
    hipLaunchKernelGGL(k_state_hash, dim3(blocks), dim3(256), 0, c->stream, g->kmaxes, g->rows,
                       g->V, g->kmax, d);
    HIPX(c, hipGetLastError());
    HIPX(c, hipMemcpyAsync(c->h_scalars + 16, d, 8, hipMemcpyDeviceToHost, c->stream));
    sync(c);
    uint64_t h = 0;
    std::memcpy(&h, c->h_scalars + 16, 8);
    return h;
}

// out[0] += cells where a and b differ, out[1] = min over them of the cell
// index (left at ~0 when none differs).
__global__ void k_compare_i32(const int32_t* a, const int32_t* b, uint64_t n,
                              unsigned long long* out) {
    unsigned long long cnt = 0, first = ~0ull;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (uint64_t)gridDim.x * blockDim.x) {
        if (a[i] != b[i]) {
            ++cnt;
            first = first < i ? first : i;
        }
    }
    for (int d = 32; d > 0; d >>= 1) {
        cnt += __shfl_xor(cnt, d, 64);
        const unsigned long long f = __shfl_xor(first, d, 64);
        first = first < f ? first : f;
    }
    if ((threadIdx.x & 63) == 0 && cnt) {
        atomicAdd(out, cnt);
        atomicMin(out + 1, first);
    }
}

// Mismatching cells of a[0..n) and b[0..n) (device arrays) and the first of
// them (~0 if none); `base` is added to the first index.
void compare_cells(dcx_ctx* c, const int32_t* a, const int32_t* b, uint64_t n, uint64_t base,
                   uint64_t* count, uint64_t* first) {
    if (!n) return;
    unsigned long long* d = reinterpret_cast<unsigned long long*>(c->d_scalars + 16);
    HIPX(c, hipMemsetAsync(d, 0, 8, c->stream));
    HIPX(c, hipMemsetAsync(d + 1, 0xff, 8, c->stream));
    const unsigned blocks = (unsigned)std::min<uint64_t>((n + 255) / 256, 8192);
    hipLaunchKernelGGL(k_compare_i32, dim3(blocks), dim3(256), 0, c->stream, a, b, n, d);
    HIPX(c, hipGetLastError());
    HIPX(c, hipMemcpyAsync(c->h_scalars + 16, d, 16, hipMemcpyDeviceToHost, c->stream));
    sync(c);
    uint64_t r[2];
    std::memcpy(r, c->h_scalars + 16, 16);
    *count += r[0];
    if (r[0] && r[1] + base < *first) *first = r[1] + base;
}

struct ServeReq {
    uint64_t id = 0;
    std::vector<uint32_t> ins, del;
    Staged L;
    size_t off = 0;
    hipEvent_t up = nullptr;
    double t_submit = 0, t_admit = 0;
};

}  // namespace

struct dcx_server {
    dcx_ctx* c = nullptr;
    dcx_graph* g = nullptr;
    dcx_serve_config cfg{};
    hipStream_t copy = nullptr;
    uint32_t* dpool = nullptr;
    uint32_t* hpool = nullptr;
    size_t pool_words = 0, head = 0;
    struct Block {
        size_t off, words;
    };
    std::deque<Block> live;  // pool blocks of admitted requests, oldest first
    std::deque<ServeReq*> waiting, admitted;
    size_t in_worker = 0;
    std::deque<dcx_serve_result> done;
    std::mutex mu;
    std::condition_variable cv_work, cv_done;
    std::thread worker;
    bool stopping = false;
    uint64_t next_id = 1;
    uint32_t groups = 0;
    dcx_serve_stats st{};
    Clock::time_point t0;
    dcx_status failed = DCX_OK;  // first failure; no device work after it
    std::string fail_text;

    double now() const { return std::chrono::duration<double, std::milli>(Clock::now() - t0).count(); }

    // FIFO ring allocation in the pool (blocks are freed oldest first).
    bool ring_alloc(size_t w, size_t* off) {
        if (w > pool_words) return false;
        if (live.empty()) head = 0;
        const size_t oldest = live.empty() ? 0 : live.front().off;
        const bool wrapped = !live.empty() && live.back().off < oldest;
        size_t at;
        if (!wrapped) {
            if (head + w <= pool_words)
                at = head;
            else if (live.empty() || w <= oldest)
                at = 0;
            else
                return false;
        } else {
            if (head + w <= oldest)
                at = head;
            else
                return false;
        }
        head = at + w;
        live.push_back(Block{at, w});
        *off = at;
        return true;
    }

    // Admits r if a slot and pool room are free (mu held).
    bool try_admit(ServeReq* r) {
        if (admitted.size() + in_worker >= cfg.max_inflight) return false;
        size_t off = 0;
        if (!ring_alloc(r->L.words, &off)) return false;
        r->off = off;
        fill_batch(r->L, hpool + off, r->ins.data(), r->del.data());
        if (hipEventCreateWithFlags(&r->up, hipEventDisableTiming) != hipSuccess ||
            (r->L.in_words &&
             hipMemcpyAsync(dpool + off, hpool + off, 4 * r->L.in_words, hipMemcpyHostToDevice,
                            copy) != hipSuccess) ||
            hipEventRecord(r->up, copy) != hipSuccess) {
            failed = DCX_ERR_HIP;
        }
        r->t_admit = now();
        admitted.push_back(r);
        return true;
    }

    void finish(ServeReq* r, dcx_status s, uint32_t group, const dcx_update_stats* u,
                double t_start, double t_done, uint64_t hash) {
        dcx_serve_result x{};
        x.id = r->id;
        x.status = s;
        x.group = group;
        x.inserted = s == DCX_OK ? (uint32_t)(r->ins.size() / 2) : 0;
        x.deleted = s == DCX_OK ? (uint32_t)(r->del.size() / 2) : 0;
        x.kmax_after = u ? u->kmax_after : -1;
        x.rows_recomputed = u ? u->rows_recomputed : 0;
        x.submit_ms = r->t_submit;
        x.admit_ms = r->t_admit;
        x.start_ms = t_start;
        x.done_ms = t_done;
        x.state_hash = hash;
        done.push_back(x);
    }

    void run() {
        (void)hipSetDevice(c->device);
        for (;;) {
            std::vector<ServeReq*> grp;
            {
                std::unique_lock<std::mutex> lk(mu);
                while (!waiting.empty() && try_admit(waiting.front())) waiting.pop_front();
                cv_work.wait(lk, [&] { return stopping || !admitted.empty(); });
                if (admitted.empty()) {
                    // stopping and drained (with nothing in flight every
                    // waiting request fits, so `waiting` is empty here; a
                    // request left over anyway is failed, not dropped)
                    for (ServeReq* r : waiting) {
                        finish(r, DCX_ERR_STATE, 0, nullptr, now(), now(), 0);
                        delete r;
                    }
                    waiting.clear();
                    return;
                }
                uint64_t edges = 0;
                while (!admitted.empty() && grp.size() < cfg.max_group) {
                    ServeReq* r = admitted.front();
                    const uint64_t e = (r->ins.size() + r->del.size()) / 2;
                    if (!grp.empty() && edges + e > cfg.max_group_edges) break;
                    edges += e;
                    grp.push_back(r);
                    admitted.pop_front();
                }
                in_worker = grp.size();
            }
            const uint32_t gid = ++groups;
            const double t_start = now();
            std::vector<dcx_status> sts(grp.size(), DCX_OK);
            dcx_update_stats U{};
            uint64_t hash = 0;
            dcx_status prior;
            {
                std::lock_guard<std::mutex> lk(mu);
                prior = failed;
            }
            // after a failure the graph is not trusted: the rest of the
            // requests complete with that status and touch nothing
            const dcx_status gs = prior != DCX_OK ? prior : guarded(c, [&]() -> dcx_status {
                U.kmax_before = g->kmax_valid ? g->kmax : -1;
                // 1. which of the group's pairs the graph holds before the
                //    group: one find pass per request, one wait for all
                for (size_t i = 0; i < grp.size(); ++i) {
                    ServeReq* r = grp[i];
                    const Staged& L = r->L;
                    uint32_t* D = dpool + r->off;
                    HIPX(c, hipStreamWaitEvent(c->stream, r->up, 0));
                    if (L.words > L.in_words)
                        HIPX(c, hipMemsetAsync(D + L.in_words, 0, 4 * (L.words - L.in_words),
                                               c->stream));
                    const Csr so{g->out_off, g->out_deg, g->out_nbr};
                    const Csr si{g->in_off, g->in_deg, g->in_nbr};
                    HIPX(c, launch_find_edges(so, si, D + L.at[L.i_del], L.n_del,
                                              D + L.at[L.i_fdel], c->stream));
                    HIPX(c, launch_find_edges(so, si, D + L.at[L.i_ins], L.n_ins,
                                              D + L.at[L.i_fins], c->stream));
                    if (L.words > L.in_words)
                        HIPX(c, hipMemcpyAsync(hpool + r->off + L.in_words, D + L.in_words,
                                               4 * (L.words - L.in_words), hipMemcpyDeviceToHost,
                                               c->stream));
                }
                sync(c);
                // 2. each request in order against the graph as the requests
                //    before it left it (as edit_batch would check it alone):
                //    its deletions must be present, its insertions absent
                //    unless it deletes them first; a request that fails
                //    changes nothing
                std::unordered_map<uint64_t, bool> now;  // pairs the accepted requests touched
                auto present = [&](uint64_t key, uint32_t before) {
                    const auto it = now.find(key);
                    return it != now.end() ? it->second : before != 0;
                };
                EditLog log;  // the accepted requests' edits in order
                for (size_t i = 0; i < grp.size(); ++i) {
                    ServeReq* r = grp[i];
                    const Staged& L = r->L;
                    const uint32_t* H = hpool + r->off;
                    const uint32_t* del = r->del.data();
                    const uint32_t* ins = r->ins.data();
                    dcx_status st = DCX_OK;
                    for (uint32_t j = 0; j < L.n_del && st == DCX_OK; ++j)
                        if (!present((uint64_t)del[2 * j] << 32 | del[2 * j + 1],
                                     H[L.at[L.i_fdel] + j]))
                            st = DCX_ERR_NOT_FOUND;
                    if (st == DCX_OK && L.n_ins) {
                        std::unordered_set<uint64_t> deleted;
                        for (uint32_t j = 0; j < L.n_del; ++j)
                            deleted.insert((uint64_t)del[2 * j] << 32 | del[2 * j + 1]);
                        for (uint32_t j = 0; j < L.n_ins && st == DCX_OK; ++j) {
                            const uint64_t key = (uint64_t)ins[2 * j] << 32 | ins[2 * j + 1];
                            if (present(key, H[L.at[L.i_fins] + j]) && !deleted.count(key))
                                st = DCX_ERR_INVALID_ARG;
                        }
                    }
                    sts[i] = st;
                    if (st != DCX_OK) continue;
                    for (uint32_t j = 0; j < L.n_del; ++j)
                        now[(uint64_t)del[2 * j] << 32 | del[2 * j + 1]] = false;
                    for (uint32_t j = 0; j < L.n_ins; ++j)
                        now[(uint64_t)ins[2 * j] << 32 | ins[2 * j + 1]] = true;
                    log.add(del, L.n_del, false);
                    log.add(ins, L.n_ins, true);
                }
                // 3. their net effect as one edit (valid by construction)
                const NetEdits N = log.net();
                const uint32_t n_ins = (uint32_t)(N.ins.size() / 2);
                const uint32_t n_del = (uint32_t)(N.del.size() / 2);
                const bool changed = n_ins + n_del > 0;
                int32_t m_del = -1;
                if (changed) {
                    const Staged L = layout_batch(N.ins.data(), n_ins, N.del.data(), n_del);
                    ensure_staging(c, L.words);
                    fill_batch(L, c->upd_host, N.ins.data(), N.del.data());
                    uint32_t rel = 0;
                    const dcx_status es = edit_batch(c, g, L, c->upd_host, c->upd_dev, false,
                                                     N.ins.data(), N.del.data(), &m_del, &rel);
                    if (es != DCX_OK) throw Fail{es};
                    U.relocated += rel;
                }
                U.net_deleted = (int32_t)n_del;
                U.net_inserted = (int32_t)n_ins;
                if (g->kmax_valid && changed && local_enabled(c, g)) {
                    U.path = local_maintain(c, g, N, &U);
                    local_result(c, g, U.path);
                } else if (g->kmax_valid) {
                    maintain(c, g, N.ins.data(), n_ins, m_del, changed, n_ins == 0, &U);
                }
                U.kmax_after = g->kmax_valid ? g->kmax : -1;
                if (cfg.hash_states) hash = state_hash(c, g);
                return DCX_OK;
            });
            const double t_done = now();
            std::lock_guard<std::mutex> lk(mu);
            if (gs != DCX_OK && failed == DCX_OK) {
                failed = gs;
                fail_text = c->err;
            }
            for (size_t i = 0; i < grp.size(); ++i) {
                finish(grp[i], gs != DCX_OK ? gs : sts[i], gid, &U, t_start, t_done, hash);
                (void)hipEventDestroy(grp[i]->up);
                live.pop_front();
                delete grp[i];
            }
            in_worker = 0;
            st.groups++;
            st.requests += grp.size();
            st.max_group = std::max<uint32_t>(st.max_group, (uint32_t)grp.size());
            st.busy_ms += t_done - t_start;
            while (!waiting.empty() && try_admit(waiting.front())) waiting.pop_front();
            cv_done.notify_all();
        }
    }
};

extern "C" {

dcx_status dcx_serve_config_init(dcx_serve_config* cfg) {
    if (!cfg) return DCX_ERR_INVALID_ARG;
    std::memset(cfg, 0, sizeof *cfg);
    cfg->hbm_budget = 64ull << 20;
    cfg->max_inflight = 4096;
    cfg->max_group = 256;
    cfg->max_group_edges = 1u << 16;
    return DCX_OK;
}

dcx_status dcx_state_hash(dcx_ctx* ctx, dcx_graph* g, uint64_t* hash) {
    dcx_status s = check_graph(ctx, g);
    if (s != DCX_OK) return s;
    if (!hash) return DCX_ERR_INVALID_ARG;
    if (!g->kmax_valid || !g->rows_valid) return DCX_ERR_STATE;
    return guarded(ctx, [&]() -> dcx_status {
        *hash = state_hash(ctx, g);
        return DCX_OK;
    });
}

dcx_status dcx_compare_i32(dcx_ctx* ctx, const int32_t* a, const int32_t* b, uint64_t n,
                           uint64_t* mismatches, uint64_t* first) {
    if (!ctx || !mismatches || (n && (!a || !b))) return DCX_ERR_INVALID_ARG;
    return guarded(ctx, [&]() -> dcx_status {
        *mismatches = 0;
        uint64_t f = ~0ull;
        compare_cells(ctx, a, b, n, 0, mismatches, &f);
        if (first) *first = f;
        return DCX_OK;
    });
}

dcx_status dcx_compare_decompositions(dcx_ctx* ctx, dcx_graph* a, dcx_graph* b,
                                      uint64_t* mismatches, uint64_t* first) {
    dcx_status s = check_graph(ctx, a);
    if (s != DCX_OK) return s;
    s = check_graph(ctx, b);
    if (s != DCX_OK) return s;
    if (!mismatches) return DCX_ERR_INVALID_ARG;
    if (a->V != b->V) {
        set_err(ctx, "dcx_compare_decompositions: %u vertices against %u", a->V, b->V);
        return DCX_ERR_INVALID_ARG;
    }
    if (!a->kmax_valid || !a->rows_valid || !b->kmax_valid || !b->rows_valid)
        return DCX_ERR_STATE;
    return guarded(ctx, [&]() -> dcx_status {
        const uint64_t V = a->V;
        uint64_t cnt = 0, f = ~0ull;
        compare_cells(ctx, a->kmaxes, b->kmaxes, V, 0, &cnt, &f);
        // rows 0..min(kmax): cell (k, v) is index V * (1 + k) + v
        const int32_t kc = std::min(a->kmax, b->kmax);
        if (kc >= 0)
            compare_cells(ctx, a->rows, b->rows, (uint64_t)(kc + 1) * V, V, &cnt, &f);
        // rows only one side has: every nonzero cell there differs
        const dcx_graph* hi = a->kmax > b->kmax ? a : b;
        for (int32_t k = kc + 1; k <= hi->kmax; ++k) {
            if (!ctx->zeros || ctx->zeros_n < V) {
                dev_free(ctx->zeros);
                ctx->zeros_n = 0;
                dev_alloc(ctx, &ctx->zeros, (size_t)V);
                ctx->zeros_n = V;
                HIPX(ctx, hipMemsetAsync(ctx->zeros, 0, 4 * V, ctx->stream));
            }
            compare_cells(ctx, hi->rows + (size_t)k * V, ctx->zeros, V, V * (1 + (uint64_t)k),
                          &cnt, &f);
        }
        *mismatches = cnt;
        if (first) *first = f;
        return DCX_OK;
    });
}

dcx_status dcx_serve_start(dcx_ctx* ctx, dcx_graph* g, const dcx_serve_config* cfg,
                           dcx_server** out) {
    dcx_status s = check_graph(ctx, g);
    if (s != DCX_OK) return s;
    if (!out) return DCX_ERR_INVALID_ARG;
    *out = nullptr;
    dcx_serve_config def;
    dcx_serve_config_init(&def);
    if (!cfg) cfg = &def;
    if (!g->kmax_valid || !g->rows_valid) {
        set_err(ctx, "dcx_serve_start: decompose the graph first");
        return DCX_ERR_STATE;
    }
    std::unique_ptr<dcx_server> sv(new (std::nothrow) dcx_server);
    if (!sv) return DCX_ERR_OUT_OF_MEMORY;
    sv->c = ctx;
    sv->g = g;
    sv->cfg = *cfg;
    if (!sv->cfg.hbm_budget) sv->cfg.hbm_budget = def.hbm_budget;
    if (!sv->cfg.max_inflight) sv->cfg.max_inflight = def.max_inflight;
    if (!sv->cfg.max_group) sv->cfg.max_group = def.max_group;
    if (!sv->cfg.max_group_edges) sv->cfg.max_group_edges = def.max_group_edges;
    sv->pool_words = sv->cfg.hbm_budget / 4;
    s = guarded(ctx, [&]() -> dcx_status {
        HIPX(ctx, hipStreamCreateWithFlags(&sv->copy, hipStreamNonBlocking));
        dev_alloc(ctx, &sv->dpool, sv->pool_words);
        HIPX(ctx, hipHostMalloc(reinterpret_cast<void**>(&sv->hpool), sv->pool_words * 4,
                                hipHostMallocDefault));
        return DCX_OK;
    });
    if (s != DCX_OK) {
        if (sv->copy) (void)hipStreamDestroy(sv->copy);
        dev_free(sv->dpool);
        if (sv->hpool) (void)hipHostFree(sv->hpool);
        return s;
    }
    sv->t0 = Clock::now();
    g->served = true;
    dcx_server* raw = sv.get();
    sv->worker = std::thread([raw] { raw->run(); });
    *out = sv.release();
    return DCX_OK;
}

dcx_status dcx_serve_submit(dcx_server* sv, const uint32_t* ins, uint32_t n_ins,
                            const uint32_t* del, uint32_t n_del, uint64_t* id) {
    if (!sv || (n_ins && !ins) || (n_del && !del)) return DCX_ERR_INVALID_ARG;
    std::unique_ptr<ServeReq> r(new (std::nothrow) ServeReq);
    if (!r) return DCX_ERR_OUT_OF_MEMORY;
    const double t_submit = sv->now();
    r->ins.assign(ins, ins + 2ull * n_ins);
    r->del.assign(del, del + 2ull * n_del);
    // shape errors are known now; they complete at once without device work
    // (checked without the context, whose lock the worker holds while it
    // maintains a group)
    dcx_status shape = validate_pairs(nullptr, sv->g->V, del, n_del, "delete");
    if (shape == DCX_OK) shape = validate_pairs(nullptr, sv->g->V, ins, n_ins, "insert");
    if (shape == DCX_OK) r->L = layout_batch(ins, n_ins, del, n_del);
    if (hipSetDevice(sv->c->device) != hipSuccess) return DCX_ERR_NO_DEVICE;
    std::lock_guard<std::mutex> lk(sv->mu);
    if (sv->stopping) return DCX_ERR_STATE;
    if (sv->failed != DCX_OK) return sv->failed;
    r->id = sv->next_id++;
    r->t_submit = t_submit;
    if (id) *id = r->id;
    if (shape != DCX_OK) {
        r->t_admit = t_submit;
        sv->finish(r.get(), shape, 0, nullptr, t_submit, sv->now(), 0);
        sv->cv_done.notify_all();
        return DCX_OK;
    }
    if (4 * r->L.words > sv->cfg.hbm_budget) {
        // larger than the whole staging pool: it could never be admitted
        sv->next_id--;
        return DCX_ERR_BUDGET;
    }
    ServeReq* raw = r.release();
    if (!(sv->waiting.empty() && sv->try_admit(raw))) sv->waiting.push_back(raw);
    sv->st.max_waiting = std::max<uint32_t>(sv->st.max_waiting, (uint32_t)sv->waiting.size());
    sv->cv_work.notify_one();
    return DCX_OK;
}

dcx_status dcx_serve_poll(dcx_server* sv, dcx_serve_result* out, uint32_t cap, uint32_t* n,
                          int32_t wait_ms) {
    if (!sv || !n || (cap && !out)) return DCX_ERR_INVALID_ARG;
    std::unique_lock<std::mutex> lk(sv->mu);
    if (sv->done.empty() && wait_ms > 0)
        sv->cv_done.wait_for(lk, std::chrono::milliseconds(wait_ms),
                             [&] { return !sv->done.empty(); });
    uint32_t m = 0;
    while (m < cap && !sv->done.empty()) {
        out[m++] = sv->done.front();
        sv->done.pop_front();
    }
    *n = m;
    return sv->failed;
}

const char* dcx_serve_last_error(dcx_server* sv) {
    if (!sv) return "";
    std::lock_guard<std::mutex> lk(sv->mu);
    return sv->fail_text.c_str();
}

dcx_status dcx_serve_stop(dcx_server* sv, dcx_serve_stats* stats) {
    if (!sv) return DCX_ERR_INVALID_ARG;
    {
        std::lock_guard<std::mutex> lk(sv->mu);
        sv->stopping = true;
    }
    sv->cv_work.notify_all();
    if (sv->worker.joinable()) sv->worker.join();
    const dcx_status s = sv->failed;
    if (stats) {
        *stats = sv->st;
        stats->elapsed_ms = sv->now();
    }
    (void)hipSetDevice(sv->c->device);
    if (sv->copy) {
        (void)hipStreamSynchronize(sv->copy);
        (void)hipStreamDestroy(sv->copy);
    }
    dev_free(sv->dpool);
    if (sv->hpool) (void)hipHostFree(sv->hpool);
    for (ServeReq* r : sv->waiting) delete r;
    sv->g->served = false;
    delete sv;
    return s;
}

}  // extern "C"
