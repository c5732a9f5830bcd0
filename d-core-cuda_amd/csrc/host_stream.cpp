// Update-stream generator for the serving workload (host only, C-ABI).
//
// Batch b draws from its own SplitMix64 stream (stream_for(seed, b)): its
// size floor(exp(U * ln(max_batch))) clipped to [1, max_batch], then one
// Bernoulli(p_insert) per slot for insert-vs-delete, then the edges.  An
// insert is an R-MAT sample (same quadrant rule as dcx_gen_rmat) that is
// not a self-loop, not in the current graph and not already in the batch;
// a delete is a uniformly chosen current edge not already in the batch.
// The graph then becomes (current - deletes) + inserts, so every batch is
// valid for dcx_update against the graph the previous batches produced.
// The reference's experiment drivers sample insert/delete edges the same
// way one at a time (getRandomInsertEdge / getRandomDeleteEdge,
// src/graph.cpp:110-144).
#include <cmath>
#include <cstring>
#include <unordered_set>
#include <vector>

#include "dcx.h"
#include "rng.hpp"

namespace dcx {
namespace {

// Open-addressing index: edge key -> position in the edge vector.
struct EdgeIndex {
    static constexpr uint64_t kEmpty = 0, kTomb = ~0ull;  // both are self-loops
    std::vector<uint64_t> key;
    std::vector<uint32_t> pos;
    uint64_t mask = 0;

    void init(uint64_t n) {
        uint64_t cap = 16;
        while (cap < 2 * n + 16) cap <<= 1;
        key.assign(cap, kEmpty);
        pos.assign(cap, 0);
        mask = cap - 1;
    }
    uint64_t slot_of(uint64_t k) const {  // slot holding k, or ~0
        for (uint64_t h = mix64(k) & mask;; h = (h + 1) & mask) {
            if (key[h] == k) return h;
            if (key[h] == kEmpty) return ~0ull;
        }
    }
    bool contains(uint64_t k) const { return slot_of(k) != ~0ull; }
    void put(uint64_t k, uint32_t p) {
        for (uint64_t h = mix64(k) & mask;; h = (h + 1) & mask)
            if (key[h] == kEmpty || key[h] == kTomb) {
                key[h] = k;
                pos[h] = p;
                return;
            }
    }
    void set_pos(uint64_t k, uint32_t p) { pos[slot_of(k)] = p; }
    uint32_t get_pos(uint64_t k) const { return pos[slot_of(k)]; }
    void erase(uint64_t k) { key[slot_of(k)] = kTomb; }
};

uint32_t batch_size(SplitMix& r, uint32_t max_batch) {
    const double x = std::floor(std::exp(r.uniform() * std::log((double)max_batch)));
    if (x < 1.0) return 1;
    if (x > (double)max_batch) return max_batch;
    return (uint32_t)x;
}

}  // namespace
}  // namespace dcx

using namespace dcx;

extern "C" dcx_status dcx_gen_update_stream(uint32_t V, uint64_t E, const uint32_t* src,
                                            const uint32_t* dst, uint32_t scale, double a,
                                            double b, double c, uint32_t nbatches,
                                            uint32_t max_batch, double p_insert, uint64_t seed,
                                            uint64_t* ins_off, uint64_t* del_off,
                                            uint32_t* ins_pairs, uint64_t ins_cap,
                                            uint32_t* del_pairs, uint64_t del_cap) {
    if (!ins_off || !del_off || max_batch == 0 || p_insert < 0 || p_insert > 1 ||
        (E && (!src || !dst)) || scale < 1 || scale > 31 || (V != (1u << scale)))
        return DCX_ERR_INVALID_ARG;
    if (a < 0 || b < 0 || c < 0 || a + b + c > 1.0) return DCX_ERR_INVALID_ARG;
    const bool full = ins_pairs != nullptr || del_pairs != nullptr;
    // sizes and kinds first (they depend on the seed alone)
    std::vector<uint32_t> n_ins(nbatches), n_del(nbatches);
    ins_off[0] = del_off[0] = 0;
    for (uint32_t bi = 0; bi < nbatches; ++bi) {
        SplitMix r = stream_for(seed, bi);
        const uint32_t n = batch_size(r, max_batch);
        uint32_t ni = 0;
        for (uint32_t j = 0; j < n; ++j) ni += r.uniform() < p_insert;
        n_ins[bi] = ni;
        n_del[bi] = n - ni;
        ins_off[bi + 1] = ins_off[bi] + ni;
        del_off[bi + 1] = del_off[bi] + (n - ni);
    }
    if (!full) return DCX_OK;
    if (!ins_pairs || !del_pairs || 2 * ins_off[nbatches] > ins_cap ||
        2 * del_off[nbatches] > del_cap)
        return DCX_ERR_CAPACITY;

    std::vector<uint64_t> ev(E);
    EdgeIndex idx;
    idx.init(E + ins_off[nbatches]);
    for (uint64_t e = 0; e < E; ++e) {
        const uint64_t k = (uint64_t)src[e] << 32 | dst[e];
        if (src[e] >= V || dst[e] >= V || src[e] == dst[e] || idx.contains(k))
            return DCX_ERR_INVALID_ARG;
        ev[e] = k;
        idx.put(k, (uint32_t)e);
    }
    const double ab = a + b, abc = a + b + c;
    std::unordered_set<uint64_t> chosen;
    std::vector<uint64_t> dels, ins;
    for (uint32_t bi = 0; bi < nbatches; ++bi) {
        SplitMix r = stream_for(seed, bi);
        const uint32_t n = batch_size(r, max_batch);
        for (uint32_t j = 0; j < n; ++j) (void)r.uniform();  // the kinds
        chosen.clear();
        dels.clear();
        ins.clear();
        const uint32_t nd = (uint32_t)std::min<uint64_t>(n_del[bi], ev.size());
        if (nd != n_del[bi]) return DCX_ERR_CAPACITY;  // graph too small to delete from
        while (dels.size() < nd) {
            const uint64_t k = ev[r.below((uint32_t)ev.size())];
            if (chosen.insert(k).second) dels.push_back(k);
        }
        while (ins.size() < n_ins[bi]) {
            uint32_t s = 0, t = 0;
            for (uint32_t lvl = 0; lvl < scale; ++lvl) {
                const double u = r.uniform();
                const uint32_t row = u >= ab, col = (u >= a && u < ab) || u >= abc;
                s = (s << 1) | row;
                t = (t << 1) | col;
            }
            const uint64_t k = (uint64_t)s << 32 | t;
            if (s == t || idx.contains(k) || !chosen.insert(k).second) continue;
            ins.push_back(k);
        }
        uint32_t* dp = del_pairs + 2 * del_off[bi];
        for (size_t j = 0; j < dels.size(); ++j) {
            dp[2 * j] = (uint32_t)(dels[j] >> 32);
            dp[2 * j + 1] = (uint32_t)dels[j];
            // swap-remove from the edge vector
            const uint32_t p = idx.get_pos(dels[j]);
            const uint64_t last = ev.back();
            ev[p] = last;
            idx.set_pos(last, p);
            ev.pop_back();
            idx.erase(dels[j]);
        }
        uint32_t* ip = ins_pairs + 2 * ins_off[bi];
        for (size_t j = 0; j < ins.size(); ++j) {
            ip[2 * j] = (uint32_t)(ins[j] >> 32);
            ip[2 * j + 1] = (uint32_t)ins[j];
            idx.put(ins[j], (uint32_t)ev.size());
            ev.push_back(ins[j]);
        }
    }
    return DCX_OK;
}
