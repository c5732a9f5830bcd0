// Device ingest of an edge list: included at the end of dcx_api.hip.
//
// Builds on the GPU the same CSR-with-slack graph that dcx_graph_upload_edges
// builds on the host: the layout of Graph::readFile (src/graph.cpp:195-243),
// off[v+1] - off[v] = deg(v) + slack, free slots 0xFFFFFFFF.
//
// Input-order mode (no DCX_INGEST_DEDUP): each list holds its neighbours in
// input order, duplicates kept, as readFile does.
//   1. One pass over the pairs counts both degrees, records the first pair
//      with an id out of range (and the first self-loop, unless self-loops
//      are dropped, as readFile drops them: src/graph.cpp:167).
//   2. Offsets are the exclusive scan of deg + slack.
//   3. Per side, the keys (owner << B) | input index are radix-sorted on the
//      owner bits only; a dropped self-loop gets the owner V, which sorts
//      after every kept pair.  The sort is stable, so an owner's edges sit
//      at sorted positions start(owner) + r in input order, where
//      start(owner) = off[owner] - owner * slack, and sorted position j goes
//      to slot j + owner * slack.
// Dedup mode (DCX_INGEST_DEDUP): each list holds every neighbour once, in
// increasing id order -- the graph the host builds from a sorted,
// deduplicated edge list such as the seeded generators write.  Per side the
// keys (owner << b) | neighbour (owner V for a dropped self-loop) are
// sorted; a key is kept when it differs from its predecessor; the exclusive
// scan of the kept flags gives each kept key its position p among the kept
// ones; an owner's degree is the span of its run of positions; kept key j
// goes to slot p + owner * slack.
#pragma once

namespace {

__global__ void k_ingest_count(const uint32_t* src, const uint32_t* dst, uint64_t E, uint32_t V,
                               int drop_loops, int count_degrees, uint32_t* odeg, uint32_t* ideg,
                               unsigned long long* first_bad, unsigned long long* loops) {
    for (uint64_t e = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; e < E;
         e += (uint64_t)gridDim.x * blockDim.x) {
        const uint32_t s = src[e], d = dst[e];
        if (s >= V || d >= V || (s == d && !drop_loops)) {
            atomicMin(first_bad, (unsigned long long)e);
            continue;
        }
        if (s == d) {
            atomicAdd(loops, 1ull);
            continue;
        }
        if (!count_degrees) continue;
        atomicAdd(&odeg[s], 1u);
        atomicAdd(&ideg[d], 1u);
    }
}

// slots[v] = deg[v] + slack for v < V and slots[V] = 0 (the scan's total
// lands in off[V]); the graph's int32 degrees and list capacities; the
// largest degree goes to *maxdeg (a degree must fit int32).
__global__ void k_ingest_slots(const uint32_t* deg, uint32_t V, uint32_t slack, uint32_t* slots,
                               int32_t* gdeg, uint32_t* capv, uint32_t* maxdeg) {
    uint32_t m = 0;
    for (uint64_t v = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; v <= V;
         v += (uint64_t)gridDim.x * blockDim.x) {
        if (v == V) {
            slots[V] = 0;
            continue;
        }
        const uint32_t n = deg[v];
        m = max(m, n);
        slots[v] = n + slack;
        capv[v] = n + slack;
        gdeg[v] = (int32_t)n;
    }
    if (m) atomicMax(maxdeg, m);
}

// Sort keys of one side: owner = src (out-lists) or dst (in-lists); a
// self-loop (only present when dropped) gets the owner V.  Input-order
// mode: (owner << B) | e.  Dedup mode: (owner << B) | neighbour.
__global__ void k_ingest_keys(const uint32_t* src, const uint32_t* dst, uint64_t E, uint32_t V,
                              int outs, int dedup, int B, uint64_t* keys) {
    for (uint64_t e = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; e < E;
         e += (uint64_t)gridDim.x * blockDim.x) {
        const uint32_t s = src[e], d = dst[e];
        const uint64_t owner = s == d ? V : (outs ? s : d);
        const uint64_t low = dedup ? (uint64_t)(outs ? d : s) : e;
        keys[e] = (owner << B) | low;
    }
}

// Input-order mode: sorted position j of a kept pair goes to slot
// j + owner * slack (dropped self-loops sort last and are skipped).
__global__ void k_ingest_place(const uint64_t* keys, uint64_t E, int B, uint32_t V,
                               const uint32_t* other, uint32_t slack, uint32_t* nbr) {
    const uint64_t mask = (1ull << B) - 1;
    for (uint64_t j = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; j < E;
         j += (uint64_t)gridDim.x * blockDim.x) {
        const uint64_t k = keys[j];
        const uint64_t owner = k >> B;
        if (owner >= V) continue;
        nbr[j + owner * slack] = other[k & mask];
    }
}

__device__ __forceinline__ bool ingest_kept(const uint64_t* keys, uint64_t j, int B, uint32_t V) {
    const uint64_t k = keys[j];
    return (k >> B) < V && (j == 0 || keys[j - 1] != k);
}

// Dedup mode: flag[j] = 1 for the first copy of each kept key.
__global__ void k_ingest_flags(const uint64_t* keys, uint64_t E, int B, uint32_t V,
                               uint32_t* flag) {
    for (uint64_t j = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; j < E;
         j += (uint64_t)gridDim.x * blockDim.x)
        flag[j] = ingest_kept(keys, j, B, V) ? 1u : 0u;
}

// Dedup mode, after the scan (pos[j] = kept keys before j): at the first
// and last key of each owner's run, its first and one-past-last positions.
__global__ void k_ingest_runs(const uint64_t* keys, const uint32_t* pos, uint64_t E, int B,
                              uint32_t V, uint32_t* first, uint32_t* end) {
    for (uint64_t j = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; j < E;
         j += (uint64_t)gridDim.x * blockDim.x) {
        const uint64_t o = keys[j] >> B;
        if (o >= V) continue;
        if (j == 0 || (keys[j - 1] >> B) != o) first[o] = pos[j];
        if (j + 1 == E || (keys[j + 1] >> B) != o)
            end[o] = pos[j] + (ingest_kept(keys, j, B, V) ? 1u : 0u);
    }
}

// Dedup mode: deg[v] = end[v] - first[v] (both zero for an owner without
// keys), written over end[].
__global__ void k_ingest_span(uint32_t V, const uint32_t* first, uint32_t* end) {
    for (uint64_t v = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; v < V;
         v += (uint64_t)gridDim.x * blockDim.x)
        end[v] -= first[v];
}

// Dedup mode: kept key j goes to slot pos[j] + owner * slack.
__global__ void k_ingest_place_unique(const uint64_t* keys, const uint32_t* pos, uint64_t E,
                                      int B, uint32_t V, uint32_t slack, uint32_t* nbr) {
    const uint64_t mask = (1ull << B) - 1;
    for (uint64_t j = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; j < E;
         j += (uint64_t)gridDim.x * blockDim.x) {
        if (!ingest_kept(keys, j, B, V)) continue;
        const uint64_t k = keys[j];
        nbr[pos[j] + (k >> B) * slack] = (uint32_t)(k & mask);
    }
}

// Sample i of the R-MAT generator: the draws of dcx_gen_rmat before its
// sort and dedup (same counter-based stream per sample, same quadrant rule).
__global__ void k_gen_rmat(uint32_t scale, uint64_t samples, double a, double ab, double abc,
                           uint64_t seed, uint32_t* src, uint32_t* dst) {
    for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < samples;
         i += (uint64_t)gridDim.x * blockDim.x) {
        SplitMix r = stream_for(seed, i);
        uint32_t s = 0, t = 0;
        for (uint32_t lvl = 0; lvl < scale; ++lvl) {
            const double u = r.uniform();
            const uint32_t row = u >= ab, col = (u >= a && u < ab) || u >= abc;
            s = (s << 1) | row;
            t = (t << 1) | col;
        }
        src[i] = s;
        dst[i] = t;
    }
}

unsigned ingest_blocks(uint64_t n) {
    return (unsigned)std::max<uint64_t>(1, std::min<uint64_t>((n + 255) / 256, 8192));
}

int bits_for(uint64_t n) {  // bits of the largest value below n (0 for n <= 1)
    int b = 0;
    while (b < 64 && n > 1 && ((n - 1) >> b) != 0) ++b;
    return b;
}

dcx_status ingest_edges(dcx_ctx* ctx, uint32_t V, uint64_t E, const uint32_t* src,
                        const uint32_t* dst, uint32_t flags, uint32_t slack, dcx_graph** out) {
    *out = nullptr;
    const bool on_device = flags & DCX_INGEST_ON_DEVICE;
    const bool drop = flags & DCX_INGEST_DROP_SELF_LOOPS;
    const bool dedup = flags & DCX_INGEST_DEDUP;
    if (V > kMaxVertices) {
        set_err(ctx, "V = %u exceeds the supported %u vertices", V, kMaxVertices);
        return DCX_ERR_CAPACITY;
    }
    if (E > 0xffffffffull) {
        set_err(ctx, "E = %llu pairs reach 2^32", (unsigned long long)E);
        return DCX_ERR_CAPACITY;
    }
    hipStream_t st = ctx->stream;
    // temporaries, freed on every way out
    std::vector<void*> tmp;
    struct Free {
        std::vector<void*>& t;
        ~Free() {
            for (void* p : t) (void)hipFree(p);
        }
    } free_tmp{tmp};
    auto alloc = [&](auto** p, size_t n) {
        dev_alloc(ctx, p, n);
        tmp.push_back(*p);
    };
    const uint32_t* d_src = src;
    const uint32_t* d_dst = dst;
    if (!on_device && E) {
        uint32_t *s = nullptr, *d = nullptr;
        alloc(&s, E);
        alloc(&d, E);
        HIPX(ctx, hipMemcpyAsync(s, src, 4 * E, hipMemcpyHostToDevice, st));
        HIPX(ctx, hipMemcpyAsync(d, dst, 4 * E, hipMemcpyHostToDevice, st));
        d_src = s;
        d_dst = d;
    }
    uint32_t *deg = nullptr, *slots = nullptr, *scan_tmp = nullptr;
    // [0] first bad pair, [1] largest degree, [2] self-loops dropped
    unsigned long long* bad = nullptr;
    alloc(&deg, 2ull * V + 1);
    alloc(&slots, (size_t)V + 1);
    alloc(&scan_tmp, scan_tmp_words(std::max<uint64_t>((uint64_t)V + 1, dedup ? E : 0)));
    alloc(&bad, 3);
    HIPX(ctx, hipMemsetAsync(deg, 0, 4ull * (2ull * V + 1), st));
    HIPX(ctx, hipMemsetAsync(bad, 0xff, 8, st));
    HIPX(ctx, hipMemsetAsync(bad + 1, 0, 16, st));
    if (E)
        hipLaunchKernelGGL(k_ingest_count, dim3(ingest_blocks(E)), dim3(256), 0, st, d_src, d_dst,
                           E, V, (int)drop, (int)!dedup, deg, deg + V, bad, bad + 2);
    HIPX(ctx, hipGetLastError());
    unsigned long long hbad[3] = {0, 0, 0};
    HIPX(ctx, hipMemcpyAsync(hbad, bad, 24, hipMemcpyDeviceToHost, st));
    sync(ctx);
    const unsigned long long first_bad = hbad[0];
    if (first_bad != ~0ull) {
        set_err(ctx, "pair %llu: vertex id out of range (V = %u)%s", first_bad, V,
                drop ? "" : " or self-loop");
        return DCX_ERR_INVALID_ARG;
    }
    uint64_t *keys = nullptr, *alt = nullptr;
    uint32_t *table = nullptr, *pos = nullptr;
    if (E) {
        alloc(&keys, E);
        alloc(&alt, E);
        alloc(&table, radix_tmp_words(E));
        if (dedup) alloc(&pos, E);
    }
    // key layout: owner (up to V inclusive) above B low bits
    const int B = dedup ? bits_for(V) : bits_for(E);
    const int key_bits = B + bits_for(drop ? (uint64_t)V + 1 : (uint64_t)V);
    // Dedup: sort both sides first, then size the graph from the kept count.
    // The sorted keys of the out side are kept while the in side sorts.
    uint64_t* sorted[2] = {nullptr, nullptr};
    uint64_t kept[2] = {E, E};
    uint64_t *keys_in = nullptr, *alt_in = nullptr;
    uint32_t* pos_in = nullptr;
    if (dedup && E) {
        alloc(&keys_in, E);
        alloc(&alt_in, E);
        alloc(&pos_in, E);
        for (int side = 0; side < 2; ++side) {
            uint64_t* k = side == 0 ? keys : keys_in;
            uint64_t* a = side == 0 ? alt : alt_in;
            uint32_t* p = side == 0 ? pos : pos_in;
            uint32_t* d_first = slots;                 // V words of scratch
            uint32_t* d_end = deg + (side == 0 ? 0 : V);
            hipLaunchKernelGGL(k_ingest_keys, dim3(ingest_blocks(E)), dim3(256), 0, st, d_src,
                               d_dst, E, V, side == 0, 1, B, k);
            HIPX(ctx, hipGetLastError());
            HIPX(ctx, launch_radix_sort_u64(k, a, E, key_bits, table, &sorted[side], st));
            hipLaunchKernelGGL(k_ingest_flags, dim3(ingest_blocks(E)), dim3(256), 0, st,
                               sorted[side], E, B, V, p);
            HIPX(ctx, hipGetLastError());
            // the last flag is needed for the total after the in-place scan
            uint32_t last_flag = 0, last_pos = 0;
            HIPX(ctx, hipMemcpyAsync(&last_flag, p + E - 1, 4, hipMemcpyDeviceToHost, st));
            HIPX(ctx, launch_excl_scan_u32(p, p, E, scan_tmp, st));
            HIPX(ctx, hipMemcpyAsync(&last_pos, p + E - 1, 4, hipMemcpyDeviceToHost, st));
            HIPX(ctx, hipMemsetAsync(d_first, 0, 4ull * V, st));
            hipLaunchKernelGGL(k_ingest_runs, dim3(ingest_blocks(E)), dim3(256), 0, st,
                               sorted[side], p, E, B, V, d_first, d_end);
            hipLaunchKernelGGL(k_ingest_span, dim3(ingest_blocks(V)), dim3(256), 0, st, V,
                               d_first, d_end);
            HIPX(ctx, hipGetLastError());
            sync(ctx);
            kept[side] = (uint64_t)last_pos + last_flag;
        }
        if (kept[0] != kept[1]) {
            set_err(ctx, "ingest: %llu distinct out-pairs but %llu in-pairs",
                    (unsigned long long)kept[0], (unsigned long long)kept[1]);
            return DCX_ERR_INTERNAL;
        }
    } else {
        kept[0] = kept[1] = E - hbad[2];
    }
    const uint64_t Ek = kept[0];
    const uint64_t cap = Ek + (uint64_t)slack * V;
    if ((double)Ek + (double)slack * V >= 4294967296.0) {
        set_err(ctx, "E + slack * V = %.0f slots reach 2^32", (double)Ek + (double)slack * V);
        return DCX_ERR_CAPACITY;
    }
    std::unique_ptr<dcx_graph> g(new dcx_graph);
    g->ctx = ctx;
    g->V = V;
    g->E = Ek;
    g->out_cap = g->in_cap = cap;
    // headroom for lists relocated by later inserts (as dcx_graph_upload_csr)
    auto headroom = [](uint64_t n) { return n + n / 16 + 4096; };
    g->out_alloc = g->in_alloc = std::min<uint64_t>(headroom(cap), 0xffffffffull);
    try {
        dev_alloc(ctx, &g->out_off, (size_t)V + 1);
        dev_alloc(ctx, &g->in_off, (size_t)V + 1);
        dev_alloc(ctx, &g->out_deg, V);
        dev_alloc(ctx, &g->in_deg, V);
        dev_alloc(ctx, &g->out_capv, V);
        dev_alloc(ctx, &g->in_capv, V);
        dev_alloc(ctx, &g->out_nbr, g->out_alloc);
        dev_alloc(ctx, &g->in_nbr, g->in_alloc);
        HIPX(ctx, hipMemsetAsync(g->out_nbr, 0xff, 4ull * g->out_alloc, st));
        HIPX(ctx, hipMemsetAsync(g->in_nbr, 0xff, 4ull * g->in_alloc, st));
        uint32_t* maxdeg = reinterpret_cast<uint32_t*>(bad + 1);
        for (int side = 0; side < 2; ++side) {
            const bool outs = side == 0;
            uint32_t* off = outs ? g->out_off : g->in_off;
            hipLaunchKernelGGL(k_ingest_slots, dim3(ingest_blocks((uint64_t)V + 1)), dim3(256), 0,
                               st, deg + (outs ? 0 : V), V, slack, slots,
                               outs ? g->out_deg : g->in_deg, outs ? g->out_capv : g->in_capv,
                               maxdeg);
            HIPX(ctx, hipGetLastError());
            HIPX(ctx, launch_excl_scan_u32(slots, off, (uint64_t)V + 1, scan_tmp, st));
            if (!E) continue;
            if (dedup) {
                hipLaunchKernelGGL(k_ingest_place_unique, dim3(ingest_blocks(E)), dim3(256), 0,
                                   st, sorted[side], outs ? pos : pos_in, E, B, V, slack,
                                   outs ? g->out_nbr : g->in_nbr);
            } else {
                hipLaunchKernelGGL(k_ingest_keys, dim3(ingest_blocks(E)), dim3(256), 0, st, d_src,
                                   d_dst, E, V, (int)outs, 0, B, keys);
                HIPX(ctx, hipGetLastError());
                uint64_t* srt = nullptr;
                HIPX(ctx, launch_radix_sort_u64(keys, alt, E, key_bits, table, &srt, st, B));
                hipLaunchKernelGGL(k_ingest_place, dim3(ingest_blocks(E)), dim3(256), 0, st, srt,
                                   E, B, V, outs ? d_dst : d_src, slack,
                                   outs ? g->out_nbr : g->in_nbr);
            }
            HIPX(ctx, hipGetLastError());
        }
        uint32_t md = 0;
        HIPX(ctx, hipMemcpyAsync(&md, maxdeg, 4, hipMemcpyDeviceToHost, st));
        sync(ctx);
        if (md > 0x7fffffffu) {
            set_err(ctx, "a vertex has %u list entries; degrees must fit int32", md);
            throw Fail{DCX_ERR_CAPACITY};
        }
    } catch (...) {
        dcx_graph* raw = g.release();
        free_graph_arrays(raw);
        delete raw;
        throw;
    }
    *out = g.release();
    return DCX_OK;
}

}  // namespace

extern "C" {

dcx_status dcx_graph_ingest_edges(dcx_ctx* ctx, uint32_t V, uint64_t E, const uint32_t* src,
                                  const uint32_t* dst, int32_t on_device, uint32_t slack,
                                  dcx_graph** out) {
    if (!ctx || !out || (E && (!src || !dst))) return DCX_ERR_INVALID_ARG;
    return guarded(ctx, [&]() {
        return ingest_edges(ctx, V, E, src, dst, on_device ? DCX_INGEST_ON_DEVICE : 0u, slack,
                            out);
    });
}

dcx_status dcx_gen_rmat_device(dcx_ctx* ctx, uint32_t scale, uint64_t samples, double a,
                               double b, double c, uint64_t seed, uint32_t* src, uint32_t* dst) {
    if (!ctx || (samples && (!src || !dst))) return DCX_ERR_INVALID_ARG;
    if (scale < 1 || scale > 31 || a < 0 || b < 0 || c < 0 || a + b + c > 1.0)
        return DCX_ERR_INVALID_ARG;
    return guarded(ctx, [&]() -> dcx_status {
        if (samples)
            hipLaunchKernelGGL(k_gen_rmat, dim3(ingest_blocks(samples)), dim3(256), 0,
                               ctx->stream, scale, samples, a, a + b, a + b + c, seed, src, dst);
        HIPX(ctx, hipGetLastError());
        sync(ctx);
        return DCX_OK;
    });
}

dcx_status dcx_graph_ingest(dcx_ctx* ctx, uint32_t V, uint64_t E, const uint32_t* src,
                            const uint32_t* dst, uint32_t flags, uint32_t slack,
                            dcx_graph** out) {
    if (!ctx || !out || (E && (!src || !dst))) return DCX_ERR_INVALID_ARG;
    if (flags & ~(uint32_t)(DCX_INGEST_ON_DEVICE | DCX_INGEST_DROP_SELF_LOOPS | DCX_INGEST_DEDUP))
        return DCX_ERR_INVALID_ARG;
    return guarded(ctx, [&]() { return ingest_edges(ctx, V, E, src, dst, flags, slack, out); });
}

}  // extern "C"
