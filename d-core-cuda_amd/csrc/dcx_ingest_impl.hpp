// Device ingest of an edge list: included at the end of dcx_api.hip.
//
// Builds on the GPU the same CSR-with-slack graph that dcx_graph_upload_edges
// builds on the host: the layout of Graph::readFile (src/graph.cpp:195-243),
// off[v+1] - off[v] = deg(v) + slack, each list's neighbours in input order,
// free slots 0xFFFFFFFF.  Duplicate pairs are kept, as there.
//   1. One pass over the pairs counts both degrees and records the first
//      pair with an id out of range or a self-loop.
//   2. Offsets are the exclusive scan of deg + slack.
//   3. Per side, the keys (owner << B) | input index, built in input order,
//      are radix-sorted on the owner bits only: the sort is stable, so the
//      owner's edges then sit at sorted positions start(owner) + r in input
//      order, where start(owner) = off[owner] - owner * slack, and edge r of
//      the list goes to slot j + owner * slack for sorted position j.
#pragma once

namespace {

__global__ void k_ingest_count(const uint32_t* src, const uint32_t* dst, uint64_t E, uint32_t V,
                               uint32_t* odeg, uint32_t* ideg, unsigned long long* first_bad) {
    for (uint64_t e = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; e < E;
         e += (uint64_t)gridDim.x * blockDim.x) {
        const uint32_t s = src[e], d = dst[e];
        if (s >= V || d >= V || s == d) {
            atomicMin(first_bad, (unsigned long long)e);
            continue;
        }
        atomicAdd(&odeg[s], 1u);
        atomicAdd(&ideg[d], 1u);
    }
}

// slots[v] = deg[v] + slack for v < V and slots[V] = 0 (the scan's total
// lands in off[V]); the graph's int32 degrees and list capacities
__global__ void k_ingest_slots(const uint32_t* deg, uint32_t V, uint32_t slack, uint32_t* slots,
                               int32_t* gdeg, uint32_t* capv) {
    for (uint64_t v = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; v <= V;
         v += (uint64_t)gridDim.x * blockDim.x) {
        if (v == V) {
            slots[V] = 0;
            continue;
        }
        const uint32_t n = deg[v];
        slots[v] = n + slack;
        capv[v] = n + slack;
        gdeg[v] = (int32_t)n;
    }
}

__global__ void k_ingest_keys(const uint32_t* owner, uint64_t E, int B, uint64_t* keys) {
    for (uint64_t e = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; e < E;
         e += (uint64_t)gridDim.x * blockDim.x)
        keys[e] = ((uint64_t)owner[e] << B) | e;
}

__global__ void k_ingest_place(const uint64_t* keys, uint64_t E, int B, const uint32_t* other,
                               uint32_t slack, uint32_t* nbr) {
    const uint64_t mask = (1ull << B) - 1;
    for (uint64_t j = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; j < E;
         j += (uint64_t)gridDim.x * blockDim.x) {
        const uint64_t k = keys[j];
        nbr[j + (k >> B) * slack] = other[k & mask];
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

}  // namespace

extern "C" {

dcx_status dcx_graph_ingest_edges(dcx_ctx* ctx, uint32_t V, uint64_t E, const uint32_t* src,
                                  const uint32_t* dst, int32_t on_device, uint32_t slack,
                                  dcx_graph** out) {
    if (!ctx || !out || (E && (!src || !dst))) return DCX_ERR_INVALID_ARG;
    return guarded(ctx, [&]() -> dcx_status {
        *out = nullptr;
        if (V > kMaxVertices) {
            set_err(ctx, "V = %u exceeds the supported %u vertices", V, kMaxVertices);
            return DCX_ERR_CAPACITY;
        }
        const uint64_t cap = E + (uint64_t)slack * V;
        if ((double)E + (double)slack * V >= 4294967296.0) {
            set_err(ctx, "E + slack * V = %.0f slots reach 2^32", (double)E + (double)slack * V);
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
        unsigned long long* bad = nullptr;
        alloc(&deg, 2ull * V + 1);
        alloc(&slots, (size_t)V + 1);
        alloc(&scan_tmp, scan_tmp_words((uint64_t)V + 1));
        alloc(&bad, 1);
        HIPX(ctx, hipMemsetAsync(deg, 0, 4ull * (2ull * V + 1), st));
        HIPX(ctx, hipMemsetAsync(bad, 0xff, 8, st));
        if (E)
            hipLaunchKernelGGL(k_ingest_count, dim3(ingest_blocks(E)), dim3(256), 0, st, d_src,
                               d_dst, E, V, deg, deg + V, bad);
        HIPX(ctx, hipGetLastError());
        unsigned long long first_bad = 0;
        HIPX(ctx, hipMemcpyAsync(&first_bad, bad, 8, hipMemcpyDeviceToHost, st));
        sync(ctx);
        if (first_bad != ~0ull) {
            set_err(ctx, "pair %llu: vertex id out of range (V = %u) or self-loop", first_bad, V);
            return DCX_ERR_INVALID_ARG;
        }
        std::unique_ptr<dcx_graph> g(new dcx_graph);
        g->ctx = ctx;
        g->V = V;
        g->E = E;
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
            uint64_t *keys = nullptr, *alt = nullptr;
            uint32_t* table = nullptr;
            if (E) {
                alloc(&keys, E);
                alloc(&alt, E);
                alloc(&table, radix_tmp_words(E));
            }
            const int B = bits_for(E), key_bits = B + bits_for(V);
            for (int side = 0; side < 2; ++side) {
                const bool outs = side == 0;
                uint32_t* off = outs ? g->out_off : g->in_off;
                hipLaunchKernelGGL(k_ingest_slots, dim3(ingest_blocks((uint64_t)V + 1)), dim3(256),
                                   0, st, deg + (outs ? 0 : V), V, slack, slots,
                                   outs ? g->out_deg : g->in_deg, outs ? g->out_capv : g->in_capv);
                HIPX(ctx, hipGetLastError());
                HIPX(ctx, launch_excl_scan_u32(slots, off, (uint64_t)V + 1, scan_tmp, st));
                if (!E) continue;
                hipLaunchKernelGGL(k_ingest_keys, dim3(ingest_blocks(E)), dim3(256), 0, st,
                                   outs ? d_src : d_dst, E, B, keys);
                HIPX(ctx, hipGetLastError());
                uint64_t* sorted = nullptr;
                HIPX(ctx, launch_radix_sort_u64(keys, alt, E, key_bits, table, &sorted, st, B));
                hipLaunchKernelGGL(k_ingest_place, dim3(ingest_blocks(E)), dim3(256), 0, st,
                                   sorted, E, B, outs ? d_dst : d_src, slack,
                                   outs ? g->out_nbr : g->in_nbr);
                HIPX(ctx, hipGetLastError());
            }
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

}  // extern "C"
