// Many graphs in one decomposition: included at the end of dcx_api.hip.
//
// The (k,l)-cores of a disjoint union of graphs are the unions of the
// graphs' (k,l)-cores, so kmaxes and every l-row of the union restricted to
// one graph's vertex range are exactly that graph's own decomposition (rows
// above the graph's kmax are zero there).  A batch of graphs of uneven size
// is therefore uploaded as one union graph and decomposed by the same
// launches as a single graph -- the kmax pass, the core ordering and one
// persistent rows launch whose teams each run a k-pass over all graphs at
// once -- instead of a launch sequence per graph and per k.
#pragma once

namespace {

__global__ void k_part_summary(const int32_t* kmaxes, const int32_t* row0, uint32_t V,
                               const uint32_t* vbase, uint32_t n, int32_t* kmax_out,
                               int32_t* lmax_out) {
    for (uint32_t v = blockIdx.x * blockDim.x + threadIdx.x; v < V; v += gridDim.x * blockDim.x) {
        // part of v: the last i with vbase[i] <= v
        uint32_t lo = 0, hi = n - 1;
        while (lo < hi) {
            const uint32_t mid = (lo + hi + 1) >> 1;
            if (vbase[mid] <= v)
                lo = mid;
            else
                hi = mid - 1;
        }
        atomicMax(kmax_out + lo, kmaxes[v]);
        // l-values are non-increasing in k, so a vertex's largest is row 0
        atomicMax(lmax_out + lo, row0[v]);
    }
}

}  // namespace

extern "C" {

dcx_status dcx_graph_upload_batch(dcx_ctx* ctx, uint32_t n, const uint32_t* V, const uint64_t* E,
                                  const uint32_t* const* src, const uint32_t* const* dst,
                                  uint32_t slack, uint32_t* vbase, dcx_graph** out) {
    if (!ctx || !out || !vbase || (n && (!V || !E || !src || !dst))) return DCX_ERR_INVALID_ARG;
    uint64_t tv = 0, te = 0;
    for (uint32_t i = 0; i < n; ++i) {
        tv += V[i];
        te += E[i];
        if (E[i] && (!src[i] || !dst[i])) return DCX_ERR_INVALID_ARG;
    }
    if (tv > kMaxVertices) return DCX_ERR_CAPACITY;
    std::vector<uint32_t> us, ud;
    try {
        us.resize(te);
        ud.resize(te);
    } catch (const std::bad_alloc&) {
        return DCX_ERR_OUT_OF_MEMORY;
    }
    uint64_t pos = 0;
    uint32_t base = 0;
    for (uint32_t i = 0; i < n; ++i) {
        vbase[i] = base;
        for (uint64_t e = 0; e < E[i]; ++e) {
            if (src[i][e] >= V[i] || dst[i][e] >= V[i]) return DCX_ERR_INVALID_ARG;
            us[pos] = base + src[i][e];
            ud[pos] = base + dst[i][e];
            ++pos;
        }
        base += V[i];
    }
    vbase[n] = base;
    return dcx_graph_upload_edges(ctx, base, te, us.data(), ud.data(), slack, out);
}

dcx_status dcx_batch_summary(dcx_ctx* ctx, const dcx_graph* g, uint32_t n, const uint32_t* vbase,
                             int32_t* kmax, int32_t* lmax) {
    dcx_status s = check_graph(ctx, g);
    if (s != DCX_OK) return s;
    if (!n || !vbase || !kmax || !lmax) return DCX_ERR_INVALID_ARG;
    if (!g->rows_valid) return DCX_ERR_STATE;
    if (vbase[0] != 0 || vbase[n] != g->V) return DCX_ERR_INVALID_ARG;
    for (uint32_t i = 0; i < n; ++i)
        if (vbase[i + 1] < vbase[i]) return DCX_ERR_INVALID_ARG;
    return guarded(ctx, [&]() -> dcx_status {
        uint32_t* d = nullptr;
        dev_alloc(ctx, &d, 3ull * n + 1);
        try {
            int32_t* dk = reinterpret_cast<int32_t*>(d + n + 1);
            int32_t* dl = dk + n;
            HIPX(ctx, hipMemcpyAsync(d, vbase, 4ull * (n + 1), hipMemcpyHostToDevice, ctx->stream));
            HIPX(ctx, launch_fill_i32(dk, -1, n, ctx->stream));
            HIPX(ctx, launch_fill_i32(dl, 0, n, ctx->stream));
            if (g->V && g->kmax >= 0) {
                hipLaunchKernelGGL(k_part_summary, dim3(std::min<uint32_t>((g->V + 255) / 256, 4096)),
                                   dim3(256), 0, ctx->stream, g->kmaxes, g->rows, g->V, d, n, dk,
                                   dl);
                HIPX(ctx, hipGetLastError());
            }
            HIPX(ctx, hipMemcpyAsync(kmax, dk, 4ull * n, hipMemcpyDeviceToHost, ctx->stream));
            HIPX(ctx, hipMemcpyAsync(lmax, dl, 4ull * n, hipMemcpyDeviceToHost, ctx->stream));
            sync(ctx);
        } catch (...) {
            dev_free(d);
            throw;
        }
        dev_free(d);
        // an empty part (no vertex) reports kmax -1 and lmax 0, like an
        // empty graph
        return DCX_OK;
    });
}

}  // extern "C"
