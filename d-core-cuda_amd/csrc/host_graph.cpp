// Host-side graph building for the dcx library: seeded generators, edge-list
// normalisation, CSR-with-slack construction and the reference's file
// formats.  No GPU is touched here; every function is plain C-ABI.
//
// Reference counterparts:
//   generateGraph          src/graph_gen.cpp:9-34   (G(n,m), rejection)
//   Graph::readFile        src/graph.cpp:146-250    (text -> CSR with gap)
//   pruneEmptyVertices     src/graph.cpp:110-144
//   GraphCPU::readFile     src/graph_cpu.cpp:38-102 (V from header, no prune)
//   writeBinary/readBinary src/graph.cpp:261-325    (<file>-binary-g=N)
//   read/writeDecompFile   src/utils.h:76-109       (.decomp)
//   writeDecompFileCPU     src/utils.h:110-124      (.decompCPU)
#include <omp.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <string>
#include <unordered_set>
#include <vector>

#include "dcx.h"
#include "rng.hpp"

namespace dcx {

// Parallel LSD radix sort of 64-bit keys; only the low `bits` bits vary.
static void radix_sort_u64(std::vector<uint64_t>& a, int bits) {
    const int R = 11, B = 1 << R;
    size_t n = a.size();
    if (n < 2) return;
    std::vector<uint64_t> tmp(n);
    int nt = omp_get_max_threads();
    std::vector<size_t> hist((size_t)nt * B);
    for (int shift = 0; shift < bits; shift += R) {
        std::fill(hist.begin(), hist.end(), 0);
#pragma omp parallel num_threads(nt)
        {
            int t = omp_get_thread_num();
            size_t lo = n * t / nt, hi = n * (t + 1) / nt;
            size_t* h = &hist[(size_t)t * B];
            for (size_t i = lo; i < hi; ++i) h[(a[i] >> shift) & (B - 1)]++;
#pragma omp barrier
#pragma omp single
            {
                size_t sum = 0;
                for (int d = 0; d < B; ++d)
                    for (int tt = 0; tt < nt; ++tt) {
                        size_t c = hist[(size_t)tt * B + d];
                        hist[(size_t)tt * B + d] = sum;
                        sum += c;
                    }
            }
            for (size_t i = lo; i < hi; ++i) tmp[h[(a[i] >> shift) & (B - 1)]++] = a[i];
        }
        a.swap(tmp);
    }
}

static int bits_for(uint64_t v) {
    int b = 0;
    while (b < 64 && (v >> b)) ++b;
    return b;
}

// Sort (src,dst) keys, drop self-loops and duplicates, write out.
static uint64_t finish_edges(std::vector<uint64_t>& keys, uint32_t V, uint32_t* src,
                             uint32_t* dst, uint64_t cap) {
    radix_sort_u64(keys, 32 + bits_for(V ? V - 1 : 0));
    uint64_t m = 0;
    uint64_t prev = ~0ULL;
    for (uint64_t k : keys) {
        uint32_t s = (uint32_t)(k >> 32), t = (uint32_t)k;
        if (s == t || k == prev) continue;
        prev = k;
        if (m < cap) {
            src[m] = s;
            dst[m] = t;
        }
        ++m;
    }
    return m;
}

}  // namespace dcx

using namespace dcx;

extern "C" {

dcx_status dcx_gen_er(uint32_t V, uint64_t E, uint64_t seed, uint32_t* src, uint32_t* dst) {
    if (V < 2 || !src || !dst) return DCX_ERR_INVALID_ARG;
    if ((double)E > (double)V * (double)(V - 1)) return DCX_ERR_INVALID_ARG;
    // Rejection sampling of distinct non-loop pairs, as generateGraph does
    // (src/graph_gen.cpp:12-19), from a seeded stream instead of time().
    SplitMix rng(mix64(seed) ^ 0x45525f67656eULL);
    std::unordered_set<uint64_t> seen;
    seen.reserve(E * 2);
    std::vector<uint64_t> keys;
    keys.reserve(E);
    while (keys.size() < E) {
        uint32_t s = rng.below(V), t = rng.below(V);
        if (s == t) continue;
        uint64_t k = ((uint64_t)s << 32) | t;
        if (seen.insert(k).second) keys.push_back(k);
    }
    uint64_t m = finish_edges(keys, V, src, dst, E);
    return m == E ? DCX_OK : DCX_ERR_INTERNAL;
}

dcx_status dcx_gen_rmat(uint32_t scale, uint64_t samples, double a, double b, double c,
                        uint64_t seed, uint32_t* src, uint32_t* dst, uint64_t cap,
                        uint64_t* E_out) {
    if (scale < 1 || scale > 31 || !E_out) return DCX_ERR_INVALID_ARG;
    if (a < 0 || b < 0 || c < 0 || a + b + c > 1.0) return DCX_ERR_INVALID_ARG;
    std::vector<uint64_t> keys(samples);
    const double ab = a + b, abc = a + b + c;
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < (int64_t)samples; ++i) {
        SplitMix r = stream_for(seed, (uint64_t)i);
        uint32_t s = 0, t = 0;
        for (uint32_t lvl = 0; lvl < scale; ++lvl) {
            double u = r.uniform();
            uint32_t row = u >= ab, col = (u >= a && u < ab) || u >= abc;
            s = (s << 1) | row;
            t = (t << 1) | col;
        }
        keys[i] = ((uint64_t)s << 32) | t;
    }
    uint64_t m = finish_edges(keys, 1u << scale, src, dst, cap);
    *E_out = m;
    return (src && dst && m > cap) ? DCX_ERR_CAPACITY : DCX_OK;
}

dcx_status dcx_gen_chunglu(uint32_t V, uint64_t samples, double gamma, uint64_t seed,
                           uint32_t* src, uint32_t* dst, uint64_t cap, uint64_t* E_out) {
    if (V < 2 || gamma <= 1.0 || !E_out) return DCX_ERR_INVALID_ARG;
    // weights w_j ~ (j+1)^(-1/(gamma-1)), independently permuted for the
    // out and in side; endpoints drawn by inverse CDF.
    const double ex = -1.0 / (gamma - 1.0);
    std::vector<double> cdf(V);
    double acc = 0;
    for (uint32_t j = 0; j < V; ++j) {
        acc += std::pow((double)j + 1.0, ex);
        cdf[j] = acc;
    }
    for (auto& x : cdf) x /= acc;
    auto perm = [&](uint64_t tag) {
        std::vector<uint32_t> p(V);
        for (uint32_t j = 0; j < V; ++j) p[j] = j;
        SplitMix r(mix64(seed ^ tag));
        for (uint32_t j = V - 1; j > 0; --j) std::swap(p[j], p[r.below(j + 1)]);
        return p;
    };
    std::vector<uint32_t> pout = perm(0x6f7574ULL), pin = perm(0x696eULL);
    std::vector<uint64_t> keys(samples);
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < (int64_t)samples; ++i) {
        SplitMix r = stream_for(seed, (uint64_t)i);
        double u1 = r.uniform(), u2 = r.uniform();
        uint32_t j1 = (uint32_t)(std::upper_bound(cdf.begin(), cdf.end(), u1) - cdf.begin());
        uint32_t j2 = (uint32_t)(std::upper_bound(cdf.begin(), cdf.end(), u2) - cdf.begin());
        if (j1 >= V) j1 = V - 1;
        if (j2 >= V) j2 = V - 1;
        keys[i] = ((uint64_t)pout[j1] << 32) | pin[j2];
    }
    uint64_t m = finish_edges(keys, V, src, dst, cap);
    *E_out = m;
    return (src && dst && m > cap) ? DCX_ERR_CAPACITY : DCX_OK;
}

dcx_status dcx_csr_build(uint32_t V, uint64_t E, const uint32_t* src, const uint32_t* dst,
                         uint32_t slack, uint32_t* out_off, int32_t* out_deg, uint32_t* out_nbr,
                         uint32_t* in_off, int32_t* in_deg, uint32_t* in_nbr) {
    // Layout of Graph::readFile (src/graph.cpp:195-243): off[v+1]-off[v] =
    // deg(v) + slack, neighbours in input order, slack slots after them.
    if (!out_off || !out_deg || !in_off || !in_deg) return DCX_ERR_INVALID_ARG;
    if ((E && (!src || !dst || !out_nbr || !in_nbr))) return DCX_ERR_INVALID_ARG;
    if ((double)E + (double)slack * V >= 4294967296.0) return DCX_ERR_CAPACITY;
    for (uint64_t e = 0; e < E; ++e)
        if (src[e] >= V || dst[e] >= V) return DCX_ERR_INVALID_ARG;
    std::vector<uint32_t> od(V, 0), id(V, 0);
    for (uint64_t e = 0; e < E; ++e) {
        od[src[e]]++;
        id[dst[e]]++;
    }
    for (uint32_t v = 0; v < V; ++v)
        if (od[v] > 0x7fffffffu || id[v] > 0x7fffffffu) return DCX_ERR_CAPACITY;  // int32 degree
    uint64_t so = 0, si = 0;
    for (uint32_t v = 0; v < V; ++v) {
        out_off[v] = (uint32_t)so;
        in_off[v] = (uint32_t)si;
        out_deg[v] = (int32_t)od[v];
        in_deg[v] = (int32_t)id[v];
        so += od[v] + slack;
        si += id[v] + slack;
    }
    out_off[V] = (uint32_t)so;
    in_off[V] = (uint32_t)si;
    if (out_nbr) std::fill(out_nbr, out_nbr + so, 0xffffffffu);
    if (in_nbr) std::fill(in_nbr, in_nbr + si, 0xffffffffu);
    std::vector<uint32_t> co(out_off, out_off + V), ci(in_off, in_off + V);
    for (uint64_t e = 0; e < E; ++e) {
        out_nbr[co[src[e]]++] = dst[e];
        in_nbr[ci[dst[e]]++] = src[e];
    }
    return DCX_OK;
}

// ---- text edge lists --------------------------------------------------

dcx_status dcx_text_count(const char* path, uint32_t* V_header, uint64_t* lines) {
    if (!path || !V_header || !lines) return DCX_ERR_INVALID_ARG;
    FILE* f = std::fopen(path, "r");
    if (!f) return DCX_ERR_IO;
    char buf[256];
    if (!std::fgets(buf, sizeof buf, f) || buf[0] != '#') {
        std::fclose(f);
        return DCX_ERR_FORMAT;
    }
    unsigned long long hv = std::strtoull(buf + 1, nullptr, 10);
    if (hv > 0xffffffffULL) {
        std::fclose(f);
        return DCX_ERR_FORMAT;
    }
    uint64_t n = 0;
    unsigned long long s, t;
    while (std::fscanf(f, "%llu %llu", &s, &t) == 2) ++n;
    std::fclose(f);
    *V_header = (uint32_t)hv;
    *lines = n;
    return DCX_OK;
}

dcx_status dcx_text_read(const char* path, int32_t prune, uint32_t* src, uint32_t* dst,
                         uint64_t cap, uint32_t* V_out, uint64_t* E_out) {
    // V from the "# V" header (src/graph_cpu.cpp:49), self-loops dropped
    // (src/graph_cpu.cpp:58, src/graph.cpp:167), input order kept.  With
    // `prune`, vertices with no edge are removed and ids renumbered in
    // increasing order (pruneEmptyVertices, src/graph.cpp:110-144).
    if (!path || !V_out || !E_out || (cap && (!src || !dst))) return DCX_ERR_INVALID_ARG;
    FILE* f = std::fopen(path, "r");
    if (!f) return DCX_ERR_IO;
    char buf[256];
    if (!std::fgets(buf, sizeof buf, f) || buf[0] != '#') {
        std::fclose(f);
        return DCX_ERR_FORMAT;
    }
    unsigned long long hv = std::strtoull(buf + 1, nullptr, 10);
    uint32_t V = (uint32_t)hv;
    uint64_t m = 0;
    unsigned long long s, t;
    dcx_status st = DCX_OK;
    while (std::fscanf(f, "%llu %llu", &s, &t) == 2) {
        if (s >= V || t >= V) {
            st = DCX_ERR_FORMAT;
            break;
        }
        if (s == t) continue;
        if (m >= cap) {
            st = DCX_ERR_CAPACITY;
            break;
        }
        src[m] = (uint32_t)s;
        dst[m] = (uint32_t)t;
        ++m;
    }
    std::fclose(f);
    if (st != DCX_OK) return st;
    if (prune) {
        std::vector<uint32_t> map(V, 0xffffffffu);
        for (uint64_t e = 0; e < m; ++e) map[src[e]] = map[dst[e]] = 0;
        uint32_t nv = 0;
        for (uint32_t v = 0; v < V; ++v)
            if (map[v] == 0) map[v] = nv++;
        for (uint64_t e = 0; e < m; ++e) {
            src[e] = map[src[e]];
            dst[e] = map[dst[e]];
        }
        V = nv;
    }
    *V_out = V;
    *E_out = m;
    return DCX_OK;
}

dcx_status dcx_text_write(const char* path, uint32_t V, uint64_t E, const uint32_t* src,
                          const uint32_t* dst) {
    // generateGraph's output format: "# V" then "s t" lines (src/graph_gen.cpp:27-29).
    if (!path || (E && (!src || !dst))) return DCX_ERR_INVALID_ARG;
    FILE* f = std::fopen(path, "w");
    if (!f) return DCX_ERR_IO;
    std::fprintf(f, "# %u\n", V);
    for (uint64_t e = 0; e < E; ++e) std::fprintf(f, "%u %u\n", src[e], dst[e]);
    return std::fclose(f) == 0 ? DCX_OK : DCX_ERR_IO;
}

// ---- binary CSR cache (<file>-binary-g=N) -----------------------------

dcx_status dcx_binary_header(const char* path, uint32_t* V, uint32_t* E) {
    if (!path || !V || !E) return DCX_ERR_INVALID_ARG;
    std::ifstream in(path, std::ios::binary);
    if (!in) return DCX_ERR_IO;
    in.read(reinterpret_cast<char*>(V), 4);
    in.read(reinterpret_cast<char*>(E), 4);
    return in ? DCX_OK : DCX_ERR_FORMAT;
}

dcx_status dcx_binary_read(const char* path, uint32_t gap, uint32_t* out_off, int32_t* out_deg,
                           uint32_t* out_nbr, uint32_t* in_off, int32_t* in_deg,
                           uint32_t* in_nbr) {
    // Layout of Graph::readBinary (src/graph.cpp:287-325): u32 V, u32 E,
    // i32 in_deg[V], i32 out_deg[V], u32 in_off[V+1], u32 out_off[V+1],
    // u32 in_nbr[E+gap*V], u32 out_nbr[E+gap*V].
    uint32_t V, E;
    dcx_status st = dcx_binary_header(path, &V, &E);
    if (st != DCX_OK) return st;
    if (!out_off || !out_deg || !out_nbr || !in_off || !in_deg || !in_nbr)
        return DCX_ERR_INVALID_ARG;
    std::ifstream in(path, std::ios::binary);
    in.seekg(8);
    uint64_t cap = (uint64_t)E + (uint64_t)gap * V;
    in.read(reinterpret_cast<char*>(in_deg), 4ull * V);
    in.read(reinterpret_cast<char*>(out_deg), 4ull * V);
    in.read(reinterpret_cast<char*>(in_off), 4ull * (V + 1));
    in.read(reinterpret_cast<char*>(out_off), 4ull * (V + 1));
    in.read(reinterpret_cast<char*>(in_nbr), 4ull * cap);
    in.read(reinterpret_cast<char*>(out_nbr), 4ull * cap);
    if (!in) return DCX_ERR_FORMAT;
    if (in_off[V] > cap || out_off[V] > cap) return DCX_ERR_FORMAT;
    for (uint32_t v = 0; v < V; ++v)
        if (in_off[v + 1] < in_off[v] || out_off[v + 1] < out_off[v] || in_deg[v] < 0 ||
            out_deg[v] < 0 || (uint32_t)in_deg[v] > in_off[v + 1] - in_off[v] ||
            (uint32_t)out_deg[v] > out_off[v + 1] - out_off[v])
            return DCX_ERR_FORMAT;
    return DCX_OK;
}

dcx_status dcx_binary_write(const char* path, uint32_t V, uint32_t E, uint32_t gap,
                            const uint32_t* out_off, const int32_t* out_deg,
                            const uint32_t* out_nbr, const uint32_t* in_off,
                            const int32_t* in_deg, const uint32_t* in_nbr) {
    if (!path || !out_off || !out_deg || !out_nbr || !in_off || !in_deg || !in_nbr)
        return DCX_ERR_INVALID_ARG;
    std::ofstream out(path, std::ios::binary | std::ios::trunc);
    if (!out) return DCX_ERR_IO;
    uint64_t cap = (uint64_t)E + (uint64_t)gap * V;
    out.write(reinterpret_cast<const char*>(&V), 4);
    out.write(reinterpret_cast<const char*>(&E), 4);
    out.write(reinterpret_cast<const char*>(in_deg), 4ull * V);
    out.write(reinterpret_cast<const char*>(out_deg), 4ull * V);
    out.write(reinterpret_cast<const char*>(in_off), 4ull * (V + 1));
    out.write(reinterpret_cast<const char*>(out_off), 4ull * (V + 1));
    out.write(reinterpret_cast<const char*>(in_nbr), 4ull * cap);
    out.write(reinterpret_cast<const char*>(out_nbr), 4ull * cap);
    return out ? DCX_OK : DCX_ERR_IO;
}

// ---- decomposition result files ---------------------------------------

dcx_status dcx_decomp_write(const char* path, uint32_t V, int32_t kmax, int32_t lmax,
                            const int32_t* kmaxes, const int32_t* rows) {
    // .decomp (writeDecompFile, src/utils.h:98-109): i32 kmax, i32 lmax,
    // i32 kmaxes[V], then kmax+1 rows of i32[V].
    if (!path || kmax < -1 || (V && !kmaxes) || (kmax >= 0 && V && !rows))
        return DCX_ERR_INVALID_ARG;
    std::ofstream out(path, std::ios::binary | std::ios::trunc);
    if (!out) return DCX_ERR_IO;
    out.write(reinterpret_cast<const char*>(&kmax), 4);
    out.write(reinterpret_cast<const char*>(&lmax), 4);
    out.write(reinterpret_cast<const char*>(kmaxes), 4ull * V);
    out.write(reinterpret_cast<const char*>(rows), 4ull * V * (uint64_t)(kmax + 1));
    return out ? DCX_OK : DCX_ERR_IO;
}

dcx_status dcx_decomp_header(const char* path, int32_t* kmax, int32_t* lmax) {
    if (!path || !kmax || !lmax) return DCX_ERR_INVALID_ARG;
    std::ifstream in(path, std::ios::binary);
    if (!in) return DCX_ERR_IO;
    in.read(reinterpret_cast<char*>(kmax), 4);
    in.read(reinterpret_cast<char*>(lmax), 4);
    return (in && *kmax >= -1) ? DCX_OK : DCX_ERR_FORMAT;
}

dcx_status dcx_decomp_read(const char* path, uint32_t V, int32_t* kmaxes, int32_t* rows) {
    // readDecompFile (src/utils.h:76-97); V comes from the graph.
    int32_t kmax, lmax;
    dcx_status st = dcx_decomp_header(path, &kmax, &lmax);
    if (st != DCX_OK) return st;
    if ((V && !kmaxes) || (kmax >= 0 && V && !rows)) return DCX_ERR_INVALID_ARG;
    std::ifstream in(path, std::ios::binary | std::ios::ate);
    uint64_t want = 8 + 4ull * V + 4ull * V * (uint64_t)(kmax + 1);
    if ((uint64_t)in.tellg() != want) return DCX_ERR_FORMAT;
    in.seekg(8);
    in.read(reinterpret_cast<char*>(kmaxes), 4ull * V);
    in.read(reinterpret_cast<char*>(rows), 4ull * V * (uint64_t)(kmax + 1));
    return in ? DCX_OK : DCX_ERR_FORMAT;
}

dcx_status dcx_decompcpu_write(const char* path, uint32_t V, const int64_t* row_ptr,
                               const int32_t* values) {
    // .decompCPU (writeDecompFileCPU, src/utils.h:110-124): "V", then one
    // line per vertex with lmaxes[0..kmaxes[v]][v] separated by spaces; no
    // newline after the last vertex.
    if (!path || (V && (!row_ptr || !values))) return DCX_ERR_INVALID_ARG;
    FILE* f = std::fopen(path, "w");
    if (!f) return DCX_ERR_IO;
    std::fprintf(f, "%u\n", V);
    for (uint32_t v = 0; v < V; ++v) {
        for (int64_t j = row_ptr[v]; j < row_ptr[v + 1]; ++j)
            std::fprintf(f, j + 1 < row_ptr[v + 1] ? "%d " : "%d", values[j]);
        if (v + 1 != V) std::fputc('\n', f);
    }
    return std::fclose(f) == 0 ? DCX_OK : DCX_ERR_IO;
}

}  // extern "C"

// ---- edge batching (getKmaxEdgeBatches, main.cu:1441-1525) ------------

extern "C" dcx_status dcx_kmax_edge_batches(uint32_t V, const int32_t* kmaxes,
                                            const uint32_t* pairs, uint32_t n, int32_t mode,
                                            uint32_t* batch_of_edge, uint32_t* n_batches) {
    // Edge root: the endpoint with the larger kmax, the target on ties
    // (main.cu:1449-1451); edges are bucketed by the root's kmax.
    // mode 0 (KEDGE_BATCHING, the reference's active mode): within each
    // bucket, in order, repeatedly take a batch of edges with distinct roots.
    // mode 1 (KMAX_BATCHING): each batch takes the first remaining edge of
    // successive buckets whose kmax differs from the batch's first edge by
    // more than one.
    if (!n_batches || (n && (!kmaxes || !pairs || !batch_of_edge)) || (mode != 0 && mode != 1))
        return DCX_ERR_INVALID_ARG;
    std::vector<uint32_t> root(n);
    std::vector<int32_t> ek(n);
    int32_t maxk = 0;
    for (uint32_t e = 0; e < n; ++e) {
        const uint32_t a = pairs[2 * e], b = pairs[2 * e + 1];
        if (a >= V || b >= V) return DCX_ERR_INVALID_ARG;
        root[e] = kmaxes[a] > kmaxes[b] ? a : b;
        ek[e] = kmaxes[root[e]];
        if (ek[e] < 0) return DCX_ERR_INVALID_ARG;
        maxk = std::max(maxk, ek[e]);
    }
    std::vector<std::vector<uint32_t>> B(n ? (size_t)maxk + 1 : 0);
    for (uint32_t e = 0; e < n; ++e) B[ek[e]].push_back(e);
    uint32_t nb = 0;
    if (mode == 0) {
        std::vector<uint32_t> mark(V, 0xffffffffu);  // batch that used the root
        for (auto& bucket : B) {
            while (!bucket.empty()) {
                std::vector<uint32_t> rest;
                bool any = false;
                for (uint32_t e : bucket) {
                    if (!any || mark[root[e]] != nb) {
                        batch_of_edge[e] = nb;
                        mark[root[e]] = nb;
                        any = true;
                    } else {
                        rest.push_back(e);
                    }
                }
                bucket.swap(rest);
                ++nb;
            }
        }
    } else {
        std::vector<size_t> head(B.size(), 0);
        for (;;) {
            bool any = false;
            int32_t first_k = 0;
            for (size_t j = 0; j < B.size(); ++j) {
                if (head[j] >= B[j].size()) continue;
                const uint32_t e = B[j][head[j]];
                if (!any || std::abs(ek[e] - first_k) > 1) {
                    if (!any) first_k = ek[e];
                    batch_of_edge[e] = nb;
                    ++head[j];
                    any = true;
                }
            }
            if (!any) break;
            ++nb;
        }
    }
    *n_batches = nb;
    return DCX_OK;
}
