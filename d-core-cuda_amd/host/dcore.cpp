// C++ host API over the dcx C-ABI; see dcore.hpp for the mapping to the
// reference's interface.
#include "dcore.hpp"

#include <algorithm>
#include <fstream>
#include <map>
#include <numeric>
#include <set>

namespace dcx_host {

void check(dcx_status s, const char* where, const dcx_ctx* ctx) {
    if (s == DCX_OK) return;
    std::string msg = std::string(where) + ": " + dcx_status_str(s);
    if (ctx) {
        const char* d = dcx_ctx_last_error(ctx);
        if (d && *d) msg += std::string(" (") + d + ")";
    }
    throw Error(s, msg);
}

Device::Device(int device, dcx_variant variant) {
    dcx_config cfg;
    check(dcx_config_init(&cfg), "dcx_config_init");
    cfg.device = device;
    cfg.variant = variant;
    check(dcx_ctx_create(&cfg, &ctx_), "dcx_ctx_create");
}

Device::~Device() {
    if (ctx_) dcx_ctx_destroy(ctx_);
}

dcx_timing Device::lastTiming() const {
    dcx_timing t;
    check(dcx_last_timing(ctx_, &t), "dcx_last_timing", ctx_);
    return t;
}

static void flat_pairs(const std::vector<std::pair<vertex, vertex>>& e, std::vector<uint32_t>& p) {
    p.resize(2 * e.size());
    for (size_t i = 0; i < e.size(); ++i) {
        p[2 * i] = e[i].first;
        p[2 * i + 1] = e[i].second;
    }
}

// ---- Graph -----------------------------------------------------------

void Graph::rebuild(unsigned nV, const std::vector<std::pair<vertex, vertex>>& e) {
    V = nV;
    E = (unsigned)e.size();
    edges = e;
    std::vector<uint32_t> src(E), dst(E);
    for (unsigned i = 0; i < E; ++i) {
        src[i] = e[i].first;
        dst[i] = e[i].second;
    }
    const uint64_t cap = (uint64_t)E + (uint64_t)gap * V;
    out_neighbors_offset.assign(V + 1, 0);
    in_neighbors_offset.assign(V + 1, 0);
    out_degrees.assign(V, 0);
    in_degrees.assign(V, 0);
    out_neighbors.assign(cap ? cap : 1, 0);
    in_neighbors.assign(cap ? cap : 1, 0);
    check(dcx_csr_build(V, E, src.data(), dst.data(), gap, out_neighbors_offset.data(),
                        out_degrees.data(), out_neighbors.data(), in_neighbors_offset.data(),
                        in_degrees.data(), in_neighbors.data()),
          "dcx_csr_build");
    kmaxes.assign(V, 0);
    lmaxes.assign(1, std::vector<degree>(V, 0));
}

Graph::Graph(unsigned nV, unsigned g) : gap(g) { rebuild(nV, {}); }

Graph::Graph(unsigned nV, const std::vector<std::pair<vertex, vertex>>& e, unsigned g) : gap(g) {
    rebuild(nV, e);
}

Graph::Graph(const std::string& inputFile, unsigned g, bool writeCache) : gap(g) {
    if (readBinary(inputFile)) return;
    readFile(inputFile, true);
    if (writeCache) writeBinary(inputFile);
}

void Graph::readFile(const std::string& inputFile, bool prune) {
    uint32_t Vh = 0;
    uint64_t lines = 0;
    check(dcx_text_count(inputFile.c_str(), &Vh, &lines), "dcx_text_count");
    std::vector<uint32_t> s(lines ? lines : 1), d(lines ? lines : 1);
    uint32_t nV = 0;
    uint64_t nE = 0;
    check(dcx_text_read(inputFile.c_str(), prune ? 1 : 0, s.data(), d.data(), lines, &nV, &nE),
          "dcx_text_read");
    std::vector<std::pair<vertex, vertex>> e(nE);
    for (uint64_t i = 0; i < nE; ++i) e[i] = {s[i], d[i]};
    rebuild(nV, e);
}

static std::string binary_name(const std::string& f, unsigned gap) {
    return gap == 0 ? f + "-binary" : f + "-binary-g=" + std::to_string(gap);
}

bool Graph::readBinary(const std::string& inputFile) {
    const std::string p = binary_name(inputFile, gap);
    uint32_t nV = 0, nE = 0;
    if (dcx_binary_header(p.c_str(), &nV, &nE) != DCX_OK) return false;
    const uint64_t cap = (uint64_t)nE + (uint64_t)gap * nV;
    V = nV;
    E = nE;
    out_neighbors_offset.assign(V + 1, 0);
    in_neighbors_offset.assign(V + 1, 0);
    out_degrees.assign(V, 0);
    in_degrees.assign(V, 0);
    out_neighbors.assign(cap ? cap : 1, 0);
    in_neighbors.assign(cap ? cap : 1, 0);
    check(dcx_binary_read(p.c_str(), gap, out_neighbors_offset.data(), out_degrees.data(),
                          out_neighbors.data(), in_neighbors_offset.data(), in_degrees.data(),
                          in_neighbors.data()),
          "dcx_binary_read");
    edges.clear();
    for (vertex v = 0; v < V; ++v)
        for (degree j = 0; j < out_degrees[v]; ++j)
            edges.emplace_back(v, out_neighbors[out_neighbors_offset[v] + j]);
    kmaxes.assign(V, 0);
    lmaxes.assign(1, std::vector<degree>(V, 0));
    return true;
}

void Graph::writeBinary(const std::string& inputFile) const {
    check(dcx_binary_write(binary_name(inputFile, gap).c_str(), V, E, gap,
                           out_neighbors_offset.data(), out_degrees.data(), out_neighbors.data(),
                           in_neighbors_offset.data(), in_degrees.data(), in_neighbors.data()),
          "dcx_binary_write");
}

void Graph::writeFile(const std::string& outputFile) const {
    std::vector<uint32_t> s(edges.size()), d(edges.size());
    for (size_t i = 0; i < edges.size(); ++i) {
        s[i] = edges[i].first;
        d[i] = edges[i].second;
    }
    check(dcx_text_write((outputFile + "-pruned").c_str(), V, edges.size(), s.data(), d.data()),
          "dcx_text_write");
}

void Graph::insertEdges(const std::vector<std::pair<vertex, vertex>>& e) {
    std::vector<std::pair<vertex, vertex>> all = edges;
    all.insert(all.end(), e.begin(), e.end());
    rebuild(V, all);
}

void Graph::insertEdgesInPlace(const std::vector<std::pair<vertex, vertex>>& e) {
    // Uses the free slots like src/graph.cpp:63-70, but rebuilds with the
    // same gap when a list lacks room for all of the batch's edges into it
    // instead of overwriting its neighbour.
    std::map<vertex, unsigned> need_out, need_in;
    for (auto& [u, v] : e) {
        if (u >= V || v >= V) throw Error(DCX_ERR_INVALID_ARG, "insertEdgesInPlace: id >= V");
        ++need_out[u];
        ++need_in[v];
    }
    auto room = [](const std::vector<offset>& off, const std::vector<degree>& deg, vertex x,
                   unsigned n) { return off[x] + deg[x] + n <= off[x + 1]; };
    for (auto& [u, n] : need_out)
        if (!room(out_neighbors_offset, out_degrees, u, n)) {
            insertEdges(e);
            return;
        }
    for (auto& [v, n] : need_in)
        if (!room(in_neighbors_offset, in_degrees, v, n)) {
            insertEdges(e);
            return;
        }
    for (auto& [u, v] : e) {
        out_neighbors[out_neighbors_offset[u] + out_degrees[u]++] = v;
        in_neighbors[in_neighbors_offset[v] + in_degrees[v]++] = u;
        edges.emplace_back(u, v);
        ++E;
    }
}

void Graph::deleteEdges(const std::vector<std::pair<vertex, vertex>>& e) {
    std::set<std::pair<vertex, vertex>> del(e.begin(), e.end());
    std::vector<std::pair<vertex, vertex>> keep;
    size_t found = 0;
    for (auto& x : edges) {
        if (del.count(x))
            ++found;
        else
            keep.push_back(x);
    }
    if (found != del.size()) throw Error(DCX_ERR_NOT_FOUND, "deleteEdges: edge not in graph");
    rebuild(V, keep);
}

void Graph::deleteEdgesInPlace(const std::vector<std::pair<vertex, vertex>>& e) {
    // The design commented out at src/graph.cpp:78-108, done as
    // swap-with-last so lists stay dense.
    for (auto& [u, v] : e) {
        auto drop = [](std::vector<vertex>& nbr, offset start, degree& deg, vertex x) {
            for (degree j = 0; j < deg; ++j)
                if (nbr[start + j] == x) {
                    nbr[start + j] = nbr[start + deg - 1];
                    --deg;
                    return true;
                }
            return false;
        };
        if (!drop(out_neighbors, out_neighbors_offset[u], out_degrees[u], v) ||
            !drop(in_neighbors, in_neighbors_offset[v], in_degrees[v], u))
            throw Error(DCX_ERR_NOT_FOUND, "deleteEdgesInPlace: edge not in graph");
        edges.erase(std::find(edges.begin(), edges.end(), std::make_pair(u, v)));
        --E;
    }
}

static uint64_t next_rand(uint64_t& s) {
    uint64_t z = (s += 0x9e3779b97f4a7c15ULL);
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ULL;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebULL;
    return z ^ (z >> 31);
}

std::pair<vertex, vertex> Graph::getRandomInsertEdge() {
    // An absent, non-loop pair (src/graph.cpp:327-366 without its
    // one-use-per-vertex restriction, which only protected the gap).
    if (V < 2) throw Error(DCX_ERR_STATE, "getRandomInsertEdge: fewer than 2 vertices");
    for (int attempt = 0; attempt < 1000000; ++attempt) {
        vertex u = (vertex)(next_rand(rng_) % V), v = (vertex)(next_rand(rng_) % V);
        if (u == v) continue;
        bool present = false;
        for (degree j = 0; j < out_degrees[u] && !present; ++j)
            present = out_neighbors[out_neighbors_offset[u] + j] == v;
        if (!present) return {u, v};
    }
    throw Error(DCX_ERR_STATE, "getRandomInsertEdge: graph is (nearly) complete");
}

std::pair<vertex, vertex> Graph::getRandomDeleteEdge() {
    if (edges.empty()) throw Error(DCX_ERR_STATE, "getRandomDeleteEdge: no edges");
    return edges[next_rand(rng_) % edges.size()];
}

// ---- GraphGPU --------------------------------------------------------

GraphGPU::GraphGPU(Device& dev, Graph& g) : dev_(&dev) {
    V = g.V;
    E = g.E;
    check(dcx_graph_upload_csr(dev.ctx(), g.V, g.E, g.out_neighbors_offset.data(),
                               g.out_degrees.data(), g.out_neighbors.data(),
                               g.in_neighbors_offset.data(), g.in_degrees.data(),
                               g.in_neighbors.data(), &g_),
          "dcx_graph_upload_csr", dev.ctx());
    kmax = g.kmax;
    lmax = g.lmax;
    kmaxes = std::move(g.kmaxes);  // as the reference's constructor does
    lmaxes = std::move(g.lmaxes);
}

GraphGPU::~GraphGPU() {
    if (g_) dcx_graph_free(g_);
}

void GraphGPU::pull() {
    uint32_t nV;
    uint64_t nE;
    int32_t K, L;
    check(dcx_graph_info(g_, &nV, &nE, &K, &L), "dcx_graph_info");
    E = (unsigned)nE;
    if (K < 0) return;
    kmax = K;
    lmax = L;
    kmaxes.resize(V);
    check(dcx_get_kmaxes(dev_->ctx(), g_, kmaxes.data()), "dcx_get_kmaxes", dev_->ctx());
    if (L < 0) return;
    std::vector<degree> rows((size_t)(K + 1) * V);
    check(dcx_get_rows(dev_->ctx(), g_, 0, K + 1, rows.data()), "dcx_get_rows", dev_->ctx());
    lmaxes.assign(K + 1, std::vector<degree>(V));
    for (int32_t k = 0; k <= K; ++k)
        std::copy(rows.begin() + (size_t)k * V, rows.begin() + (size_t)(k + 1) * V,
                  lmaxes[k].begin());
}

void GraphGPU::insertEdges(const std::vector<std::pair<vertex, vertex>>& e) {
    std::vector<uint32_t> p;
    flat_pairs(e, p);
    check(dcx_graph_insert_edges(dev_->ctx(), g_, p.data(), (uint32_t)e.size()),
          "dcx_graph_insert_edges", dev_->ctx());
    E += (unsigned)e.size();
}

void GraphGPU::insertEdgesInPlace(const std::vector<std::pair<vertex, vertex>>& e) {
    insertEdges(e);
}

void GraphGPU::deleteEdges(const std::vector<std::pair<vertex, vertex>>& e) {
    std::vector<uint32_t> p;
    flat_pairs(e, p);
    check(dcx_graph_delete_edges(dev_->ctx(), g_, p.data(), (uint32_t)e.size()),
          "dcx_graph_delete_edges", dev_->ctx());
    E -= (unsigned)e.size();
}

void GraphGPU::deleteEdgesInPlace(const std::vector<std::pair<vertex, vertex>>& e) {
    deleteEdges(e);
}

std::pair<vertex, vertex> GraphGPU::getRandomInsertEdge() {
    throw Error(DCX_ERR_STATE, "GraphGPU::getRandomInsertEdge: use the host Graph");
}

std::pair<vertex, vertex> GraphGPU::getRandomDeleteEdge() {
    throw Error(DCX_ERR_STATE, "GraphGPU::getRandomDeleteEdge: use the host Graph");
}

// ---- decomposition ---------------------------------------------------

std::vector<std::vector<degree>>& dcore(Device& dev, Graph& g, dcx_variant variant) {
    GraphGPU gg(dev, g);
    int32_t K, L;
    check(dcx_decompose(dev.ctx(), gg.handle(), variant, &K, &L), "dcx_decompose", dev.ctx());
    gg.pull();
    g.kmax = gg.kmax;
    g.lmax = gg.lmax;
    g.kmaxes = std::move(gg.kmaxes);
    g.lmaxes = std::move(gg.lmaxes);
    return g.lmaxes;
}

std::pair<degree, std::vector<degree>> KList(GraphGPU& g, unsigned k, dcx_variant variant) {
    std::vector<degree> row(g.V);
    int32_t lk = 0;
    check(dcx_klist(g.device().ctx(), g.handle(), (int32_t)k, variant, row.data(), &lk),
          "dcx_klist", g.device().ctx());
    return {lk, row};
}

degree KListInplace(GraphGPU& g, std::vector<degree>& arr, unsigned k, dcx_variant variant) {
    arr.resize(g.V);
    int32_t lk = 0;
    check(dcx_klist(g.device().ctx(), g.handle(), (int32_t)k, variant, arr.data(), &lk),
          "dcx_klist", g.device().ctx());
    return lk;
}

std::vector<degree> checkKmax(GraphGPU& g) {
    std::vector<degree> km(g.V);
    int32_t K;
    check(dcx_kmax(g.device().ctx(), g.handle(), km.data(), &K), "dcx_kmax", g.device().ctx());
    return km;
}

std::pair<std::vector<degree>, std::vector<std::vector<degree>>> checkDCore(GraphGPU& g) {
    int32_t K, L;
    check(dcx_decompose(g.device().ctx(), g.handle(), DCX_VARIANT_AUTO, &K, &L), "dcx_decompose",
          g.device().ctx());
    g.pull();
    return {g.kmaxes, g.lmaxes};
}

// ---- maintenance -----------------------------------------------------

void maintenance(GraphGPU& g, const std::vector<std::pair<vertex, vertex>>& modifiedEdges,
                 bool isDelete) {
    std::vector<uint32_t> p;
    flat_pairs(modifiedEdges, p);
    dcx_update_stats st;
    const uint32_t n = (uint32_t)modifiedEdges.size();
    check(dcx_update(g.device().ctx(), g.handle(), isDelete ? nullptr : p.data(), isDelete ? 0 : n,
                     isDelete ? p.data() : nullptr, isDelete ? n : 0, &st),
          "dcx_update", g.device().ctx());
    g.pull();
}

std::tuple<std::chrono::duration<double, std::milli>, std::chrono::duration<double, std::milli>,
           std::chrono::duration<double, std::milli>>
maintenanceFast(GraphGPU& g, const std::vector<std::pair<vertex, vertex>>& modifiedEdges) {
    std::vector<uint32_t> p;
    flat_pairs(modifiedEdges, p);
    dcx_update_stats st;
    check(dcx_update(g.device().ctx(), g.handle(), p.data(), (uint32_t)modifiedEdges.size(),
                     nullptr, 0, &st),
          "dcx_update", g.device().ctx());
    using ms = std::chrono::duration<double, std::milli>;
    return {ms(st.insert_ms), ms(st.kmax_ms), ms(st.rows_ms)};
}

std::vector<std::vector<std::pair<vertex, vertex>>> getKmaxEdgeBatches(
    const std::vector<degree>& kmaxes, const std::vector<std::pair<vertex, vertex>>& edges,
    bool kmaxBatching) {
    std::vector<uint32_t> p, batch(edges.size());
    flat_pairs(edges, p);
    uint32_t nb = 0;
    check(dcx_kmax_edge_batches((uint32_t)kmaxes.size(), kmaxes.data(), p.data(),
                                (uint32_t)edges.size(), kmaxBatching ? 1 : 0, batch.data(), &nb),
          "dcx_kmax_edge_batches");
    std::vector<std::vector<std::pair<vertex, vertex>>> out(nb);
    for (size_t i = 0; i < edges.size(); ++i) out[batch[i]].push_back(edges[i]);
    return out;
}

// ---- files -----------------------------------------------------------

void generateGraph(const std::string& outputFile, unsigned V, unsigned E, uint64_t seed) {
    std::vector<uint32_t> s(E), d(E);
    check(dcx_gen_er(V, E, seed, s.data(), d.data()), "dcx_gen_er");
    check(dcx_text_write(outputFile.c_str(), V, E, s.data(), d.data()), "dcx_text_write");
}

bool readDecompFile(GraphInterface& g, const std::string& filename) {
    const std::string p = filename + ".decomp";
    int32_t K, L;
    if (dcx_decomp_header(p.c_str(), &K, &L) != DCX_OK) return false;
    std::vector<degree> rows((size_t)(K + 1) * g.V);
    g.kmaxes.resize(g.V);
    check(dcx_decomp_read(p.c_str(), g.V, g.kmaxes.data(), rows.data()), "dcx_decomp_read");
    g.kmax = K;
    g.lmax = L;
    g.lmaxes.assign(K + 1, std::vector<degree>(g.V));
    for (int32_t k = 0; k <= K; ++k)
        std::copy(rows.begin() + (size_t)k * g.V, rows.begin() + (size_t)(k + 1) * g.V,
                  g.lmaxes[k].begin());
    return true;
}

void writeDecompFile(const GraphInterface& g, const std::string& filename) {
    std::vector<degree> rows;
    for (auto& r : g.lmaxes) rows.insert(rows.end(), r.begin(), r.end());
    check(dcx_decomp_write((filename + ".decomp").c_str(), g.V, g.kmax, g.lmax, g.kmaxes.data(),
                           rows.data()),
          "dcx_decomp_write");
}

void writeDecompFileCPU(const GraphInterface& g, const std::string& filename) {
    std::vector<int64_t> rp(g.V + 1, 0);
    std::vector<int32_t> vals;
    for (vertex v = 0; v < g.V; ++v) {
        for (degree k = 0; k <= g.kmaxes[v]; ++k) vals.push_back(g.lmaxes[k][v]);
        rp[v + 1] = (int64_t)vals.size();
    }
    check(dcx_decompcpu_write((filename + ".decompCPU").c_str(), g.V, rp.data(), vals.data()),
          "dcx_decompcpu_write");
}

}  // namespace dcx_host
