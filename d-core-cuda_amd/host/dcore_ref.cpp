// The reference's global-namespace interface over the dcx C-ABI (see
// dcore_ref.hpp for what each pointer struct means here).
#include "dcore_ref.hpp"

#include <ctime>

using dcx_host::check;
using dcx_host::Error;

namespace {

// One context for the process (device 0), never destroyed: the HIP runtime
// reclaims it at exit.
dcx_host::Device& device() {
    static dcx_host::Device* d = new dcx_host::Device(0);
    return *d;
}

dcx_ctx* ctx() { return device().ctx(); }

std::vector<uint32_t> flat(const std::vector<std::pair<vertex, vertex>>& e) {
    std::vector<uint32_t> p(2 * e.size());
    for (size_t i = 0; i < e.size(); ++i) {
        p[2 * i] = e[i].first;
        p[2 * i + 1] = e[i].second;
    }
    return p;
}

dcx_graph* upload(Graph& g) {
    dcx_graph* h = nullptr;
    check(dcx_graph_upload_csr(ctx(), g.V, g.E, g.out_neighbors_offset.data(),
                               g.out_degrees.data(), g.out_neighbors.data(),
                               g.in_neighbors_offset.data(), g.in_degrees.data(),
                               g.in_neighbors.data(), &h),
          "dcx_graph_upload_csr", ctx());
    return h;
}

dcx_graph* need(const device_graph_pointers& g_p, const char* where) {
    if (!g_p.graph)
        throw Error(DCX_ERR_STATE, std::string(where) + ": no device graph "
                                                        "(allocateDeviceGraphMemory first)");
    return g_p.graph;
}

// The resident decomposition of h, computed if the graph has none.
void ensure_decomposed(dcx_graph* h) {
    uint32_t V;
    uint64_t E;
    int32_t K, L;
    check(dcx_graph_info(h, &V, &E, &K, &L), "dcx_graph_info");
    if (K >= 0 && L >= 0) return;
    check(dcx_decompose(ctx(), h, DCX_VARIANT_AUTO, &K, &L), "dcx_decompose", ctx());
}

// kmax/kmaxes/lmax/lmaxes of g from the device graph h.
void pull(dcx_graph* h, GraphInterface& g) {
    uint32_t V;
    uint64_t E;
    int32_t K, L;
    check(dcx_graph_info(h, &V, &E, &K, &L), "dcx_graph_info");
    g.E = (unsigned)E;
    if (K < 0) return;
    g.kmax = K;
    g.lmax = L;
    g.kmaxes.resize(V);
    check(dcx_get_kmaxes(ctx(), h, g.kmaxes.data()), "dcx_get_kmaxes", ctx());
    std::vector<degree> rows((size_t)(K + 1) * V);
    check(dcx_get_rows(ctx(), h, 0, K + 1, rows.data()), "dcx_get_rows", ctx());
    g.lmaxes.assign(K + 1, std::vector<degree>(V));
    for (int32_t k = 0; k <= K; ++k)
        std::copy(rows.begin() + (size_t)k * V, rows.begin() + (size_t)(k + 1) * V,
                  g.lmaxes[k].begin());
}

uint64_t device_bytes(const Graph& g) {
    // adjacency with gap, offsets, degrees (what the reference allocates)
    return 2ull * 4 * (g.E + (uint64_t)g.gap * g.V) + 2ull * 4 * (g.V + 1) + 4ull * 4 * g.V;
}

}  // namespace

// ---- device memory -----------------------------------------------------

unsigned allocateDeviceGraphMemory(Graph& g, device_graph_pointers& g_p) {
    if (g_p.graph) dcx_graph_free(g_p.graph);
    g_p.graph = upload(g);
    g_p.V = g.V;
    return (unsigned)device_bytes(g);
}

void deallocateDeviceGraphMemory(device_graph_pointers& g_p) {
    if (g_p.graph) dcx_graph_free(g_p.graph);
    g_p.graph = nullptr;
}

void moveGraphToDevice(Graph& g, device_graph_pointers& g_p) { allocateDeviceGraphMemory(g, g_p); }

void refreshGraphOnGPU(Graph&, device_graph_pointers&) {}
void refreshGraphOnGPU(unsigned, device_graph_pointers&) {}

unsigned allocateDeviceAccessoryMemory(Graph& g, device_accessory_pointers& a_p) {
    a_p.V = g.V;
    return 0;
}
void deallocateDeviceAccessoryMemory(device_accessory_pointers&) {}

unsigned allocateDeviceMaintenanceMemory(Graph& g, device_maintenance_pointers& m_p) {
    m_p.V = g.V;
    return 0;
}
void deallocateDeviceMaintenanceMemory(device_maintenance_pointers& m_p) { m_p.graph = nullptr; }

void putKmaxInDeviceMemory(device_maintenance_pointers&, std::vector<degree>&) {}
void putLmaxInDeviceMemory(device_maintenance_pointers&, std::vector<degree>&) {}

void getKmaxFromDeviceMemory(device_maintenance_pointers& m_p, std::vector<degree>& kmax) {
    if (!m_p.graph) throw Error(DCX_ERR_STATE, "getKmaxFromDeviceMemory: nothing maintained yet");
    kmax.resize(m_p.V);
    check(dcx_get_kmaxes(ctx(), m_p.graph, kmax.data()), "dcx_get_kmaxes", ctx());
}

void getLmaxFromDeviceMemory(device_maintenance_pointers& m_p, std::vector<degree>& lmax) {
    if (!m_p.graph) throw Error(DCX_ERR_STATE, "getLmaxFromDeviceMemory: nothing maintained yet");
    lmax.resize(m_p.V);
    check(dcx_get_rows(ctx(), m_p.graph, 0, 1, lmax.data()), "dcx_get_rows", ctx());
}

void initializeDeviceMaintenanceMemory(unsigned, device_maintenance_pointers&) {}

void putModifiedEdgesInDeviceMemory(device_graph_pointers&, std::vector<std::pair<vertex, vertex>>) {
}

// ---- GraphGPU ------------------------------------------------------------

GraphGPU::GraphGPU() {}

GraphGPU::GraphGPU(Graph& g, device_graph_pointers dgp) {
    V = g.V;
    E = g.E;
    kmax = g.kmax;
    lmax = g.lmax;
    kmaxes = std::move(g.kmaxes);
    lmaxes = std::move(g.lmaxes);
    g_p = dgp;
}

GraphGPU::~GraphGPU() {}

void GraphGPU::insertEdges(const std::vector<std::pair<vertex, vertex>>& e) {
    auto p = flat(e);
    check(dcx_graph_insert_edges(ctx(), need(g_p, "GraphGPU::insertEdges"), p.data(),
                                 (uint32_t)e.size()),
          "dcx_graph_insert_edges", ctx());
    E += (unsigned)e.size();
}

void GraphGPU::insertEdgesInPlace(const std::vector<std::pair<vertex, vertex>>& e) {
    insertEdges(e);
}

void GraphGPU::deleteEdges(const std::vector<std::pair<vertex, vertex>>& e) {
    auto p = flat(e);
    check(dcx_graph_delete_edges(ctx(), need(g_p, "GraphGPU::deleteEdges"), p.data(),
                                 (uint32_t)e.size()),
          "dcx_graph_delete_edges", ctx());
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

// ---- decomposition -------------------------------------------------------

std::pair<degree, std::vector<degree>> KList(unsigned V, device_graph_pointers& g_p,
                                             device_accessory_pointers&, unsigned k) {
    std::vector<degree> row(V);
    int32_t lk = 0;
    check(dcx_klist(ctx(), need(g_p, "KList"), (int32_t)k, DCX_VARIANT_AUTO, row.data(), &lk),
          "dcx_klist", ctx());
    return {lk, row};
}

degree KListInplace(unsigned V, std::vector<degree>& arr, device_graph_pointers& g_p,
                    device_accessory_pointers&, unsigned k) {
    arr.resize(V);
    int32_t lk = 0;
    check(dcx_klist(ctx(), need(g_p, "KListInplace"), (int32_t)k, DCX_VARIANT_AUTO, arr.data(),
                    &lk),
          "dcx_klist", ctx());
    return lk;
}

std::vector<std::vector<degree>>& dcore(Graph& g) { return dcx_host::dcore(device(), g); }

std::vector<degree> checkKmax(Graph& g, device_graph_pointers&, device_accessory_pointers&) {
    dcx_graph* h = upload(g);
    std::vector<degree> km(g.V);
    int32_t K;
    dcx_status s = dcx_kmax(ctx(), h, km.data(), &K);
    dcx_graph_free(h);
    check(s, "dcx_kmax", ctx());
    return km;
}

std::pair<std::vector<degree>, std::vector<std::vector<degree>>> checkDCore(
    Graph& g, device_graph_pointers&, device_accessory_pointers&) {
    dcx_graph* h = upload(g);
    dcx_host::Graph out;  // holder of the downloaded values
    out.V = g.V;
    try {
        int32_t K, L;
        check(dcx_decompose(ctx(), h, DCX_VARIANT_AUTO, &K, &L), "dcx_decompose", ctx());
        pull(h, out);
    } catch (...) {
        dcx_graph_free(h);
        throw;
    }
    dcx_graph_free(h);
    return {std::move(out.kmaxes), std::move(out.lmaxes)};
}

// ---- maintenance -----------------------------------------------------------

std::vector<std::vector<std::pair<vertex, vertex>>> getKmaxEdgeBatches(
    std::vector<degree>& kmaxes, const std::vector<std::pair<vertex, vertex>>& edges) {
    return dcx_host::getKmaxEdgeBatches(kmaxes, edges, false);
}

void maintenance(GraphInterface& g, device_graph_pointers& g_p, device_maintenance_pointers& m_p,
                 device_accessory_pointers&, std::vector<std::pair<vertex, vertex>>& modifiedEdges,
                 bool isDelete, bool inPlace) {
    dcx_graph* h = need(g_p, "maintenance");
    ensure_decomposed(h);
    auto p = flat(modifiedEdges);
    const uint32_t n = (uint32_t)modifiedEdges.size();
    dcx_update_stats st;
    check(dcx_update(ctx(), h, isDelete ? nullptr : p.data(), isDelete ? 0 : n,
                     isDelete ? p.data() : nullptr, isDelete ? n : 0, &st),
          "dcx_update", ctx());
    m_p.graph = h;
    m_p.V = g_p.V;
    // the host graph follows, as in the reference (main.cu:1569-1576)
    if (dynamic_cast<Graph*>(&g)) {
        if (inPlace) {
            if (!isDelete)
                g.insertEdgesInPlace(modifiedEdges);
            else
                g.deleteEdgesInPlace(modifiedEdges);
        } else {
            if (!isDelete)
                g.insertEdges(modifiedEdges);
            else
                g.deleteEdges(modifiedEdges);
        }
    }
    pull(h, g);
}

std::tuple<std::chrono::duration<double, std::milli>, std::chrono::duration<double, std::milli>,
           std::chrono::duration<double, std::milli>>
maintenanceFast(GraphGPU& g, device_graph_pointers& g_p, device_maintenance_pointers& m_p,
                device_accessory_pointers&, std::vector<std::pair<vertex, vertex>>& modifiedEdges) {
    dcx_graph* h = need(g_p, "maintenanceFast");
    ensure_decomposed(h);
    auto p = flat(modifiedEdges);
    dcx_update_stats st;
    check(dcx_update(ctx(), h, p.data(), (uint32_t)modifiedEdges.size(), nullptr, 0, &st),
          "dcx_update", ctx());
    m_p.graph = h;
    m_p.V = g_p.V;
    pull(h, g);
    using ms = std::chrono::duration<double, std::milli>;
    return {ms(st.insert_ms), ms(st.kmax_ms), ms(st.rows_ms)};
}

// ---- files and generation -------------------------------------------------

bool readDecompFile(Graph& g, const std::string& filename) {
    return dcx_host::readDecompFile(g, filename);
}
void writeDecompFile(Graph& g, const std::string& filename) {
    dcx_host::writeDecompFile(g, filename);
}
void writeDecompFileCPU(Graph& g, const std::string& filename) {
    dcx_host::writeDecompFileCPU(g, filename);
}
void generateGraph(const std::string& filename, unsigned V, unsigned E) {
    dcx_host::generateGraph(filename, V, E, (uint64_t)std::time(nullptr));
}
