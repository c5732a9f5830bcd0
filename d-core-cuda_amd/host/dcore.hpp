// C++ host API of the D-core library: the reference's interface for the
// decomposition and maintenance path, implemented on the dcx C-ABI
// (include/dcx.h).  Names, argument meaning and result shapes follow the
// reference so that a caller of llysdal/d-core-cuda can switch over:
//
//   reference                                     here
//   GraphInterface (src/common.h:64-79)           dcx_host::GraphInterface
//   Graph (src/graph.h:11-37)                     dcx_host::Graph (host CSR + files)
//   GraphGPU (src/graph_gpu.h:10-22)              dcx_host::GraphGPU (device-resident)
//   dcore(Graph&) (main.cu:197-262)               dcx_host::dcore(Device&, Graph&)
//   KList / KListInplace (main.cu:144-195)        dcx_host::KList / KListInplace
//   checkKmax / checkDCore (main.cu:264-293)      dcx_host::checkKmax / checkDCore
//   maintenance / maintenanceFast (main.cu:1536-1751)  dcx_host::maintenance / maintenanceFast
//   getKmaxEdgeBatches (main.cu:1441-1525)        dcx_host::getKmaxEdgeBatches
//   generateGraph (src/graph_gen.h:8)             dcx_host::generateGraph (seeded)
//   read/writeDecompFile(CPU) (src/utils.h:76-124) dcx_host::readDecompFile / writeDecompFile(CPU)
//
// This namespace is the library's own C++ API: device state lives behind
// GraphGPU (no device_*_pointers structs to pass around); every failure
// throws dcx_host::Error carrying the dcx_status instead of going unchecked;
// edge deletion works; generateGraph takes a seed.  dcore_ref.hpp puts the
// reference's exact global-namespace signatures (dcore(Graph&), KList(V,
// g_p, a_p, k), maintenance(g, g_p, m_p, a_p, ...), the allocate* family)
// on top of it, for callers written against the reference.
#pragma once

#include <chrono>
#include <memory>
#include <stdexcept>
#include <string>
#include <tuple>
#include <utility>
#include <vector>

#include "dcx.h"

namespace dcx_host {

typedef int degree;       // src/common.h:59
typedef unsigned vertex;  // src/common.h:61
typedef unsigned offset;  // src/common.h:62

class Error : public std::runtime_error {
public:
    Error(dcx_status s, const std::string& what) : std::runtime_error(what), status(s) {}
    dcx_status status;
};

// Throws Error if s != DCX_OK (detail from the context when given).
void check(dcx_status s, const char* where, const dcx_ctx* ctx = nullptr);

// One device context (stream, scratch) shared by the graphs of a process.
class Device {
public:
    explicit Device(int device = 0, dcx_variant variant = DCX_VARIANT_AUTO);
    ~Device();
    Device(const Device&) = delete;
    Device& operator=(const Device&) = delete;
    dcx_ctx* ctx() const { return ctx_; }
    dcx_timing lastTiming() const;

private:
    dcx_ctx* ctx_ = nullptr;
};

class GraphInterface {
public:
    unsigned V = 0;
    unsigned E = 0;
    degree kmax = 0;
    std::vector<degree> kmaxes;
    degree lmax = 0;
    std::vector<std::vector<degree>> lmaxes;
    virtual void insertEdges(const std::vector<std::pair<vertex, vertex>>& edges) = 0;
    virtual void insertEdgesInPlace(const std::vector<std::pair<vertex, vertex>>& edges) = 0;
    virtual void deleteEdges(const std::vector<std::pair<vertex, vertex>>& edges) = 0;
    virtual void deleteEdgesInPlace(const std::vector<std::pair<vertex, vertex>>& edges) = 0;
    virtual std::pair<vertex, vertex> getRandomInsertEdge() = 0;
    virtual std::pair<vertex, vertex> getRandomDeleteEdge() = 0;
    virtual ~GraphInterface() {}
};

// Host graph: CSR with `gap` free slots per vertex, the layout of
// Graph::readFile (src/graph.cpp:180-243).
class Graph : public GraphInterface {
public:
    unsigned gap = 1;  // OFFSET_GAP (src/common.h:37)
    std::vector<std::pair<vertex, vertex>> edges;
    std::vector<vertex> in_neighbors, out_neighbors;
    std::vector<offset> in_neighbors_offset, out_neighbors_offset;
    std::vector<degree> in_degrees, out_degrees;

    Graph() = default;
    // Loads "<file>-binary-g=<gap>" if present, else the text edge list
    // (pruning isolated vertices like the reference), then writes the
    // binary cache (src/graph.cpp:381-395).
    explicit Graph(const std::string& inputFile, unsigned gap = 1, bool writeCache = true);
    explicit Graph(unsigned V, unsigned gap = 1);  // no edges (src/graph.cpp:397-412)
    Graph(unsigned V, const std::vector<std::pair<vertex, vertex>>& edges, unsigned gap = 1);

    void readFile(const std::string& inputFile, bool prune = true);
    bool readBinary(const std::string& inputFile);
    void writeBinary(const std::string& inputFile) const;
    void writeFile(const std::string& outputFile) const;

    void insertEdges(const std::vector<std::pair<vertex, vertex>>& edges) override;
    void insertEdgesInPlace(const std::vector<std::pair<vertex, vertex>>& edges) override;
    void deleteEdges(const std::vector<std::pair<vertex, vertex>>& edges) override;
    void deleteEdgesInPlace(const std::vector<std::pair<vertex, vertex>>& edges) override;
    std::pair<vertex, vertex> getRandomInsertEdge() override;
    std::pair<vertex, vertex> getRandomDeleteEdge() override;

    void seed(uint64_t s) { rng_ = s; }

private:
    void rebuild(unsigned V, const std::vector<std::pair<vertex, vertex>>& e);
    uint64_t rng_ = 0x5eed;
};

// Device-resident graph with its resident decomposition.  Constructed
// from a host Graph; like the reference's GraphGPU constructor
// (src/graph_gpu.cu:25-33) it takes over kmax/kmaxes/lmax/lmaxes.
class GraphGPU : public GraphInterface {
public:
    GraphGPU(Device& dev, Graph& g);
    ~GraphGPU();
    GraphGPU(const GraphGPU&) = delete;
    GraphGPU& operator=(const GraphGPU&) = delete;

    Device& device() const { return *dev_; }
    dcx_graph* handle() const { return g_; }
    // Refreshes kmax/kmaxes/lmax/lmaxes from the device.
    void pull();

    void insertEdges(const std::vector<std::pair<vertex, vertex>>& edges) override;
    void insertEdgesInPlace(const std::vector<std::pair<vertex, vertex>>& edges) override;
    void deleteEdges(const std::vector<std::pair<vertex, vertex>>& edges) override;
    void deleteEdgesInPlace(const std::vector<std::pair<vertex, vertex>>& edges) override;
    std::pair<vertex, vertex> getRandomInsertEdge() override;
    std::pair<vertex, vertex> getRandomDeleteEdge() override;

private:
    Device* dev_;
    dcx_graph* g_ = nullptr;
};

// ---- decomposition --------------------------------------------------

// Full decomposition on the device; fills g.kmax, g.kmaxes, g.lmax,
// g.lmaxes and returns g.lmaxes (dcore, main.cu:197-262).
std::vector<std::vector<degree>>& dcore(Device& dev, Graph& g,
                                        dcx_variant variant = DCX_VARIANT_AUTO);
// One row k on the whole graph (KList main.cu:144-169): {lmax_k, row}.
std::pair<degree, std::vector<degree>> KList(GraphGPU& g, unsigned k,
                                             dcx_variant variant = DCX_VARIANT_AUTO);
// Same, written into arr (KListInplace main.cu:171-195).
degree KListInplace(GraphGPU& g, std::vector<degree>& arr, unsigned k,
                    dcx_variant variant = DCX_VARIANT_AUTO);
// In-core numbers (checkKmax main.cu:264-274).
std::vector<degree> checkKmax(GraphGPU& g);
// Recomputation from scratch for verification (checkDCore main.cu:276-293).
std::pair<std::vector<degree>, std::vector<std::vector<degree>>> checkDCore(GraphGPU& g);

// ---- maintenance ----------------------------------------------------

// Applies an insert (or delete) batch and maintains the decomposition
// (maintenance main.cu:1536-1663); afterwards g's host copies are current.
void maintenance(GraphGPU& g, const std::vector<std::pair<vertex, vertex>>& modifiedEdges,
                 bool isDelete = false);
// Same with the {insert, kmax, klist} time split (maintenanceFast
// main.cu:1665-1751).
std::tuple<std::chrono::duration<double, std::milli>, std::chrono::duration<double, std::milli>,
           std::chrono::duration<double, std::milli>>
maintenanceFast(GraphGPU& g, const std::vector<std::pair<vertex, vertex>>& modifiedEdges);
// Edge batches by root kmax (main.cu:1441-1525); kmaxBatching selects the
// KMAX_BATCHING mode instead of KEDGE_BATCHING.
std::vector<std::vector<std::pair<vertex, vertex>>> getKmaxEdgeBatches(
    const std::vector<degree>& kmaxes, const std::vector<std::pair<vertex, vertex>>& edges,
    bool kmaxBatching = false);

// ---- files and generation -------------------------------------------

// Seeded G(n,m) written as "# V" + sorted "s t" lines (src/graph_gen.cpp).
void generateGraph(const std::string& outputFile, unsigned V, unsigned E, uint64_t seed = 1);
bool readDecompFile(GraphInterface& g, const std::string& filename);  // <filename>.decomp
void writeDecompFile(const GraphInterface& g, const std::string& filename);
void writeDecompFileCPU(const GraphInterface& g, const std::string& filename);  // .decompCPU

}  // namespace dcx_host
