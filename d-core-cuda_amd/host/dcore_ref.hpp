// The reference's own C++ interface, with its exact global-namespace
// signatures, on top of the dcx library: a caller written against
// llysdal/d-core-cuda's main.cu compiles against this header with its
// includes replaced by this one, and runs on the GPU through the dcx C-ABI.
//
//   reference (file:line)                        here
//   typedefs degree/vertex/offset (src/common.h:59-62)      same
//   GraphInterface (src/common.h:64-79), Graph (src/graph.h:11-37)
//                                                dcx_host::GraphInterface / Graph
//   device_graph_pointers (src/common.h:81-91)   opaque handle of a device graph
//   device_accessory_pointers (:93-99)           opaque (peel scratch is the library's)
//   device_maintenance_pointers (:101-112)       opaque, bound to the graph it maintains
//   allocate*/deallocate*/moveGraphToDevice/refreshGraphOnGPU/put*/get*
//                                                (src/cuda_memory.h:5-133)
//   GraphGPU(Graph&, device_graph_pointers) (src/graph_gpu.h:10-22)
//   dcore(Graph&) (main.cu:197), KList (:144), KListInplace (:171),
//   checkKmax(Graph&, ...) (:264), checkDCore (:276), maintenance (:1536),
//   maintenanceFast (:1665), getKmaxEdgeBatches (:1441)
//   readDecompFile/writeDecompFile/writeDecompFileCPU (src/utils.h:76-124),
//   generateGraph (src/graph_gen.h:8)
//
// What the pointer structs mean here.  The library owns all device memory:
//   * allocateDeviceGraphMemory(g, g_p) uploads g (the reference allocates,
//     then moveGraphToDevice copies); moveGraphToDevice re-uploads g.
//   * The accessory and maintenance structs need no memory of their own
//     (the library grows its scratch on demand); their allocate functions
//     return 0 bytes.  refreshGraphOnGPU is a no-op: the library's passes
//     never clobber the device degrees, which the reference restores after
//     every KList.
//   * putKmaxInDeviceMemory / putLmaxInDeviceMemory record nothing: the
//     device graph keeps its own resident decomposition (computed on first
//     use by maintenance when the graph has none), which equals the host
//     copies the reference would upload.  getKmaxFromDeviceMemory /
//     getLmaxFromDeviceMemory download it.
// Every failure throws dcx_host::Error with the dcx_status instead of
// going unchecked.  The GPU is device 0, one context per process.
#pragma once

#include <chrono>
#include <string>
#include <tuple>
#include <utility>
#include <vector>

#include "dcore.hpp"

typedef dcx_host::degree degree;
typedef dcx_host::vertex vertex;
typedef dcx_host::offset offset;
using GraphInterface = dcx_host::GraphInterface;
using Graph = dcx_host::Graph;

struct device_graph_pointers {
    dcx_graph* graph = nullptr;  // device graph with its resident decomposition
    unsigned V = 0;
};
struct device_accessory_pointers {
    unsigned V = 0;
};
struct device_maintenance_pointers {
    dcx_graph* graph = nullptr;  // the graph last maintained through this struct
    unsigned V = 0;
};

// ---- device memory (src/cuda_memory.h:5-133) --------------------------
unsigned allocateDeviceGraphMemory(Graph& g, device_graph_pointers& g_p);
void deallocateDeviceGraphMemory(device_graph_pointers& g_p);
void moveGraphToDevice(Graph& g, device_graph_pointers& g_p);
void refreshGraphOnGPU(Graph& g, device_graph_pointers& g_p);
void refreshGraphOnGPU(unsigned V, device_graph_pointers& g_p);
unsigned allocateDeviceAccessoryMemory(Graph& g, device_accessory_pointers& a_p);
void deallocateDeviceAccessoryMemory(device_accessory_pointers& a_p);
unsigned allocateDeviceMaintenanceMemory(Graph& g, device_maintenance_pointers& m_p);
void deallocateDeviceMaintenanceMemory(device_maintenance_pointers& m_p);
void putKmaxInDeviceMemory(device_maintenance_pointers& m_p, std::vector<degree>& kmax);
void getKmaxFromDeviceMemory(device_maintenance_pointers& m_p, std::vector<degree>& kmax);
void putLmaxInDeviceMemory(device_maintenance_pointers& m_p, std::vector<degree>& lmax);
// Row 0 of the resident decomposition (the reference's l_max buffer holds
// one row at a time).
void getLmaxFromDeviceMemory(device_maintenance_pointers& m_p, std::vector<degree>& lmax);
void initializeDeviceMaintenanceMemory(unsigned V, device_maintenance_pointers& m_p);
void putModifiedEdgesInDeviceMemory(device_graph_pointers& g_p,
                                    std::vector<std::pair<vertex, vertex>> edges);

// ---- the device graph class (src/graph_gpu.h:10-22) -------------------
// Takes over g's kmax/kmaxes/lmax/lmaxes like the reference's constructor
// (src/graph_gpu.cu:25-33) and shares dgp's device graph.  insertEdges is a
// structural insert on the device (graphInsertEdges, src/graph_gpu.cu:49-74),
// which drops the resident decomposition; maintenanceFast inserts and
// maintains in one call.
class GraphGPU : public GraphInterface {
public:
    device_graph_pointers g_p;
    GraphGPU();
    explicit GraphGPU(Graph& g, device_graph_pointers dgp);
    void insertEdges(const std::vector<std::pair<vertex, vertex>>& edges) override;
    void insertEdgesInPlace(const std::vector<std::pair<vertex, vertex>>& edges) override;
    std::pair<vertex, vertex> getRandomInsertEdge() override;
    void deleteEdges(const std::vector<std::pair<vertex, vertex>>& edges) override;
    void deleteEdgesInPlace(const std::vector<std::pair<vertex, vertex>>& edges) override;
    std::pair<vertex, vertex> getRandomDeleteEdge() override;
    ~GraphGPU();
};

// ---- decomposition (main.cu:144-293) -----------------------------------
std::pair<degree, std::vector<degree>> KList(unsigned V, device_graph_pointers& g_p,
                                             device_accessory_pointers& a_p, unsigned k);
degree KListInplace(unsigned V, std::vector<degree>& arr, device_graph_pointers& g_p,
                    device_accessory_pointers& a_p, unsigned k);
std::vector<std::vector<degree>>& dcore(Graph& g);
// Recompute from the host graph g, as the reference does (it re-uploads g
// first, main.cu:264-293); here on a temporary device graph, so the
// maintained decomposition on g_p stays resident.
std::vector<degree> checkKmax(Graph& g, device_graph_pointers& g_p,
                              device_accessory_pointers& a_p);
std::pair<std::vector<degree>, std::vector<std::vector<degree>>> checkDCore(
    Graph& g, device_graph_pointers& g_p, device_accessory_pointers& a_p);

// ---- maintenance (main.cu:1441-1751) -----------------------------------
std::vector<std::vector<std::pair<vertex, vertex>>> getKmaxEdgeBatches(
    std::vector<degree>& kmaxes, const std::vector<std::pair<vertex, vertex>>& edgesToBeInserted);
// Applies the batch to the device graph and maintains its decomposition,
// then applies it to g (in place or by rebuild, as the reference does) and
// refreshes g.kmax/kmaxes/lmax/lmaxes from the device.
void maintenance(GraphInterface& g, device_graph_pointers& g_p, device_maintenance_pointers& m_p,
                 device_accessory_pointers& a_p,
                 std::vector<std::pair<vertex, vertex>>& modifiedEdges, bool isDelete = false,
                 bool inPlace = false);
// Insert batch on a GraphGPU: {insert, kmax, klist} times.
std::tuple<std::chrono::duration<double, std::milli>, std::chrono::duration<double, std::milli>,
           std::chrono::duration<double, std::milli>>
maintenanceFast(GraphGPU& g, device_graph_pointers& g_p, device_maintenance_pointers& m_p,
                device_accessory_pointers& a_p,
                std::vector<std::pair<vertex, vertex>>& modifiedEdges);

// ---- files and generation (src/utils.h:76-124, src/graph_gen.h:8) ------
bool readDecompFile(Graph& g, const std::string& filename);
void writeDecompFile(Graph& g, const std::string& filename);
void writeDecompFileCPU(Graph& g, const std::string& filename);
// Seeded from the clock like the reference's generator (src/graph_gen.cpp:9-11).
void generateGraph(const std::string& filename, unsigned V, unsigned E);
