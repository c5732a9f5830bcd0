// A caller written against the reference's own interface, in the shape of
// its main() (main.cu:1784-2218): the global-namespace types and functions,
// the device_*_pointers structs, allocate*Memory, dcore(Graph&), KList,
// checkKmax/checkDCore(Graph&, g_p, a_p), maintenance(g, g_p, m_p, a_p,
// edges, isDelete, inPlace), GraphGPU(g, g_p) and maintenanceFast.  The only
// change from such a caller is its include: this library's dcore_ref.hpp.
//
//   ref_caller <graph file> [golden .decomp base name]
//
// With a golden base name, dcore(g) must equal the reference's .decomp.
// Then random single inserts and deletes are maintained and checked against
// checkDCore after each (the SINGLE_INSERT_RANDOM / SINGLE_DELETE_RANDOM
// modes), and insert batches through GraphGPU + maintenanceFast against a
// second host graph (the SPEED_TEST mode with SINGlE_INSERT_CHECK).  Prints
// "found N errors" and returns 1 if N > 0.
#include "dcore_ref.hpp"

#include <iostream>
#include <string>

using namespace std;

static unsigned compareWith(GraphInterface& g, pair<vector<degree>, vector<vector<degree>>>& comp) {
    unsigned errors = 0;
    for (vertex v = 0; v < g.V; ++v) {
        if (g.kmaxes[v] != comp.first[v]) {
            cout << "mismatching kmax[" << v << "] " << g.kmaxes[v] << "!=" << comp.first[v] << endl;
            errors++;
        }
    }
    if (g.lmaxes.size() != comp.second.size()) {
        cout << "mismatching k-max sizes! " << g.lmaxes.size() << "!=" << comp.second.size()
             << endl;
        return errors + 1;
    }
    for (degree k = 0; k < (degree)g.lmaxes.size(); ++k) {
        for (vertex v = 0; v < g.V; ++v) {
            if (g.lmaxes[k][v] != comp.second[k][v]) {
                cout << "mismatching lmax[" << k << "][" << v << "] " << g.lmaxes[k][v]
                     << "!=" << comp.second[k][v] << endl;
                errors++;
            }
        }
    }
    return errors;
}

int main(int argc, char* argv[]) {
    if (argc < 2) {
        cout << "usage: ref_caller <graph file> [golden .decomp base name]" << endl;
        return 2;
    }
    const string filename = argv[1];
    unsigned totalErrors = 0;

    Graph g(filename);
    cout << "> " << filename << " V: " << g.V << " E: " << g.E << endl;

    if (!readDecompFile(g, filename)) {
        cout << "Calculating decomp file..." << endl;
        auto& res = dcore(g);
        cout << "kmax " << g.kmax << " lmax " << g.lmax << " rows " << res.size() << endl;
        writeDecompFile(g, filename);
    }
    if (argc > 2) {
        Graph golden(filename);
        if (!readDecompFile(golden, argv[2])) {
            cout << "cannot read the golden decomp " << argv[2] << endl;
            return 2;
        }
        pair<vector<degree>, vector<vector<degree>>> gold{golden.kmaxes, golden.lmaxes};
        if (g.kmax != golden.kmax || g.lmax != golden.lmax) {
            cout << "kmax/lmax " << g.kmax << "/" << g.lmax << " != golden " << golden.kmax << "/"
                 << golden.lmax << endl;
            totalErrors++;
        }
        totalErrors += compareWith(g, gold);
        cout << "compared with the golden decomposition" << endl;
    }

    device_graph_pointers g_p;
    unsigned long long graph_mem_size = allocateDeviceGraphMemory(g, g_p);
    device_accessory_pointers a_p;
    unsigned long long accessory_mem_size = allocateDeviceAccessoryMemory(g, a_p);
    device_maintenance_pointers m_p;
    unsigned long long maintenance_mem_size = allocateDeviceMaintenanceMemory(g, m_p);
    cout << "Allocated memory graph: " << graph_mem_size << " accessory: " << accessory_mem_size
         << " maintenance: " << maintenance_mem_size << endl;

    // single rows and the kmax pass on the uploaded graph
    for (degree k = 0; k <= g.kmax; ++k) {
        auto [l, core] = KList(g.V, g_p, a_p, k);
        vector<degree> arr;
        degree l2 = KListInplace(g.V, arr, g_p, a_p, k);
        for (vertex v = 0; v < g.V; ++v)
            if (core[v] != g.lmaxes[k][v] || arr[v] != g.lmaxes[k][v]) {
                cout << "KList row " << k << " differs at " << v << endl;
                totalErrors++;
                break;
            }
        if (l != l2) totalErrors++;
    }
    auto km = checkKmax(g, g_p, a_p);
    for (vertex v = 0; v < g.V; ++v)
        if (km[v] != g.kmaxes[v]) {
            cout << "checkKmax differs at " << v << endl;
            totalErrors++;
            break;
        }

    // SINGLE_INSERT_RANDOM
    unsigned inserts = 10;
    cout << "Doing " << inserts << " random insertions..." << endl;
    for (unsigned i = 0; i < inserts; ++i) {
        auto edgeToAdd = vector<pair<vertex, vertex>>();
        edgeToAdd.push_back(g.getRandomInsertEdge());
        cout << "Inserting " << edgeToAdd[0].first << "->" << edgeToAdd[0].second << endl;
        maintenance(g, g_p, m_p, a_p, edgeToAdd, false, true);
        auto comp = checkDCore(g, g_p, a_p);
        unsigned errors = compareWith(g, comp);
        cout << "found " << errors << " errors" << endl;
        totalErrors += errors;
    }

    // SINGLE_DELETE_RANDOM
    unsigned deletes = 5;
    cout << "Doing " << deletes << " random deletions..." << endl;
    for (unsigned i = 0; i < deletes; ++i) {
        auto edgeToDelete = vector<pair<vertex, vertex>>();
        edgeToDelete.push_back(g.getRandomDeleteEdge());
        cout << "Deleting " << edgeToDelete[0].first << "->" << edgeToDelete[0].second << endl;
        maintenance(g, g_p, m_p, a_p, edgeToDelete, true, true);
        auto comp = checkDCore(g, g_p, a_p);
        unsigned errors = compareWith(g, comp);
        cout << "found " << errors << " errors" << endl;
        totalErrors += errors;
    }
    vector<degree> kmaxBack, row0;
    getKmaxFromDeviceMemory(m_p, kmaxBack);
    getLmaxFromDeviceMemory(m_p, row0);
    if (kmaxBack != g.kmaxes || row0 != g.lmaxes[0]) {
        cout << "device copies differ from the host's" << endl;
        totalErrors++;
    }

    // SPEED_TEST with SINGlE_INSERT_CHECK: a second host graph follows the
    // same inserts and is recomputed by checkDCore
    Graph gc = g;
    device_graph_pointers gc_p;
    allocateDeviceGraphMemory(gc, gc_p);
    device_accessory_pointers ac_p;
    allocateDeviceAccessoryMemory(gc, ac_p);
    GraphGPU m(g, g_p);
    unsigned batches = 3, batchSize = 4;
    cout << "Doing " << batches << " insert batches of batchsize " << batchSize << "..." << endl;
    for (unsigned i = 0; i < batches; ++i) {
        vector<pair<vertex, vertex>> edgeToAdd;
        while (edgeToAdd.size() < batchSize) {
            auto e = gc.getRandomInsertEdge();
            bool dup = false;
            for (auto& x : edgeToAdd) dup |= x == e;
            if (!dup) edgeToAdd.push_back(e);
        }
        auto times = maintenanceFast(m, g_p, m_p, a_p, edgeToAdd);
        cout << "insert: " << get<0>(times).count() << "ms kmax: " << get<1>(times).count()
             << "ms klist: " << get<2>(times).count() << "ms" << endl;
        gc.insertEdgesInPlace(edgeToAdd);
        auto comp = checkDCore(gc, gc_p, ac_p);
        unsigned errors = compareWith(m, comp);
        cout << "found " << errors << " errors" << endl;
        totalErrors += errors;
    }
    auto batchesByKmax = getKmaxEdgeBatches(m.kmaxes, gc.edges);
    size_t batched = 0;
    for (auto& b : batchesByKmax) batched += b.size();
    if (batched != gc.edges.size()) totalErrors++;

    deallocateDeviceGraphMemory(gc_p);
    deallocateDeviceAccessoryMemory(ac_p);
    deallocateDeviceMaintenanceMemory(m_p);
    deallocateDeviceAccessoryMemory(a_p);
    deallocateDeviceGraphMemory(g_p);
    cout << "found " << totalErrors << " errors" << endl;
    return totalErrors ? 1 : 0;
}
