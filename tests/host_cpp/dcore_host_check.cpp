// Drives the C++ host API (d-core-cuda_amd/host/dcore.hpp) the way the
// reference's main.cu drives its own (main.cu:1820-1925): load a graph,
// decompose, read and write the result files, run single-row passes,
// maintain under inserts and deletes, and check against a recomputation.
//
//   dcore_host_check files  <dir>   file formats and host-only functions
//   dcore_host_check device <dir>   everything that needs the GPU
//
// <dir> holds copies of the golden files digraph, digraph-binary-g=1 and
// digraph.decomp (written by the reference itself).  Prints one line per
// failed check and returns the number of failures (0 = all passed).
#include <algorithm>
#include <cstdio>
#include <fstream>
#include <iterator>
#include <set>
#include <sstream>
#include <string>
#include <vector>

#include "dcore.hpp"

using namespace dcx_host;

static int failures = 0, checks = 0;
#define CHECK(cond)                                                        \
    do {                                                                   \
        ++checks;                                                          \
        if (!(cond)) {                                                     \
            ++failures;                                                    \
            std::printf("FAIL %s:%d: %s\n", __FILE__, __LINE__, #cond);    \
        }                                                                  \
    } while (0)

static std::string slurp(const std::string& p) {
    std::ifstream f(p, std::ios::binary);
    return std::string(std::istreambuf_iterator<char>(f), std::istreambuf_iterator<char>());
}

static void copy_file(const std::string& from, const std::string& to) {
    std::ofstream(to, std::ios::binary) << slurp(from);
}

// Degree sums, offsets and edge list of a host graph agree.
static bool consistent(const Graph& g) {
    if (g.edges.size() != g.E || g.out_degrees.size() != g.V) return false;
    uint64_t so = 0, si = 0;
    for (unsigned v = 0; v < g.V; ++v) {
        so += g.out_degrees[v];
        si += g.in_degrees[v];
    }
    if (so != g.E || si != g.E) return false;
    // every list within its slots and holding exactly the graph's edges
    std::multiset<std::pair<vertex, vertex>> want(g.edges.begin(), g.edges.end()), outs, ins;
    for (unsigned v = 0; v < g.V; ++v) {
        if (g.out_neighbors_offset[v] + g.out_degrees[v] > g.out_neighbors_offset[v + 1] ||
            g.in_neighbors_offset[v] + g.in_degrees[v] > g.in_neighbors_offset[v + 1])
            return false;
        for (degree j = 0; j < g.out_degrees[v]; ++j)
            outs.emplace(v, g.out_neighbors[g.out_neighbors_offset[v] + j]);
        for (degree j = 0; j < g.in_degrees[v]; ++j)
            ins.emplace(g.in_neighbors[g.in_neighbors_offset[v] + j], v);
    }
    return outs == want && ins == want;
}

static std::vector<std::pair<vertex, vertex>> random_new_edges(Graph& g, unsigned n) {
    std::set<std::pair<vertex, vertex>> have(g.edges.begin(), g.edges.end()), picked;
    std::vector<std::pair<vertex, vertex>> out;
    while (out.size() < n) {
        const auto e = g.getRandomInsertEdge();
        if (e.first == e.second || have.count(e) || picked.count(e)) continue;
        picked.insert(e);
        out.push_back(e);
    }
    return out;
}

static void files(const std::string& dir) {
    // text loading (pruning isolated vertices) writes the binary cache; it
    // has the reference's layout (same size; the reference leaves the gap
    // slots uninitialised, so those bytes differ) and reads back the same
    const std::string t = dir + "/text_only_digraph";
    copy_file(dir + "/digraph", t);
    Graph gt(t, 1, true);
    CHECK(slurp(t + "-binary-g=1").size() == slurp(dir + "/digraph-binary-g=1").size());
    Graph gt2(t, 1, false);  // now from the cache just written
    Graph gb(dir + "/digraph");  // the reference's binary cache
    for (const Graph* x : {&gt2, &gb}) {
        CHECK(x->V == gt.V && x->E == gt.E);
        CHECK(x->out_neighbors_offset == gt.out_neighbors_offset);
        CHECK(x->in_neighbors_offset == gt.in_neighbors_offset);
        CHECK(x->in_degrees == gt.in_degrees && x->out_degrees == gt.out_degrees);
        bool lists = true;
        for (unsigned v = 0; v < gt.V; ++v)
            for (int j = 0; j < gt.out_degrees[v]; ++j)
                lists = lists && x->out_neighbors[gt.out_neighbors_offset[v] + j] ==
                                     gt.out_neighbors[gt.out_neighbors_offset[v] + j];
        CHECK(lists);
    }
    CHECK(consistent(gb));

    // the reference's decomposition file: read, write back byte for byte
    CHECK(readDecompFile(gb, dir + "/digraph"));
    CHECK(gb.kmaxes.size() == gb.V && (int)gb.lmaxes.size() == gb.kmax + 1);
    writeDecompFile(gb, dir + "/out");
    CHECK(slurp(dir + "/out.decomp") == slurp(dir + "/digraph.decomp"));
    writeDecompFileCPU(gb, dir + "/out");
    {
        std::ifstream f(dir + "/out.decompCPU");
        std::string line;
        unsigned lines = 0;
        while (std::getline(f, line)) ++lines;
        CHECK(lines >= gb.V);
    }
    CHECK(!readDecompFile(gb, dir + "/absent"));

    // seeded generator: E distinct pairs, reloaded through the text path
    generateGraph(dir + "/gen", 300, 2000, 7);
    Graph gg(dir + "/gen", 1, false);
    CHECK(gg.E == 2000 && consistent(gg));
    generateGraph(dir + "/gen2", 300, 2000, 7);
    CHECK(slurp(dir + "/gen") == slurp(dir + "/gen2"));

    // edge batches partition the edges
    std::vector<degree> km(gg.V);
    for (unsigned v = 0; v < gg.V; ++v) km[v] = (degree)(v % 7);
    for (bool byKmax : {false, true}) {
        const auto b = getKmaxEdgeBatches(km, gg.edges, byKmax);
        using EdgeSet = std::multiset<std::pair<vertex, vertex>>;
        EdgeSet all, want(gg.edges.begin(), gg.edges.end());
        for (auto& x : b) all.insert(x.begin(), x.end());
        CHECK(all == want);
    }

    // host-side structural edits
    gg.seed(11);
    const auto add = random_new_edges(gg, 5);
    gg.insertEdges(add);
    CHECK(gg.E == 2005 && consistent(gg));
    gg.deleteEdges(add);
    CHECK(gg.E == 2000 && consistent(gg));
    gg.insertEdgesInPlace(add);
    CHECK(gg.E == 2005 && consistent(gg));
    gg.deleteEdgesInPlace(add);
    CHECK(gg.E == 2000 && consistent(gg));
    // several edges into one list in one in-place batch: the gap of 1 holds
    // only one, so the graph is rebuilt instead of overrunning the next list
    {
        std::vector<std::pair<vertex, vertex>> fan;
        for (vertex x = 0; x < gg.V && fan.size() < 3; ++x) {
            bool present = false;
            for (auto& e : gg.edges) present |= e == std::make_pair((vertex)5, x);
            if (x != 5 && !present) fan.emplace_back(5, x);
        }
        gg.insertEdgesInPlace(fan);
        CHECK(gg.E == 2003 && consistent(gg));
        gg.deleteEdgesInPlace(fan);
        CHECK(gg.E == 2000 && consistent(gg));
    }
}

static void device(const std::string& dir) {
    Device dev(0);
    // dcore() on the reference's graph equals the reference's own result
    Graph g(dir + "/digraph");
    Graph golden(dir + "/digraph");
    CHECK(readDecompFile(golden, dir + "/digraph"));
    dcx_host::dcore(dev, g);
    CHECK(g.kmax == golden.kmax && g.lmax == golden.lmax);
    CHECK(g.kmaxes == golden.kmaxes && g.lmaxes == golden.lmaxes);

    // single rows and in-core numbers on the device-resident graph
    GraphGPU gg(dev, g);
    CHECK(checkKmax(gg) == golden.kmaxes);
    for (degree k = 0; k <= golden.kmax; ++k) {
        const auto r = KList(gg, (unsigned)k);
        CHECK(r.second == golden.lmaxes[k]);
        std::vector<degree> arr(gg.V, -7);
        const degree lk = KListInplace(gg, arr, (unsigned)k, DCX_VARIANT_PEEL);
        CHECK(lk == r.first && arr == golden.lmaxes[k]);
    }

    // a bigger graph: every variant gives the same result
    generateGraph(dir + "/big", 4000, 40000, 3);
    Graph b1(dir + "/big", 1, false), b2(dir + "/big", 1, false);
    dcx_host::dcore(dev, b1, DCX_VARIANT_BATCHED);
    dcx_host::dcore(dev, b2, DCX_VARIANT_PEEL);
    CHECK(b1.kmax == b2.kmax && b1.kmaxes == b2.kmaxes && b1.lmaxes == b2.lmaxes);

    // maintenance (the SINGLE_INSERT / SPEED_TEST flow): inserts, deletes,
    // each followed by checkDCore against a recomputation
    b1.seed(5);
    const auto ins = random_new_edges(b1, 64);
    GraphGPU gb(dev, b1);
    maintenance(gb, ins, false);
    auto comp = checkDCore(gb);
    CHECK(comp.first == gb.kmaxes && comp.second == gb.lmaxes);
    maintenance(gb, {ins.begin(), ins.begin() + 20}, true);
    comp = checkDCore(gb);
    CHECK(comp.first == gb.kmaxes && comp.second == gb.lmaxes);
    const std::vector<std::pair<vertex, vertex>> again(ins.begin(), ins.begin() + 20);
    const auto t = maintenanceFast(gb, again);
    CHECK(std::get<0>(t).count() >= 0.0 && std::get<2>(t).count() >= 0.0);
    gb.pull();
    comp = checkDCore(gb);
    CHECK(comp.first == gb.kmaxes && comp.second == gb.lmaxes);

    // deleting an absent edge fails loudly and leaves the state as it was
    const auto before = gb.lmaxes;
    bool threw = false;
    try {
        maintenance(gb, {{0u, 0u}}, true);
    } catch (const Error&) {
        threw = true;
    }
    CHECK(threw);
    gb.pull();
    CHECK(gb.lmaxes == before);
}

int main(int argc, char** argv) {
    if (argc != 3) {
        std::fprintf(stderr, "usage: %s files|device <dir>\n", argv[0]);
        return 2;
    }
    const std::string mode = argv[1], dir = argv[2];
    try {
        if (mode == "files")
            files(dir);
        else if (mode == "device")
            device(dir);
        else
            return 2;
    } catch (const std::exception& e) {
        std::printf("FAIL exception: %s\n", e.what());
        ++failures;
    }
    std::printf("%s: %d checks, %d failed\n", mode.c_str(), checks, failures);
    return failures ? 1 : 0;
}
