// Reference entry points (louvain.hpp, rebuild.hpp, coloring.hpp, distgraph.hpp) over the
// vite_amd C ABI, at MPI world size 1.
#include "vite_louvain.hpp"

#include <cinttypes>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <iostream>
#include <map>
#include <random>
#include <sstream>
#include <algorithm>
#include <numeric>

#include "../../include/vite_amd.h"

namespace {

[[noreturn]] void fatal(const char* what, const char* detail)
{
    std::fprintf(stderr, "vite_amd: %s: %s\n", what, detail ? detail : "");
    std::fflush(stderr);
    std::exit(EXIT_FAILURE);
}

vt_ctx* g_ctx = nullptr;

void check(int rc, const char* what)
{
    if (rc != VT_OK) fatal(what, vt_last_error(g_ctx));
}

void require_one_rank(int me, int nprocs, const char* fn)
{
    if (nprocs != 1 || me != 0) {
        char buf[128];
        std::snprintf(buf, sizeof(buf), "me=%d nprocs=%d; this build runs at world size 1", me, nprocs);
        fatal(fn, buf);
    }
}

// Device copies of host graphs, keyed by address and checked against a hash of the whole
// CSR (every offset, tail and weight bit), so a reused address or a graph edited in place
// between calls never reuses a stale device copy.  The hash reads the CSR once at host
// memory speed (four independent multiply-rotate lanes), about 1 s for a 9 GB graph.
struct Cached {
    vt_graph* g = nullptr;
    GraphElem nv = -1, ne = -1;
    uint64_t fp = 0;
};
std::map<const DistGraph*, Cached> g_cache;

inline uint64_t rotl64(uint64_t x, int r) { return (x << r) | (x >> (64 - r)); }

uint64_t hash_words(const void* data, size_t nwords, uint64_t seed)
{
    const uint64_t K1 = 0x9E3779B185EBCA87ull, K2 = 0xC2B2AE3D27D4EB4Full;
    uint64_t lane[4] = {seed + K1 + K2, seed + K2, seed, seed - K1};
    const unsigned char* p = static_cast<const unsigned char*>(data);
    size_t i = 0;
    for (; i + 4 <= nwords; i += 4)
        for (int k = 0; k < 4; k++) {
            uint64_t w;
            std::memcpy(&w, p + 8 * (i + k), 8);
            lane[k] = rotl64(lane[k] + w * K2, 31) * K1;
        }
    uint64_t h = rotl64(lane[0], 1) + rotl64(lane[1], 7) + rotl64(lane[2], 12) + rotl64(lane[3], 18);
    for (; i < nwords; i++) {
        uint64_t w;
        std::memcpy(&w, p + 8 * i, 8);
        h = rotl64(h ^ (rotl64(w * K2, 31) * K1), 27) * K1 + 0x85EBCA77C2B2AE63ull;
    }
    h ^= (uint64_t)nwords;
    h ^= h >> 33;
    h *= K2;
    h ^= h >> 29;
    return h;
}

uint64_t fingerprint(const Graph& g)
{
    const GraphElem nv = g.getNumVertices(), ne = g.getNumEdges();
    uint64_t h = hash_words(g.edgeListIndexes.data(), (size_t)(nv + 1), (uint64_t)nv);
    return hash_words(g.edgeList.data(), 2 * (size_t)ne, h ^ (uint64_t)ne);
}

void forget(const DistGraph* dg)
{
    auto it = g_cache.find(dg);
    if (it == g_cache.end()) return;
    if (it->second.g) vt_graph_free(g_ctx, it->second.g);
    g_cache.erase(it);
}

vt_graph* device_graph(const DistGraph& dg)
{
    vite_amd_context();
    const Graph& g = dg.getLocalGraph();
    const GraphElem nv = g.getNumVertices(), ne = g.getNumEdges();
    const uint64_t fp = fingerprint(g);
    auto it = g_cache.find(&dg);
    if (it != g_cache.end() && it->second.nv == nv && it->second.ne == ne && it->second.fp == fp) return it->second.g;
    forget(&dg);
    Cached c;
    check(vt_graph_upload(g_ctx, nv, ne, g.edgeListIndexes.data(), reinterpret_cast<const vt_edge*>(g.edgeList.data()),
                          &c.g),
          "graph upload");
    c.nv = nv;
    c.ne = ne;
    c.fp = fp;
    g_cache[&dg] = c;
    return c.g;
}

// exchangeVertexReqs at world size 1 (louvain.cpp:3249-3397): no ghosts.
void no_ghosts(int nprocs, size_t& ssz, size_t& rsz, std::vector<GraphElem>& ssizes, std::vector<GraphElem>& rsizes,
               std::vector<GraphElem>& svdata, std::vector<GraphElem>& rvdata)
{
    ssz = rsz = 0;
    ssizes.assign(nprocs, 0);
    rsizes.assign(nprocs, 0);
    svdata.clear();
    rvdata.clear();
}

GraphWeight phase(const int me, const int nprocs, const DistGraph& dg, size_t& ssz, size_t& rsz,
                  std::vector<GraphElem>& ssizes, std::vector<GraphElem>& rsizes, std::vector<GraphElem>& svdata,
                  std::vector<GraphElem>& rvdata, CommunityVector& cvect, GraphWeight lower, GraphWeight thresh,
                  int& iters, int mode, int et_type, GraphWeight et_delta, long numColor, const ColorVector* colors)
{
    require_one_rank(me, nprocs, "distLouvainMethod");
    no_ghosts(nprocs, ssz, rsz, ssizes, rsizes, svdata, rvdata);
    vt_graph* g = device_graph(dg);
    vt_phase_params p;
    std::memset(&p, 0, sizeof(p));
    p.lower = lower;
    p.thresh = thresh;
    p.et_type = et_type;
    p.mode = mode;
    p.et_delta = et_delta;
    p.num_color = numColor;
    p.colors = colors ? colors->data() : nullptr;
    const char* seq = std::getenv("VITE_AMD_SEQUENTIAL_SUMS");
    if (seq && seq[0] == '1') p.flags |= VT_FLAG_SEQUENTIAL_SUMS;
    cvect.resize(dg.getLocalGraph().getNumVertices());
    int32_t it = 0;
    double q = 0.0;
    check(vt_louvain_phase(g_ctx, g, &p, cvect.data(), &it, &q), "distLouvainMethod");
    iters = it;
    return q;
}

}  // namespace

vt_ctx* vite_amd_context()
{
    if (!g_ctx) {
        const char* d = std::getenv("VITE_AMD_DEVICE");
        const int dev = d ? std::atoi(d) : 0;
        if (vt_ctx_create(dev, 0, &g_ctx) != VT_OK) fatal("context", "no usable GPU");
    }
    return g_ctx;
}

// ---- the nine Louvain overloads -----------------------------------------------------------

GraphWeight distLouvainMethod(const int me, const int nprocs, const DistGraph& dg, size_t& ssz, size_t& rsz,
                              std::vector<GraphElem>& ssizes, std::vector<GraphElem>& rsizes,
                              std::vector<GraphElem>& svdata, std::vector<GraphElem>& rvdata, CommunityVector& cvect,
                              const GraphWeight lower, const GraphWeight thresh, int& iters)
{
    return phase(me, nprocs, dg, ssz, rsz, ssizes, rsizes, svdata, rvdata, cvect, lower, thresh, iters, VT_MODE_PLAIN, 0,
                 1.0, 0, nullptr);
}

GraphWeight distLouvainMethod(const int me, const int nprocs, const DistGraph& dg, size_t& ssz, size_t& rsz,
                              std::vector<GraphElem>& ssizes, std::vector<GraphElem>& rsizes,
                              std::vector<GraphElem>& svdata, std::vector<GraphElem>& rvdata, CommunityVector& cvect,
                              const GraphWeight lower, const GraphWeight thresh, int& iters, bool ETLocalOrRemote)
{
    return phase(me, nprocs, dg, ssz, rsz, ssizes, rsizes, svdata, rvdata, cvect, lower, thresh, iters, VT_MODE_PLAIN,
                 ETLocalOrRemote ? 1 : 3, 1.0, 0, nullptr);
}

GraphWeight distLouvainMethod(const int me, const int nprocs, const DistGraph& dg, size_t& ssz, size_t& rsz,
                              std::vector<GraphElem>& ssizes, std::vector<GraphElem>& rsizes,
                              std::vector<GraphElem>& svdata, std::vector<GraphElem>& rvdata, CommunityVector& cvect,
                              const GraphWeight lower, const GraphWeight thresh, int& iters, GraphWeight ETDelta,
                              bool ETLocalOrRemote)
{
    return phase(me, nprocs, dg, ssz, rsz, ssizes, rsizes, svdata, rvdata, cvect, lower, thresh, iters, VT_MODE_PLAIN,
                 ETLocalOrRemote ? 2 : 4, ETDelta, 0, nullptr);
}

GraphWeight distLouvainMethodWithColoring(const int me, const int nprocs, const DistGraph& dg, const long numColor,
                                          const ColorVector& vertexColor, size_t& ssz, size_t& rsz,
                                          std::vector<GraphElem>& ssizes, std::vector<GraphElem>& rsizes,
                                          std::vector<GraphElem>& svdata, std::vector<GraphElem>& rvdata,
                                          CommunityVector& cvect, const GraphWeight lower, const GraphWeight thresh,
                                          int& iters)
{
    return phase(me, nprocs, dg, ssz, rsz, ssizes, rsizes, svdata, rvdata, cvect, lower, thresh, iters, VT_MODE_COLOR, 0,
                 1.0, numColor, &vertexColor);
}

GraphWeight distLouvainMethodWithColoring(const int me, const int nprocs, const DistGraph& dg, const long numColor,
                                          const ColorVector& vertexColor, size_t& ssz, size_t& rsz,
                                          std::vector<GraphElem>& ssizes, std::vector<GraphElem>& rsizes,
                                          std::vector<GraphElem>& svdata, std::vector<GraphElem>& rvdata,
                                          CommunityVector& cvect, const GraphWeight lower, const GraphWeight thresh,
                                          int& iters, bool ETLocalOrRemote)
{
    return phase(me, nprocs, dg, ssz, rsz, ssizes, rsizes, svdata, rvdata, cvect, lower, thresh, iters, VT_MODE_COLOR,
                 ETLocalOrRemote ? 1 : 3, 1.0, numColor, &vertexColor);
}

GraphWeight distLouvainMethodWithColoring(const int me, const int nprocs, const DistGraph& dg, const long numColor,
                                          const ColorVector& vertexColor, size_t& ssz, size_t& rsz,
                                          std::vector<GraphElem>& ssizes, std::vector<GraphElem>& rsizes,
                                          std::vector<GraphElem>& svdata, std::vector<GraphElem>& rvdata,
                                          CommunityVector& cvect, const GraphWeight lower, const GraphWeight thresh,
                                          int& iters, GraphWeight ETDelta, bool ETLocalOrRemote)
{
    return phase(me, nprocs, dg, ssz, rsz, ssizes, rsizes, svdata, rvdata, cvect, lower, thresh, iters, VT_MODE_COLOR,
                 ETLocalOrRemote ? 2 : 4, ETDelta, numColor, &vertexColor);
}

GraphWeight distLouvainMethodVertexOrder(const int me, const int nprocs, const DistGraph& dg, const long numColor,
                                         const ColorVector& vertexColor, size_t& ssz, size_t& rsz,
                                         std::vector<GraphElem>& ssizes, std::vector<GraphElem>& rsizes,
                                         std::vector<GraphElem>& svdata, std::vector<GraphElem>& rvdata,
                                         CommunityVector& cvect, const GraphWeight lower, const GraphWeight thresh,
                                         int& iters)
{
    return phase(me, nprocs, dg, ssz, rsz, ssizes, rsizes, svdata, rvdata, cvect, lower, thresh, iters, VT_MODE_VORDER, 0,
                 1.0, numColor, &vertexColor);
}

GraphWeight distLouvainMethodVertexOrder(const int me, const int nprocs, const DistGraph& dg, const long numColor,
                                         const ColorVector& vertexColor, size_t& ssz, size_t& rsz,
                                         std::vector<GraphElem>& ssizes, std::vector<GraphElem>& rsizes,
                                         std::vector<GraphElem>& svdata, std::vector<GraphElem>& rvdata,
                                         CommunityVector& cvect, const GraphWeight lower, const GraphWeight thresh,
                                         int& iters, bool ETLocalOrRemote)
{
    return phase(me, nprocs, dg, ssz, rsz, ssizes, rsizes, svdata, rvdata, cvect, lower, thresh, iters, VT_MODE_VORDER,
                 ETLocalOrRemote ? 1 : 3, 1.0, numColor, &vertexColor);
}

GraphWeight distLouvainMethodVertexOrder(const int me, const int nprocs, const DistGraph& dg, const long numColor,
                                         const ColorVector& vertexColor, size_t& ssz, size_t& rsz,
                                         std::vector<GraphElem>& ssizes, std::vector<GraphElem>& rsizes,
                                         std::vector<GraphElem>& svdata, std::vector<GraphElem>& rvdata,
                                         CommunityVector& cvect, const GraphWeight lower, const GraphWeight thresh,
                                         int& iters, GraphWeight ETDelta, bool ETLocalOrRemote)
{
    return phase(me, nprocs, dg, ssz, rsz, ssizes, rsizes, svdata, rvdata, cvect, lower, thresh, iters, VT_MODE_VORDER,
                 ETLocalOrRemote ? 2 : 4, ETDelta, numColor, &vertexColor);
}

// ---- rebuild / coloring ---------------------------------------------------------------------

static DistGraph* host_from_device(vt_graph* g)
{
    int64_t nv = 0, ne = 0;
    vt_graph_info(g, &nv, &ne);
    DistGraph* dg = new DistGraph(nv, ne);
    PartRanges parts{0, nv};
    dg->createLocalGraph(nv, ne, &parts);
    Graph& lg = dg->getLocalGraph();
    check(vt_graph_download(g_ctx, g, lg.edgeListIndexes.data(), reinterpret_cast<vt_edge*>(lg.edgeList.data())),
          "graph download");
    Cached c;
    c.g = g;
    c.nv = nv;
    c.ne = ne;
    c.fp = fingerprint(lg);
    g_cache[dg] = c;
    return dg;
}

void distbuildNextLevelGraph(int nprocs, int me, DistGraph*& dg, const size_t&, const size_t&,
                             const std::vector<GraphElem>&, const std::vector<GraphElem>&,
                             const std::vector<GraphElem>&, const std::vector<GraphElem>&, CommunityVector& cvect)
{
    require_one_rank(me, nprocs, "distbuildNextLevelGraph");
    vt_graph* g = device_graph(*dg);
    if ((GraphElem)cvect.size() != dg->getLocalGraph().getNumVertices())
        fatal("distbuildNextLevelGraph", "cvect size differs from the number of vertices");
    vt_graph* out = nullptr;
    check(vt_rebuild(g_ctx, g, cvect.data(), &out), "distbuildNextLevelGraph");
    forget(dg);
    delete dg;  // the reference replaces *dg (rebuild.cpp:334,351)
    dg = host_from_device(out);
}

ColorElem distColoringMultiHashMinMax(const int me, const int nprocs, const DistGraph& dg, ColorVector& vertexColor,
                                      const ColorElem nHash, const int target_percent, const bool singleIteration)
{
    require_one_rank(me, nprocs, "distColoringMultiHashMinMax");
    vt_graph* g = device_graph(dg);
    vertexColor.resize(dg.getLocalGraph().getNumVertices());
    int64_t nc = 0;
    check(vt_color(g_ctx, g, nHash, target_percent, singleIteration ? 1 : 0, vertexColor.data(), &nc),
          "distColoringMultiHashMinMax");
    return nc;
}

// ---- graph input / generation ---------------------------------------------------------------

// The .bin layout (distgraph.cpp:119-253): {nv, ne} as two GraphElem, nv+1 offsets, then
// 16-byte edge records from offset 16 + 8(nv+1).  At world size 1 the one rank reads
// every row; its edge count is offsets[nv] - offsets[0] and the offsets are rebased to 0,
// as the reference does after its read (distgraph.cpp:208-247).
static void read_bin(const char* fn, std::string& fileName, DistGraph*& dg, bool announce)
{
    std::FILE* fp = std::fopen(fileName.c_str(), "rb");
    if (!fp) fatal(fn, (" Error opening file! " + fileName).c_str());
    GraphElem hdr[2] = {-1, -1};
    if (std::fread(hdr, sizeof(GraphElem), 2, fp) != 2 || hdr[0] < 0 || hdr[1] < 0) {
        std::fclose(fp);
        fatal(fn, ("bad header in " + fileName).c_str());
    }
    const GraphElem nv = hdr[0];
    if (announce) std::cout << "Reading file of " << nv << " vertices and " << hdr[1] << " edges." << std::endl;
    dg = new DistGraph(nv, hdr[1]);
    PartRanges parts{0, nv};
    dg->createLocalGraph(nv, 0, &parts);
    Graph& g = dg->getLocalGraph();
    if (std::fread(g.edgeListIndexes.data(), sizeof(GraphElem), (size_t)nv + 1, fp) != (size_t)nv + 1) {
        std::fclose(fp);
        fatal(fn, "short read of the offsets");
    }
    const GraphElem base = g.edgeListIndexes[0];
    const GraphElem lne = g.edgeListIndexes[nv] - base;
    if (base < 0 || lne < 0) {
        std::fclose(fp);
        fatal(fn, "offsets out of order");
    }
    g.setNumEdges(lne);
    const long off = (long)(2 + nv + 1) * (long)sizeof(GraphElem) + (long)base * (long)sizeof(Edge);
    if (std::fseek(fp, off, SEEK_SET) != 0 ||
        std::fread(g.edgeList.data(), sizeof(Edge), (size_t)lne, fp) != (size_t)lne) {
        std::fclose(fp);
        fatal(fn, "short read of the edge records");
    }
    std::fclose(fp);
    for (auto& o : g.edgeListIndexes) o -= base;
}

void loadDistGraphMPIIO(int me, int nprocs, int, DistGraph*& dg, std::string& fileName)
{
    require_one_rank(me, nprocs, "loadDistGraphMPIIO");
    read_bin("loadDistGraphMPIIO", fileName, dg, true);
}

void balanceEdges(int nprocs, std::string& fileName, std::vector<GraphElem>& mbins)
{
    // distgraph.cpp:72-116: scan the offsets once, filling part p until it holds
    // ne / nprocs edges (the last part takes the rest); mbins gets the parts' first rows.
    std::FILE* fp = std::fopen(fileName.c_str(), "rb");
    if (!fp) fatal("balanceEdges", (" Error opening file! " + fileName).c_str());
    GraphElem hdr[2] = {0, 0};
    if (std::fread(hdr, sizeof(GraphElem), 2, fp) != 2) {
        std::fclose(fp);
        fatal("balanceEdges", "bad header");
    }
    std::vector<GraphElem> nbins(nprocs, 0);
    if ((int)mbins.size() < nprocs + 1) mbins.assign(nprocs + 1, 0);
    const GraphElem cap = hdr[1] / nprocs;
    GraphElem prev = 0, cur = 0;
    int p = 0;
    for (GraphElem m = 0; m < hdr[0]; m++) {
        if (std::fread(&cur, sizeof(GraphElem), 1, fp) != 1) {
            std::fclose(fp);
            fatal("balanceEdges", "short read of the offsets");
        }
        if (nbins[p] < cap || p == nprocs - 1) nbins[p] += cur - prev;
        if (nbins[p] >= cap && p < nprocs - 1) p++;
        mbins[p + 1]++;
        prev = cur;
    }
    std::fclose(fp);
    for (int k = 1; k < nprocs + 1; k++) mbins[k] += mbins[k - 1];
}

void loadDistGraphMPIIOBalanced(int me, int nprocs, int, DistGraph*& dg, std::string& fileName)
{
    // distgraph.cpp:256-387: the edge-balanced split has one part at world size 1, so the
    // rows read are those of the plain read.
    require_one_rank(me, nprocs, "loadDistGraphMPIIOBalanced");
    std::vector<GraphElem> mbins(nprocs + 1, 0);
    balanceEdges(nprocs, fileName, mbins);
    std::cout << "Trying to achieve equal edge distribution across processes." << std::endl;
    read_bin("loadDistGraphMPIIOBalanced", fileName, dg, false);
    if (mbins[1] != dg->getLocalGraph().getNumVertices()) fatal("loadDistGraphMPIIOBalanced", "part bounds differ");
}

void writeGraph(int me, int nprocs, DistGraph*& dg, std::vector<GraphElem>& edgeCount, std::string& fileName)
{
    // distgraph.cpp:986-1064 at world size 1: edgeCount holds each row's degree at [v+1]
    // (and 0 at [0]); its prefix sums are the offsets written.
    require_one_rank(me, nprocs, "writeGraph");
    const Graph& g = dg->getLocalGraph();
    const GraphElem nv = dg->getTotalNumVertices();
    if ((GraphElem)edgeCount.size() < nv + 1) fatal("writeGraph", "edgeCount needs nv + 1 entries");
    std::vector<GraphElem> offs(nv + 1);
    std::partial_sum(edgeCount.begin(), edgeCount.begin() + nv + 1, offs.begin());
    GraphElem hdr[2] = {nv, (GraphElem)g.edgeList.size()};
    std::FILE* fp = std::fopen(fileName.c_str(), "wb");
    if (!fp) fatal("writeGraph", (" Error opening output binary file for storing graph data! " + fileName).c_str());
    const bool ok = std::fwrite(hdr, sizeof(GraphElem), 2, fp) == 2 &&
                    std::fwrite(offs.data(), sizeof(GraphElem), (size_t)nv + 1, fp) == (size_t)nv + 1 &&
                    std::fwrite(g.edgeList.data(), sizeof(Edge), (size_t)hdr[1], fp) == (size_t)hdr[1];
    if (std::fclose(fp) != 0 || !ok) fatal("writeGraph", ("short write to " + fileName).c_str());
}

void writeGraphBin(const DistGraph& dg, const std::string& fileName)
{
    const Graph& g = dg.getLocalGraph();
    const GraphElem nv = g.getNumVertices();
    std::vector<GraphElem> edgeCount(nv + 1, 0);
    for (GraphElem v = 0; v < nv; v++) edgeCount[v + 1] = g.edgeListIndexes[v + 1] - g.edgeListIndexes[v];
    DistGraph* p = const_cast<DistGraph*>(&dg);
    std::string fn = fileName;
    writeGraph(0, 1, p, edgeCount, fn);
}

// gatherAllComm at world size 1 (louvain.cpp:3439-3480): the root's own vector.
void gatherAllComm(int root, int me, int nprocs, std::vector<GraphElem>& commAll, const CommunityVector& localComm)
{
    require_one_rank(me, nprocs, "gatherAllComm");
    if (root != 0) fatal("gatherAllComm", "root must be rank 0");
    commAll.assign(localComm.begin(), localComm.end());
}

void vite_amd_release(const DistGraph* dg)
{
    if (g_ctx) forget(dg);
}

// ---- ground truth ---------------------------------------------------------------------------

void loadGroundTruthFile(std::vector<GraphElem>& commGroundTruth, std::string const& groundTruthFileName,
                         bool isGroundTruthZeroBased)
{
    std::ifstream f(groundTruthFileName);
    if (!f) {
        std::cerr << "Error opening ground truth file: " << groundTruthFileName << std::endl;
        std::exit(EXIT_FAILURE);
    }
    std::string line;
    while (std::getline(f, line)) {
        // as louvain.cpp:3419-3428: the second field of each line, in line order; a line
        // without it still adds an entry (-1, or -2 from a one-based file)
        GraphElem v = -1, c = -1;
        std::istringstream iss(line);
        iss >> v >> c;
        if (!isGroundTruthZeroBased) c--;
        commGroundTruth.push_back(c);
    }
    std::cout << "Loaded ground-truth file: " << groundTruthFileName << ", containing community information for "
              << commGroundTruth.size() << " vertices." << std::endl;
}

void compare_communities(std::vector<GraphElem> const& C1, std::vector<GraphElem> const& C2)
{
    if (C1.empty() || C1.size() != C2.size()) fatal("compare_communities", "partitions differ in size or are empty");
    vt_compare_stats st;
    check(vt_compare_communities(vite_amd_context(), (int64_t)C1.size(), C1.data(), C2.data(), &st),
          "compare_communities");
    // the report of compare.cpp:274-293
    std::cout << "Number of unique communities in C1 = " << st.nc_truth << ", and C2 = " << st.nc_found << std::endl;
    std::cout << "*******************************************" << std::endl;
    std::cout << "Communities comparison statistics:" << std::endl;
    std::cout << "*******************************************" << std::endl;
    std::cout << "|C1| (truth)       : " << st.n << std::endl;
    std::cout << "#communities in C1 : " << st.nc_truth << std::endl;
    std::cout << "|C2| (output)      : " << st.n << std::endl;
    std::cout << "#communities in C2 : " << st.nc_found << std::endl;
    std::cout << "-------------------------------------------" << std::endl;
    std::cout << "Same-Same (True positive)  : " << st.same_same << std::endl;
    std::cout << "Same-Diff (False negative) : " << st.same_diff << std::endl;
    std::cout << "Diff-Same (False positive) : " << st.diff_same << std::endl;
    std::cout << "-------------------------------------------" << std::endl;
    std::cout << "Precision :  " << st.precision << " (" << (st.precision * 100) << ")" << std::endl;
    std::cout << "Recall    :  " << st.recall << " (" << (st.recall * 100) << ")" << std::endl;
    std::cout << "F-score   :  " << st.f_score << std::endl;
    std::cout << "-------------------------------------------" << std::endl;
    std::cout << "Gini coefficient, C1  :  " << st.gini_truth << std::endl;
    std::cout << "Gini coefficient, C2  :  " << st.gini_found << std::endl;
    std::cout << "*******************************************" << std::endl;
}

// compute_gini_coeff (compare.cpp:310-337): sorts colorSize in place, then the one-thread
// left folds and the closing expression.
double compute_gini_coeff(GraphElem* colorSize, int numColors)
{
    std::sort(colorSize, colorSize + numColors);
    double numFunc = 0.0, denFunc = 0.0;
    for (GraphElem i = 0; i < numColors; i++) {
        numFunc = numFunc + (double)((i + 1) * colorSize[i]);
        denFunc = denFunc + (double)colorSize[i];
    }
    return ((2 * numFunc) / (numColors * denFunc)) - ((double)(numColors + 1) / (double)numColors);
}

extern "C" int vt_gen_vite_rgg(vt_ctx* ctx, int64_t n, uint32_t x0, vt_graph** out);

// generateInMemGraph at one rank (distgraph.cpp:391-407): Vite's LCG-seeded RGG, built on the
// device.  x0 is std::seed_seq{1}'s first word, as LCG::reseeder draws it (utils.hpp:163-170).
void generateInMemGraph(int rank, int nprocs, DistGraph*& dg, GraphElem nv, GraphWeight randomEdgePercent,
                        std::string fileOut)
{
    require_one_rank(rank, nprocs, "generateInMemGraph");
    if (randomEdgePercent > 0.0)
        fatal("generateInMemGraph", "random extra edges use a time seed in the reference; use generateSeededRGG");
    vite_amd_context();
    std::seed_seq seq({1u});
    std::vector<std::uint32_t> seeds(1);
    seq.generate(seeds.begin(), seeds.end());
    vt_graph* g = nullptr;
    check(vt_gen_vite_rgg(g_ctx, nv, seeds[0], &g), "generateInMemGraph");
    dg = host_from_device(g);
    if (!fileOut.empty()) {
        writeGraphBin(*dg, fileOut);
        std::cout << "Written binary file: " << fileOut << std::endl;
    }
}

DistGraph* generateSeededRGG(unsigned long long seed, GraphElem n, GraphWeight extraPercent)
{
    vite_amd_context();
    vt_graph* g = nullptr;
    check(vt_gen_rgg(g_ctx, seed, n, extraPercent, &g), "generateSeededRGG");
    return host_from_device(g);
}

DistGraph* generateSeededBA(unsigned long long seed, GraphElem n, GraphElem m)
{
    vite_amd_context();
    vt_graph* g = nullptr;
    check(vt_gen_ba(g_ctx, seed, n, m, &g), "generateSeededBA");
    return host_from_device(g);
}

void createCommunityMPIType() {}
void destroyCommunityMPIType() {}
void createEdgeMPIType() {}
void destroyEdgeMPIType() {}
