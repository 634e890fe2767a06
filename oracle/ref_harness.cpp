// Test-oracle harness around the compiled reference (ECP-ExaGraph/vite).
//
// This file is test infrastructure.  It is compiled by oracle/Makefile.ref against the
// reference sources where they lie (it #includes louvain.cpp, rebuild.cpp, coloring.cpp,
// distgraph.cpp, compare.cpp and, for the trace binary, main.cpp) so that the
// reference's internal-linkage functions (louvain.hpp:139-190) can be called on a given
// state.  Nothing in the product links it.
//
// Two build products:
//   * libvite_ref.so  : extern "C" replay entry points (ref_sweep, ref_phase, ref_rebuild,
//                       ref_color, ref_gen_rgg) used by the oracle and GPU parity tests.
//   * ref_trace       : the reference's own main() (main.cpp:108-607) with every call to
//                       distLouvainMethod* / distbuildNextLevelGraph recorded and printed
//                       at 17 significant digits (REF_TRACE_MAIN).

#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <vector>

#ifdef REF_TRACE_MAIN
#define distLouvainMethod distLouvainMethod_ref
#define distLouvainMethodWithColoring distLouvainMethodWithColoring_ref
#define distLouvainMethodVertexOrder distLouvainMethodVertexOrder_ref
#define distbuildNextLevelGraph distbuildNextLevelGraph_ref
#endif

#include "louvain.cpp"
#include "rebuild.cpp"
#include "coloring.cpp"
#include "distgraph.cpp"
#include "compare.cpp"

#ifdef REF_TRACE_MAIN
#undef distLouvainMethod
#undef distLouvainMethodWithColoring
#undef distLouvainMethodVertexOrder
#undef distbuildNextLevelGraph
#endif

static std::uint64_t fnv_vec(const std::vector<GraphElem>& v)
{
    std::uint64_t h = 1469598103934665603ull;
    for (GraphElem x : v) {
        std::uint64_t u = (std::uint64_t)x;
        for (int b = 0; b < 8; b++) { h ^= (u >> (8 * b)) & 0xff; h *= 1099511628211ull; }
    }
    return h;
}

#ifdef REF_TRACE_MAIN
// ---------------------------------------------------------------------------------
// Trace wrappers: same signatures as louvain.hpp:80-137 and rebuild.hpp:109-112.
// Every call made by main.cpp's phase loop lands here, is forwarded to the reference
// implementation, and is recorded.
// ---------------------------------------------------------------------------------
#define TRACE_CALL(kind, call)                                                        \
    do {                                                                              \
        const GraphElem tnv_ = dg.getTotalNumVertices();                              \
        const GraphElem ne_ = dg.getLocalGraph().getNumEdges();                       \
        GraphWeight q_ = call;                                                        \
        std::printf("TRACE %s nv=%lld ne=%lld thresh=%.17g lower=%.17g q=%.17g iters=%d " \
                    "cvect_fnv=%016llx\n",                                            \
                    kind, (long long)tnv_, (long long)ne_, thresh, lower, q_, iters,  \
                    (unsigned long long)fnv_vec(cvect));                              \
        std::fflush(stdout);                                                          \
        return q_;                                                                    \
    } while (0)

GraphWeight distLouvainMethod(const int me, const int nprocs, const DistGraph &dg,
        size_t &ssz, size_t &rsz, std::vector<GraphElem> &ssizes,
        std::vector<GraphElem> &rsizes, std::vector<GraphElem> &svdata,
        std::vector<GraphElem> &rvdata, CommunityVector &cvect, const GraphWeight lower,
        const GraphWeight thresh, int& iters)
{ TRACE_CALL("plain", distLouvainMethod_ref(me, nprocs, dg, ssz, rsz, ssizes, rsizes, svdata, rvdata, cvect, lower, thresh, iters)); }

GraphWeight distLouvainMethod(const int me, const int nprocs, const DistGraph &dg,
        size_t &ssz, size_t &rsz, std::vector<GraphElem> &ssizes,
        std::vector<GraphElem> &rsizes, std::vector<GraphElem> &svdata,
        std::vector<GraphElem> &rvdata, CommunityVector &cvect, const GraphWeight lower,
        const GraphWeight thresh, int& iters, bool ETLocalOrRemote)
{ TRACE_CALL("et", distLouvainMethod_ref(me, nprocs, dg, ssz, rsz, ssizes, rsizes, svdata, rvdata, cvect, lower, thresh, iters, ETLocalOrRemote)); }

GraphWeight distLouvainMethod(const int me, const int nprocs, const DistGraph &dg,
        size_t &ssz, size_t &rsz, std::vector<GraphElem> &ssizes,
        std::vector<GraphElem> &rsizes, std::vector<GraphElem> &svdata,
        std::vector<GraphElem> &rvdata, CommunityVector &cvect, const GraphWeight lower,
        const GraphWeight thresh, int& iters, GraphWeight ETDelta, bool ETLocalOrRemote)
{ TRACE_CALL("etprob", distLouvainMethod_ref(me, nprocs, dg, ssz, rsz, ssizes, rsizes, svdata, rvdata, cvect, lower, thresh, iters, ETDelta, ETLocalOrRemote)); }

GraphWeight distLouvainMethodWithColoring(const int me, const int nprocs, const DistGraph &dg,
        const long numColor, const ColorVector &vertexColor, size_t &ssz, size_t &rsz,
        std::vector<GraphElem> &ssizes, std::vector<GraphElem> &rsizes,
        std::vector<GraphElem> &svdata, std::vector<GraphElem> &rvdata,
        CommunityVector &cvect, const GraphWeight lower, const GraphWeight thresh, int& iters)
{ TRACE_CALL("color", distLouvainMethodWithColoring_ref(me, nprocs, dg, numColor, vertexColor, ssz, rsz, ssizes, rsizes, svdata, rvdata, cvect, lower, thresh, iters)); }

GraphWeight distLouvainMethodWithColoring(const int me, const int nprocs, const DistGraph &dg,
        const long numColor, const ColorVector &vertexColor, size_t &ssz, size_t &rsz,
        std::vector<GraphElem> &ssizes, std::vector<GraphElem> &rsizes,
        std::vector<GraphElem> &svdata, std::vector<GraphElem> &rvdata,
        CommunityVector &cvect, const GraphWeight lower, const GraphWeight thresh,
        int& iters, bool ETLocalOrRemote)
{ TRACE_CALL("color_et", distLouvainMethodWithColoring_ref(me, nprocs, dg, numColor, vertexColor, ssz, rsz, ssizes, rsizes, svdata, rvdata, cvect, lower, thresh, iters, ETLocalOrRemote)); }

GraphWeight distLouvainMethodWithColoring(const int me, const int nprocs, const DistGraph &dg,
        const long numColor, const ColorVector &vertexColor,
        size_t &ssz, size_t &rsz, std::vector<GraphElem> &ssizes,
        std::vector<GraphElem> &rsizes, std::vector<GraphElem> &svdata,
        std::vector<GraphElem> &rvdata, CommunityVector &cvect,
        const GraphWeight lower, const GraphWeight thresh, int& iters, GraphWeight ETDelta,
        bool ETLocalOrRemote)
{ TRACE_CALL("color_etprob", distLouvainMethodWithColoring_ref(me, nprocs, dg, numColor, vertexColor, ssz, rsz, ssizes, rsizes, svdata, rvdata, cvect, lower, thresh, iters, ETDelta, ETLocalOrRemote)); }

GraphWeight distLouvainMethodVertexOrder(const int me, const int nprocs, const DistGraph &dg,
        const long numColor, const ColorVector &vertexColor, size_t &ssz, size_t &rsz,
        std::vector<GraphElem> &ssizes, std::vector<GraphElem> &rsizes,
        std::vector<GraphElem> &svdata, std::vector<GraphElem> &rvdata,
        CommunityVector &cvect, const GraphWeight lower, const GraphWeight thresh, int& iters)
{ TRACE_CALL("vorder", distLouvainMethodVertexOrder_ref(me, nprocs, dg, numColor, vertexColor, ssz, rsz, ssizes, rsizes, svdata, rvdata, cvect, lower, thresh, iters)); }

GraphWeight distLouvainMethodVertexOrder(const int me, const int nprocs, const DistGraph &dg,
        const long numColor, const ColorVector &vertexColor, size_t &ssz, size_t &rsz,
        std::vector<GraphElem> &ssizes, std::vector<GraphElem> &rsizes,
        std::vector<GraphElem> &svdata, std::vector<GraphElem> &rvdata,
        CommunityVector &cvect, const GraphWeight lower, const GraphWeight thresh, int& iters,
        bool ETLocalOrRemote)
{ TRACE_CALL("vorder_et", distLouvainMethodVertexOrder_ref(me, nprocs, dg, numColor, vertexColor, ssz, rsz, ssizes, rsizes, svdata, rvdata, cvect, lower, thresh, iters, ETLocalOrRemote)); }

GraphWeight distLouvainMethodVertexOrder(const int me, const int nprocs, const DistGraph &dg,
        const long numColor, const ColorVector &vertexColor, size_t &ssz, size_t &rsz,
        std::vector<GraphElem> &ssizes, std::vector<GraphElem> &rsizes,
        std::vector<GraphElem> &svdata, std::vector<GraphElem> &rvdata,
        CommunityVector &cvect, const GraphWeight lower, const GraphWeight thresh, int& iters,
        GraphWeight ETDelta, bool ETLocalOrRemote)
{ TRACE_CALL("vorder_etprob", distLouvainMethodVertexOrder_ref(me, nprocs, dg, numColor, vertexColor, ssz, rsz, ssizes, rsizes, svdata, rvdata, cvect, lower, thresh, iters, ETDelta, ETLocalOrRemote)); }

void distbuildNextLevelGraph(int nprocs, int me, DistGraph*& dg,
        const size_t &ssz, const size_t &rsz, const std::vector<GraphElem> &ssizes,
        const std::vector<GraphElem> &rsizes, const std::vector<GraphElem> &svdata,
        const std::vector<GraphElem> &rvdata, CommunityVector &cvect)
{
    distbuildNextLevelGraph_ref(nprocs, me, dg, ssz, rsz, ssizes, rsizes, svdata, rvdata, cvect);
    const Graph& g = dg->getLocalGraph();
    std::uint64_t h = 1469598103934665603ull;
    for (GraphElem i = 0; i <= g.getNumVertices(); i++) { h ^= (std::uint64_t)g.edgeListIndexes[i]; h *= 1099511628211ull; }
    for (GraphElem e = 0; e < g.getNumEdges(); e++) { h ^= (std::uint64_t)g.edgeList[e].tail; h *= 1099511628211ull; }
    std::printf("TRACE rebuild nv=%lld ne=%lld csr_fnv=%016llx\n", (long long)dg->getTotalNumVertices(),
                (long long)g.getNumEdges(), (unsigned long long)h);
    std::fflush(stdout);
}

#define main vite_ref_main
#include "main.cpp"
#undef main

int main(int argc, char* argv[])
{
    std::cout.precision(17);
    return vite_ref_main(argc, argv);
}

#else  // shared-library replay API

// Every replay entry point runs the reference on one OpenMP thread: its reductions for 2m and
// modularity are ordered by the thread count, and the oracle restates the one-thread order.
static void ensure_mpi()
{
    omp_set_num_threads(1);
    int inited = 0;
    MPI_Initialized(&inited);
    if (!inited) {
        int provided = 0;
        MPI_Init_thread(nullptr, nullptr, MPI_THREAD_SINGLE, &provided);
        createCommunityMPIType();
        createEdgeMPIType();
    }
}

static DistGraph* make_dg(std::int64_t nv, std::int64_t ne, const std::int64_t* offs, const Edge* edges)
{
    DistGraph* dg = new DistGraph(nv, ne);
    PartRanges parts(2);
    parts[0] = 0;
    parts[1] = nv;
    dg->createLocalGraph(nv, ne, &parts);
    Graph& g = dg->getLocalGraph();
    std::memcpy(g.edgeListIndexes.data(), offs, sizeof(std::int64_t) * (nv + 1));
    if (ne) std::memcpy((void*)g.edgeList.data(), (const void*)edges, sizeof(Edge) * ne);
    return dg;
}

std::ofstream ofs;

extern "C" {

// One local-move sweep of the reference over `order` (distExecuteLouvainIteration,
// louvain.cpp:2281-2417), from a caller-provided state.  localCupdate starts at zero and
// receives the deltas in processing order, exactly as one OpenMP thread does.
void ref_sweep(std::int64_t nv, std::int64_t ne, const std::int64_t* offs, const Edge* edges,
               const std::int64_t* curr, const std::int64_t* csize, const double* cdeg,
               const double* vdeg, double constant, const std::int64_t* order, std::int64_t norder,
               std::int64_t* target_out, double* cw_out, std::int64_t* dsize_out, double* ddeg_out)
{
    ensure_mpi();
    DistGraph* dg = make_dg(nv, ne, offs, edges);
    CommunityVector currComm(curr, curr + nv), targetComm(nv, -1);
    GraphWeightVector vDegree(vdeg, vdeg + nv), clusterWeight(nv, 0.0);
    CommVector localCinfo(nv), localCupdate(nv);
    for (std::int64_t i = 0; i < nv; i++) { localCinfo[i].size = csize[i]; localCinfo[i].degree = cdeg[i]; }
    VertexCommMap remoteComm;
    CommMap remoteCinfo, remoteCupdate;
    for (std::int64_t k = 0; k < norder; k++)
        distExecuteLouvainIteration(order[k], *dg, currComm, targetComm, vDegree, localCinfo,
                                    localCupdate, remoteComm, remoteCinfo, remoteCupdate,
                                    constant, clusterWeight, 0);
    for (std::int64_t i = 0; i < nv; i++) {
        target_out[i] = targetComm[i];
        cw_out[i] = clusterWeight[i];
        dsize_out[i] = localCupdate[i].size;
        ddeg_out[i] = localCupdate[i].degree;
    }
    delete dg;
}

// The reference's sweep loop (louvain.cpp:497-514: omp parallel, distCleanCWandCU, omp for
// schedule(guided) over distExecuteLouvainIteration) on vertices [v0, v1) with `nthreads`
// threads, `reps` times from the same state.  Returns the mean seconds of the loop alone;
// targets of the last repetition go to target_out (nv entries; untouched vertices get -1).
double ref_sweep_omp(std::int64_t nv, std::int64_t ne, const std::int64_t* offs, const Edge* edges,
                     const std::int64_t* curr, const std::int64_t* csize, const double* cdeg, const double* vdeg,
                     double constant, std::int64_t v0, std::int64_t v1, int nthreads, int reps,
                     std::int64_t* target_out)
{
    ensure_mpi();
    DistGraph* dg = make_dg(nv, ne, offs, edges);
    CommunityVector currComm(curr, curr + nv), targetComm(nv, -1);
    GraphWeightVector vDegree(vdeg, vdeg + nv), clusterWeight(nv, 0.0);
    CommVector localCinfo(nv), localCupdate(nv);
    for (std::int64_t i = 0; i < nv; i++) { localCinfo[i].size = csize[i]; localCinfo[i].degree = cdeg[i]; }
    VertexCommMap remoteComm;
    CommMap remoteCinfo, remoteCupdate;
    omp_set_num_threads(nthreads);
    double total = 0.0;
    for (int r = 0; r < reps; r++) {
        const double t0 = omp_get_wtime();
#pragma omp parallel shared(clusterWeight, localCupdate, currComm, targetComm, vDegree, localCinfo, remoteCinfo, \
                                remoteComm, dg, remoteCupdate) firstprivate(constant)
        {
            distCleanCWandCU(nv, clusterWeight, localCupdate);
#pragma omp for schedule(guided)
            for (std::int64_t i = v0; i < v1; i++)
                distExecuteLouvainIteration(i, *dg, currComm, targetComm, vDegree, localCinfo, localCupdate,
                                            remoteComm, remoteCinfo, remoteCupdate, constant, clusterWeight, 0);
        }
        total += omp_get_wtime() - t0;
    }
    for (std::int64_t i = 0; i < nv; i++) target_out[i] = targetComm[i];
    delete dg;
    return total / (reps > 0 ? reps : 1);
}

// distInitLouvain pieces (louvain.cpp:2161-2218): vertex degrees and 1/2m.
double ref_init_degrees(std::int64_t nv, std::int64_t ne, const std::int64_t* offs, const Edge* edges,
                        double* vdeg_out)
{
    ensure_mpi();
    DistGraph* dg = make_dg(nv, ne, offs, edges);
    GraphWeightVector vDegree(nv);
    CommVector localCinfo(nv);
    distSumVertexDegree(dg->getLocalGraph(), vDegree, localCinfo);
    double c = distCalcConstantForSecondTerm(vDegree);
    for (std::int64_t i = 0; i < nv; i++) vdeg_out[i] = vDegree[i];
    delete dg;
    return c;
}

// distComputeModularity (louvain.cpp:2468-2532).
double ref_modularity(std::int64_t nv, const double* cw, const std::int64_t* csize, const double* cdeg,
                      double constant)
{
    ensure_mpi();
    Graph g(nv, 0);
    CommVector localCinfo(nv);
    for (std::int64_t i = 0; i < nv; i++) { localCinfo[i].size = csize[i]; localCinfo[i].degree = cdeg[i]; }
    GraphWeightVector clusterWeight(cw, cw + nv);
    return distComputeModularity(g, localCinfo, clusterWeight, constant, 0);
}

// One Louvain phase through the reference overloads (louvain.hpp:80-137).
// mode 0 = plain, 1 = WithColoring, 2 = VertexOrder.  et_type 0 = none, 1..4 as main.cpp:350-365.
double ref_phase(std::int64_t nv, std::int64_t ne, const std::int64_t* offs, const Edge* edges,
                 int mode, int et_type, double et_delta, long num_color, const std::int64_t* colors,
                 double lower, double thresh, std::int64_t* cvect_out, int* iters_out)
{
    ensure_mpi();
    DistGraph* dg = make_dg(nv, ne, offs, edges);
    std::vector<GraphElem> ssizes, rsizes, svdata, rvdata;
    size_t ssz = 0, rsz = 0;
    CommunityVector cvect;
    ColorVector vc;
    if (colors) vc.assign(colors, colors + nv);
    int iters = 0;
    double q = 0.0;
    const bool local = (et_type == 1 || et_type == 2);
    if (mode == 0) {
        if (et_type == 0) q = distLouvainMethod(0, 1, *dg, ssz, rsz, ssizes, rsizes, svdata, rvdata, cvect, lower, thresh, iters);
        else if (et_type == 1 || et_type == 3) q = distLouvainMethod(0, 1, *dg, ssz, rsz, ssizes, rsizes, svdata, rvdata, cvect, lower, thresh, iters, local);
        else q = distLouvainMethod(0, 1, *dg, ssz, rsz, ssizes, rsizes, svdata, rvdata, cvect, lower, thresh, iters, et_delta, local);
    } else if (mode == 1) {
        if (et_type == 0) q = distLouvainMethodWithColoring(0, 1, *dg, num_color, vc, ssz, rsz, ssizes, rsizes, svdata, rvdata, cvect, lower, thresh, iters);
        else if (et_type == 1 || et_type == 3) q = distLouvainMethodWithColoring(0, 1, *dg, num_color, vc, ssz, rsz, ssizes, rsizes, svdata, rvdata, cvect, lower, thresh, iters, local);
        else q = distLouvainMethodWithColoring(0, 1, *dg, num_color, vc, ssz, rsz, ssizes, rsizes, svdata, rvdata, cvect, lower, thresh, iters, et_delta, local);
    } else {
        if (et_type == 0) q = distLouvainMethodVertexOrder(0, 1, *dg, num_color, vc, ssz, rsz, ssizes, rsizes, svdata, rvdata, cvect, lower, thresh, iters);
        else if (et_type == 1 || et_type == 3) q = distLouvainMethodVertexOrder(0, 1, *dg, num_color, vc, ssz, rsz, ssizes, rsizes, svdata, rvdata, cvect, lower, thresh, iters, local);
        else q = distLouvainMethodVertexOrder(0, 1, *dg, num_color, vc, ssz, rsz, ssizes, rsizes, svdata, rvdata, cvect, lower, thresh, iters, et_delta, local);
    }
    for (std::int64_t i = 0; i < nv; i++) cvect_out[i] = cvect[i];
    *iters_out = iters;
    delete dg;
    return q;
}

// distbuildNextLevelGraph (rebuild.cpp:480-504).  Output arrays are malloc'd; free with ref_free.
int ref_rebuild(std::int64_t nv, std::int64_t ne, const std::int64_t* offs, const Edge* edges,
                const std::int64_t* cvect, std::int64_t* nv_out, std::int64_t* ne_out,
                std::int64_t** offs_out, Edge** edges_out)
{
    ensure_mpi();
    DistGraph* dg = make_dg(nv, ne, offs, edges);
    std::vector<GraphElem> ssizes(1, 0), rsizes(1, 0), svdata, rvdata;
    size_t ssz = 0, rsz = 0;
    CommunityVector cv(cvect, cvect + nv);
    distbuildNextLevelGraph(1, 0, dg, ssz, rsz, ssizes, rsizes, svdata, rvdata, cv);
    const Graph& g = dg->getLocalGraph();
    const std::int64_t nv2 = g.getNumVertices(), ne2 = g.getNumEdges();
    *nv_out = nv2;
    *ne_out = ne2;
    *offs_out = (std::int64_t*)std::malloc(sizeof(std::int64_t) * (nv2 + 1));
    *edges_out = (Edge*)std::malloc(sizeof(Edge) * (ne2 ? ne2 : 1));
    std::memcpy(*offs_out, g.edgeListIndexes.data(), sizeof(std::int64_t) * (nv2 + 1));
    if (ne2) std::memcpy((void*)*edges_out, (const void*)g.edgeList.data(), sizeof(Edge) * ne2);
    delete dg;
    return 0;
}

// distColoringMultiHashMinMax (coloring.cpp:52-118).
long ref_color(std::int64_t nv, std::int64_t ne, const std::int64_t* offs, const Edge* edges,
               long nhash, int target_pct, int single, std::int64_t* colors_out)
{
    ensure_mpi();
    DistGraph* dg = make_dg(nv, ne, offs, edges);
    ColorVector vc;
    long nc = distColoringMultiHashMinMax(0, 1, *dg, vc, nhash, target_pct, single != 0);
    for (std::int64_t i = 0; i < nv; i++) colors_out[i] = vc[i];
    delete dg;
    return nc;
}

// generateInMemGraph (distgraph.cpp:391-407) at one rank, no random edges.
int ref_gen_rgg(std::int64_t n, std::int64_t* nv_out, std::int64_t* ne_out,
                std::int64_t** offs_out, Edge** edges_out)
{
    ensure_mpi();
    DistGraph* dg = nullptr;
    generateInMemGraph(0, 1, dg, n, 0.0, std::string());
    const Graph& g = dg->getLocalGraph();
    const std::int64_t nv2 = g.getNumVertices(), ne2 = g.getNumEdges();
    *nv_out = nv2;
    *ne_out = ne2;
    *offs_out = (std::int64_t*)std::malloc(sizeof(std::int64_t) * (nv2 + 1));
    *edges_out = (Edge*)std::malloc(sizeof(Edge) * (ne2 ? ne2 : 1));
    std::memcpy(*offs_out, g.edgeListIndexes.data(), sizeof(std::int64_t) * (nv2 + 1));
    if (ne2) std::memcpy((void*)*edges_out, (const void*)g.edgeList.data(), sizeof(Edge) * ne2);
    delete dg;
    return 0;
}

void ref_free(void* p) { std::free(p); }

}  // extern "C"
#endif
