// The reference's C++ entry points for the Louvain path, with the reference signatures,
// implemented on one MI355X through the C ABI (include/vite_amd.h).  A program written
// against Vite's louvain.hpp / rebuild.hpp / coloring.hpp / distgraph.hpp links against
// this instead, at MPI world size 1 (nprocs must be 1; me must be 0).
//
// Errors follow the reference's convention: a message on stderr, then the process exits
// (the reference calls MPI_Abort(-99), main.cpp:714-768).
//
// The device copy of a DistGraph is cached by address and reused by later calls (the
// Louvain phase, then the rebuild) while a hash of its whole CSR is unchanged; a graph
// edited in place between calls is uploaded again.
#pragma once

#include <string>
#include <vector>

#include "vite_graph.hpp"

// ---- louvain.hpp:80-137 -------------------------------------------------------------
GraphWeight distLouvainMethod(const int me, const int nprocs, const DistGraph& dg, size_t& ssz, size_t& rsz,
                              std::vector<GraphElem>& ssizes, std::vector<GraphElem>& rsizes,
                              std::vector<GraphElem>& svdata, std::vector<GraphElem>& rvdata,
                              CommunityVector& cvect, const GraphWeight lower, const GraphWeight thresh, int& iters);

GraphWeight distLouvainMethod(const int me, const int nprocs, const DistGraph& dg, size_t& ssz, size_t& rsz,
                              std::vector<GraphElem>& ssizes, std::vector<GraphElem>& rsizes,
                              std::vector<GraphElem>& svdata, std::vector<GraphElem>& rvdata,
                              CommunityVector& cvect, const GraphWeight lower, const GraphWeight thresh, int& iters,
                              bool ETLocalOrRemote);

GraphWeight distLouvainMethod(const int me, const int nprocs, const DistGraph& dg, size_t& ssz, size_t& rsz,
                              std::vector<GraphElem>& ssizes, std::vector<GraphElem>& rsizes,
                              std::vector<GraphElem>& svdata, std::vector<GraphElem>& rvdata,
                              CommunityVector& cvect, const GraphWeight lower, const GraphWeight thresh, int& iters,
                              GraphWeight ETDelta, bool ETLocalOrRemote);

GraphWeight distLouvainMethodWithColoring(const int me, const int nprocs, const DistGraph& dg, const long numColor,
                                          const ColorVector& vertexColor, size_t& ssz, size_t& rsz,
                                          std::vector<GraphElem>& ssizes, std::vector<GraphElem>& rsizes,
                                          std::vector<GraphElem>& svdata, std::vector<GraphElem>& rvdata,
                                          CommunityVector& cvect, const GraphWeight lower, const GraphWeight thresh,
                                          int& iters);

GraphWeight distLouvainMethodWithColoring(const int me, const int nprocs, const DistGraph& dg, const long numColor,
                                          const ColorVector& vertexColor, size_t& ssz, size_t& rsz,
                                          std::vector<GraphElem>& ssizes, std::vector<GraphElem>& rsizes,
                                          std::vector<GraphElem>& svdata, std::vector<GraphElem>& rvdata,
                                          CommunityVector& cvect, const GraphWeight lower, const GraphWeight thresh,
                                          int& iters, bool ETLocalOrRemote);

GraphWeight distLouvainMethodWithColoring(const int me, const int nprocs, const DistGraph& dg, const long numColor,
                                          const ColorVector& vertexColor, size_t& ssz, size_t& rsz,
                                          std::vector<GraphElem>& ssizes, std::vector<GraphElem>& rsizes,
                                          std::vector<GraphElem>& svdata, std::vector<GraphElem>& rvdata,
                                          CommunityVector& cvect, const GraphWeight lower, const GraphWeight thresh,
                                          int& iters, GraphWeight ETDelta, bool ETLocalOrRemote);

GraphWeight distLouvainMethodVertexOrder(const int me, const int nprocs, const DistGraph& dg, const long numColor,
                                         const ColorVector& vertexColor, size_t& ssz, size_t& rsz,
                                         std::vector<GraphElem>& ssizes, std::vector<GraphElem>& rsizes,
                                         std::vector<GraphElem>& svdata, std::vector<GraphElem>& rvdata,
                                         CommunityVector& cvect, const GraphWeight lower, const GraphWeight thresh,
                                         int& iters);

GraphWeight distLouvainMethodVertexOrder(const int me, const int nprocs, const DistGraph& dg, const long numColor,
                                         const ColorVector& vertexColor, size_t& ssz, size_t& rsz,
                                         std::vector<GraphElem>& ssizes, std::vector<GraphElem>& rsizes,
                                         std::vector<GraphElem>& svdata, std::vector<GraphElem>& rvdata,
                                         CommunityVector& cvect, const GraphWeight lower, const GraphWeight thresh,
                                         int& iters, bool ETLocalOrRemote);

GraphWeight distLouvainMethodVertexOrder(const int me, const int nprocs, const DistGraph& dg, const long numColor,
                                         const ColorVector& vertexColor, size_t& ssz, size_t& rsz,
                                         std::vector<GraphElem>& ssizes, std::vector<GraphElem>& rsizes,
                                         std::vector<GraphElem>& svdata, std::vector<GraphElem>& rvdata,
                                         CommunityVector& cvect, const GraphWeight lower, const GraphWeight thresh,
                                         int& iters, GraphWeight ETDelta, bool ETLocalOrRemote);

// ---- rebuild.hpp:109-112 ---------------------------------------------------------------
void distbuildNextLevelGraph(int nprocs, int me, DistGraph*& dg, const size_t& ssz, const size_t& rsz,
                             const std::vector<GraphElem>& ssizes, const std::vector<GraphElem>& rsizes,
                             const std::vector<GraphElem>& svdata, const std::vector<GraphElem>& rvdata,
                             CommunityVector& cvect);

// ---- coloring.hpp:86 ---------------------------------------------------------------------
ColorElem distColoringMultiHashMinMax(const int me, const int nprocs, const DistGraph& dg, ColorVector& vertexColor,
                                      const ColorElem nHash, const int target_percent, const bool singleIteration);

// ---- distgraph.hpp:109-119 ------------------------------------------------------------
void balanceEdges(int nprocs, std::string& fileName, std::vector<GraphElem>& mbins);
void loadDistGraphMPIIO(int me, int nprocs, int ranks_per_node, DistGraph*& dg, std::string& fileName);
void loadDistGraphMPIIOBalanced(int me, int nprocs, int ranks_per_node, DistGraph*& dg, std::string& fileName);
void generateInMemGraph(int rank, int nprocs, DistGraph*& dg, GraphElem nv, GraphWeight randomEdgePercent,
                        std::string fileOut);
void writeGraph(int me, int nprocs, DistGraph*& dg, std::vector<GraphElem>& edgeCount, std::string& fileName);
// The same .bin file from the graph's own offsets (no per-row counts needed).
void writeGraphBin(const DistGraph& dg, const std::string& fileName);

// ---- louvain.hpp:206-208 -----------------------------------------------------------------
void gatherAllComm(int root, int me, int nprocs, std::vector<GraphElem>& commAll, const CommunityVector& localComm);

// ---- ground truth: louvain.hpp:201-203, compare.hpp:56-57 -----------------------------
// LFR-style "vertex community" lines; community ids one-based unless zero-based is set.
void loadGroundTruthFile(std::vector<GraphElem>& commGroundTruth, std::string const& groundTruthFileName,
                         bool isGroundTruthZeroBased);
// Prints the reference's comparison report for C1 (truth) against C2 (computed); the pair
// counts and scores come from vt_compare_communities on the device.
void compare_communities(std::vector<GraphElem> const& C1, std::vector<GraphElem> const& C2);
double compute_gini_coeff(GraphElem* colorSize, int numColors);

// The workload generators (seeded RGG with extra edges, Barabasi-Albert) on the device.
DistGraph* generateSeededRGG(unsigned long long seed, GraphElem n, GraphWeight extraPercent);
DistGraph* generateSeededBA(unsigned long long seed, GraphElem n, GraphElem m);

// No-ops kept for source compatibility (louvain.hpp:192-193, rebuild.hpp:93-94).
void createCommunityMPIType();
void destroyCommunityMPIType();
void createEdgeMPIType();
void destroyEdgeMPIType();

// Library context used by the functions above (device 0 unless VITE_AMD_DEVICE is set).
struct vt_ctx;
vt_ctx* vite_amd_context();
// Frees the device copy of dg, if one is cached (call before deleting a graph whose device
// copy should not outlive it; a later call on the same address uploads it again).
void vite_amd_release(const DistGraph* dg);
