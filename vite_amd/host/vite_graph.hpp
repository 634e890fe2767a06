// Host graph types, object-layout compatible with the reference's (edge.hpp:60-98,
// graph.hpp:62-231, distgraph.hpp:77-286), so a translation unit compiled against Vite's
// own headers can hand its Graph / DistGraph objects to this library and take back ones
// this library allocated:
//
//   Edge      {tail, weight}                                         16 bytes
//   Comm      {size, degree}                                         16 bytes
//   Graph     {numVertices, numEdges, edgeListIndexes, edgeList}     64 bytes
//   DistGraph {totalNumVertices, totalNumEdges, localGraph, parts}   32 bytes
//
// Member order, access and the inline methods' behaviour follow the reference, so the
// mangled entry points in vite_louvain.hpp bind to the same objects either way.  The
// only behavioural difference is MPI: these types live at world size 1, so printStats
// reports the one rank's figures without MPI reductions (the same lines the reference
// prints at one rank), and createLocalGraph accepts a null part list (one part).
#pragma once

#include <algorithm>
#include <cassert>
#include <cmath>
#include <cstdint>
#include <iostream>
#include <ostream>
#include <vector>

typedef int64_t GraphElem;
typedef double GraphWeight;
typedef int64_t ColorElem;

// 16-byte edge record (edge.hpp:72-80).
struct Edge {
    GraphElem tail;
    GraphWeight weight;
    Edge() : tail(-1), weight(0.0) {}
    Edge(GraphElem t, GraphWeight w) : tail(t), weight(w) {}
};
static_assert(sizeof(Edge) == 16, "Edge must stay 16 bytes");

// Community record (graph.hpp:62-67).
struct Comm {
    GraphElem size;
    GraphWeight degree;
    Comm() : size(0), degree(0) {}
};

typedef std::vector<GraphElem> EdgeIndexes;
typedef std::vector<GraphElem> CommunityVector;
typedef std::vector<GraphWeight> GraphWeightVector;
typedef std::vector<ColorElem> ColorVector;
typedef std::vector<GraphElem> PartRanges;
typedef std::vector<Comm> CommVector;

// CSR graph (graph.hpp:77-106): counts first (protected), then the public CSR vectors.
class Graph {
protected:
    typedef std::vector<Edge> EdgeList;
    GraphElem numVertices;
    GraphElem numEdges;

public:
    EdgeIndexes edgeListIndexes;
    EdgeList edgeList;

    Graph(const GraphElem onv, const GraphElem one)
        : numVertices(onv), numEdges(one), edgeListIndexes(onv + 1, 0), edgeList(one)
    {
    }
    Graph(const Graph& other)
        : numVertices(other.numVertices), numEdges(other.numEdges), edgeListIndexes(other.edgeListIndexes),
          edgeList(other.edgeList)
    {
    }
    ~Graph() {}

    GraphElem getNumVertices() const { return numVertices; }
    GraphElem getNumEdges() const { return numEdges; }
    void setEdgeWeightstoOne()
    {
        for (GraphElem e = 0; e < numEdges; e++) edgeList[e].weight = 1.0;
    }
    void setNumEdges(GraphElem ne)
    {
        edgeList.resize(ne);
        numEdges = ne;
    }
    void getEdgeRangeForVertex(const GraphElem v, GraphElem& e0, GraphElem& e1) const
    {
        assert(v >= 0 && v < numVertices);
        e0 = edgeListIndexes[v];
        e1 = edgeListIndexes[v + 1];
    }
    const Edge& getEdge(const GraphElem e) const { return edgeList[e]; }
    void setEdgeStartForVertex(const GraphElem v, const GraphElem e0) { edgeListIndexes[v] = e0; }
    Edge& getEdge(const GraphElem e) { return edgeList[e]; }

    friend std::ostream& operator<<(std::ostream& os, const Graph&) { return os; }

protected:
    Graph() : numVertices(0), numEdges(0) {}
    Graph& operator=(const Graph&) = default;
};

// 1D-partitioned graph (distgraph.hpp:77-107): counts and the local graph (protected),
// then the public part list.  At world size 1 one part covers every vertex.
class DistGraph {
protected:
    GraphElem totalNumVertices;
    GraphElem totalNumEdges;
    Graph* localGraph;

public:
    DistGraph(const GraphElem tnv, const GraphElem tne)
        : totalNumVertices(tnv), totalNumEdges(tne), localGraph(nullptr), parts(nullptr)
    {
    }
    DistGraph(const DistGraph& other)
        : totalNumVertices(other.totalNumVertices), totalNumEdges(other.totalNumEdges),
          localGraph(new Graph(*other.localGraph)), parts(new PartRanges(*other.parts))
    {
    }
    ~DistGraph()
    {
        delete localGraph;
        delete parts;
    }

    GraphElem getTotalNumVertices() const { return totalNumVertices; }
    GraphElem getTotalNumEdges() const { return totalNumEdges; }

    void createLocalGraph(const GraphElem lnv, const GraphElem lne, const PartRanges* oparts = nullptr)
    {
        localGraph = new Graph(lnv, lne);
        parts = oparts ? new PartRanges(*oparts) : new PartRanges{0, lnv};
    }
    Graph& getLocalGraph() { return *localGraph; }
    const Graph& getLocalGraph() const { return *localGraph; }
    GraphElem getBase(const int me) const { return (*parts)[me]; }
    GraphElem getBound(const int me) const { return (*parts)[me + 1]; }
    GraphElem localToGlobal(GraphElem idx, const int me) const { return idx + getBase(me); }
    GraphElem globalToLocal(GraphElem idx, const int me) const { return idx - getBase(me); }
    int getOwner(const GraphElem v) const
    {
        // the last part whose start is <= v (distgraph.hpp:249-262)
        return (int)(std::upper_bound(parts->begin(), parts->end(), v) - parts->begin()) - 1;
    }
    PartRanges* parts;
    void setNumEdges(GraphElem ne) { totalNumEdges = ne; }

    // distgraph.hpp:143-189 at world size 1: the one rank's edge count is the sum, maximum
    // and minimum, and the variance is zero.
    void printStats()
    {
        const long lne = (long)getLocalGraph().getNumEdges();
        const long nv = (long)totalNumVertices;
        const double average = (double)lne;
        const double avg_sq = (double)(lne * lne);
        const double var = avg_sq - average * average;
        std::cout << std::endl;
        std::cout << "-------------------------------------------------------" << std::endl;
        std::cout << "Graph edge distribution characteristics" << std::endl;
        std::cout << "-------------------------------------------------------" << std::endl;
        std::cout << "Number of vertices: " << nv << std::endl;
        std::cout << "Number of edges: " << lne << std::endl;
        std::cout << "Maximum number of edges: " << lne << std::endl;
        std::cout << "Average number of edges: " << average << std::endl;
        std::cout << "Minimum number of edges: " << lne << std::endl;
        std::cout << "Expected value of X^2: " << avg_sq << std::endl;
        std::cout << "Variance: " << var << std::endl;
        std::cout << "Standard deviation: " << std::sqrt(var) << std::endl;
        std::cout << "-------------------------------------------------------" << std::endl;
    }
    GraphElem getNumGhosts(const int me) const
    {
        const Graph& g = getLocalGraph();
        GraphElem ghosts = 0;
        for (GraphElem e = 0; e < g.getNumEdges(); e++)
            if (getOwner(g.edgeList[e].tail) != me) ghosts++;
        return ghosts;
    }

protected:
    DistGraph() : totalNumVertices(0), totalNumEdges(0), localGraph(nullptr), parts(nullptr) {}
    DistGraph& operator=(const DistGraph&) = default;
};
