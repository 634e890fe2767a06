// Host graph types with the reference's layout and public members, so code written
// against Vite's Graph / DistGraph (edge.hpp:60-98, graph.hpp:62-231,
// distgraph.hpp:77-286) compiles unchanged against this library at world size 1.
// The CSR vectors are public, as in the reference, so their data() and size() can be
// handed to the C ABI without copies.
#pragma once

#include <cassert>
#include <cstdint>
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

// CSR graph (graph.hpp:77-106).
class Graph {
public:
    EdgeIndexes edgeListIndexes;
    std::vector<Edge> edgeList;

    Graph(GraphElem nv, GraphElem ne) : edgeListIndexes(nv + 1, 0), edgeList(ne), numVertices(nv), numEdges(ne) {}
    GraphElem getNumVertices() const { return numVertices; }
    GraphElem getNumEdges() const { return numEdges; }
    void setNumEdges(GraphElem ne)
    {
        edgeList.resize(ne);
        numEdges = ne;
    }
    void getEdgeRangeForVertex(GraphElem v, GraphElem& e0, GraphElem& e1) const
    {
        e0 = edgeListIndexes[v];
        e1 = edgeListIndexes[v + 1];
    }
    const Edge& getEdge(GraphElem e) const { return edgeList[e]; }
    Edge& getEdge(GraphElem e) { return edgeList[e]; }
    void setEdgeStartForVertex(GraphElem v, GraphElem e0) { edgeListIndexes[v] = e0; }

private:
    GraphElem numVertices, numEdges;
};

// 1D-partitioned graph (distgraph.hpp:77-107); at world size 1 one part covers everything.
class DistGraph {
public:
    PartRanges* parts = nullptr;

    DistGraph(GraphElem tnv, GraphElem tne) : totalNumVertices(tnv), totalNumEdges(tne) {}
    DistGraph(const DistGraph&) = delete;
    DistGraph& operator=(const DistGraph&) = delete;
    ~DistGraph()
    {
        delete localGraph;
        delete parts;
    }
    void createLocalGraph(GraphElem lnv, GraphElem lne, const PartRanges* oparts = nullptr)
    {
        localGraph = new Graph(lnv, lne);
        parts = oparts ? new PartRanges(*oparts) : new PartRanges{0, lnv};
    }
    GraphElem getTotalNumVertices() const { return totalNumVertices; }
    GraphElem getTotalNumEdges() const { return totalNumEdges; }
    void setNumEdges(GraphElem ne) { totalNumEdges = ne; }
    Graph& getLocalGraph() { return *localGraph; }
    const Graph& getLocalGraph() const { return *localGraph; }
    GraphElem getBase(int me) const { return (*parts)[me]; }
    GraphElem getBound(int me) const { return (*parts)[me + 1]; }
    GraphElem localToGlobal(GraphElem idx, int me) const { return idx + getBase(me); }
    GraphElem globalToLocal(GraphElem idx, int me) const { return idx - getBase(me); }
    int getOwner(GraphElem) const { return 0; }
    void printStats();

private:
    GraphElem totalNumVertices, totalNumEdges;
    Graph* localGraph = nullptr;
};
