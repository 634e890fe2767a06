// Prints the object layout of Edge, Comm, Graph and DistGraph as seen through one header:
// compiled once against the reference's distgraph.hpp and once against vite_amd's
// vite_graph.hpp, the two outputs must be identical for objects to cross between code
// compiled from either.  It also exercises the public surface a caller of the reference
// uses (copy construction, createLocalGraph, the range and owner queries).
#include <iostream>  // the reference header uses std::cout without including it

#include GRAPH_HEADER

#include <cstddef>
#include <cstdio>

namespace {

template <class T, class M>
long off(const T* base, const M* member)
{
    return (long)((const char*)member - (const char*)base);
}

struct GraphView : Graph {
    GraphView() : Graph(3, 5) {}
    void dump() const
    {
        const Graph* b = this;
        std::printf("Graph size=%zu numVertices=%ld numEdges=%ld edgeListIndexes=%ld edgeList=%ld\n", sizeof(Graph),
                    off(b, &numVertices), off(b, &numEdges), off(b, &edgeListIndexes), off(b, &edgeList));
    }
};

struct DistGraphView : DistGraph {
    DistGraphView() : DistGraph(7, 9) {}
    void dump() const
    {
        const DistGraph* b = this;
        std::printf("DistGraph size=%zu totalNumVertices=%ld totalNumEdges=%ld localGraph=%ld parts=%ld\n",
                    sizeof(DistGraph), off(b, &totalNumVertices), off(b, &totalNumEdges), off(b, &localGraph),
                    off(b, &parts));
    }
};

}  // namespace

int main()
{
    Edge e;
    Comm c;
    std::printf("Edge size=%zu tail=%ld weight=%ld\n", sizeof(Edge), off(&e, &e.tail), off(&e, &e.weight));
    std::printf("Comm size=%zu size=%ld degree=%ld\n", sizeof(Comm), off(&c, &c.size), off(&c, &c.degree));
    GraphView().dump();
    DistGraphView().dump();

    // behaviour of the inline surface on a 3-vertex path 0-1-2
    PartRanges parts{0, 3};
    DistGraph dg(3, 4);
    dg.createLocalGraph(3, 4, &parts);
    Graph& g = dg.getLocalGraph();
    const GraphElem offs[4] = {0, 1, 3, 4};
    const GraphElem tails[4] = {1, 0, 2, 1};
    for (int v = 0; v <= 3; v++) g.setEdgeStartForVertex(v, offs[v]);
    for (int k = 0; k < 4; k++) {
        g.getEdge(k).tail = tails[k];
        g.getEdge(k).weight = 0.5 * (k + 1);
    }
    DistGraph copy(dg);
    const Graph& cg = copy.getLocalGraph();
    GraphElem e0 = 0, e1 = 0;
    cg.getEdgeRangeForVertex(1, e0, e1);
    std::printf("copy nv=%ld ne=%ld tnv=%ld tne=%ld row1=[%ld,%ld) w2=%.2f base=%ld bound=%ld owner2=%d owner3=%d\n",
                (long)cg.getNumVertices(), (long)cg.getNumEdges(), (long)copy.getTotalNumVertices(),
                (long)copy.getTotalNumEdges(), (long)e0, (long)e1, cg.getEdge(2).weight, (long)copy.getBase(0),
                (long)copy.getBound(0), copy.getOwner(2), copy.getOwner(3));
    g.setNumEdges(2);
    std::printf("after setNumEdges nv=%ld ne=%ld size=%zu copy_ne=%ld\n", (long)g.getNumVertices(),
                (long)g.getNumEdges(), g.edgeList.size(), (long)cg.getNumEdges());
    return 0;
}
