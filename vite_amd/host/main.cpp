// vite_amd_cluster: command-line driver with the reference's flags and output lines
// (main.cpp:108-607, getopt string main.cpp:614), running every Louvain phase, rebuild and
// coloring on one MI355X through the reference-shaped entry points in vite_louvain.hpp.
//
// Supported: -f -n -e(0 only) -s -j -i -t -a -c -d -p -h -o -g -z -b -r (-b reads through
// the edge-balanced loader, whose one part at world size 1 is the whole graph; -r, the
// ranks per node for MPI-IO aggregators, is accepted and has no effect).
// Workload graphs:  -x "rgg <seed> <n> <extra_pct>"  or  -x "ba <seed> <n> <m>".
#include <getopt.h>

#include <algorithm>
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <iostream>
#include <map>
#include <sstream>
#include <string>
#include <vector>

#include "vite_louvain.hpp"

namespace {

struct Options {
    std::string input, output_graph, workload;
    bool coloring = false, vertex_order = false, one_phase = false, single_color_iter = false;
    bool generate = false, just_process = false, output_communities = false, threshold_cycling = false;
    bool early_term = false, balanced = false;
    bool compare = false, truth_zero_based = true;  // -g FILE, -z (main.cpp:678-684)
    std::string ground_truth;
    int max_colors = 8, et_type = 0;
    GraphElem gen_nv = 0;
    GraphWeight random_edge_pct = 0.0, et_delta = 1.0, threshold = 1.0E-06;
};

double now()
{
    return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

[[noreturn]] void usage_error(const std::string& m)
{
    std::cerr << m << std::endl;
    std::exit(EXIT_FAILURE);
}

void parse_colors(const char* arg, Options& o)
{
    std::stringstream ss(arg);
    std::string tok;
    std::vector<std::string> parts;
    while (std::getline(ss, tok, ' ')) parts.push_back(tok);
    if (parts.empty()) usage_error("-c/-d needs a color count");
    o.max_colors = std::stoi(parts[0]);
    if (parts.size() > 1) {
        std::string b = parts[1];
        std::transform(b.begin(), b.end(), b.begin(), ::tolower);
        o.single_color_iter = (b == "true" || b == "1");
    }
}

Options parse(int argc, char** argv)
{
    Options o;
    int c;
    while ((c = getopt(argc, argv, "f:bc:od:r:t:a:ig:zpn:e:s:jh:x:")) != -1) {
        switch (c) {
            case 'f': o.input = optarg; break;
            case 'b': o.balanced = true; break;
            case 'r': break;  // ranks per node: no MPI-IO aggregators here
            case 'c': parse_colors(optarg, o); o.coloring = true; break;
            case 'd': parse_colors(optarg, o); o.vertex_order = true; break;
            case 'o': o.output_communities = true; break;
            case 't': o.early_term = true; o.et_type = std::atoi(optarg); break;
            case 'a': o.et_delta = std::atof(optarg); break;
            case 'i': o.threshold_cycling = true; break;
            case 'g': o.ground_truth = optarg; o.compare = true; break;
            case 'z': o.truth_zero_based = false; break;
            case 'p': o.one_phase = true; break;
            case 'n': o.generate = true; o.gen_nv = std::atol(optarg); break;
            case 'e': o.random_edge_pct = std::atof(optarg); break;
            case 's': o.output_graph = optarg; break;
            case 'j': o.just_process = true; break;
            case 'h': o.threshold = std::atof(optarg); break;
            case 'x': o.workload = optarg; break;
            default: usage_error("unknown option");
        }
    }
    // main.cpp:710-769
    if (o.early_term && (o.et_type < 1 || o.et_type > 4))
        usage_error("early-term-type parameter is between 1-4 : -t 1 OR -t 2 OR -t 3 OR -t 4");
    if (o.early_term && (o.et_type == 2 || o.et_type == 4) && (o.et_delta > 1.0 || o.et_delta < 0.0))
        usage_error("early-term-alpha must be between 0 and 1: -t 2 -a 0.5");
    if (!o.generate && o.workload.empty() && o.input.empty()) usage_error("Must specify a binary file name with -f");
    if (o.compare && o.ground_truth.empty()) usage_error("Must specify a ground-truth file name with -g");
    if (o.coloring && o.vertex_order)
        usage_error("Cannot turn on both coloring (-c) and vertex ordering (-d). Re-run using one of the options.");
    if (o.one_phase && o.threshold_cycling)
        usage_error("Cannot turn on threshold-cycling (-i) with a single phase run (-p). Re-run using one of the options.");
    if (!o.generate && o.random_edge_pct > 0.0)
        usage_error("Must specify -n <...> for graph generation first and then -p <...> to add random edges to it.");
    return o;
}

DistGraph* make_graph(const Options& o)
{
    if (!o.workload.empty()) {
        std::stringstream ss(o.workload);
        std::string kind;
        unsigned long long seed = 0;
        GraphElem n = 0;
        ss >> kind >> seed >> n;
        if (kind == "rgg") {
            double pct = 0.0;
            ss >> pct;
            return generateSeededRGG(seed, n, pct);
        }
        if (kind == "ba") {
            GraphElem m = 8;
            ss >> m;
            return generateSeededBA(seed, n, m);
        }
        usage_error("-x expects \"rgg <seed> <n> <extra_pct>\" or \"ba <seed> <n> <m>\"");
    }
    DistGraph* dg = nullptr;
    if (o.generate) {
        generateInMemGraph(0, 1, dg, o.gen_nv, o.random_edge_pct, o.output_graph);
    } else {
        std::string f = o.input;
        if (o.balanced) loadDistGraphMPIIOBalanced(0, 1, 1, dg, f);
        else loadDistGraphMPIIO(0, 1, 1, dg, f);
    }
    return dg;
}

// Renumbers a level's partition by ascending community id and composes it into the
// per-vertex map (main.cpp:384-425).
void accumulate_communities(const CommunityVector& cvect, int phase, std::vector<GraphElem>& commAll)
{
    std::vector<GraphElem> sorted(cvect);
    std::sort(sorted.begin(), sorted.end());
    sorted.erase(std::unique(sorted.begin(), sorted.end()), sorted.end());
    std::vector<GraphElem> ren(cvect.size());
    for (size_t i = 0; i < cvect.size(); i++)
        ren[i] = std::lower_bound(sorted.begin(), sorted.end(), cvect[i]) - sorted.begin();
    if (phase == 0) std::copy(ren.begin(), ren.end(), commAll.begin());
    else
        for (auto& c : commAll) c = ren[c];
}

}  // namespace

int main(int argc, char** argv)
{
    const Options o = parse(argc, argv);
    // main.cpp:119-137: the ground truth is read before the graph
    std::vector<GraphElem> commGroundTruth;
    if (o.compare) {
        const double t1 = now();
        loadGroundTruthFile(commGroundTruth, o.ground_truth, o.truth_zero_based);
        std::cout << "Time taken to load ground truth file (" << o.ground_truth << "): " << (now() - t1) << std::endl;
    }
    const double td0 = now();
    DistGraph* dg = make_graph(o);
    dg->printStats();
    const double td1 = now();
    std::cout << "Time to create distributed graph: " << (td1 - td0) << std::endl;
    if (!o.output_graph.empty() && !o.generate) writeGraphBin(*dg, o.output_graph);
    if (o.just_process) {
        std::cout << "Option to just process the graph is turned on, so exiting without invoking community detection."
                  << std::endl;
        delete dg;
        return 0;
    }

    const GraphElem nv0 = dg->getTotalNumVertices();
    std::vector<GraphElem> commAll;
    const bool track = o.output_communities || o.compare;
    if (track) commAll.assign(nv0, -1);
    std::vector<GraphElem> ssizes, rsizes, svdata, rvdata;
    size_t ssz = 0, rsz = 0;
    CommunityVector cvect;
    ColorVector colors;
    GraphWeight currMod = -1.0, prevMod = -1.0, threshold = o.threshold;
    int phase = 0, short_phase = 0, iters = 0, tot_iters = 0;
    double total = td1 - td0, ctime = 0.0;

    for (;;) {
        double ptotal = 0.0;
        if (o.threshold_cycling && !o.one_phase) {
            if (short_phase > 12) short_phase = 0;
            threshold = short_phase <= 2 ? 1.0E-3 : short_phase <= 6 ? 1.0E-4 : short_phase <= 9 ? 1.0E-5 : o.threshold;
        } else {
            threshold = o.threshold;
        }
        double t1 = now();
        const bool colored_phase = (o.coloring || o.vertex_order) && phase == 0;
        if (colored_phase) {
            const ColorElem numColors =
                distColoringMultiHashMinMax(0, 1, *dg, colors, o.max_colors / 2, 70, o.single_color_iter);
            std::cout << "Number of colors (2*nHash): " << numColors << std::endl;
            const double tc = now();
            std::cout << "Coloring Time: " << tc - t1 << std::endl;
            t1 = now();
            auto call = [&](auto fn) {
                if (o.early_term && o.et_type == 1)
                    return fn(true, 0.0, 1);
                if (o.early_term && o.et_type == 2)
                    return fn(true, o.et_delta, 2);
                if (o.early_term && o.et_type == 3)
                    return fn(false, 0.0, 1);
                if (o.early_term && o.et_type == 4)
                    return fn(false, o.et_delta, 2);
                return fn(false, 0.0, 0);
            };
            currMod = call([&](bool local, double delta, int kind) {
                if (o.coloring) {
                    if (kind == 0)
                        return distLouvainMethodWithColoring(0, 1, *dg, numColors + 1, colors, ssz, rsz, ssizes, rsizes,
                                                             svdata, rvdata, cvect, currMod, threshold, iters);
                    if (kind == 1)
                        return distLouvainMethodWithColoring(0, 1, *dg, numColors + 1, colors, ssz, rsz, ssizes, rsizes,
                                                             svdata, rvdata, cvect, currMod, threshold, iters, local);
                    return distLouvainMethodWithColoring(0, 1, *dg, numColors + 1, colors, ssz, rsz, ssizes, rsizes,
                                                         svdata, rvdata, cvect, currMod, threshold, iters, delta, local);
                }
                if (kind == 0)
                    return distLouvainMethodVertexOrder(0, 1, *dg, numColors + 1, colors, ssz, rsz, ssizes, rsizes,
                                                        svdata, rvdata, cvect, currMod, threshold, iters);
                if (kind == 1)
                    return distLouvainMethodVertexOrder(0, 1, *dg, numColors + 1, colors, ssz, rsz, ssizes, rsizes,
                                                        svdata, rvdata, cvect, currMod, threshold, iters, local);
                return distLouvainMethodVertexOrder(0, 1, *dg, numColors + 1, colors, ssz, rsz, ssizes, rsizes, svdata,
                                                    rvdata, cvect, currMod, threshold, iters, delta, local);
            });
        } else {
            if (o.early_term && (o.et_type == 1 || o.et_type == 3))
                currMod = distLouvainMethod(0, 1, *dg, ssz, rsz, ssizes, rsizes, svdata, rvdata, cvect, currMod,
                                            threshold, iters, o.et_type == 1);
            else if (o.early_term && (o.et_type == 2 || o.et_type == 4))
                currMod = distLouvainMethod(0, 1, *dg, ssz, rsz, ssizes, rsizes, svdata, rvdata, cvect, currMod,
                                            threshold, iters, o.et_delta, o.et_type == 2);
            else
                currMod = distLouvainMethod(0, 1, *dg, ssz, rsz, ssizes, rsizes, svdata, rvdata, cvect, currMod,
                                            threshold, iters);
        }
        const double t0 = now();
        tot_iters += iters;
        ptotal += t0 - t1;
        ctime += t0 - t1;

        if ((currMod - prevMod) > threshold) {
            if (track) accumulate_communities(cvect, phase, commAll);
            if (!o.one_phase) {
                const double t3 = now();
                distbuildNextLevelGraph(1, 0, dg, ssz, rsz, ssizes, rsizes, svdata, rvdata, cvect);
                const double t2 = now();
                std::cout << "**************************" << std::endl;
                std::cout << "Rebuild Time: " << t2 - t3 << std::endl;
                ptotal += t2 - t3;
            }
        } else {
            if (o.threshold_cycling && !o.one_phase && phase < 10) {
                const double tx = now();
                currMod = distLouvainMethod(0, 1, *dg, ssz, rsz, ssizes, rsizes, svdata, rvdata, cvect, currMod,
                                            1.0E-6, iters);
                ptotal += now() - tx;
            }
            break;
        }
        std::cout << "Level " << phase << ", Modularity: " << currMod << ", Clustering time: " << t0 - t1
                  << ", Iterations: " << tot_iters << std::endl;
        prevMod = currMod;
        currMod = -1.0;
        std::cout << "Phase Total time: " << ptotal << std::endl;
        total += ptotal;
        phase++;
        if (o.one_phase) {
            std::cout << "Exiting at the end of first phase." << std::endl;
            break;
        }
        if (o.threshold_cycling && !o.one_phase) short_phase++;
        if (phase == 200 || tot_iters > 10000) {
            std::cout << "Program ran for 200 phases or 10K iterations, exiting..." << std::endl;
            break;
        }
    }

    std::cout.precision(17);
    std::cout << "-------------------------------------------------------" << std::endl;
    std::cout << "Average total time (secs.): " << total << std::endl;
    std::cout << "Average time for clustering (secs.): " << ctime << std::endl;
    std::cout << "Convergence threshold (for fixed-threshold workloads): " << threshold << std::endl;
    std::cout << "Final modularity: " << currMod << std::endl;
    std::cout << "MODS (final modularity * average clustering time): " << currMod * ctime << std::endl;
    std::cout << "-------------------------------------------------------" << std::endl;

    if (o.output_communities) {
        const std::string out = (o.input.empty() ? std::string("generated") : o.input) + ".communities";
        std::ofstream f(out);
        if (!f) usage_error("Error creating community file: " + out);
        for (GraphElem i = 0; i < nv0; i++) f << commAll[i] << "\n";
        std::cout << "Saved community information (" << commAll.size() << " vertices)  on file: " << out << std::endl;
    }
    // main.cpp:574-580
    if (o.compare) {
        if (commGroundTruth.size() != commAll.size())
            usage_error("ground truth holds " + std::to_string(commGroundTruth.size()) + " vertices, the graph " +
                        std::to_string(commAll.size()));
        compare_communities(commGroundTruth, commAll);
    }
    delete dg;
    return 0;
}
