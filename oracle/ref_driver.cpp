// ORACLE — TEST INFRASTRUCTURE ONLY.
//
// A C entry layer over the parts of the reference that compile from their own sources
// with g++ alone: the edge-list loader (src/RPGraph.cpp), the common helpers
// (src/RPCommon.cpp) and the CPU Barnes-Hut (src/RPBarnesHutApproximator.cpp). The
// recipe in oracle/Makefile builds them where they lie in the reference checkout into
// oracle/_ref/libref.so; nothing of the reference is copied here. Tests use it to pin
// the engine's loader and the CPU-baseline restatement.
#include <cstdint>
#include <cstdlib>
#include <string>
#include <vector>
#include "RPGraph.hpp"
#include "RPBarnesHutApproximator.hpp"

static RPGraph::UGraph* g_graph = nullptr;

extern "C" {

// Loads an edge list with the reference's UGraph. Returns 0 on success.
int ref_load(const char* path, int* n, long long* m) {
    delete g_graph;
    g_graph = new RPGraph::UGraph(std::string(path));
    *n = (int)g_graph->num_nodes();
    *m = (long long)g_graph->num_edges();
    return 0;
}

// node_map_r: internal id -> id in the file.
void ref_node_ids(long long* ids) {
    for (uint32_t v = 0; v < g_graph->num_nodes(); ++v) ids[v] = g_graph->node_map_r[v];
}

void ref_degrees(int* deg) {
    for (uint32_t v = 0; v < g_graph->num_nodes(); ++v) deg[v] = (int)g_graph->degree(v);
}

// The COO list CUDAForceAtlas2 stages (src/RPGPUForceAtlas2.cu:73-84).
void ref_coo(int* src, int* dst) {
    long long k = 0;
    for (uint32_t v = 0; v < g_graph->num_nodes(); ++v)
        for (uint32_t t : g_graph->neighbors_with_geq_id(v)) { src[k] = (int)v; dst[k] = (int)t; ++k; }
}

// The reference's initial layout draw: srandom(seed) (src/graph_viewer.cpp:49), then per
// node x in [-w/2, w/2] and y in [-h/2, h/2] from RPGraph::get_random, as
// GraphLayout::randomizePositions does (src/RPGraphLayout.cpp:45-52).
void ref_random_positions(unsigned seed, int n, float w, float h, float* xy) {
    srandom(seed);
    for (int i = 0; i < n; ++i) {
        xy[2 * i] = RPGraph::get_random(-w / 2.0, w / 2.0);
        xy[2 * i + 1] = RPGraph::get_random(-h / 2.0, h / 2.0);
    }
}

// The reference's CPU Barnes-Hut: insert n particles into a root (cx, cy, len) in order,
// then the force on each query particle.
void ref_bh_forces(int n, const float* pos, const float* mass, float cx, float cy, float len,
                   float theta, int nq, const float* qpos, const float* qmass, float* out) {
    RPGraph::BarnesHutApproximator bh(RPGraph::Coordinate(cx, cy), len, theta);
    for (int i = 0; i < n; ++i)
        bh.insertParticle(RPGraph::Coordinate(pos[2 * i], pos[2 * i + 1]), mass[i]);
    for (int k = 0; k < nq; ++k) {
        RPGraph::Real2DVector f =
            bh.approximateForce(RPGraph::Coordinate(qpos[2 * k], qpos[2 * k + 1]), qmass[k], theta);
        out[2 * k] = f.x;
        out[2 * k + 1] = f.y;
    }
}

}
