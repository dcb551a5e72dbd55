// C entry points of the host layer (include/gfa2host.h).
#include "gfa2host.h"

#include <cstring>
#include <string>
#include <vector>

#include "RPGraph.hpp"
#include "RPGraphLayout.hpp"

struct gfa2host_graph
{
    RPGraph::UGraph g;
};

extern "C" {

int gfa2host_load(const char* path, gfa2host_graph_t* out, int32_t* n, int64_t* m)
{
    if (!path || !out) return -1;
    gfa2host_graph* h = new gfa2host_graph{RPGraph::UGraph(std::string(path))};
    if (!h->g.ok())
    {
        delete h;
        return -1;
    }
    *out = h;
    if (n) *n = (int32_t)h->g.num_nodes();
    if (m) *m = (int64_t)h->g.num_edges();
    return 0;
}

int gfa2host_from_edges(int64_t count, const uint32_t* s, const uint32_t* t, gfa2host_graph_t* out, int32_t* n,
                        int64_t* m)
{
    if (!out || count < 0 || (count > 0 && (!s || !t))) return -1;
    std::vector<RPGraph::nid_t> a(s, s + count), b(t, t + count);
    gfa2host_graph* h = new gfa2host_graph{RPGraph::UGraph(a, b)};
    *out = h;
    if (n) *n = (int32_t)h->g.num_nodes();
    if (m) *m = (int64_t)h->g.num_edges();
    return 0;
}

void gfa2host_graph_arrays(gfa2host_graph_t h, int64_t* ids, int32_t* deg, int32_t* src, int32_t* dst)
{
    const RPGraph::UGraph& g = h->g;
    for (RPGraph::nid_t v = 0; v < g.num_nodes(); ++v)
    {
        if (ids) ids[v] = g.node_map_r[v];
        if (deg) deg[v] = (int32_t)g.degree(v);
    }
    if (src || dst)
    {
        std::vector<int32_t> s, d;
        g.coo(s, d);
        if (src) memcpy(src, s.data(), s.size() * 4);
        if (dst) memcpy(dst, d.data(), d.size() * 4);
    }
}

int gfa2host_save_binary(gfa2host_graph_t h, const char* path) { return h->g.save_binary(path) ? 0 : -1; }

int gfa2host_write_layout(gfa2host_graph_t h, const float* xy, const char* path, const char* fmt, int32_t image_w,
                          int32_t image_h)
{
    RPGraph::GraphLayout layout(h->g);
    memcpy(layout.data(), xy, sizeof(float) * 2 * h->g.num_nodes());
    const std::string f = fmt;
    bool ok;
    if (f == "csv") ok = layout.writeToCSV(path);
    else if (f == "bin") ok = layout.writeToBin(path);
    else if (f == "png") ok = layout.writeToPNG(image_w, image_h, path);
    else return -2;
    return ok ? 0 : -1;
}

void gfa2host_free(gfa2host_graph_t h) { delete h; }
}
