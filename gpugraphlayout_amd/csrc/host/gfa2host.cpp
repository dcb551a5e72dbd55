// C entry points of the host layer (include/gfa2host.h).
#include "gfa2host.h"

#include <algorithm>
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

namespace
{
    struct SplitMix64
    {
        uint64_t x;
        uint64_t next()
        {
            uint64_t z = (x += 0x9E3779B97F4A7C15ull);
            z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
            z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
            return z ^ (z >> 31);
        }
        uint64_t below(uint64_t r) { return next() % r; }
    };
}

extern "C" int64_t gfa2host_barabasi_albert(int32_t n, int32_t m, uint64_t seed, int32_t* s, int32_t* t)
{
    if (m < 1 || n <= m) return 0;
    SplitMix64 rng{seed};
    std::vector<int32_t> pool;
    pool.reserve((size_t)2 * m * n);
    int64_t e = 0;
    for (int32_t u = 1; u <= m; ++u)
    {
        s[e] = 0;
        t[e] = u;
        ++e;
        pool.push_back(0);
        pool.push_back(u);
    }
    std::vector<int32_t> chosen;
    chosen.reserve(m);
    for (int32_t v = m + 1; v < n; ++v)
    {
        chosen.clear();
        while ((int32_t)chosen.size() < m)
        {
            const int32_t u = pool[rng.below(pool.size())];
            bool dup = false;
            for (int32_t c : chosen) dup |= c == u;
            if (!dup) chosen.push_back(u);
        }
        std::sort(chosen.begin(), chosen.end());
        for (int32_t u : chosen)
        {
            s[e] = v;
            t[e] = u;
            ++e;
            pool.push_back(v);
            pool.push_back(u);
        }
    }
    return e;
}
