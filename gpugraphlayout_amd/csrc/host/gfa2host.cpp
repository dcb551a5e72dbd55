// C entry points of the host layer (include/gfa2host.h).
#include "gfa2host.h"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <string>
#include <thread>
#include <vector>

#include "RPGraph.hpp"
#include "RPGraphLayout.hpp"
#include "hashing.hpp"

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

void gfa2host_png_frame(float minx, float maxx, float miny, float maxy, int32_t image_w, int32_t image_h,
                        int64_t nodes, int64_t edges, gfa2_raster* out)
{
    if (out)
        *out = RPGraph::png_frame(minx, maxx, miny, maxy, image_w, image_h, (RPGraph::nid_t)nodes,
                                  (RPGraph::nid_t)edges);
}

int gfa2host_render_gray(gfa2host_graph_t h, const float* xy, const gfa2_raster* f, uint8_t* gray)
{
    if (!h || !xy || !f || !gray || f->width <= 0 || f->height <= 0) return -1;
    std::vector<unsigned char> g;
    RPGraph::render_gray(*f, h->g, xy, g);
    memcpy(gray, g.data(), g.size());
    return 0;
}

int gfa2host_write_png_gray(const char* path, int32_t image_w, int32_t image_h, const uint8_t* gray)
{
    if (!path || !gray || image_w <= 0 || image_h <= 0) return -1;
    return RPGraph::write_png_gray(path, image_w, image_h, gray) ? 0 : -1;
}
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

namespace
{
    // SplitMix64 output number k of a stream seeded with `seed` (counter form).
    inline uint64_t splitmix_at(uint64_t seed, uint64_t k)
    {
        uint64_t z = seed + (k + 1) * 0x9E3779B97F4A7C15ull;
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
        return z ^ (z >> 31);
    }
}

extern "C" int64_t gfa2host_rmat(int32_t scale, int64_t n_edges, double a, double b, double c, uint64_t seed,
                                 int32_t* s, int32_t* t)
{
    if (scale < 1 || scale > 31 || n_edges < 0 || !(a >= 0 && b >= 0 && c >= 0 && a + b + c <= 1.0)) return -1;
    const double pairs = std::ldexp(1.0, scale) * (std::ldexp(1.0, scale) - 1) / 2;
    if ((double)n_edges > 0.5 * pairs) return -1;
    const double ab = a + b, abc = a + b + c;
    gfa2host::KeySet seen((uint64_t)n_edges);
    const uint64_t max_draws = 64 * (uint64_t)n_edges + (1u << 20);   // ends skewed parameter sets
    // Draws are computed in parallel a chunk at a time; the first-occurrence filter runs in
    // draw order on one thread, so the result does not depend on the thread count.
    constexpr uint64_t kChunk = 1u << 20;
    std::vector<uint32_t> cu(kChunk), cv(kChunk);
    unsigned nthreads = std::thread::hardware_concurrency();
    nthreads = nthreads < 1 ? 1 : (nthreads > 16 ? 16 : nthreads);
    int64_t got = 0;
    for (uint64_t k0 = 0; got < n_edges && k0 < max_draws; k0 += kChunk)
    {
        auto work = [&](unsigned tid) {
            const uint64_t lo = kChunk * tid / nthreads, hi = kChunk * (tid + 1) / nthreads;
            for (uint64_t i = lo; i < hi; ++i)
            {
                const uint64_t k = k0 + i;
                uint64_t u = 0, v = 0;
                for (int l = 0; l < scale; ++l)
                {
                    const double x =
                        (double)(splitmix_at(seed, k * (uint64_t)scale + (uint64_t)l) >> 11) * 0x1.0p-53;
                    u = (u << 1) | (uint64_t)(x >= ab);
                    v = (v << 1) | (uint64_t)(((x >= a) && (x < ab)) || (x >= abc));
                }
                cu[i] = (uint32_t)u;
                cv[i] = (uint32_t)v;
            }
        };
        std::vector<std::thread> pool;
        for (unsigned q = 1; q < nthreads; ++q) pool.emplace_back(work, q);
        work(0);
        for (auto& th : pool) th.join();
        const uint64_t m = (max_draws - k0) < kChunk ? (max_draws - k0) : kChunk;
        for (uint64_t i = 0; i < m && got < n_edges; ++i)
        {
            const uint64_t u = cu[i], v = cv[i];
            if (u == v) continue;
            const uint64_t lo = u < v ? u : v, hi = u < v ? v : u;
            if (!seen.insert(((lo << scale) | hi) + 1)) continue;
            s[got] = (int32_t)u;
            t[got] = (int32_t)v;
            ++got;
        }
    }
    return got;
}
