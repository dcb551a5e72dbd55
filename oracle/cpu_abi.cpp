// ORACLE — TEST INFRASTRUCTURE ONLY. A CPU implementation of the part of the engine's C ABI
// (include/gfa2.h) that the C++ host layer calls, built on the oracle's reference-semantics
// restatement (fa2_oracle.cpp). It lets tests drive the host classes and graph_viewer on a
// machine without a GPU and compare their outputs with the oracle's own trajectory. It is
// linked only into the test build of graph_viewer under oracle/_build/; the product links
// libgfa2.so, whose entry points fail without a GPU.
//
// Iteration (src/RPGPUForceAtlas2.cu:206-248 order): gravity, attraction over the COO list
// (linear, or linlog as src/RPCPUForceAtlas2.cpp:63-67), Barnes-Hut repulsion with the group
// vote, speed update, displacement. Exact repulsion and rendering are not provided
// (GFA2_ERR_UNSUPPORTED), prevent_overlap is refused as by the engine.
#include "../include/gfa2.h"

#include <cmath>
#include <cstring>
#include <new>
#include <vector>

extern "C" {
void orc_gravity(int n, const float* pos, const float* mass, int strong, float k_g, float* f);
void orc_attraction_coo(long long m, const int* src, const int* dst, const float* pos, float* f);
void orc_attraction_general(long long m, const int* src, const int* dst, const float* w, float delta, int linlog,
                            const float* pos, float* f);
int orc_bh_repulsion(int n, float* pos, const float* mass, float k_r, float itolsq, float epssq, int G,
                     int max_levels, float* rep, int* perm, long long* interactions);
void orc_speed(int n, const float* f, const float* fprev, const float* mass, float* state, double* out_st);
void orc_displace(int n, float gs, const float* f, float* fprev, float* pos);
}

struct gfa2_engine {
    gfa2_params prm{};
    int n = 0;
    std::vector<int> src, dst;
    std::vector<float> mass, pos, fprev, f, fga, rep;
    float state[4] = {1.0f, 1.0f, 1.0f, 10.0f};   // speed, efficiency, jitter tolerance, k_s_max
    bool has_graph = false, has_pos = false;
    int iteration = 0;
    double swing = 0.0, traction = 0.0;
    long long interactions = 0, interactions_total = 0;
    int max_depth = 0;
};

namespace {

int validate(const gfa2_params& p) {
    if (p.prevent_overlap) return GFA2_ERR_UNSUPPORTED;
    if (!p.use_barneshut) return GFA2_ERR_UNSUPPORTED;
    if (p.group_size != 32 && p.group_size != 64) return GFA2_ERR_INVALID_ARG;
    if (!(p.theta > 0.0f) || !(p.epssq >= 0.0f)) return GFA2_ERR_INVALID_ARG;
    return GFA2_OK;
}

void reset_state(gfa2_engine* h) {
    h->state[0] = h->prm.global_speed;
    h->state[1] = h->prm.speed_efficiency;
    h->state[2] = h->prm.jitter_tolerance;
    h->state[3] = h->prm.k_s_max;
    std::fill(h->fprev.begin(), h->fprev.end(), 0.0f);
    h->iteration = 0;
    h->interactions_total = 0;
}

int one_step(gfa2_engine* h) {
    const int n = h->n;
    const long long m = (long long)h->src.size();
    std::vector<float> f(2 * (size_t)n, 0.0f), rep(2 * (size_t)n, 0.0f);
    orc_gravity(n, h->pos.data(), h->mass.data(), h->prm.strong_gravity, h->prm.gravity, f.data());
    if (h->prm.linlog)
        orc_attraction_general(m, h->src.data(), h->dst.data(), nullptr, 0.0f, 1, h->pos.data(), f.data());
    else
        orc_attraction_coo(m, h->src.data(), h->dst.data(), h->pos.data(), f.data());
    h->fga = f;
    long long inter = 0;
    // as orc_fa2_step: the tree build may nudge coincident bodies in place, as the reference does
    const int md = orc_bh_repulsion(n, h->pos.data(), h->mass.data(), h->prm.scale,
                                    1.0f / (h->prm.theta * h->prm.theta), h->prm.epssq, h->prm.group_size, 128,
                                    rep.data(), nullptr, &inter);
    if (md < 0) return GFA2_ERR_TREE_OVERFLOW;
    for (size_t i = 0; i < f.size(); ++i) f[i] += rep[i];
    double st[2];
    orc_speed(n, f.data(), h->fprev.data(), h->mass.data(), h->state, st);
    orc_displace(n, h->state[0], f.data(), h->fprev.data(), h->pos.data());
    h->f = f;
    h->rep = rep;
    h->swing = st[0];
    h->traction = st[1];
    h->interactions = inter;
    h->interactions_total += inter;
    h->max_depth = md;
    ++h->iteration;
    return GFA2_OK;
}

}  // namespace

extern "C" {

void gfa2_default_params(gfa2_params* p) {
    if (!p) return;
    p->use_barneshut = 1;
    p->strong_gravity = 0;
    p->gravity = 1.0f;
    p->scale = 1.0f;
    p->theta = 1.0f;
    p->epssq = (float)(0.05 * 0.05);
    p->jitter_tolerance = 1.0f;
    p->k_s_max = 10.0f;
    p->global_speed = 1.0f;
    p->speed_efficiency = 1.0f;
    p->group_size = 32;
    p->cell_factor = 0;
    p->linlog = 0;
    p->edge_weight_influence = 0.0f;
    p->prevent_overlap = 0;
}

int gfa2_create(const gfa2_params* params, int device, gfa2_handle* out) {
    if (!out) return GFA2_ERR_INVALID_ARG;
    *out = nullptr;
    gfa2_params prm;
    if (params) prm = *params;
    else gfa2_default_params(&prm);
    const int v = validate(prm);
    if (v != GFA2_OK) return v;
    if (device != 0) return GFA2_ERR_INVALID_ARG;
    gfa2_engine* h = new (std::nothrow) gfa2_engine();
    if (!h) return GFA2_ERR_OOM;
    h->prm = prm;
    reset_state(h);
    *out = h;
    return GFA2_OK;
}

int gfa2_destroy(gfa2_handle h) {
    if (!h) return GFA2_ERR_INVALID_ARG;
    delete h;
    return GFA2_OK;
}

int gfa2_set_params(gfa2_handle h, const gfa2_params* prm) {
    if (!h || !prm) return GFA2_ERR_INVALID_ARG;
    const int v = validate(*prm);
    if (v != GFA2_OK) return v;
    h->prm = *prm;
    return GFA2_OK;
}

int gfa2_set_graph_coo(gfa2_handle h, int32_t n, int64_t m, const int32_t* src, const int32_t* dst,
                       const float* mass) {
    if (!h || n < 1 || m < 0 || (m > 0 && (!src || !dst))) return n < 1 ? GFA2_ERR_INVALID_GRAPH : GFA2_ERR_INVALID_ARG;
    std::vector<int> deg((size_t)n, 0);
    for (int64_t e = 0; e < m; ++e) {
        if (src[e] < 0 || src[e] >= n || dst[e] < 0 || dst[e] >= n || src[e] == dst[e]) return GFA2_ERR_INVALID_GRAPH;
        ++deg[src[e]];
        ++deg[dst[e]];
    }
    h->n = n;
    h->src.assign(src, src + m);
    h->dst.assign(dst, dst + m);
    h->mass.resize(n);
    for (int v = 0; v < n; ++v) h->mass[v] = mass ? mass[v] : (float)(deg[v] + 1.0);
    h->pos.assign(2 * (size_t)n, 0.0f);
    h->fprev.assign(2 * (size_t)n, 0.0f);
    h->f.assign(2 * (size_t)n, 0.0f);
    h->fga.assign(2 * (size_t)n, 0.0f);
    h->rep.assign(2 * (size_t)n, 0.0f);
    h->has_graph = true;
    h->has_pos = false;
    reset_state(h);
    return GFA2_OK;
}

int gfa2_set_positions(gfa2_handle h, const float* xy) {
    if (!h || !xy) return GFA2_ERR_INVALID_ARG;
    if (!h->has_graph) return GFA2_ERR_NO_GRAPH;
    std::memcpy(h->pos.data(), xy, sizeof(float) * 2 * (size_t)h->n);
    h->has_pos = true;
    return GFA2_OK;
}

int gfa2_get_positions(gfa2_handle h, float* xy) {
    if (!h || !xy) return GFA2_ERR_INVALID_ARG;
    if (!h->has_graph) return GFA2_ERR_NO_GRAPH;
    std::memcpy(xy, h->pos.data(), sizeof(float) * 2 * (size_t)h->n);
    return GFA2_OK;
}

int gfa2_reset(gfa2_handle h) {
    if (!h) return GFA2_ERR_INVALID_ARG;
    if (!h->has_graph) return GFA2_ERR_NO_GRAPH;
    reset_state(h);
    return GFA2_OK;
}

int gfa2_step(gfa2_handle h, int32_t iters) {
    if (!h || iters < 0) return GFA2_ERR_INVALID_ARG;
    if (!h->has_graph) return GFA2_ERR_NO_GRAPH;
    if (!h->has_pos) return GFA2_ERR_INVALID_ARG;
    for (int i = 0; i < iters; ++i) {
        const int rc = one_step(h);
        if (rc != GFA2_OK) return rc;
    }
    return GFA2_OK;
}

int gfa2_sync(gfa2_handle h) { return h ? GFA2_OK : GFA2_ERR_INVALID_ARG; }

int gfa2_set_graph_capture(gfa2_handle h, int32_t on) {
    (void)on;
    return h ? GFA2_OK : GFA2_ERR_INVALID_ARG;
}

int gfa2_get_forces(gfa2_handle h, int32_t kind, float* out) {
    if (!h || !out) return GFA2_ERR_INVALID_ARG;
    if (!h->has_graph) return GFA2_ERR_NO_GRAPH;
    const std::vector<float>* src = kind == GFA2_FORCE_TOTAL ? &h->f
                                    : kind == GFA2_FORCE_REPULSION ? &h->rep
                                    : kind == GFA2_FORCE_ATTRACTION_GRAVITY ? &h->fga
                                    : nullptr;
    if (!src) return GFA2_ERR_INVALID_ARG;
    std::memcpy(out, src->data(), sizeof(float) * 2 * (size_t)h->n);
    return GFA2_OK;
}

int gfa2_get_stats(gfa2_handle h, gfa2_stats* s) {
    if (!h || !s) return GFA2_ERR_INVALID_ARG;
    std::memset(s, 0, sizeof(*s));
    s->iteration = h->iteration;
    s->nodes = h->n;
    s->edges = (int64_t)h->src.size();
    s->global_speed = h->state[0];
    s->speed_efficiency = h->state[1];
    s->swing = h->swing;
    s->traction = h->traction;
    s->interactions = h->interactions;
    s->interactions_total = h->interactions_total;
    s->max_level = h->max_depth > 0 ? h->max_depth - 1 : 0;
    return GFA2_OK;
}

int gfa2_get_bbox(gfa2_handle h, float* out) {
    if (!h || !out) return GFA2_ERR_INVALID_ARG;
    if (!h->has_graph || !h->has_pos) return GFA2_ERR_NO_GRAPH;
    float x0 = h->pos[0], x1 = h->pos[0], y0 = h->pos[1], y1 = h->pos[1];
    for (int i = 1; i < h->n; ++i) {
        x0 = std::fmin(x0, h->pos[2 * i]);
        x1 = std::fmax(x1, h->pos[2 * i]);
        y0 = std::fmin(y0, h->pos[2 * i + 1]);
        y1 = std::fmax(y1, h->pos[2 * i + 1]);
    }
    out[0] = x0;
    out[1] = x1;
    out[2] = y0;
    out[3] = y1;
    return GFA2_OK;
}

int gfa2_render(gfa2_handle h, const gfa2_raster* r, uint8_t* gray) {
    (void)r;
    (void)gray;
    return h ? GFA2_ERR_UNSUPPORTED : GFA2_ERR_INVALID_ARG;
}

const char* gfa2_strerror(int code) {
    switch (code) {
        case GFA2_OK: return "ok";
        case GFA2_ERR_NO_DEVICE: return "no HIP device found";
        case GFA2_ERR_OOM: return "out of device memory";
        case GFA2_ERR_INVALID_ARG: return "invalid argument";
        case GFA2_ERR_INVALID_GRAPH: return "invalid graph (empty, malformed CSR, self-loop or id out of range)";
        case GFA2_ERR_TREE_OVERFLOW: return "quadtree cell capacity exceeded";
        case GFA2_ERR_HIP: return "HIP runtime error";
        case GFA2_ERR_NO_GRAPH: return "no graph loaded";
        case GFA2_ERR_UNSUPPORTED: return "unsupported configuration";
        default: return "unknown error";
    }
}

}  // extern "C"
