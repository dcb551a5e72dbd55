// ORACLE — TEST INFRASTRUCTURE ONLY (the timed CPU baseline).
//
// A restatement of the reference's serial CPU backend in approximate (Barnes-Hut) and
// exact modes: CPUForceAtlas2 (src/RPCPUForceAtlas2.cpp:34-246) over the pointer
// quadtree of BarnesHutApproximator (src/RPBarnesHutApproximator.cpp:31-173), with the
// layout helpers it calls (src/RPGraphLayout.cpp:45-146) and the vector helpers of
// src/RPCommon.cpp:50-217. It keeps the reference's data-structure choices (hash-map
// degrees and adjacency, neighbour lists returned by copy, heap-allocated cells, a FIFO
// queue for the force walk) so its speed is representative of the reference's CPU path.
// It is only ever the CPU baseline beside the engine, never the engine's parity oracle:
// its Barnes-Hut differs from the GPU one (root box, quadrant numbering, unweighted
// centroids, acceptance test).
//
// Pinning: the quadtree insertion and force walk are compared bit-for-bit with the
// reference's own BarnesHutApproximator compiled from its sources (oracle/_ref, in the
// container that holds the reference); the FA2 step by hand cases on two- and
// three-node graphs.

#include <cmath>
#include <cstdint>
#include <cstdlib>
#include <limits>
#include <queue>
#include <unordered_map>
#include <vector>

namespace {

typedef uint32_t nid_t;

struct Vec { float x, y; };

inline Vec operator-(Vec a, Vec b) { return {a.x - b.x, a.y - b.y}; }

// Coordinate::quadrant, src/RPCommon.cpp:175-188 (of a vector relative to (0,0)).
inline int quadrant(Vec v) {
    if (v.x <= 0) return v.y >= 0 ? 0 : 3;
    return v.y >= 0 ? 1 : 2;
}

// BarnesHutCell, src/RPBarnesHutApproximator.cpp:31-60.
struct Cell {
    float lb, rb, ub, bb;
    Vec center, mass_center;
    nid_t num_sub = 0;
    float total_mass;
    float length;
    Cell* sub[4] = {nullptr, nullptr, nullptr, nullptr};

    Cell(Vec c, float len, Vec p, float m) : center(c), mass_center(p), total_mass(m), length(len) {
        lb = (float)(c.x - len / 2.0);
        rb = (float)(c.x + len / 2.0);
        bb = (float)(c.y - len / 2.0);
        ub = (float)(c.y + len / 2.0);
    }
    ~Cell() { for (int q = 0; q < 4; ++q) delete sub[q]; }

    // add_leafcell: child centre offsets mix a double (length/4.0) x-offset with a float
    // (length/4) y-offset, as the reference does (:44-52).
    void add_leaf(int q, float m, Vec p) {
        Vec c{0, 0};
        if (q == 0) c = {(float)(center.x - length / 4.0), center.y + length / 4};
        else if (q == 1) c = {(float)(center.x + length / 4.0), center.y + length / 4};
        else if (q == 2) c = {(float)(center.x + length / 4.0), center.y - length / 4};
        else if (q == 3) c = {(float)(center.x - length / 4.0), center.y - length / 4};
        sub[q] = new Cell(c, (float)(length / 2.0), p, m);
        num_sub += 1;
    }
};

struct BH {
    Cell* root = nullptr;
    Vec root_center{0, 0};
    float root_length = 0;

    void reset(Vec c, float len) {
        delete root;
        root = nullptr;
        root_center = c;
        root_length = len;
    }

    // insertParticle, src/RPBarnesHutApproximator.cpp:111-173.
    void insert(Vec p, float m) {
        if (!root) { root = new Cell(root_center, root_length, p, m); return; }
        Cell* cur = root;
        for (;;) {
            int qn = quadrant(p - cur->center);
            if (p.y > cur->ub || p.x > cur->rb || p.x < cur->lb || p.y < cur->bb) return;
            if (cur->num_sub == 0) {
                if (p.x == cur->mass_center.x && p.y == cur->mass_center.y) {
                    cur->total_mass += m;
                    return;
                }
                int qe = quadrant(cur->mass_center - cur->center);
                cur->add_leaf(qe, cur->total_mass, cur->mass_center);
            }
            cur->total_mass += m;
            float k = (float)cur->num_sub;
            cur->mass_center = {cur->mass_center.x * k, cur->mass_center.y * k};
            cur->mass_center.x += p.x;
            cur->mass_center.y += p.y;
            float d = (float)(cur->num_sub + 1);
            cur->mass_center.x /= d;
            cur->mass_center.y /= d;
            if (cur->sub[qn] == nullptr) { cur->add_leaf(qn, m, p); return; }
            cur->num_sub += 1;
            cur = cur->sub[qn];
        }
    }

    // approximateForce, src/RPBarnesHutApproximator.cpp:62-109: breadth-first walk;
    // accept when length^2/D^2 < theta^2 or the cell holds one particle.
    Vec force(Vec p, float pm, float theta) {
        Vec f{0.0f, 0.0f};
        std::queue<Cell*> q;
        q.push(root);
        while (!q.empty()) {
            Cell* c = q.front();
            q.pop();
            float dx = p.x - c->mass_center.x, dy = p.y - c->mass_center.y;
            float D2 = dx * dx + dy * dy;
            if (D2 == 0) {
                if (c->num_sub == 0) continue;
                return {(float)rand(), (float)rand()};
            }
            if (c->length * c->length / D2 < theta * theta || c->num_sub == 0) {
                float s = pm * c->total_mass / D2;
                f.x += dx * s;
                f.y += dy * s;
            } else {
                for (int i = 0; i < 4; ++i)
                    if (c->sub[i] != nullptr) q.push(c->sub[i]);
            }
        }
        return f;
    }
};

// UGraph storage (src/RPGraph.hpp:43-81): hash-map degrees and min-id-keyed adjacency.
struct Graph {
    nid_t n = 0, m = 0;
    std::unordered_map<nid_t, nid_t> degrees;
    std::unordered_map<nid_t, std::vector<nid_t>> adj;
    nid_t degree(nid_t v) { return degrees[v]; }
    std::vector<nid_t> neighbors_with_geq_id(nid_t v) { return adj[v]; }  // a copy, as :128-131
};

struct CPUFA2 {
    Graph g;
    std::vector<Vec> pos, forces, prev;
    bool use_bh = true, strong = false;
    float k_g = 1, k_r = 1, theta = 1.0f;
    float global_speed = 1.0f, speed_efficiency = 1.0f, jitter_tolerance = 1.0f;
    float k_s = 0.1f, k_s_max = 10.0f;
    float last_swing = 0.0f, last_traction = 0.0f;   // S and T of the last updateSpeeds
    BH bh;

    float mass(nid_t v) { return (float)(g.degree(v) + 1.0); }  // src/RPForceAtlas2.cpp:74-77

    // GraphLayout extent helpers, src/RPGraphLayout.cpp:64-109 and :141-146. maxX/maxY
    // start from numeric_limits<float>::min(), as the reference does.
    float minX() { float v = std::numeric_limits<float>::max(); for (auto& p : pos) if (p.x < v) v = p.x; return v; }
    float maxX() { float v = std::numeric_limits<float>::min(); for (auto& p : pos) if (p.x > v) v = p.x; return v; }
    float minY() { float v = std::numeric_limits<float>::max(); for (auto& p : pos) if (p.y < v) v = p.y; return v; }
    float maxY() { float v = std::numeric_limits<float>::min(); for (auto& p : pos) if (p.y > v) v = p.y; return v; }
    float xrange() { return maxX() - minX(); }
    float yrange() { return maxY() - minY(); }
    float span() { return ceilf(fmaxf(xrange(), yrange())); }
    Vec center() { return {(float)(minX() + xrange() / 2.0), (float)(minY() + yrange() / 2.0)}; }

    // rebuild_bh, src/RPCPUForceAtlas2.cpp:215-224.
    void rebuild() {
        bh.reset(center(), span() + 10);
        for (nid_t v = 0; v < g.n; ++v) bh.insert(pos[v], (float)(g.degree(v) + 1));
    }

    // apply_gravity, :104-124.
    void gravity(nid_t v) {
        float x = pos[v].x, y = pos[v].y;
        float d = std::sqrt(x * x + y * y);
        if (d == 0.0) return;
        float fg = strong ? k_g * mass(v) : k_g * mass(v) / d;
        forces[v].x += -x * fg;
        forces[v].y += -y * fg;
    }

    // apply_attract, :55-83 (linear attraction, linlog off).
    void attract(nid_t v) {
        Vec f{0.0f, 0.0f};
        for (nid_t t : g.neighbors_with_geq_id(v)) {
            float dx = pos[t].x - pos[v].x, dy = pos[t].y - pos[v].y;
            f.x += dx * 1.0f;
            f.y += dy * 1.0f;
            forces[t].x += dx * -1.0f;
            forces[t].y += dy * -1.0f;
        }
        forces[v].x += f.x;
        forces[v].y += f.y;
    }

    // apply_repulsion, :85-102. The exact branch keeps the reference's direction
    // (getDistanceVector(n,t) = p_t - p_n), which makes it attractive.
    void repulsion(nid_t v) {
        if (use_bh) {
            Vec f = bh.force(pos[v], mass(v), theta);
            forces[v].x += f.x * k_r;
            forces[v].y += f.y * k_r;
        } else {
            for (nid_t t = 0; t < g.n; ++t) {
                if (t == v) continue;
                float dx = pos[v].x - pos[t].x, dy = pos[v].y - pos[t].y;
                float dist = std::sqrt(dx * dx + dy * dy);
                float fr = dist == 0.0 ? std::numeric_limits<float>::max()
                                       : k_r * mass(v) * mass(t) / dist / dist;
                forces[v].x += (pos[t].x - pos[v].x) * fr;
                forces[v].y += (pos[t].y - pos[v].y) * fr;
            }
        }
    }

    float swg(nid_t v) {
        float dx = forces[v].x - prev[v].x, dy = forces[v].y - prev[v].y;
        return std::sqrt(dx * dx + dy * dy);
    }
    float tra(nid_t v) {
        float dx = forces[v].x + prev[v].x, dy = forces[v].y + prev[v].y;
        return (float)(std::sqrt(dx * dx + dy * dy) / 2.0);
    }

    // updateSpeeds, :144-197.
    void update_speeds() {
        float S = 0.0f, T = 0.0f;
        for (nid_t v = 0; v < g.n; ++v) {
            S += mass(v) * swg(v);
            T += mass(v) * tra(v);
        }
        last_swing = S;
        last_traction = T;
        float eojt = (float)(0.05 * std::sqrt((double)g.n));
        float minjt = std::sqrt(eojt);
        float jt = jitter_tolerance * fmaxf(minjt, fminf(k_s_max, eojt * T / powf((float)g.n, 2.0f)));
        float min_se = (float)0.05;
        if (S / T > 2.0) {
            if (speed_efficiency > min_se) speed_efficiency = (float)(speed_efficiency * 0.5);
            jt = fmaxf(jt, jitter_tolerance);
        }
        float target = jt * speed_efficiency * T / S;
        if (S > jt * T) {
            if (speed_efficiency > min_se) speed_efficiency = (float)(speed_efficiency * 0.7);
        } else if (global_speed < 1000) {
            speed_efficiency = (float)(speed_efficiency * 1.3);
        }
        global_speed += fminf(target - global_speed, (float)0.5 * global_speed);
    }

    // apply_displacement, :199-213.
    void displace(nid_t v) {
        float factor = (float)(global_speed / (1.0 + std::sqrt(global_speed * swg(v))));
        pos[v].x = pos[v].x + forces[v].x * factor;
        pos[v].y = pos[v].y + forces[v].y * factor;
    }

    // doStep, :226-246.
    void step() {
        if (use_bh) rebuild();
        for (nid_t v = 0; v < g.n; ++v) {
            gravity(v);
            attract(v);
            repulsion(v);
        }
        update_speeds();
        for (nid_t v = 0; v < g.n; ++v) {
            displace(v);
            prev[v] = forces[v];
            forces[v] = {0.0f, 0.0f};
        }
    }
};

}  // namespace

extern "C" {

// Creates a CPU baseline instance. Edges are given as the reference's COO (mapped ids,
// src < dst, in UGraph insertion order) and inserted into the hash-map adjacency in
// that order. pos: interleaved x,y.
void* orc_cpu_create(int n, long long m, const int* src, const int* dst, const float* pos,
                     int strong, float k_g, float k_r, int use_bh) {
    CPUFA2* h = new CPUFA2();
    h->g.n = (nid_t)n;
    h->g.m = (nid_t)m;
    for (long long e = 0; e < m; ++e) {
        nid_t s = (nid_t)src[e], t = (nid_t)dst[e];
        nid_t a = s < t ? s : t, b = s < t ? t : s;
        h->g.adj[a].push_back(b);
        h->g.degrees[s] += 1;
        h->g.degrees[t] += 1;
    }
    h->pos.resize(n);
    h->forces.assign(n, Vec{0.0f, 0.0f});
    h->prev.assign(n, Vec{0.0f, 0.0f});
    for (int i = 0; i < n; ++i) h->pos[i] = {pos[2 * i], pos[2 * i + 1]};
    h->strong = strong != 0;
    h->k_g = k_g;
    h->k_r = k_r;
    h->use_bh = use_bh != 0;
    return h;
}

void orc_cpu_step(void* h, int iters) {
    CPUFA2* c = (CPUFA2*)h;
    for (int i = 0; i < iters; ++i) c->step();
}

void orc_cpu_get_positions(void* h, float* pos) {
    CPUFA2* c = (CPUFA2*)h;
    for (nid_t i = 0; i < c->g.n; ++i) { pos[2 * i] = c->pos[i].x; pos[2 * i + 1] = c->pos[i].y; }
}

// state = {global_speed, speed_efficiency, jitter_tolerance, k_s_max, swing, traction}
// (swing and traction: S and T of the last iteration)
void orc_cpu_get_state(void* h, float* state) {
    CPUFA2* c = (CPUFA2*)h;
    state[0] = c->global_speed; state[1] = c->speed_efficiency;
    state[2] = c->jitter_tolerance; state[3] = c->k_s_max;
    state[4] = c->last_swing; state[5] = c->last_traction;
}

void orc_cpu_destroy(void* h) { delete (CPUFA2*)h; }

// Stand-alone CPU Barnes-Hut (the pieces above), for pinning against the reference's own
// BarnesHutApproximator: insert n particles into a root (cx, cy, len) in order, then
// write the force on each of the nq query particles into out.
void orc_cpu_bh_forces(int n, const float* pos, const float* mass, float cx, float cy, float len,
                       float theta, int nq, const float* qpos, const float* qmass, float* out) {
    BH bh;
    bh.reset(Vec{cx, cy}, len);
    for (int i = 0; i < n; ++i) bh.insert(Vec{pos[2 * i], pos[2 * i + 1]}, mass[i]);
    for (int k = 0; k < nq; ++k) {
        Vec f = bh.force(Vec{qpos[2 * k], qpos[2 * k + 1]}, qmass[k], theta);
        out[2 * k] = f.x;
        out[2 * k + 1] = f.y;
    }
    bh.reset(Vec{0, 0}, 0);
}

}  // extern "C"
