// ORACLE — TEST INFRASTRUCTURE ONLY.
//
// A serial CPU restatement of the reference's GPU ForceAtlas2 iteration
// (govertb/GPUGraphLayout, CUDAForceAtlas2::doStep, src/RPGPUForceAtlas2.cu:206-248).
// Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may load this
// library, and only as the checker. The engine never links it.
//
// Each function names the reference lines it restates. The quadtree is built here by
// serial insertion (the semantics of TreeBuildingKernel run by one thread), which is an
// independent construction from the engine's sort-based one, so agreement between the
// two is a real check.
//
// Floating point: compiled with -ffp-contract=off, so every expression rounds exactly as
// written. Where the reference mixes double literals into float arithmetic, the double
// promotion is reproduced. nvcc's FMA contraction, rsqrtf and atomic ordering are not
// reproducible on a CPU; parity with them is checked under a stated tolerance.
//
// Pinning (what checks this file, besides the engine):
//   gravity / attraction / displacement / speed: closed forms and hand cases (tests)
//   bbox: numpy min/max
//   tree shape / counts / masses / permutation: an independent numpy key-prefix builder
//     replaying the reference's FP32 centre arithmetic, insertion-order independence,
//     hand-built point sets, root count = V and root mass = sum of masses
//   walk: theta -> 0 equals brute-force pairwise; group size 1 equals a per-body walk
//     written independently in Python; symmetric configurations
//   exact pairwise FP64: closed forms on two- and three-body cases

#include <cmath>
#include <cstdint>
#include <cstdlib>
#include <cstring>
#include <vector>
#include <algorithm>
#include <thread>

namespace {

// Reference constants: src/RPForceAtlas2.cpp:38-49. epssq = 0.05*0.05 evaluated in
// double and stored as float; itolsq = 1/theta^2 in float.
inline float ref_epssq() { return (float)(0.05 * 0.05); }

struct Tree {
    // Cells are numbered in creation order; cell 0 is the root. Child links: >= 0 body
    // index, <= -2 cell (-(cell)-2), -1 empty. This numbering is internal: everything
    // exported is in pre-order, which does not depend on insertion order.
    std::vector<int> child;       // 4 per cell
    std::vector<int> level;       // root = 0
    std::vector<int> count;
    std::vector<float> mass, cx, cy;
    int max_depth = 1;            // the reference's maxdepthd: deepest body, root children = 1
    float rootx = 0, rooty = 0, radius = 0;

    int new_cell(int lvl) {
        int id = (int)level.size();
        level.push_back(lvl);
        child.insert(child.end(), {-1, -1, -1, -1});
        count.push_back(0);
        mass.push_back(-1.0f);
        cx.push_back(0.0f);
        cy.push_back(0.0f);
        return id;
    }
};

inline int enc_cell(int c) { return -c - 2; }
inline bool is_cell(int link) { return link <= -2; }
inline int dec_cell(int link) { return -link - 2; }

// BoundingBoxKernel, src/RPBHKernels.cu:70-157 (min/max; radius and root centre at :140-148).
void bbox(int n, const float* pos, float* rootx, float* rooty, float* radius) {
    float minx = pos[0], maxx = pos[0], miny = pos[1], maxy = pos[1];
    for (int i = 0; i < n; ++i) {
        minx = fminf(minx, pos[2 * i]);
        maxx = fmaxf(maxx, pos[2 * i]);
        miny = fminf(miny, pos[2 * i + 1]);
        maxy = fmaxf(maxy, pos[2 * i + 1]);
    }
    *radius = fmaxf(maxx - minx, maxy - miny) * 0.5f;
    *rootx = (minx + maxx) * 0.5f;
    *rooty = (miny + maxy) * 0.5f;
}

// TreeBuildingKernel, src/RPBHKernels.cu:184-329, executed by one thread over bodies in
// index order. Quadrant at a cell centred (x,y): bit0 = x < px, bit1 = y < py (:217-224).
// The child centre is x +/- r with r halved per level by r *= 0.5f, accumulated along
// the path in float (:228-241, :292-308). A leaf slot holding another body is split
// until the two bodies part (:259-317). A body whose position equals the one it meets
// is scaled by 0.99 (a double multiply) and inserted again (:271-278); run serially,
// the body inserted later is the one moved. Returns 0, or -1 when max_levels is hit.
int build_tree(int n, float* pos, Tree& t, int max_levels) {
    t = Tree();
    bbox(n, pos, &t.rootx, &t.rooty, &t.radius);
    t.new_cell(0);
    for (int i = 0; i < n; ++i) {
    restart:
        float px = pos[2 * i], py = pos[2 * i + 1];
        int cell = 0;
        int depth = 1;
        float r = t.radius * 0.5f;
        float dx = -r, dy = -r;
        int j = 0;
        if (t.rootx < px) { j = 1; dx = r; }
        if (t.rooty < py) { j |= 2; dy = r; }
        float x = t.rootx + dx, y = t.rooty + dy;
        int ch = t.child[4 * cell + j];
        while (is_cell(ch)) {
            cell = dec_cell(ch);
            depth++;
            r *= 0.5f;
            dx = dy = -r;
            j = 0;
            if (x < px) { j = 1; dx = r; }
            if (y < py) { j |= 2; dy = r; }
            x += dx;
            y += dy;
            ch = t.child[4 * cell + j];
        }
        if (ch == -1) {
            t.child[4 * cell + j] = i;
            t.max_depth = std::max(t.max_depth, depth);
            continue;
        }
        // ch is a body occupying the slot.
        if (pos[2 * ch] == px && pos[2 * ch + 1] == py) {
            pos[2 * i] = (float)(pos[2 * i] * .99);
            pos[2 * i + 1] = (float)(pos[2 * i + 1] * .99);
            goto restart;
        }
        int slot_cell = cell, slot_j = j;
        for (;;) {
            if (depth >= max_levels) return -1;
            int nc = t.new_cell(t.level[slot_cell] + 1);
            t.child[4 * slot_cell + slot_j] = enc_cell(nc);
            depth++;
            r *= 0.5f;
            int jo = 0;
            if (x < pos[2 * ch]) jo = 1;
            if (y < pos[2 * ch + 1]) jo |= 2;
            t.child[4 * nc + jo] = ch;
            j = 0;
            dx = dy = -r;
            if (x < px) { j = 1; dx = r; }
            if (y < py) { j |= 2; dy = r; }
            x += dx;
            y += dy;
            if (t.child[4 * nc + j] >= 0) { slot_cell = nc; slot_j = j; continue; }
            t.child[4 * nc + j] = i;
            break;
        }
        t.max_depth = std::max(t.max_depth, depth);
    }
    return 0;
}

// SummarizationKernel, src/RPBHKernels.cu:357-506: count = sum of child counts (a body
// counts 1); mass accumulated from 0.0f over quadrants 0..3 in order; centroid
// (sum m*p) * (1/mass) (:388-421, :468-494). Post-order recursion gives the same
// per-cell operation sequence as the reference's readiness-driven passes.
void summarize(Tree& t, int c, const float* pos, const float* mass) {
    float cm = 0.0f, px = 0.0f, py = 0.0f;
    int cnt = 0;
    for (int q = 0; q < 4; ++q) {
        int ch = t.child[4 * c + q];
        if (ch == -1) continue;
        float m;
        if (is_cell(ch)) {
            int cc = dec_cell(ch);
            summarize(t, cc, pos, mass);
            m = t.mass[cc];
            cnt += t.count[cc];
            px += t.cx[cc] * m;
            py += t.cy[cc] * m;
        } else {
            m = mass[ch];
            cnt += 1;
            px += pos[2 * ch] * m;
            py += pos[2 * ch + 1] * m;
        }
        cm += m;
    }
    t.count[c] = cnt;
    float inv = 1.0f / cm;
    t.cx[c] = px * inv;
    t.cy[c] = py * inv;
    t.mass[c] = cm;
}

// SortKernel, src/RPBHKernels.cu:513-559: in-order (quadrant order) body permutation;
// also lists cells in pre-order, the canonical order used for comparisons.
void order(const Tree& t, int c, std::vector<int>& perm, std::vector<int>& pre) {
    pre.push_back(c);
    for (int q = 0; q < 4; ++q) {
        int ch = t.child[4 * c + q];
        if (ch == -1) continue;
        if (is_cell(ch)) order(t, dec_cell(ch), perm, pre);
        else perm.push_back(ch);
    }
}

// Per-level opening thresholds, src/RPBHKernels.cu:576-590: dq[0] = (2*radius)^2*itolsq,
// dq[d] = dq[d-1]*0.25f, then epssq added to each entry.
void thresholds(float radius, float itolsq, float epssq, int levels, std::vector<float>& dq) {
    dq.assign(levels, 0.0f);
    float tmp = radius * 2;
    dq[0] = tmp * tmp * itolsq;
    for (int i = 1; i < levels; ++i) dq[i] = dq[i - 1] * 0.25f;
    for (int i = 0; i < levels; ++i) dq[i] += epssq;
}

// ForceCalculationKernel, src/RPBHKernels.cu:566-693. Groups are G consecutive entries
// of the sorted order (a 32-lane warp in the reference). Children of a node are taken in
// quadrant order, empty slots skipped (the reference compacts them, :535-540). A body is
// always interacted with directly (itself included, contributing 0 via epssq). A cell
// at stack depth d (children of the root are at d = 0) is accepted only when every
// lane of the group has dx^2+dy^2+epssq >= dq[d] (:661); otherwise it is opened.
// Accumulation: a += k_r*dx*m_i*m_n/tmp, in that operation order (:654-665).
// starts: the first sorted position of each group to walk (nstarts of them), or null for all
// groups (0, G, 2G, ...); only the walked groups' bodies get their rep entries written.
void walk(const Tree& t, int n, const float* pos, const float* mass, const std::vector<int>& perm,
          float k_r, float itolsq, float epssq, int G, float* rep, long long* interactions,
          const int* starts = nullptr, int nstarts = 0) {
    std::vector<float> dq;
    thresholds(t.radius, itolsq, epssq, 128, dq);
    long long inter = 0;
    std::vector<float> ax(G), ay(G), px(G), py(G), mi(G);
    struct Frame { int cell, q; };
    std::vector<Frame> stack;
    const int ngroups = starts ? nstarts : (n + G - 1) / G;
    for (int gi = 0; gi < ngroups; ++gi) {
        const int g0 = starts ? starts[gi] : gi * G;
        if (g0 < 0 || g0 >= n) continue;
        int L = std::min(G, n - g0);
        for (int l = 0; l < L; ++l) {
            int b = perm[g0 + l];
            px[l] = pos[2 * b]; py[l] = pos[2 * b + 1]; mi[l] = mass[b];
            ax[l] = 0.0f; ay[l] = 0.0f;
        }
        stack.clear();
        stack.push_back({0, 0});
        while (!stack.empty()) {
            Frame& f = stack.back();
            if (f.q >= 4) { stack.pop_back(); continue; }
            int ch = t.child[4 * f.cell + f.q];
            f.q++;
            if (ch == -1) continue;
            int d = (int)stack.size() - 1;
            if (!is_cell(ch)) {
                float bx = pos[2 * ch], by = pos[2 * ch + 1], bm = mass[ch];
                for (int l = 0; l < L; ++l) {
                    float dx = px[l] - bx, dy = py[l] - by;
                    float tmp = dx * dx + dy * dy + epssq;
                    ax[l] += k_r * dx * mi[l] * bm / tmp;
                    ay[l] += k_r * dy * mi[l] * bm / tmp;
                }
                inter += L;
                continue;
            }
            int c = dec_cell(ch);
            bool all = true;
            for (int l = 0; l < L && all; ++l) {
                float dx = px[l] - t.cx[c], dy = py[l] - t.cy[c];
                float tmp = dx * dx + dy * dy + epssq;
                if (!(tmp >= dq[d])) all = false;
            }
            if (all) {
                for (int l = 0; l < L; ++l) {
                    float dx = px[l] - t.cx[c], dy = py[l] - t.cy[c];
                    float tmp = dx * dx + dy * dy + epssq;
                    ax[l] += k_r * dx * mi[l] * t.mass[c] / tmp;
                    ay[l] += k_r * dy * mi[l] * t.mass[c] / tmp;
                }
                inter += L;
            } else {
                stack.push_back({c, 0});
            }
        }
        for (int l = 0; l < L; ++l) {
            int b = perm[g0 + l];
            rep[2 * b] = ax[l];
            rep[2 * b + 1] = ay[l];
        }
    }
    if (interactions) *interactions = inter;
}

}  // namespace

extern "C" {

// GravityKernel, src/RPFA2Kernels.cu:35-73. Accumulates into f (interleaved x,y).
// Strong: f_g = k_g*m. Weak: f_g = k_g*m/|p| (rsqrtf in the reference; 1/sqrtf here), 0 at
// the origin. F -= p*f_g.
void orc_gravity(int n, const float* pos, const float* mass, int strong, float k_g, float* f) {
    for (int i = 0; i < n; ++i) {
        float px = pos[2 * i], py = pos[2 * i + 1];
        float fg;
        if (strong) fg = k_g * mass[i];
        else if (px != 0.0 || py != 0.0) fg = k_g * mass[i] * (1.0f / sqrtf(px * px + py * py));
        else fg = 0.0f;
        f[2 * i] += (-px * fg);
        f[2 * i + 1] += (-py * fg);
    }
}

// AttractiveForceKernel, src/RPFA2Kernels.cu:75-109, over the COO list the reference
// builds (each undirected edge once, source < target; src/RPGPUForceAtlas2.cu:73-84),
// applied in list order: F_s += p_t - p_s, F_t += -(p_t - p_s).
void orc_attraction_coo(long long m, const int* src, const int* dst, const float* pos, float* f) {
    for (long long e = 0; e < m; ++e) {
        int s = src[e], t = dst[e];
        float dx = pos[2 * t] - pos[2 * s];
        float dy = pos[2 * t + 1] - pos[2 * s + 1];
        f[2 * s] += dx;
        f[2 * s + 1] += dy;
        f[2 * t] += -dx;
        f[2 * t + 1] += -dy;
    }
}

// Attraction with the options the reference declares: linlog as its CPU backend computes it
// (src/RPCPUForceAtlas2.cpp:55-82: f_a/d = log(1+d)/d, and the zero vector times FLT_MAX,
// i.e. nothing, at d = 0) and edge weights w^delta (delta: src/RPForceAtlas2.hpp:49, read
// nowhere in the reference; Gephi's ForceAtlas2 multiplies the attraction by w^delta). w may
// be null (weight 1). Applied in COO order like orc_attraction_coo.
void orc_attraction_general(long long m, const int* src, const int* dst, const float* w, float delta, int linlog,
                            const float* pos, float* f) {
    for (long long e = 0; e < m; ++e) {
        int s = src[e], t = dst[e];
        float dx = pos[2 * t] - pos[2 * s];
        float dy = pos[2 * t + 1] - pos[2 * s + 1];
        float fa = 1.0f;
        if (linlog) {
            float d = std::sqrt(dx * dx + dy * dy);
            fa = d == 0.0f ? 0.0f : std::log(1.0f + d) / d;
        }
        if (w) fa *= std::pow(w[e], delta);
        f[2 * s] += dx * fa;
        f[2 * s + 1] += dy * fa;
        f[2 * t] += -(dx * fa);
        f[2 * t + 1] += -(dy * fa);
    }
}

// Bounding box: out[0]=root x, out[1]=root y, out[2]=radius.
void orc_bbox(int n, const float* pos, float* out) { bbox(n, pos, &out[0], &out[1], &out[2]); }

// Builds the reference quadtree for pos (which may be modified: coincident bodies are
// scaled as in the reference), summarises it, and returns, in pre-order, every cell's
// level, count, mass and centroid, plus the in-order body permutation.
// Returns the number of cells, -1 if max_levels is exceeded, -2 if cap is too small.
int orc_tree(int n, float* pos, const float* mass, int max_levels, int cap, int* perm,
             int* cell_level, int* cell_count, float* cell_mass, float* cell_pos, int* max_depth,
             float* root) {
    Tree t;
    if (build_tree(n, pos, t, max_levels) != 0) return -1;
    summarize(t, 0, pos, mass);
    std::vector<int> p, pre;
    order(t, 0, p, pre);
    if ((int)pre.size() > cap) return -2;
    for (int i = 0; i < n; ++i) perm[i] = p[i];
    for (size_t k = 0; k < pre.size(); ++k) {
        int c = pre[k];
        cell_level[k] = t.level[c];
        cell_count[k] = t.count[c];
        cell_mass[k] = t.mass[c];
        cell_pos[2 * k] = t.cx[c];
        cell_pos[2 * k + 1] = t.cy[c];
    }
    *max_depth = t.max_depth;
    root[0] = t.rootx; root[1] = t.rooty; root[2] = t.radius;
    return (int)pre.size();
}

// Barnes-Hut repulsion with the reference's group vote. rep is overwritten (2n floats).
// Returns the reference's maxdepthd, or -1 if max_levels was exceeded.
int orc_bh_repulsion(int n, float* pos, const float* mass, float k_r, float itolsq, float epssq,
                     int G, int max_levels, float* rep, int* perm, long long* interactions) {
    Tree t;
    if (build_tree(n, pos, t, max_levels) != 0) return -1;
    summarize(t, 0, pos, mass);
    std::vector<int> p, pre;
    order(t, 0, p, pre);
    if (perm) for (int i = 0; i < n; ++i) perm[i] = p[i];
    walk(t, n, pos, mass, p, k_r, itolsq, epssq, G, rep, interactions);
    return t.max_depth;
}

// The same repulsion for the listed groups only (each a first sorted position, a multiple of
// G): for checks at sizes where walking every group on one core takes too long. rep (2n
// floats) gets the walked groups' bodies only; interactions counts those groups'.
int orc_bh_repulsion_groups(int n, float* pos, const float* mass, float k_r, float itolsq, float epssq,
                            int G, int max_levels, int nstarts, const int* starts, float* rep, int* perm,
                            long long* interactions) {
    Tree t;
    if (build_tree(n, pos, t, max_levels) != 0) return -1;
    summarize(t, 0, pos, mass);
    std::vector<int> p, pre;
    order(t, 0, p, pre);
    if (perm) for (int i = 0; i < n; ++i) perm[i] = p[i];
    walk(t, n, pos, mass, p, k_r, itolsq, epssq, G, rep, interactions, starts, nstarts);
    return t.max_depth;
}

// SpeedKernel, src/RPFA2Kernels.cu:111-221: S = sum m*|F-F'|, T = sum m*|F+F'|/2, summed
// in double here (the reference's block-tree order is not reproducible). state =
// {global_speed, speed_efficiency, jitter_tolerance, k_s_max}; the update restates
// :186-218 including its double literals. out_st = {S, T}.
void orc_speed(int n, const float* f, const float* fprev, const float* mass, float* state,
               double* out_st) {
    double S = 0.0, T = 0.0;
    for (int i = 0; i < n; ++i) {
        float dx = f[2 * i] - fprev[2 * i], dy = f[2 * i + 1] - fprev[2 * i + 1];
        float swg = sqrtf(dx * dx + dy * dy);
        dx = f[2 * i] + fprev[2 * i];
        dy = f[2 * i + 1] + fprev[2 * i + 1];
        float tra = (float)(sqrtf(dx * dx + dy * dy) / 2.0);
        S += (double)(mass[i] * swg);
        T += (double)(mass[i] * tra);
    }
    out_st[0] = S;
    out_st[1] = T;
    float swg = (float)S, etra = (float)T;
    float& gs = state[0];
    float& se = state[1];
    float jitter = state[2], ksmax = state[3];
    float eojt = (float)(0.05 * sqrtf((float)n));
    float minjt = sqrtf(eojt);
    float jt = jitter * fmaxf(minjt, fminf(ksmax, eojt * etra / powf((float)n, 2.0f)));
    float min_se = (float)0.05;
    if (swg / etra > 2.0) {
        if (se > min_se) se = (float)(se * 0.5);
        jt = fmaxf(jt, jitter);
    }
    float target = jt * se * etra / swg;
    if (swg > jt * etra) {
        if (se > min_se) se = (float)(se * 0.7);
    } else if (gs < 1000) {
        se = (float)(se * 1.3);
    }
    float max_rise = (float)0.5;
    gs += fminf(target - gs, max_rise * gs);
}

// Speed update alone, from given S and T (same restatement as above).
void orc_speed_update(int n, double S, double T, float* state) {
    float swg = (float)S, etra = (float)T;
    float& gs = state[0];
    float& se = state[1];
    float jitter = state[2], ksmax = state[3];
    float eojt = (float)(0.05 * sqrtf((float)n));
    float minjt = sqrtf(eojt);
    float jt = jitter * fmaxf(minjt, fminf(ksmax, eojt * etra / powf((float)n, 2.0f)));
    float min_se = (float)0.05;
    if (swg / etra > 2.0) {
        if (se > min_se) se = (float)(se * 0.5);
        jt = fmaxf(jt, jitter);
    }
    float target = jt * se * etra / swg;
    if (swg > jt * etra) {
        if (se > min_se) se = (float)(se * 0.7);
    } else if (gs < 1000) {
        se = (float)(se * 1.3);
    }
    gs += fminf(target - gs, (float)0.5 * gs);
}

// DisplacementKernel, src/RPFA2Kernels.cu:223-251: factor = gs/(1.0+sqrtf(gs*swg))
// (evaluated in double because of the 1.0 literal), p += F*factor, F' = F.
void orc_displace(int n, float gs, const float* f, float* fprev, float* pos) {
    for (int i = 0; i < n; ++i) {
        float fx = f[2 * i], fy = f[2 * i + 1];
        float dx = fx - fprev[2 * i], dy = fy - fprev[2 * i + 1];
        float swg = sqrtf(dx * dx + dy * dy);
        float factor = (float)(gs / (1.0 + sqrtf(gs * swg)));
        pos[2 * i] += fx * factor;
        pos[2 * i + 1] += fy * factor;
        fprev[2 * i] = fx;
        fprev[2 * i + 1] = fy;
    }
}

// One full iteration in the reference's order (src/RPGPUForceAtlas2.cu:206-248):
// gravity, attraction, tree + repulsion, speed, displacement. Attraction runs over the
// COO list. out_f receives the total force of this iteration, out_rep the repulsion
// alone, out_st {S, T}. Returns maxdepth or -1.
int orc_fa2_step(int n, long long m, const int* src, const int* dst, const float* mass,
                 int strong, float k_g, float k_r, float itolsq, float epssq, int G, int max_levels,
                 float* pos, float* fprev, float* state, float* out_f, float* out_rep,
                 double* out_st, long long* interactions) {
    std::vector<float> f(2 * (size_t)n, 0.0f), rep(2 * (size_t)n, 0.0f);
    orc_gravity(n, pos, mass, strong, k_g, f.data());
    orc_attraction_coo(m, src, dst, pos, f.data());
    int md = orc_bh_repulsion(n, pos, mass, k_r, itolsq, epssq, G, max_levels, rep.data(), nullptr,
                              interactions);
    if (md < 0) return md;
    for (size_t i = 0; i < 2 * (size_t)n; ++i) f[i] += rep[i];
    if (out_f) memcpy(out_f, f.data(), sizeof(float) * 2 * n);
    if (out_rep) memcpy(out_rep, rep.data(), sizeof(float) * 2 * n);
    double st[2];
    orc_speed(n, f.data(), fprev, mass, state, st);
    if (out_st) { out_st[0] = st[0]; out_st[1] = st[1]; }
    orc_displace(n, state[0], f.data(), fprev, pos);
    return md;
}

// FP64 exact pairwise repulsion with the same softening: for each sampled body i,
// k_r * m_i * sum_j m_j (p_i - p_j) / (|p_i - p_j|^2 + epssq). Self terms are 0.
// The sample is split over up to `threads` host threads; each sum runs over j in ascending
// order in one thread, so the result does not depend on the thread count.
void orc_exact_repulsion_f64(int n, const float* pos, const float* mass, double k_r, double epssq,
                             int nsample, const int* sample, double* out, int threads) {
    auto run = [=](int s0, int s1) {
        for (int s = s0; s < s1; ++s) {
            int i = sample[s];
            double xi = pos[2 * i], yi = pos[2 * i + 1], mi = mass[i];
            double ax = 0.0, ay = 0.0;
            for (int j = 0; j < n; ++j) {
                double dx = xi - (double)pos[2 * j], dy = yi - (double)pos[2 * j + 1];
                double w = (double)mass[j] / (dx * dx + dy * dy + epssq);
                ax += dx * w;
                ay += dy * w;
            }
            out[2 * s] = k_r * mi * ax;
            out[2 * s + 1] = k_r * mi * ay;
        }
    };
    if (threads < 1) threads = 1;
    if (threads > nsample) threads = nsample > 0 ? nsample : 1;
    std::vector<std::thread> pool;
    for (int k = 1; k < threads; ++k)
        pool.emplace_back(run, (int)((long long)nsample * k / threads), (int)((long long)nsample * (k + 1) / threads));
    run(0, (int)((long long)nsample / threads));
    for (std::thread& th : pool) th.join();
}

}  // extern "C"
