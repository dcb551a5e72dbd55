// ORACLE — TEST INFRASTRUCTURE ONLY.
//
// Layout-quality reference computations with no counterpart in the reference project:
// breadth-first hop distances and normalized stress over a list of node pairs.
// NS = (1/P) * sum (alpha*D_ij/d_ij - 1)^2 with alpha = sum(D/d) / sum(D^2/d^2), the
// scale that minimises the sum, D the layout distance and d the hop distance. All in
// double. Pinned by brute force on small graphs (tests) and by scale invariance.
#include <cmath>
#include <vector>

extern "C" {

// Hop distances from src over a symmetric CSR; unreachable nodes get -1.
void orc_bfs(int n, const int* rowptr, const int* col, int src, int* dist) {
    for (int i = 0; i < n; ++i) dist[i] = -1;
    std::vector<int> q;
    q.reserve(n);
    dist[src] = 0;
    q.push_back(src);
    for (size_t h = 0; h < q.size(); ++h) {
        int v = q[h];
        for (int e = rowptr[v]; e < rowptr[v + 1]; ++e) {
            int u = col[e];
            if (dist[u] < 0) { dist[u] = dist[v] + 1; q.push_back(u); }
        }
    }
}

// Normalized stress of positions pos (interleaved x,y) over P pairs (a[k], b[k]) with hop
// distances d[k] > 0. out = {NS, alpha}.
void orc_normalized_stress(long long P, const int* a, const int* b, const int* d, const float* pos,
                           double* out) {
    double s1 = 0.0, s2 = 0.0;
    for (long long k = 0; k < P; ++k) {
        double dx = (double)pos[2 * a[k]] - pos[2 * b[k]], dy = (double)pos[2 * a[k] + 1] - pos[2 * b[k] + 1];
        double r = std::sqrt(dx * dx + dy * dy) / d[k];
        s1 += r;
        s2 += r * r;
    }
    double alpha = s2 > 0 ? s1 / s2 : 0.0;
    double acc = 0.0;
    for (long long k = 0; k < P; ++k) {
        double dx = (double)pos[2 * a[k]] - pos[2 * b[k]], dy = (double)pos[2 * a[k] + 1] - pos[2 * b[k] + 1];
        double e = alpha * std::sqrt(dx * dx + dy * dy) / d[k] - 1.0;
        acc += e * e;
    }
    out[0] = P > 0 ? acc / (double)P : 0.0;
    out[1] = alpha;
}

}
