// Layout quality: hop distances of sampled node pairs and normalized stress (no counterpart
// in the reference).
//
// Pairs: K sources, T targets each (targets drawn uniformly among the other nodes).
// Hop distances come from a multi-source breadth-first search run 64 sources at a time:
// every node keeps a 64-bit mask of the sources that have reached it and the mask of those
// that reached it on the last level; a level is one pull over the CSR (each node ORs its
// neighbours' last-level masks), so no atomics and a deterministic result. After each
// level the pairs of the current 64 sources whose target bit just appeared get that level
// as their distance. Pairs whose target is never reached keep distance 0 and are left out.
//
// Normalized stress of positions p over the P reachable pairs, with r = |p_a - p_b| / d:
//   alpha = sum r / sum r^2,  NS = (1/P) sum (alpha r - 1)^2 = 1 - (sum r)^2 / (P sum r^2)
// summed in double in one pass (per-block partials, fixed-order final sum).
#include "engine_internal.h"

#include <vector>

namespace gfa2 {
namespace {

__global__ __launch_bounds__(256) void k_msbfs_init(unsigned long long* __restrict__ seen,
                                                    unsigned long long* __restrict__ front, int n,
                                                    const int* __restrict__ src, int k0, int nk) {
    for (int v = blockIdx.x * 256 + threadIdx.x; v < n; v += gridDim.x * 256) {
        seen[v] = 0ull;
        front[v] = 0ull;
    }
}

__global__ void k_msbfs_seed(unsigned long long* __restrict__ seen, unsigned long long* __restrict__ front,
                             const int* __restrict__ src, int k0, int nk) {
    // one thread: sources may repeat across passes but not within one
    if (threadIdx.x == 0 && blockIdx.x == 0)
        for (int k = 0; k < nk; ++k) {
            seen[src[k0 + k]] |= 1ull << k;
            front[src[k0 + k]] |= 1ull << k;
        }
}

__global__ __launch_bounds__(256) void k_msbfs_level(const int* __restrict__ rowptr, const int* __restrict__ col,
                                                     int n, const unsigned long long* __restrict__ front,
                                                     unsigned long long* __restrict__ next,
                                                     unsigned long long* __restrict__ seen, int* __restrict__ changed) {
    int any = 0;
    for (int v = blockIdx.x * 256 + threadIdx.x; v < n; v += gridDim.x * 256) {
        const unsigned long long s = seen[v];
        unsigned long long acc = 0ull;
        if (~s != 0ull) {
            const int e1 = rowptr[v + 1];
            for (int e = rowptr[v]; e < e1; ++e) acc |= front[col[e]];
        }
        const unsigned long long nw = acc & ~s;
        next[v] = nw;
        if (nw) {
            seen[v] = s | nw;
            any = 1;
        }
    }
    if (__syncthreads_or(any) && threadIdx.x == 0) atomicOr(changed, 1);
}

// Pairs of sources k0..k0+nk-1 (pair index = source index * T + j).
__global__ __launch_bounds__(256) void k_msbfs_pairs(const int* __restrict__ tgt, int T, int k0, int nk,
                                                     const unsigned long long* __restrict__ next, int level,
                                                     int* __restrict__ dist) {
    const int total = nk * T;
    for (int q = blockIdx.x * 256 + threadIdx.x; q < total; q += gridDim.x * 256) {
        const int k = q / T;
        const int p = (k0 + k) * T + (q - k * T);
        if (dist[p] == 0 && ((next[tgt[p]] >> k) & 1ull)) dist[p] = level;
    }
}

__global__ __launch_bounds__(256) void k_stress_partials(const float2* __restrict__ pos, const int* __restrict__ src,
                                                         const int* __restrict__ tgt, const int* __restrict__ dist,
                                                         int T, long long np, double* __restrict__ part) {
    __shared__ double r1[256], r2[256], rc[256];
    double s1 = 0.0, s2 = 0.0, c = 0.0;
    for (long long p = blockIdx.x * 256ll + threadIdx.x; p < np; p += (long long)gridDim.x * 256) {
        const int d = dist[p];
        if (d <= 0) continue;
        const float2 a = pos[src[p / T]], b = pos[tgt[p]];
        const double dx = (double)a.x - b.x, dy = (double)a.y - b.y;
        const double r = sqrt(dx * dx + dy * dy) / d;
        s1 += r;
        s2 += r * r;
        c += 1.0;
    }
    const int t = threadIdx.x;
    r1[t] = s1;
    r2[t] = s2;
    rc[t] = c;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if (t < o) {
            r1[t] += r1[t + o];
            r2[t] += r2[t + o];
            rc[t] += rc[t + o];
        }
        __syncthreads();
    }
    if (t == 0) {
        part[3 * blockIdx.x] = r1[0];
        part[3 * blockIdx.x + 1] = r2[0];
        part[3 * blockIdx.x + 2] = rc[0];
    }
}

}  // namespace

int stress_hop_distances(const Bufs& b, const int* d_src, int K, const int* d_tgt, int T, int* d_dist,
                         unsigned long long* seen, unsigned long long* front, unsigned long long* next, int* d_flag,
                         hipStream_t s) {
    const int n = b.n;
    int blocks = (n + 255) / 256;
    if (blocks > 4096) blocks = 4096;
    if (hipMemsetAsync(d_dist, 0, sizeof(int) * (size_t)K * T, s) != hipSuccess) return -1;
    for (int k0 = 0; k0 < K; k0 += 64) {
        const int nk = K - k0 < 64 ? K - k0 : 64;
        k_msbfs_init<<<blocks, 256, 0, s>>>(seen, front, n, d_src, k0, nk);
        k_msbfs_seed<<<1, 64, 0, s>>>(seen, front, d_src, k0, nk);
        unsigned long long* cur = front;
        unsigned long long* nxt = next;
        for (int level = 1; level <= n; ++level) {
            if (hipMemsetAsync(d_flag, 0, sizeof(int), s) != hipSuccess) return -1;
            k_msbfs_level<<<blocks, 256, 0, s>>>(b.rowptr, b.col, n, cur, nxt, seen, d_flag);
            int pb = (nk * T + 255) / 256;
            k_msbfs_pairs<<<pb > 1024 ? 1024 : pb, 256, 0, s>>>(d_tgt, T, k0, nk, nxt, level, d_dist);
            int flag = 0;
            if (hipMemcpyAsync(&flag, d_flag, sizeof(int), hipMemcpyDeviceToHost, s) != hipSuccess) return -1;
            if (hipStreamSynchronize(s) != hipSuccess) return -1;
            unsigned long long* t = cur;
            cur = nxt;
            nxt = t;
            if (!flag) break;
        }
    }
    return hipGetLastError() == hipSuccess ? 0 : -1;
}

int stress_reduce(const float2* pos, const int* d_src, const int* d_tgt, const int* d_dist, int T, long long np,
                  double* d_part, int nblocks, double out[3], hipStream_t s) {
    k_stress_partials<<<nblocks, 256, 0, s>>>(pos, d_src, d_tgt, d_dist, T, np, d_part);
    std::vector<double> h((size_t)3 * nblocks);
    if (hipMemcpyAsync(h.data(), d_part, sizeof(double) * 3 * nblocks, hipMemcpyDeviceToHost, s) != hipSuccess)
        return -1;
    if (hipStreamSynchronize(s) != hipSuccess) return -1;
    out[0] = out[1] = out[2] = 0.0;
    for (int k = 0; k < nblocks; ++k) {
        out[0] += h[3 * k];
        out[1] += h[3 * k + 1];
        out[2] += h[3 * k + 2];
    }
    return 0;
}

}  // namespace gfa2
