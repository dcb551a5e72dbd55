// Batched layout of many graphs in one launch chain (no counterpart in the reference,
// whose FA2 state lives in device globals and allows one graph per process).
//
// A batch concatenates its graphs' nodes (graph g owns nodes [node_off[g], node_off[g+1]))
// into one block-diagonal CSR. Per iteration: one block per graph reduces its bounding
// box; every node's path key is taken against its own graph's root; the keys are sorted by
// path and then stably by graph id, which leaves each graph's bodies in its own in-order
// sequence on the graph's own range of sorted positions; the tree kernels treat a graph
// change as the end of shared prefixes, so every graph gets its own root and pre-order
// record range. Walk groups, speed state and displacement are per graph, so a graph laid
// out in a batch follows exactly the trajectory it follows alone.
#include "engine_internal.h"

#include <algorithm>

namespace gfa2 {
namespace {

__global__ __launch_bounds__(256) void k_seg_bbox(const float2* __restrict__ pos, const int* __restrict__ node_off,
                                                  float4* __restrict__ seg_root) {
    __shared__ float4 red[256];
    const int g = blockIdx.x, tid = threadIdx.x;
    const int a = node_off[g], b = node_off[g + 1];
    const float2 p0 = pos[a];
    float4 acc = make_float4(p0.x, p0.y, p0.x, p0.y);
    for (int i = a + tid; i < b; i += 256) {
        const float2 p = pos[i];
        acc.x = fminf(acc.x, p.x);
        acc.y = fminf(acc.y, p.y);
        acc.z = fmaxf(acc.z, p.x);
        acc.w = fmaxf(acc.w, p.y);
    }
    red[tid] = acc;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if (tid < o) {
            const float4 v = red[tid + o];
            acc.x = fminf(acc.x, v.x);
            acc.y = fminf(acc.y, v.y);
            acc.z = fmaxf(acc.z, v.z);
            acc.w = fmaxf(acc.w, v.w);
            red[tid] = acc;
        }
        __syncthreads();
    }
    if (tid == 0) {
        const float4 t = red[0];
        // root as BoundingBoxKernel sets it, src/RPBHKernels.cu:140-148
        seg_root[g] = make_float4((t.x + t.z) * 0.5f, (t.y + t.w) * 0.5f, fmaxf(t.z - t.x, t.w - t.y) * 0.5f, 0.0f);
    }
}

__global__ __launch_bounds__(256) void k_seg_keys(const float2* __restrict__ pos, const float* __restrict__ mass, int n,
                                                  const int* __restrict__ seg_of,
                                                  const float4* __restrict__ seg_root,
                                                  unsigned long long* __restrict__ keys,
                                                  unsigned int* __restrict__ vals,
                                                  unsigned long long* __restrict__ pkey, float4* __restrict__ pm) {
    for (int i = blockIdx.x * 256 + threadIdx.x; i < n; i += gridDim.x * 256) {
        const float4 root = seg_root[seg_of[i]];
        const float2 p = pos[i];
        float r = root.z * 0.5f;
        bool bx = root.x < p.x, by = root.y < p.y;
        float x = root.x + (bx ? r : -r);
        float y = root.y + (by ? r : -r);
        unsigned long long key = (unsigned long long)((unsigned)bx | ((unsigned)by << 1)) << 62;
#pragma unroll
        for (int l = 1; l < kKeyLevels; ++l) {
            r *= 0.5f;
            bx = x < p.x;
            by = y < p.y;
            key |= (unsigned long long)((unsigned)bx | ((unsigned)by << 1)) << (62 - 2 * l);
            x += bx ? r : -r;
            y += by ? r : -r;
        }
        keys[i] = key;
        if (pkey) pkey[i] = key;      // kept for the regather after a global graph sort
        vals[i] = (unsigned)i;
        pm[i] = make_float4(p.x, p.y, mass[i], 0.0f);
    }
}

// The graph owning node v: the last g with node_off[g] <= v, by binary search over the node
// offsets staged in LDS (up to kSegLds graphs; beyond, over global memory), instead of a
// gather from the per-node graph table, which after the path sort is random over all nodes.
constexpr int kSegLds = 8192;

__device__ __forceinline__ int graph_of(const int* __restrict__ off, int nseg, unsigned int v) {
    int lo = 0, hi = nseg;                  // off[lo] <= v < off[hi]
    while (hi - lo > 1) {
        const int mid = (lo + hi) >> 1;
        if ((unsigned int)off[mid] <= v) lo = mid;
        else hi = mid;
    }
    return lo;
}

__device__ __forceinline__ const int* stage_offsets(const int* __restrict__ node_off, int nseg, int* lds) {
    if (nseg + 1 > kSegLds) return node_off;
    for (int i = threadIdx.x; i <= nseg; i += blockDim.x) lds[i] = node_off[i];
    __syncthreads();
    return lds;
}

__global__ __launch_bounds__(256) void k_seg_gkeys(const unsigned int* __restrict__ vals, int n,
                                                   const int* __restrict__ node_off, int nseg,
                                                   unsigned long long* __restrict__ keys) {
    __shared__ int off_lds[kSegLds];
    const int* off = stage_offsets(node_off, nseg, off_lds);
    for (int k = blockIdx.x * 256 + threadIdx.x; k < n; k += gridDim.x * 256)
        keys[k] = (unsigned long long)graph_of(off, nseg, vals[k]);
}

__global__ __launch_bounds__(256) void k_seg_regather(const unsigned int* __restrict__ vals, int n,
                                                      const int* __restrict__ node_off, int nseg,
                                                      const unsigned long long* __restrict__ pkey,
                                                      unsigned long long* __restrict__ keys, int* __restrict__ sg) {
    __shared__ int off_lds[kSegLds];
    const int* off = stage_offsets(node_off, nseg, off_lds);
    for (int k = blockIdx.x * 256 + threadIdx.x; k < n; k += gridDim.x * 256) {
        const unsigned int v = vals[k];
        keys[k] = pkey[v];
        sg[k] = graph_of(off, nseg, v);
    }
}

// One element per thread up to 2^24 elements (these kernels gather at random: more waves in
// flight hide more of the latency), grid-stride beyond.
int grid_for(int n) {
    int b = (n + 255) / 256;
    return b < 1 ? 1 : (b > 65536 ? 65536 : b);
}

}  // namespace

void launch_seg_bbox(const Bufs& b, hipStream_t s) {
    k_seg_bbox<<<b.nseg, 256, 0, s>>>(b.pos, b.node_off, b.seg_root);
}

void launch_seg_keys(const Bufs& b, hipStream_t s) {
    k_seg_keys<<<grid_for(b.n), 256, 0, s>>>(b.pos, b.mass, b.n, b.seg_of, b.seg_root, b.sort.keys[0], b.sort.vals[0],
                                             b.max_seg <= kSegSortMax ? nullptr : b.pkey, b.pm);
}

void launch_seg_graph_sort(Bufs& b, hipStream_t s) {
    // each block stages the graph offsets once: a grid of at most 2,048 blocks, grid-stride
    const int gb = std::min(grid_for(b.n), 2048);
    k_seg_gkeys<<<gb, 256, 0, s>>>(b.sort.vals[0], b.n, b.node_off, b.nseg, b.sort.keys[0]);
    radix_sort_pairs(b.sort, b.n, 0, (b.seg_bits + 7) & ~7, s);
    k_seg_regather<<<gb, 256, 0, s>>>(b.sort.vals[0], b.n, b.node_off, b.nseg, b.pkey, b.sort.keys[0], b.sg);
}

}  // namespace gfa2
