// Barnes-Hut quadtree construction for gfx950, deterministic and lock-free.
//
// The reference builds its quadtree by concurrent insertion with CAS locks and then
// summarises and sorts it with spin-waiting kernels (src/RPBHKernels.cu:184-559). Here
// the same tree is derived from sorted path keys instead:
//  1. bounding box (the reference's BoundingBoxKernel, src/RPBHKernels.cu:70-157), from
//     per-block partials every key block reduces redundantly;
//  2. per body a 64-bit key of 32 quadrant digits (bit0: x, bit1: y), obtained by
//     replaying the reference's float centre arithmetic along the body's path
//     (r *= 0.5f, x += +/-r; src/RPBHKernels.cu:217-241, :292-308). A quadrant tree is
//     fixed by the bodies' paths, whatever the insertion order;
//  3. radix sort of the keys: the sorted order is the reference's in-order body
//     permutation (SortKernel, src/RPBHKernels.cu:513-559);
//  4. with lcp_i the common digits of sorted neighbours i and i+1, the cells are exactly
//     the prefixes of length <= lcp_i; the cells starting at position i have levels
//     lcp_{i-1}+1 .. lcp_i. Their pre-order record index follows from an exclusive scan
//     of those counts, and each cell's extent from a galloping search over the keys;
//  5. summaries. Counts are extents. For whole-number masses (the default deg+1) a cell
//     lighter than 2^24 has the same float mass in every summation order, so its mass and
//     centroid come from exact integer prefix sums over the sorted bodies (fixed-point
//     moments, see Moment); heavier cells are summed in quadrant order. Other masses:
//     bottom-up, one launch per level, children in quadrant order, mass accumulated from
//     0.0f and centroid (sum m*p)*(1/mass), as the reference's SummarizationKernel
//     (src/RPBHKernels.cu:357-506).
// Nothing waits on another workgroup; every dependency crosses a launch boundary.
#include "engine_internal.h"

namespace gfa2 {
namespace {

__device__ __forceinline__ int lcp_levels(unsigned long long a, unsigned long long b) {
    unsigned long long x = a ^ b;
    return x == 0ull ? kKeyLevels : (__clzll((long long)x) >> 1);
}

// Reduce bounding-box partials {minx, miny, maxx, maxy} in LDS (every block does it).
__device__ void reduce_bbox(const float4* __restrict__ part, int np, float* root /*[3] LDS*/) {
    __shared__ float4 red[256];
    const int tid = threadIdx.x;
    float4 acc = part[0];
    for (int k = tid; k < np; k += blockDim.x) {
        float4 v = part[k];
        acc.x = fminf(acc.x, v.x);
        acc.y = fminf(acc.y, v.y);
        acc.z = fmaxf(acc.z, v.z);
        acc.w = fmaxf(acc.w, v.w);
    }
    red[tid] = acc;
    __syncthreads();
    for (int o = blockDim.x / 2; o > 0; o >>= 1) {
        if (tid < o) {
            float4 v = red[tid + o];
            acc.x = fminf(acc.x, v.x);
            acc.y = fminf(acc.y, v.y);
            acc.z = fmaxf(acc.z, v.z);
            acc.w = fmaxf(acc.w, v.w);
            red[tid] = acc;
        }
        __syncthreads();
    }
    if (tid == 0) {
        float4 t = red[0];
        root[2] = fmaxf(t.z - t.x, t.w - t.y) * 0.5f;   // src/RPBHKernels.cu:140
        root[0] = (t.x + t.z) * 0.5f;                    // :147
        root[1] = (t.y + t.w) * 0.5f;                    // :148
    }
    __syncthreads();
}

__global__ __launch_bounds__(256) void k_bbox_partials(const float2* __restrict__ pos, int n,
                                                       float4* __restrict__ part, DevState* st) {
    __shared__ float4 red[256];
    const int tid = threadIdx.x;
    float2 p0 = pos[0];
    float4 acc = make_float4(p0.x, p0.y, p0.x, p0.y);
    for (int i = blockIdx.x * 256 + tid; i < n; i += gridDim.x * 256) {
        float2 p = pos[i];
        acc.x = fminf(acc.x, p.x);
        acc.y = fminf(acc.y, p.y);
        acc.z = fmaxf(acc.z, p.x);
        acc.w = fmaxf(acc.w, p.y);
    }
    red[tid] = acc;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if (tid < o) {
            float4 v = red[tid + o];
            acc.x = fminf(acc.x, v.x);
            acc.y = fminf(acc.y, v.y);
            acc.z = fmaxf(acc.z, v.z);
            acc.w = fmaxf(acc.w, v.w);
            red[tid] = acc;
        }
        __syncthreads();
    }
    if (tid == 0) {
        part[blockIdx.x] = red[0];
        if (blockIdx.x == 0) st->nbbox = gridDim.x;
    }
}

// Root box + path keys. vals = identity; pm = {x, y, mass} per node, so that k_lcp gathers one
// 16-byte record per sorted body instead of a position and a mass. With dhist (a sort of one
// launch per pass), also the digit counts of all 8 passes, added to the table the previous
// iteration's fix-up cleared (the sort's own histogram launches are then skipped).
__global__ __launch_bounds__(256) void k_keys(const float2* __restrict__ pos, const float* __restrict__ mass, int n,
                                              const float4* __restrict__ part, DevState* st,
                                              float4* __restrict__ seg_root,
                                              unsigned long long* __restrict__ k0, unsigned int* __restrict__ v0,
                                              unsigned long long* __restrict__ k1, unsigned int* __restrict__ v1,
                                              float4* __restrict__ pm, int first, unsigned int* __restrict__ dhist) {
    __shared__ float root[3];
    __shared__ unsigned int h[8][256];
    if (dhist)
        for (int p = 0; p < 8; ++p) h[p][threadIdx.x] = 0u;
    // the buffer from which the sort passes that will run (8 less the skipped or unlaunched low
    // ones) end in buffer 0
    const bool odd = ((8 - max(st->sort_skip, first)) & 1) != 0;
    unsigned long long* __restrict__ keys = odd ? k1 : k0;
    unsigned int* __restrict__ vals = odd ? v1 : v0;
    reduce_bbox(part, st->nbbox, root);
    const float rootx = root[0], rooty = root[1], radius = root[2];
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        st->rootx = rootx;
        st->rooty = rooty;
        st->radius = radius;
        seg_root[0] = make_float4(rootx, rooty, radius, 0.0f);
    }
    for (int i = blockIdx.x * 256 + threadIdx.x; i < n; i += gridDim.x * 256) {
        const float2 p = pos[i];
        float r = radius * 0.5f;
        bool bx = rootx < p.x, by = rooty < p.y;
        float x = rootx + (bx ? r : -r);
        float y = rooty + (by ? r : -r);
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
        vals[i] = (unsigned)i;
        pm[i] = make_float4(p.x, p.y, mass[i], 0.0f);   // one gather per body for k_lcp
        if (dhist)
#pragma unroll
            for (int d = 0; d < 8; ++d) atomicAdd(&h[d][(unsigned)(key >> (8 * d)) & 255u], 1u);
    }
    if (dhist) {
        __syncthreads();
        for (int d = 0; d < 8; ++d) {
            const unsigned int c = h[d][threadIdx.x];
            if (c) atomicAdd(&dhist[d * 256 + threadIdx.x], c);
        }
    }
}

// A graph of a batch starts at sorted position i (always true for i == 0).
__device__ __forceinline__ bool seg_start(const int* __restrict__ sg, int i) {
    return i == 0 || (sg != nullptr && sg[i - 1] != sg[i]);
}

// Levels shared by sorted positions i and i+1; -1 at the end of a graph, except for a
// graph of one body, whose root (level 0) opens at its only position.
__device__ __forceinline__ int lcp_at(const unsigned long long* __restrict__ keys, const int* __restrict__ sg, int n,
                                      int i) {
    if (i < 0) return -1;
    const bool end = i >= n - 1 || (sg != nullptr && sg[i] != sg[i + 1]);
    if (end) return seg_start(sg, i) ? 0 : -1;
    return lcp_levels(keys[i], keys[i + 1]);
}

// The level below which cells opened at i are shared with position i-1.
__device__ __forceinline__ int lcp_prev(const signed char* __restrict__ lcp, const int* __restrict__ sg, int i) {
    return seg_start(sg, i) ? -1 : (int)lcp[i - 1];
}

// Block-wide exclusive scan of one int per thread (256 threads). Returns the exclusive
// prefix; *total gets the block sum.
__device__ int block_exclusive_scan(int v, int* total) {
    __shared__ int wsum[kTreeThreads / 64];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    int inc = v;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        int t = __shfl_up(inc, o, 64);
        if (lane >= o) inc += t;
    }
    if (lane == 63) wsum[wave] = inc;
    __syncthreads();
    int before = 0, all = 0;
#pragma unroll
    for (int w = 0; w < kTreeThreads / 64; ++w) {
        int s = wsum[w];
        if (w < wave) before += s;
        all += s;
    }
    __syncthreads();
    *total = all;
    return before + inc - v;
}

// Exact moments (integer-mass path). Sums are taken modulo 2^128 on two 64-bit limbs.
using u128 = unsigned __int128;

__device__ __forceinline__ u128 mo_x(const Moment& a) { return ((u128)a.xh << 64) | a.xl; }
__device__ __forceinline__ u128 mo_y(const Moment& a) { return ((u128)a.yh << 64) | a.yl; }

__device__ __forceinline__ Moment mo_make(u128 x, u128 y, unsigned long long m) {
    return Moment{(unsigned long long)x, (unsigned long long)(x >> 64), (unsigned long long)y,
                  (unsigned long long)(y >> 64), m};
}

__device__ __forceinline__ Moment mo_add(const Moment& a, const Moment& b) {
    return mo_make(mo_x(a) + mo_x(b), mo_y(a) + mo_y(b), a.m + b.m);
}

__device__ __forceinline__ Moment mo_sub(const Moment& a, const Moment& b) {
    return mo_make(mo_x(a) - mo_x(b), mo_y(a) - mo_y(b), a.m - b.m);
}

__device__ __forceinline__ Moment mo_shfl_up(const Moment& v, int o) {
    return Moment{__shfl_up(v.xl, o, 64), __shfl_up(v.xh, o, 64), __shfl_up(v.yl, o, 64), __shfl_up(v.yh, o, 64),
                  __shfl_up(v.m, o, 64)};
}

// Fixed-point exponent of a graph whose root radius is R < 2^e: a coordinate relative to the
// root centre lies within a few R (the centre and radius are rounded floats), so scaled by
// 2^(58-e) it stays below 2^61 in magnitude. R = 0 (one point) gives coordinates 0.
__device__ __forceinline__ int moment_shift(float radius) {
    int e = 0;
    (void)frexpf(radius, &e);
    return 58 - e;
}

__device__ __forceinline__ long long fixed_coord(float p, float c, int s) {
    double t = ldexp((double)p - (double)c, s);
    t = fmin(fmax(t, -0x1p62), 0x1p62);
    return (long long)rint(t);
}

// The signed 128-bit value hi*2^64 + lo as a double (magnitude converted, so small negative
// values keep their precision).
__device__ __forceinline__ double i128_to_double(unsigned long long lo, unsigned long long hi) {
    u128 v = ((u128)hi << 64) | lo;
    const bool neg = (long long)hi < 0;
    if (neg) v = (u128)0 - v;
    const unsigned long long vh = (unsigned long long)(v >> 64), vl = (unsigned long long)v;
    const double mag = vh == 0ull ? (double)vl : (double)vh * 0x1p64 + (double)vl;
    return neg ? -mag : mag;
}

// Block-wide exclusive scan of one Moment per thread (NT threads); *total gets the sum.
template <int NT>
__device__ Moment block_exclusive_scan_mo(const Moment& v, Moment* total) {
    __shared__ Moment ws[NT / 64];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    Moment inc = v;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        const Moment t = mo_shfl_up(inc, o);
        if (lane >= o) inc = mo_add(inc, t);
    }
    if (lane == 63) ws[wave] = inc;
    __syncthreads();
    Moment before{0ull, 0ull, 0ull, 0ull, 0ull}, all{0ull, 0ull, 0ull, 0ull, 0ull};
    for (int w = 0; w < NT / 64; ++w) {
        const Moment s = ws[w];
        if (w < wave) before = mo_add(before, s);
        all = mo_add(all, s);
    }
    __syncthreads();
    *total = all;
    return mo_sub(mo_add(before, inc), v);
}

// Mass and mass-weighted fixed-point position (relative to the graph's root centre) of a
// sorted body {x, y, mass} of graph g; the mass is a whole number below 2^24.
__device__ __forceinline__ Moment body_moment(const float4 bd, int g, const float4* __restrict__ seg_root) {
    const float4 c = seg_root[g];
    const int s = moment_shift(c.z);
    const unsigned long long m = (unsigned long long)bd.z;
    const u128 x = (u128)(__int128)fixed_coord(bd.x, c.x, s) * (u128)m;
    const u128 y = (u128)(__int128)fixed_coord(bd.y, c.y, s) * (u128)m;
    return mo_make(x, y, m);
}

// lcp per sorted position and the number of cells opened there; block sums (of the
// cell counts and, on the integer-mass path, of body mass sums). Each sorted body's
// {x, y, mass} is gathered here once and written in sorted order (sbody) for the kernels
// that follow.
__global__ __launch_bounds__(kTreeThreads) void k_lcp(const unsigned long long* __restrict__ keys,
                                                      const int* __restrict__ sg, int n,
                                                      signed char* __restrict__ lcp,
                                                      int* __restrict__ blocksum,
                                                      const unsigned int* __restrict__ vals,
                                                      const float4* __restrict__ pm,
                                                      const float4* __restrict__ seg_root,
                                                      Moment* __restrict__ bmass, float4* __restrict__ sbody) {
    // the tile's keys (with one on each side), graph ids and body ids, read coalesced into
    // LDS; each thread then takes its kTreeItems consecutive positions from there
    __shared__ unsigned long long skey[kTreeTile + 2];
    __shared__ int ssg[kTreeTile + 2];
    __shared__ unsigned int sval[kTreeTile];
    const int base = blockIdx.x * kTreeTile;
    for (int t = threadIdx.x; t < kTreeTile + 2; t += kTreeThreads) {
        const int gi = base - 1 + t;
        const bool in = gi >= 0 && gi < n;
        skey[t] = in ? keys[gi] : 0ull;
        if (sg) ssg[t] = in ? sg[gi] : -1;
        if (t < kTreeTile) sval[t] = base + t < n ? vals[base + t] : 0u;
    }
    __syncthreads();
    // seg_start and lcp_at over the staged tile (position i at index i - base + 1); the
    // position before the tile is only asked for its lcp when no graph starts at the tile
    auto seg_start_l = [&](int i) { return i == 0 || (sg != nullptr && ssg[i - base] != ssg[i - base + 1]); };
    auto lcp_at_l = [&](int i) -> int {
        if (i < 0) return -1;
        const bool end = i >= n - 1 || (sg != nullptr && ssg[i - base + 1] != ssg[i - base + 2]);
        if (end) return seg_start_l(i) ? 0 : -1;
        return lcp_levels(skey[i - base + 1], skey[i - base + 2]);
    };
    const int i0 = base + threadIdx.x * kTreeItems;
    // all gathers of the thread's items first, outside the per-item branch, so they are in
    // flight together (ids past n are 0, a valid node)
    float4 pp[kTreeItems];
#pragma unroll
    for (int k = 0; k < kTreeItems; ++k) pp[k] = pm[sval[i0 + k - base]];
    int sum = 0;
    Moment ms{0ull, 0ull, 0ull, 0ull, 0ull};
#pragma unroll
    for (int k = 0; k < kTreeItems; ++k) {
        int i = i0 + k;
        if (i < n) {
            const int prev = seg_start_l(i) ? -1 : lcp_at_l(i - 1);
            int l = lcp_at_l(i);
            lcp[i] = (signed char)l;
            sum += max(0, l - prev);
            const float4 bd = make_float4(pp[k].x, pp[k].y, pp[k].z, 0.0f);
            sbody[i] = bd;
            if (bmass) ms = mo_add(ms, body_moment(bd, sg ? ssg[i - base + 1] : 0, seg_root));
        }
    }
    int total;
    block_exclusive_scan(sum, &total);
    if (threadIdx.x == 0) blocksum[blockIdx.x] = total;
    if (bmass) {
        Moment mt;
        block_exclusive_scan_mo<kTreeThreads>(ms, &mt);
        if (threadIdx.x == 0) bmass[blockIdx.x] = mt;
    }
}

// Two-level scan of the tile totals, for more tiles than one block scans quickly: per chunk
// of kTileChunk tiles its totals (up), then one block scans the chunk totals (k_scan_blocks,
// which also sets the tree's totals), then every chunk scans its tiles from its offset (down).
// The sums are integers, so the result is the one-level scan's.
__global__ __launch_bounds__(kTileChunk) void k_scan_up(const int* __restrict__ blocksum,
                                                        const Moment* __restrict__ bmass, int nb,
                                                        int* __restrict__ chunk_sum,
                                                        Moment* __restrict__ chunk_mass) {
    const int t = blockIdx.x * kTileChunk + threadIdx.x;
    int total;
    (void)block_exclusive_scan(t < nb ? blocksum[t] : 0, &total);
    if (threadIdx.x == 0) chunk_sum[blockIdx.x] = total;
    if (bmass) {
        Moment mt;
        (void)block_exclusive_scan_mo<kTileChunk>(t < nb ? bmass[t] : Moment{0ull, 0ull, 0ull, 0ull, 0ull}, &mt);
        if (threadIdx.x == 0) chunk_mass[blockIdx.x] = mt;
    }
}

__global__ __launch_bounds__(kTileChunk) void k_scan_down(int* __restrict__ blocksum, Moment* __restrict__ bmass,
                                                          int nb, const int* __restrict__ chunk_sum,
                                                          const Moment* __restrict__ chunk_mass) {
    const int t = blockIdx.x * kTileChunk + threadIdx.x;
    int total;
    const int off = block_exclusive_scan(t < nb ? blocksum[t] : 0, &total) + chunk_sum[blockIdx.x];
    Moment moff{0ull, 0ull, 0ull, 0ull, 0ull};
    if (bmass) {
        Moment mt;
        moff = mo_add(block_exclusive_scan_mo<kTileChunk>(t < nb ? bmass[t] : Moment{0ull, 0ull, 0ull, 0ull, 0ull}, &mt),
                      chunk_mass[blockIdx.x]);
    }
    if (t < nb) {
        blocksum[t] = off;
        if (bmass) bmass[t] = moff;
    }
}

// Single block: exclusive scan of block sums (in place), cell total, capacity check,
// per-iteration counters reset.
__global__ __launch_bounds__(1024) void k_scan_blocks(int* __restrict__ blocksum, int nb, int n,
                                                      int cap_records, int* __restrict__ rec_end, DevState* st,
                                                      Moment* __restrict__ bmass) {
    __shared__ int wsum[16];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int per = (nb + 1023) / 1024;
    const int b0 = tid * per;
    // each thread's run of `per` tiles is read kAhead entries at a time with independent loads,
    // then taken in order
    constexpr int kAhead = 8;
    auto valid = [&](int k0, int j) { return k0 + j < per && b0 + k0 + j < nb; };
    if (bmass) {
        Moment loc{0ull, 0ull, 0ull, 0ull, 0ull};
        for (int k0 = 0; k0 < per; k0 += kAhead) {
            Moment c[kAhead];
#pragma unroll
            for (int j = 0; j < kAhead; ++j) c[j] = valid(k0, j) ? bmass[b0 + k0 + j] : Moment{0ull, 0ull, 0ull, 0ull, 0ull};
#pragma unroll
            for (int j = 0; j < kAhead; ++j)
                if (valid(k0, j)) loc = mo_add(loc, c[j]);
        }
        Moment tot;
        Moment run = block_exclusive_scan_mo<1024>(loc, &tot);
        for (int k0 = 0; k0 < per; k0 += kAhead) {
            Moment c[kAhead];
#pragma unroll
            for (int j = 0; j < kAhead; ++j) c[j] = valid(k0, j) ? bmass[b0 + k0 + j] : Moment{0ull, 0ull, 0ull, 0ull, 0ull};
#pragma unroll
            for (int j = 0; j < kAhead; ++j)
                if (valid(k0, j)) {
                    bmass[b0 + k0 + j] = run;
                    run = mo_add(run, c[j]);
                }
        }
    }
    int local = 0;
    for (int k0 = 0; k0 < per; k0 += kAhead) {
        int c[kAhead];
#pragma unroll
        for (int j = 0; j < kAhead; ++j) c[j] = valid(k0, j) ? blocksum[b0 + k0 + j] : 0;
#pragma unroll
        for (int j = 0; j < kAhead; ++j) local += c[j];
    }
    int inc = local;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        int t = __shfl_up(inc, o, 64);
        if (lane >= o) inc += t;
    }
    if (lane == 63) wsum[wave] = inc;
    __syncthreads();
    int before = 0, all = 0;
    for (int w = 0; w < 16; ++w) {
        if (w < wave) before += wsum[w];
        all += wsum[w];
    }
    int run = before + inc - local;
    for (int k0 = 0; k0 < per; k0 += kAhead) {
        int c[kAhead];
#pragma unroll
        for (int j = 0; j < kAhead; ++j) c[j] = valid(k0, j) ? blocksum[b0 + k0 + j] : 0;
#pragma unroll
        for (int j = 0; j < kAhead; ++j)
            if (valid(k0, j)) {
                blocksum[b0 + k0 + j] = run;
                run += c[j];
            }
    }
    if (tid == 0) {
        st->ncells = all;
        st->nrecords = n + all;
        *rec_end = n + all;
        if (n + all > cap_records) st->err |= kErrTreeOverflow;
        st->max_level = 0;
        st->nbig = 0;
    }
    if (tid <= kMaxLevels) {
        st->level_count[tid] = 0;
        st->level_cursor[tid] = 0;
    }
}

// S (exclusive prefix of opened cells), body records and brec. With `fused` (few tiles), the
// tile offsets are not scanned by a launch of their own: every block sums the totals of the
// tiles before it (k_lcp's, at most kFuseTiles of them, read from L2), its writes are bounded
// by its own extent, and the blocks at the ends take over the scan launch's state updates.
__global__ __launch_bounds__(kTreeThreads) void k_prefix(const signed char* __restrict__ lcp,
                                                         const int* __restrict__ sg, int n,
                                                         const int* __restrict__ blockoff,
                                                         int* __restrict__ S,
                                                         const float4* __restrict__ sbody,
                                                         float4* __restrict__ rec, int* __restrict__ brec,
                                                         DevState* __restrict__ st,
                                                         const float4* __restrict__ seg_root,
                                                         const Moment* __restrict__ bmass, Moment* __restrict__ pmass,
                                                         int* __restrict__ cell_pos, int max_cells, int fused,
                                                         int cap_records, int* __restrict__ rec_end) {
    // the tile's sorted bodies and graph ids (one before the tile), read coalesced into LDS
    __shared__ float4 sbd[kTreeTile];
    __shared__ int ssg[kTreeTile + 1];
    const int base = blockIdx.x * kTreeTile;
    for (int t = threadIdx.x; t < kTreeTile + 1; t += kTreeThreads) {
        const int gi = base - 1 + t;
        if (sg) ssg[t] = gi >= 0 && gi < n ? sg[gi] : -1;
        if (t < kTreeTile) sbd[t] = base + t < n ? sbody[base + t] : make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    }
    __syncthreads();
    const int i0 = base + threadIdx.x * kTreeItems;
    int s[kTreeItems];
    float4 bd[kTreeItems];
    int sum = 0;
#pragma unroll
    for (int k = 0; k < kTreeItems; ++k) {
        int i = i0 + k;
        // lcp_prev over the staged graph ids: no graph starts at i
        const bool start = i == 0 || (sg != nullptr && ssg[i - base] != ssg[i - base + 1]);
        s[k] = i < n ? max(0, (int)lcp[i] - (start ? -1 : (int)lcp[i - 1])) : 0;
        bd[k] = sbd[i - base];
        sum += s[k];
    }
    int total;
    int run = block_exclusive_scan(sum, &total);
    int off;
    Moment moff{0ull, 0ull, 0ull, 0ull, 0ull};
    if (fused) {
        int acc = 0;
        Moment macc{0ull, 0ull, 0ull, 0ull, 0ull};
        for (int t = threadIdx.x; t < (int)blockIdx.x; t += kTreeThreads) {
            acc += blockoff[t];
            if (bmass) macc = mo_add(macc, bmass[t]);
        }
        (void)block_exclusive_scan(acc, &off);
        if (bmass) (void)block_exclusive_scan_mo<kTreeThreads>(macc, &moff);
    } else {
        off = blockoff[blockIdx.x];
        if (bmass) moff = bmass[blockIdx.x];
    }
    run += off;
    // records written here: below min(base + tile, n) + off + total (the cells before and in
    // this tile), within capacity whenever the whole tree is
    const bool ok = st->err == 0 && (!fused || min(base + kTreeTile, n) + off + total <= cap_records);
    if (fused && threadIdx.x == 0 && blockIdx.x == gridDim.x - 1) {
        const int all = off + total;
        st->ncells = all;
        st->nrecords = n + all;
        *rec_end = n + all;
        if (n + all > cap_records) st->err |= kErrTreeOverflow;
    }
    if (fused && blockIdx.x == 0) {
        if (threadIdx.x == 0) {
            st->max_level = 0;
            st->nbig = 0;
        }
        if (threadIdx.x <= kMaxLevels) {
            st->level_count[threadIdx.x] = 0;
            st->level_cursor[threadIdx.x] = 0;
        }
    }
    if (pmass) {
        // exclusive prefix of body mass sums over sorted positions
        Moment mine{0ull, 0ull, 0ull, 0ull, 0ull};
#pragma unroll
        for (int k = 0; k < kTreeItems; ++k)
            if (i0 + k < n) mine = mo_add(mine, body_moment(bd[k], sg ? ssg[i0 + k - base + 1] : 0, seg_root));
        Moment tot;
        Moment pre = mo_add(block_exclusive_scan_mo<kTreeThreads>(mine, &tot), moff);
#pragma unroll
        for (int k = 0; k < kTreeItems; ++k) {
            const int i = i0 + k;
            if (i < n) {
                pmass[i] = pre;
                pre = mo_add(pre, body_moment(bd[k], sg ? ssg[i - base + 1] : 0, seg_root));
                if (i == n - 1) pmass[n] = pre;
            }
        }
    }
#pragma unroll
    for (int k = 0; k < kTreeItems; ++k) {
        int i = i0 + k;
        if (i < n) {
            S[i] = run;
            if (ok)
                for (int j = 0; j < s[k] && run + j < max_cells; ++j) cell_pos[run + j] = i;
            run += s[k];
            int rb = i + run;          // after this position's own cells
            brec[i] = rb;
            if (ok) rec[rb] = make_float4(bd[k].x, bd[k].y, bd[k].z, __int_as_float(-1));
        }
    }
    if (i0 + kTreeItems >= n && i0 < n) S[n] = run;   // total cells
}

// Cell records (skip, level, count) and the cell list, one thread per cell (grid-stride over
// the S[n] cells). cell_rec[ord] holds the cell's first sorted position on entry (written by
// k_prefix) and the cell's record index on exit (the list k_kids walks; on the fractional-mass
// path, with cell_lvl, also the per-level summaries' list). The cells opened at
// position i have levels lprev+1 .. lcp_i in ord order; each finds its extent end by a
// galloping search over the keys (sorted within a graph, so a cell's bodies are contiguous),
// bounded by the graph's end (a graph's root needs no search: it ends where the graph does).
// One thread per cell keeps the long searches of the top levels in parallel.
__global__ __launch_bounds__(kTreeThreads) void k_nodes(const unsigned long long* __restrict__ keys,
                                                        const signed char* __restrict__ lcp,
                                                        const int* __restrict__ sg, int n,
                                                        const int* __restrict__ S, float4* __restrict__ rec,
                                                        int* __restrict__ rcount, int* __restrict__ cell_rec,
                                                        unsigned char* __restrict__ cell_lvl, int* __restrict__ rec_begin,
                                                        DevState* st, const Moment* __restrict__ pmass,
                                                        const float4* __restrict__ seg_root, int* __restrict__ big_rc,
                                                        unsigned char* __restrict__ big_lvl,
                                                        const int* __restrict__ node_off) {
    __shared__ int hist[kMaxLevels + 1];
    __shared__ int lmax;
    if (threadIdx.x <= kMaxLevels) hist[threadIdx.x] = 0;
    if (threadIdx.x == 0) lmax = 0;
    __syncthreads();
    const bool ok = st->err == 0;
    const int ncells = ok ? S[n] : 0;
    for (int ord = blockIdx.x * blockDim.x + threadIdx.x; ord < ncells; ord += gridDim.x * blockDim.x) {
        const int i = cell_rec[ord];
        const int Si = S[i];
        const int L = lcp_prev(lcp, sg, i) + 1 + (ord - Si);
        const unsigned long long ki = keys[i];
        const int gi = sg ? sg[i] : 0;
        // extent end: largest j < lim (the graph's end) with lcp(k_i, k_j) >= L; the test is
        // true up to the end and false after it
        const int lim = sg ? node_off[gi + 1] : n;
        int lo = i;                   // inside
        if (L == 0) {
            lo = lim - 1;             // a graph's root holds the whole graph
        } else {
            int step = 1;
            int hi = lo + 1;
            while (hi < lim && lcp_levels(ki, keys[hi]) >= L) {
                lo = hi;
                step <<= 1;
                hi = lo + step;
            }
            if (hi > lim) hi = lim;
            while (hi - lo > 1) {
                const int mid = lo + ((hi - lo) >> 1);
                if (lcp_levels(ki, keys[mid]) >= L) lo = mid;
                else hi = mid;
            }
        }
        const int b = lo;
        const int rc = i + ord;
        const unsigned int skip = (unsigned int)(b + S[b + 1] + 1 - rc);
        const unsigned int info = (skip << kInfoLevelBits) | (unsigned int)((L > 0 ? L - 1 : 0) & 31);
        if (pmass) {
            // whole-number masses: a cell lighter than 2^24 has this exact float mass in any
            // summation order; heavier cells are summed in quadrant order later
            const Moment d = mo_sub(pmass[b + 1], pmass[i]);
            const double M = (double)d.m;
            if (M < 16777216.0) {
                const float4 c = seg_root[gi];
                const int sh = moment_shift(c.z);
                rec[rc] = make_float4((float)(ldexp(i128_to_double(d.xl, d.xh) / M, -sh) + c.x),
                                      (float)(ldexp(i128_to_double(d.yl, d.yh) / M, -sh) + c.y), (float)M,
                                      __uint_as_float(info));
            } else {
                rec[rc].w = __uint_as_float(info);
                const int slot = atomicAdd(&st->nbig, 1);
                if (slot < kMaxBig) {
                    big_rc[slot] = rc;
                    big_lvl[slot] = (unsigned char)L;
                } else {
                    st->err |= kErrTreeOverflow;
                }
            }
        } else {
            rec[rc].w = __uint_as_float(info);
        }
        rcount[rc] = b - i + 1;
        cell_rec[ord] = rc;       // the cell list of k_kids (and of the summary launches)
        if (!pmass) {
            // the per-level cell lists of the summary launches (not needed for integer masses)
            cell_lvl[ord] = (unsigned char)L;
            atomicAdd(&hist[L], 1);
        }
        if (L == 0 && rec_begin) rec_begin[gi] = rc;
        atomicMax(&lmax, L);
    }
    __syncthreads();
    if (!pmass && threadIdx.x <= kMaxLevels && hist[threadIdx.x] != 0)
        atomicAdd(&st->level_count[threadIdx.x], hist[threadIdx.x]);
    // every block would hit this one word: only blocks above the value seen so far update it
    // (the maximum only grows, so skipping on a stale value is still correct)
    if (threadIdx.x == 0 && lmax > 0 &&
        lmax > __hip_atomic_load(&st->max_level, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT))
        atomicMax(&st->max_level, lmax);
}

// Children of every cell record, for the walk: a cell's subtree is the record run
// [r, r + skip); its children are r + 1 and then each next record past the previous child's
// subtree (a body spans one record), at most four. Unused slots hold -1.
__global__ __launch_bounds__(256) void k_kids(const float4* __restrict__ rec, const int* __restrict__ rcount,
                                              const int* __restrict__ S, int n,
                                              const int* __restrict__ cell_rec, int4* __restrict__ kids,
                                              int* __restrict__ kids_part, const DevState* __restrict__ st) {
    __shared__ int rmax[8];          // largest cell at level 4, 8, ..., 32 seen by the block
    if (st->err != 0) return;        // uniform
    if (threadIdx.x < 8) rmax[threadIdx.x] = 0;
    __syncthreads();
    const int ncells = S[n];
    for (int ord = blockIdx.x * 256 + threadIdx.x; ord < ncells; ord += gridDim.x * 256) {
        const int r = cell_rec[ord];
        const unsigned int info = __float_as_uint(rec[r].w);
        // the level bits hold L - 1 (L >= 1): levels 4, 8, ..., 32 are 3, 7, ..., 31
        const unsigned int lm1 = info & 31u;
        if ((lm1 & 3u) == 3u) atomicMax(&rmax[lm1 >> 2], rcount[r]);
        const int end = r + (int)(info >> kInfoLevelBits);
        int k[4] = {-1, -1, -1, -1};
        int c = r + 1;
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            if (c < end) {
                k[q] = c;
                const unsigned int ci = __float_as_uint(rec[c].w);
                c += ci == 0xffffffffu ? 1 : max((int)(ci >> kInfoLevelBits), 1);
            }
        }
        kids[r] = make_int4(k[0], k[1], k[2], k[3]);
    }
    __syncthreads();
    // per-block maxima, reduced by the speed step (no word shared by the blocks: thousands of
    // device-scope reads or atomics on one line cost more than the kernel itself)
    if (threadIdx.x < 8) kids_part[blockIdx.x * 8 + threadIdx.x] = rmax[threadIdx.x];
}

// Bucket cells by level (order inside a level is irrelevant).
__global__ __launch_bounds__(256) void k_levels(const int* __restrict__ cell_rec,
                                                const unsigned char* __restrict__ cell_lvl,
                                                int* __restrict__ level_list, DevState* st) {
    __shared__ int off[kMaxLevels + 2];
    __shared__ int hist[kMaxLevels + 1];
    __shared__ int base[kMaxLevels + 1];
    const int tid = threadIdx.x;
    if (tid == 0) {
        int run = 0;
        for (int L = 0; L <= kMaxLevels; ++L) {
            off[L] = run;
            run += st->level_count[L];
        }
        off[kMaxLevels + 1] = run;
        if (blockIdx.x == 0)
            for (int L = 0; L <= kMaxLevels + 1; ++L) st->level_offset[L] = off[L];
    }
    if (tid <= kMaxLevels) hist[tid] = 0;
    __syncthreads();
    const int C = st->err == 0 ? st->ncells : 0;
    const int t = blockIdx.x * 256 + tid;
    int L = -1, local = 0, rc = 0;
    if (t < C) {
        L = cell_lvl[t];
        rc = cell_rec[t];
        local = atomicAdd(&hist[L], 1);
    }
    __syncthreads();
    if (tid <= kMaxLevels && hist[tid] != 0) base[tid] = atomicAdd(&st->level_cursor[tid], hist[tid]);
    __syncthreads();
    if (L >= 0) level_list[off[L] + base[L] + local] = rc;
}

// Summaries of all cells at one level; children are the records between rc+1 and
// rc+skip, hopped by their own skips, in quadrant order.
__global__ __launch_bounds__(256) void k_summarize(int L, const int* __restrict__ level_list,
                                                   float4* __restrict__ rec, const DevState* __restrict__ st) {
    if (st->err != 0) return;
    const int beg = st->level_offset[L];
    const int cnt = st->level_count[L];
    for (int t = blockIdx.x * 256 + threadIdx.x; t < cnt; t += gridDim.x * 256) {
        const int rc = level_list[beg + t];
        const unsigned int info = __float_as_uint(rec[rc].w);
        const int end = min(rc + (int)(info >> kInfoLevelBits), st->nrecords);
        float cm = 0.0f, px = 0.0f, py = 0.0f;
        int r = rc + 1;
        while (r < end) {
            const float4 c = rec[r];
            const unsigned int ci = __float_as_uint(c.w);
            px = fmaf(c.x, c.z, px);
            py = fmaf(c.y, c.z, py);
            cm += c.z;
            r += ci == 0xffffffffu ? 1 : max((int)(ci >> kInfoLevelBits), 1);
        }
        const float inv = 1.0f / cm;
        rec[rc].x = px * inv;
        rec[rc].y = py * inv;
        rec[rc].z = cm;
    }
}

// Heavy cells (mass >= 2^24, listed by k_nodes) of the integer-mass path, in one
// workgroup: sorted by record index in LDS, then summarised deepest level first, children
// in quadrant order (mass from 0.0f, centroid (sum m*p)*(1/mass), as
// src/RPBHKernels.cu:468-494). A heavy child's summary is read from LDS; light children
// and bodies were written by earlier launches.
__global__ __launch_bounds__(256) void k_big(float4* __restrict__ rec, const int* __restrict__ big_rc,
                                             const unsigned char* __restrict__ big_lvl,
                                             const DevState* __restrict__ st) {
    __shared__ int brc[kMaxBig];
    __shared__ int blv[kMaxBig];
    __shared__ float4 res[kMaxBig];
    __shared__ int maxl;
    if (st->err != 0) return;
    const int nb = min(st->nbig, kMaxBig);
    if (nb == 0) return;
    const int tid = threadIdx.x;
    int P = 1;
    while (P < nb) P <<= 1;
    if (tid == 0) maxl = 0;
    for (int t = tid; t < P; t += 256) {
        brc[t] = t < nb ? big_rc[t] : 0x7fffffff;
        blv[t] = t < nb ? (int)big_lvl[t] : -1;
    }
    __syncthreads();
    // bitonic sort by record index
    for (int k = 2; k <= P; k <<= 1)
        for (int j = k >> 1; j > 0; j >>= 1) {
            for (int t = tid; t < P; t += 256) {
                const int u = t ^ j;
                if (u > t) {
                    const bool up = (t & k) == 0;
                    if ((brc[t] > brc[u]) == up) {
                        const int a = brc[t], b = blv[t];
                        brc[t] = brc[u];
                        blv[t] = blv[u];
                        brc[u] = a;
                        blv[u] = b;
                    }
                }
            }
            __syncthreads();
        }
    for (int t = tid; t < nb; t += 256) atomicMax(&maxl, blv[t]);
    __syncthreads();
    for (int L = maxl; L >= 0; --L) {
        for (int t = tid; t < nb; t += 256) {
            if (blv[t] != L) continue;
            const int rc = brc[t];
            const int end = min(rc + (int)(__float_as_uint(rec[rc].w) >> kInfoLevelBits), st->nrecords);
            float cm = 0.0f, px = 0.0f, py = 0.0f;
            int r = rc + 1;
            while (r < end) {
                float4 c = rec[r];
                const unsigned int ci = __float_as_uint(c.w);
                if (ci != 0xffffffffu) {
                    int lo = 0, hi = nb;   // is r heavy?
                    while (hi - lo > 1) {
                        const int mid = (lo + hi) >> 1;
                        if (brc[mid] <= r) lo = mid;
                        else hi = mid;
                    }
                    if (brc[lo] == r) c = res[lo];
                }
                px = fmaf(c.x, c.z, px);
                py = fmaf(c.y, c.z, py);
                cm += c.z;
                r += ci == 0xffffffffu ? 1 : max((int)(ci >> kInfoLevelBits), 1);
            }
            const float inv = 1.0f / cm;
            res[t] = make_float4(px * inv, py * inv, cm, 0.0f);
        }
        __syncthreads();
    }
    for (int t = tid; t < nb; t += 256) {
        const float4 v = res[t];
        rec[brc[t]].x = v.x;
        rec[brc[t]].y = v.y;
        rec[brc[t]].z = v.z;
    }
}

}  // namespace

int displacement_blocks(const Bufs& b) {
    int cap = b.grid_cap < kPartBlocks ? b.grid_cap : kPartBlocks;
    int g = (b.n + 255) / 256;
    return g < 1 ? 1 : (g > cap ? cap : g);
}

void launch_bbox_partials(const Bufs& b, hipStream_t s) {
    k_bbox_partials<<<displacement_blocks(b), 256, 0, s>>>(b.pos, b.n, b.bbox_part, b.st);
}

void launch_keys(const Bufs& b, int first, hipStream_t s) {
    int blocks = (b.n + 255) / 256;
    if (blocks > b.grid_cap) blocks = b.grid_cap;
    k_keys<<<blocks, 256, 0, s>>>(b.pos, b.mass, b.n, b.bbox_part, b.st, b.seg_root, b.sort.keys[0], b.sort.vals[0],
                                  b.sort.keys[1], b.sort.vals[1], b.pm, first,
                                  b.sort.hist_in_keys ? b.sort.digit_hist : nullptr);
}

void launch_tree(Bufs& b, hipStream_t s) {
    const int n = b.n;
    const int tiles = (n + kTreeTile - 1) / kTreeTile;
    const int* sg = b.nseg > 1 ? b.sg : nullptr;
    Moment* bm = b.int_mass ? b.bmass : nullptr;
    Moment* pm = b.int_mass ? b.pmass : nullptr;
    k_lcp<<<tiles, kTreeThreads, 0, s>>>(b.sort.keys[0], sg, n, b.lcp, b.blocksum, b.sort.vals[0], b.pm,
                                         b.seg_root, bm, b.sbody);
    // up to kFuseTiles tiles, k_prefix sums the tile totals itself
    const bool fused = tiles <= kFuseTiles;
    if (tiles > 8 * kTileChunk) {
        const int chunks = (tiles + kTileChunk - 1) / kTileChunk;
        Moment* cm = b.int_mass ? b.chunk_mass : nullptr;
        k_scan_up<<<chunks, kTileChunk, 0, s>>>(b.blocksum, bm, tiles, b.chunk_sum, cm);
        k_scan_blocks<<<1, 1024, 0, s>>>(b.chunk_sum, chunks, n, b.cap_records, b.rec_begin + b.nseg, b.st, cm);
        k_scan_down<<<chunks, kTileChunk, 0, s>>>(b.blocksum, bm, tiles, b.chunk_sum, cm);
    } else if (!fused) {
        k_scan_blocks<<<1, 1024, 0, s>>>(b.blocksum, tiles, n, b.cap_records, b.rec_begin + b.nseg, b.st, bm);
    }
    const int max_cells = b.cap_records - n;
    k_prefix<<<tiles, kTreeThreads, 0, s>>>(b.lcp, sg, n, b.blocksum, b.S, b.sbody, b.rec, b.brec, b.st, b.seg_root,
                                            bm, pm, b.cell_rec, max_cells, fused ? 1 : 0, b.cap_records,
                                            b.rec_begin + b.nseg);
    // grid-stride over the cells (their number is known on the device only; at most
    // max_cells, typically below n)
    const int node_blocks = (n + kTreeThreads - 1) / kTreeThreads;
    k_nodes<<<node_blocks, kTreeThreads, 0, s>>>(b.sort.keys[0], b.lcp, sg, n, b.S, b.rec, b.rcount, b.cell_rec,
                                                 b.cell_lvl, b.rec_begin, b.st, pm, b.seg_root, b.big_rc,
                                                 b.big_lvl, b.node_off);
    // every record's info word is in place (cells: k_nodes, bodies: k_prefix)
    k_kids<<<b.kids_blocks, 256, 0, s>>>(b.rec, b.rcount, b.S, n, b.cell_rec, b.kids, b.kids_part, b.st);
    if (b.int_mass) {
        if (!b.no_big) k_big<<<1, 256, 0, s>>>(b.rec, b.big_rc, b.big_lvl, b.st);
        return;
    }
    k_levels<<<(max_cells + 255) / 256, 256, 0, s>>>(b.cell_rec, b.cell_lvl, b.level_list, b.st);
    int sum_blocks = (max_cells / 4 + 255) / 256;
    if (sum_blocks < 1) sum_blocks = 1;
    if (sum_blocks > 1024) sum_blocks = 1024;
    for (int L = kKeyLevels; L >= 0; --L)
        k_summarize<<<sum_blocks, 256, 0, s>>>(L, b.level_list, b.rec, b.st);
}

}  // namespace gfa2
