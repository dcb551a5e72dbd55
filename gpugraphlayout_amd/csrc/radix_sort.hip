// LSD radix sort of (u64 key, u32 value) pairs for gfx950.
//
// Each 8-bit pass is three launches:
//   upsweep   per-tile digit histograms (LDS atomics), stored digit-major
//   rowscan   one wave per digit: exclusive prefix of that digit over tiles + total
//   scatter   per tile: digit start = scan of the 256 totals (done redundantly in
//             LDS), stable rank within the tile from wave64 ballot matching, then the
//             scatter. Keys of one digit from one tile land contiguously.
// Every cross-tile dependency goes through a launch boundary; no inter-workgroup
// waiting, no global atomics. Stable, hence LSD-correct.
#include "engine_internal.h"

namespace gfa2 {
namespace {

constexpr int kWaves = kSortThreads / 64;

__global__ __launch_bounds__(kSortThreads) void k_upsweep(const unsigned long long* __restrict__ keys,
                                                          int n, int shift, unsigned int* __restrict__ hist,
                                                          int tiles) {
    __shared__ unsigned int h[256];
    const int tid = threadIdx.x;
    h[tid] = 0;
    __syncthreads();
    const int base = blockIdx.x * kSortTile;
#pragma unroll
    for (int r = 0; r < kSortItems; ++r) {
        int e = base + r * kSortThreads + tid;
        if (e < n) atomicAdd(&h[(unsigned)(keys[e] >> shift) & 255u], 1u);
    }
    __syncthreads();
    hist[tid * tiles + blockIdx.x] = h[tid];
}

// One wave per digit row: exclusive scan over tiles, row total into rowtot.
__global__ __launch_bounds__(256) void k_rowscan(unsigned int* __restrict__ hist, int tiles,
                                                 unsigned int* __restrict__ rowtot) {
    const int lane = threadIdx.x & 63;
    const int row = blockIdx.x * 4 + (threadIdx.x >> 6);
    if (row >= 256) return;
    unsigned int* h = hist + (size_t)row * tiles;
    const int per = (tiles + 63) / 64;
    const int b0 = lane * per;
    unsigned int sum = 0;
    for (int k = 0; k < per; ++k) {
        int t = b0 + k;
        if (t < tiles) sum += h[t];
    }
    // inclusive wave scan of per-lane sums
    unsigned int inc = sum;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        unsigned int v = __shfl_up(inc, o, 64);
        if (lane >= o) inc += v;
    }
    unsigned int run = inc - sum;
    for (int k = 0; k < per; ++k) {
        int t = b0 + k;
        if (t < tiles) {
            unsigned int c = h[t];
            h[t] = run;
            run += c;
        }
    }
    if (lane == 63) rowtot[row] = inc;
}

__global__ __launch_bounds__(kSortThreads) void k_scatter(const unsigned long long* __restrict__ kin,
                                                          const unsigned int* __restrict__ vin,
                                                          unsigned long long* __restrict__ kout,
                                                          unsigned int* __restrict__ vout, int n, int shift,
                                                          const unsigned int* __restrict__ hist, int tiles,
                                                          const unsigned int* __restrict__ rowtot) {
    __shared__ unsigned int dstart[256];   // global start of each digit for this tile
    __shared__ unsigned int wcnt[kWaves][256];
    __shared__ unsigned int scan_tmp[256];
    const int tid = threadIdx.x;
    const int lane = tid & 63;
    const int wave = tid >> 6;

    // Exclusive scan of digit totals (256 entries, one per thread), Hillis-Steele in LDS.
    unsigned int tot = rowtot[tid];
    scan_tmp[tid] = tot;
    __syncthreads();
    for (int o = 1; o < 256; o <<= 1) {
        unsigned int v = tid >= o ? scan_tmp[tid - o] : 0u;
        __syncthreads();
        scan_tmp[tid] += v;
        __syncthreads();
    }
    dstart[tid] = scan_tmp[tid] - tot + hist[(size_t)tid * tiles + blockIdx.x];
#pragma unroll
    for (int w = 0; w < kWaves; ++w) wcnt[w][tid] = 0;
    __syncthreads();

    const unsigned long long lt = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
    const int base = blockIdx.x * kSortTile;
    for (int r = 0; r < kSortItems; ++r) {
        const int e = base + r * kSortThreads + tid;
        const bool valid = e < n;
        unsigned long long key = valid ? kin[e] : 0ull;
        unsigned int val = valid ? vin[e] : 0u;
        unsigned int d = (unsigned)(key >> shift) & 255u;
        unsigned long long peers = __ballot(valid);
#pragma unroll
        for (int b = 0; b < 8; ++b) {
            unsigned long long m = __ballot((d >> b) & 1u);
            peers &= ((d >> b) & 1u) ? m : ~m;
        }
        const unsigned int rank = (unsigned)__popcll(peers & lt);
        if (valid && rank == 0) wcnt[wave][d] = (unsigned)__popcll(peers);
        __syncthreads();
        // per digit: offsets of each wave's run, then advance the digit's start
        {
            unsigned int run = dstart[tid];
#pragma unroll
            for (int w = 0; w < kWaves; ++w) {
                unsigned int c = wcnt[w][tid];
                wcnt[w][tid] = run;
                run += c;
            }
            dstart[tid] = run;
        }
        __syncthreads();
        if (valid) {
            unsigned int dst = wcnt[wave][d] + rank;
            kout[dst] = key;
            vout[dst] = val;
        }
        __syncthreads();
#pragma unroll
        for (int w = 0; w < kWaves; ++w) wcnt[w][tid] = 0;
        __syncthreads();
    }
}

}  // namespace

void radix_sort_pairs(SortBuffers& buf, int n, int begin_bit, int end_bit, hipStream_t s) {
    if (n <= 1) return;
    const int tiles = (n + kSortTile - 1) / kSortTile;
    int cur = 0;
    for (int shift = begin_bit; shift < end_bit; shift += 8) {
        k_upsweep<<<tiles, kSortThreads, 0, s>>>(buf.keys[cur], n, shift, buf.hist, tiles);
        k_rowscan<<<64, 256, 0, s>>>(buf.hist, tiles, buf.rowtot);
        k_scatter<<<tiles, kSortThreads, 0, s>>>(buf.keys[cur], buf.vals[cur], buf.keys[cur ^ 1],
                                                 buf.vals[cur ^ 1], n, shift, buf.hist, tiles, buf.rowtot);
        cur ^= 1;
    }
    if (cur != 0) {
        hipMemcpyAsync(buf.keys[0], buf.keys[1], sizeof(unsigned long long) * n, hipMemcpyDeviceToDevice, s);
        hipMemcpyAsync(buf.vals[0], buf.vals[1], sizeof(unsigned int) * n, hipMemcpyDeviceToDevice, s);
    }
}

}  // namespace gfa2
