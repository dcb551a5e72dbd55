// LSD radix sort of (u64 key, u32 value) pairs for gfx950.
//
// Up to 64 tiles (131,072 pairs): one launch per 8-bit pass ("onesweep"), after one histogram
// launch for all passes.
//   k_sort_hist  digit histograms of every pass at once (order-independent), LDS then global
//                atomics; the table is cleared by k_sort_zero first
//   k_onesweep   per tile (2048 pairs): tiles are taken in the order of a global ticket, the
//                tile's digit counts are published, the exclusive prefix of each digit over the
//                earlier tiles is found by looking back over their published counts (decoupled
//                look-back), then the stable rank within the tile (wave64 ballot matching) gives
//                each pair's destination. A tile only ever waits for tiles with smaller tickets,
//                which started before it and publish their counts before they wait themselves,
//                so the waits end; every wait is also bounded and raises an error flag.
// Larger inputs (or SortBuffers::onesweep = false): three launches per pass (per-tile histograms,
// a scan over tiles per digit, scatter), with every cross-tile dependency through a launch
// boundary.
// Both are stable (hence LSD-correct) and give the same result.
#include "engine_internal.h"

namespace gfa2 {
namespace {

constexpr int kWaves = kSortThreads / 64;

// A pass this iteration's sort skips (see radix_sort_pairs).
__device__ __forceinline__ bool pass_skipped(const DevState* __restrict__ st, int pass) {
    return st != nullptr && pass < st->sort_skip;
}

// The low passes that do not run: those the device skips and those the host did not launch.
__device__ __forceinline__ int passes_below(const DevState* __restrict__ st, int first) {
    return max(st != nullptr ? st->sort_skip : 0, first);
}

// The buffer the first pass that runs reads: buffer 0, or with `lead` (the total number of
// passes, when the keys were written where an odd count of running passes needs them) the one
// that makes the last pass end in buffer 0.
__device__ __forceinline__ int start_buffer(const DevState* __restrict__ st, int first, int lead) {
    return lead ? ((lead - passes_below(st, first)) & 1) : 0;
}

// The buffer a pass reads (it writes the other one): every pass that runs flips it (skipped
// passes move nothing).
__device__ __forceinline__ int pass_src(const DevState* __restrict__ st, int first, int lead, int pass) {
    return (start_buffer(st, first, lead) + pass - passes_below(st, first)) & 1;
}

__global__ __launch_bounds__(kSortThreads) void k_upsweep(const unsigned long long* __restrict__ k0,
                                                          const unsigned long long* __restrict__ k1,
                                                          int n, int shift, unsigned int* __restrict__ hist,
                                                          int tiles, const DevState* __restrict__ st, int first,
                                                          int lead, int pass) {
    if (pass_skipped(st, pass)) return;
    const unsigned long long* __restrict__ keys = pass_src(st, first, lead, pass) ? k1 : k0;
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

// One block per digit row: exclusive scan over the tiles, row total into rowtot. The row is
// read in chunks of 256 tiles (one per thread, coalesced), up to kRowChunks chunks requested
// before any is scanned; each chunk is one block scan (shuffles within the waves, then the
// four wave totals) carried into the next.
constexpr int kRowChunks = 16;

__global__ __launch_bounds__(256) void k_rowscan(unsigned int* __restrict__ hist, int tiles,
                                                 unsigned int* __restrict__ rowtot, const DevState* __restrict__ st,
                                                 int pass) {
    if (pass_skipped(st, pass)) return;
    __shared__ unsigned int sh[2][4];
    const int tid = threadIdx.x;
    const int lane = tid & 63;
    const int wave = tid >> 6;
    unsigned int* h = hist + (size_t)blockIdx.x * tiles;
    unsigned int carry = 0u;
    int parity = 0;
    for (int c0 = 0; c0 < tiles; c0 += 256 * kRowChunks) {
        unsigned int v[kRowChunks];
#pragma unroll
        for (int c = 0; c < kRowChunks; ++c) {
            const int t = c0 + c * 256 + tid;
            v[c] = t < tiles ? h[t] : 0u;
        }
#pragma unroll
        for (int c = 0; c < kRowChunks; ++c) {
            if (c0 + c * 256 >= tiles) break;            // uniform
            unsigned int inc = v[c];
#pragma unroll
            for (int o = 1; o < 64; o <<= 1) {
                const unsigned int t = __shfl_up(inc, o, 64);
                if (lane >= o) inc += t;
            }
            if (lane == 63) sh[parity][wave] = inc;
            __syncthreads();
            unsigned int off = 0u, total = 0u;
#pragma unroll
            for (int w = 0; w < 4; ++w) {
                const unsigned int x = sh[parity][w];
                if (w < wave) off += x;
                total += x;
            }
            const int t = c0 + c * 256 + tid;
            if (t < tiles) h[t] = carry + off + inc - v[c];
            carry += total;
            parity ^= 1;
        }
    }
    if (tid == 0) rowtot[blockIdx.x] = carry;
}

// Stable ranks within a tile, three block barriers per tile: each wave owns a contiguous
// quarter of the tile (kSortItems rows of 64 pairs), ranks its pairs row by row with
// a wave64 ballot match against its own running digit counts in LDS, then one pass over the
// digits turns the per-wave counts into destinations (wave order is tile order, so the sort
// stays stable).
__device__ __forceinline__ void rank_rows(const unsigned long long (&key)[kSortItems], int wbase, int n, int shift,
                                          unsigned int* __restrict__ wcnt, unsigned int (&rank)[kSortItems]) {
    const int lane = threadIdx.x & 63;
    const unsigned long long lt = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
#pragma unroll
    for (int r = 0; r < kSortItems; ++r) {
        const bool valid = wbase + r * 64 + lane < n;
        const unsigned int d = (unsigned)(key[r] >> shift) & 255u;
        unsigned long long peers = __ballot(valid);
#pragma unroll
        for (int b = 0; b < 8; ++b) {
            const unsigned long long m = __ballot((d >> b) & 1u);
            peers &= ((d >> b) & 1u) ? m : ~m;
        }
        const unsigned int rr = (unsigned)__popcll(peers & lt);
        // every lane reads the digit's count before its leader writes it (one wave, LDS in order)
        const unsigned int before = valid ? wcnt[d] : 0u;
        rank[r] = before + rr;
        if (valid && rr == 0) wcnt[d] = before + (unsigned)__popcll(peers);
    }
}

// Exclusive scan of one value per thread over the block's 256 threads: shuffle scans within
// the waves, then the four wave totals (one barrier). sh: 4 words of LDS, not reused before
// another barrier.
__device__ __forceinline__ unsigned int block_exscan256(unsigned int v, unsigned int* __restrict__ sh) {
    const int lane = threadIdx.x & 63;
    const int wave = threadIdx.x >> 6;
    unsigned int inc = v;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        const unsigned int t = __shfl_up(inc, o, 64);
        if (lane >= o) inc += t;
    }
    if (lane == 63) sh[wave] = inc;
    __syncthreads();
    unsigned int off = 0u;
#pragma unroll
    for (int w = 0; w < kWaves; ++w)
        if (w < wave) off += sh[w];
    return off + inc - v;
}

// LDS of the tile step: per-wave digit counts, the staged tile, scan scratch.
struct TileLds {
    unsigned int wcnt[kWaves][256];
    unsigned long long skey[kSortTile];
    unsigned int sval[kSortTile];
    int delta[256];                 // global destination minus tile position, per digit
    unsigned int sh[2][kWaves];
};

// Writes one tile's pairs (held in registers, each wave a contiguous quarter) to their sorted
// places. dglobal: this thread's digit's first destination for the tile. The pairs are ranked
// stably (rank_rows), placed in LDS in digit order, then written out in tile order, so the
// pairs of a digit leave as one run of consecutive addresses.
__device__ __forceinline__ unsigned int scatter_tile(const unsigned long long (&key)[kSortItems],
                                             const unsigned int (&val)[kSortItems], int tile, int n, int shift,
                                             unsigned int dglobal, TileLds& L,
                                             unsigned long long* __restrict__ kout, unsigned int* __restrict__ vout) {
    const int tid = threadIdx.x;
    const int lane = tid & 63;
    const int wave = tid >> 6;
    const int wbase = tile * kSortTile + wave * (kSortItems * 64);
    unsigned int rank[kSortItems];
    rank_rows(key, wbase, n, shift, L.wcnt[wave], rank);
    __syncthreads();
    // per digit (one per thread): the tile's count, its start within the tile, each wave's start
    unsigned int c[kWaves], tcount = 0u;
#pragma unroll
    for (int w = 0; w < kWaves; ++w) {
        c[w] = L.wcnt[w][tid];
        tcount += c[w];
    }
    const unsigned int tstart = block_exscan256(tcount, L.sh[1]);
    {
        unsigned int run = tstart;
#pragma unroll
        for (int w = 0; w < kWaves; ++w) {
            L.wcnt[w][tid] = run;
            run += c[w];
        }
    }
    L.delta[tid] = (int)dglobal - (int)tstart;
    __syncthreads();
#pragma unroll
    for (int r = 0; r < kSortItems; ++r) {
        if (wbase + r * 64 + lane < n) {
            const unsigned int at = L.wcnt[wave][(unsigned)(key[r] >> shift) & 255u] + rank[r];
            L.skey[at] = key[r];
            L.sval[at] = val[r];
        }
    }
    __syncthreads();
    const int count = min(kSortTile, n - tile * kSortTile);
    for (int j = tid; j < count; j += kSortThreads) {
        const unsigned long long k = L.skey[j];
        const unsigned int dst = (unsigned int)(L.delta[(unsigned)(k >> shift) & 255u] + j);
        kout[dst] = k;
        vout[dst] = L.sval[j];
    }
    return tcount;
}

__global__ __launch_bounds__(kSortThreads) void k_scatter(unsigned long long* __restrict__ k0,
                                                          unsigned int* __restrict__ v0,
                                                          unsigned long long* __restrict__ k1,
                                                          unsigned int* __restrict__ v1, int n, int shift,
                                                          const unsigned int* __restrict__ hist, int tiles,
                                                          const unsigned int* __restrict__ rowtot,
                                                          const DevState* __restrict__ st, int first, int lead,
                                                          int pass) {
    if (pass_skipped(st, pass)) return;
    const int src = pass_src(st, first, lead, pass);
    const unsigned long long* __restrict__ kin = src ? k1 : k0;
    const unsigned int* __restrict__ vin = src ? v1 : v0;
    unsigned long long* __restrict__ kout = src ? k0 : k1;
    unsigned int* __restrict__ vout = src ? v0 : v1;
    __shared__ TileLds L;
    const int tid = threadIdx.x;
    const int lane = tid & 63;
    const int wave = tid >> 6;
    const int wbase = blockIdx.x * kSortTile + wave * (kSortItems * 64);
    unsigned long long key[kSortItems];
    unsigned int val[kSortItems];
#pragma unroll
    for (int r = 0; r < kSortItems; ++r) {
        const int e = wbase + r * 64 + lane;
        key[r] = e < n ? kin[e] : 0ull;
        val[r] = e < n ? vin[e] : 0u;
    }
#pragma unroll
    for (int w = 0; w < kWaves; ++w) L.wcnt[w][tid] = 0;
    // digit starts: exclusive scan of the digit totals, plus this tile's offset within the digit
    const unsigned int dglobal = block_exscan256(rowtot[tid], L.sh[0]) + hist[(size_t)tid * tiles + blockIdx.x];
    scatter_tile(key, val, blockIdx.x, n, shift, dglobal, L, kout, vout);
}

// ---- one block per segment ----------------------------------------------------------------
// Segment g is [node_off[g], node_off[g+1]) of both buffers; its pairs are sorted in place by
// the full key, stable, by one block: the digit histograms of all passes first, then per pass
// the segment's tiles in order, each ranked stably and written to its sorted place in the
// other buffer (the running digit starts carry from tile to tile). Eight passes end in buffer
// 0. Blocks take segments in `order` (largest first), so the longest ones start first.
__global__ __launch_bounds__(kSortThreads) void k_seg_sort(unsigned long long* __restrict__ k0,
                                                           unsigned int* __restrict__ v0,
                                                           unsigned long long* __restrict__ k1,
                                                           unsigned int* __restrict__ v1,
                                                           const int* __restrict__ node_off,
                                                           const int* __restrict__ order,
                                                           const DevState* __restrict__ st) {
    __shared__ TileLds L;
    const int tid = threadIdx.x;
    const int lane = tid & 63;
    const int wave = tid >> 6;
    const int g = order[blockIdx.x];
    // the low passes the previous trees' cells let the sort skip (k_sort_fixup, told the graph
    // of each position, orders what they would have ordered)
    const int p0 = st ? st->sort_skip : 0;
    const int a = node_off[g];
    const int m = node_off[g + 1] - a;
    // the 8 digit histograms (8 x 256 words), counted in the not yet used tile staging area when
    // it holds them, then kept per thread for its own digit: hc[0] is the current pass's count,
    // shifted down after each pass
    constexpr bool kAlias = kSortTile >= 8 * 256;
    __shared__ unsigned int hist_own[kAlias ? 1 : 8 * 256];
    unsigned int* hist = kAlias ? L.sval : hist_own;
#pragma unroll
    for (int p = 0; p < 8; ++p) hist[p * 256 + tid] = 0u;
    __syncthreads();
    for (int e = tid; e < m; e += kSortThreads) {
        const unsigned long long k = k0[a + e];
#pragma unroll
        for (int p = 0; p < 8; ++p) atomicAdd(&hist[p * 256 + ((unsigned)(k >> (8 * p)) & 255u)], 1u);
    }
    __syncthreads();
    unsigned int hc[8];
#pragma unroll
    for (int p = 0; p < 8; ++p) hc[p] = hist[p * 256 + tid];
    __syncthreads();                             // the staging area is the tiles' again
    for (int p = 0; p < p0; ++p) {
#pragma unroll
        for (int q = 0; q < 7; ++q) hc[q] = hc[q + 1];
    }
    const int tiles = (m + kSortTile - 1) / kSortTile;
    for (int p = p0; p < 8; ++p) {
        const bool from1 = ((p - p0) & 1) != 0;
        const unsigned long long* kin = (from1 ? k1 : k0) + a;
        const unsigned int* vin = (from1 ? v1 : v0) + a;
        unsigned long long* kout = (from1 ? k0 : k1) + a;
        unsigned int* vout = (from1 ? v0 : v1) + a;
        // this thread's digit: its first destination in the segment, advanced tile by tile
        unsigned int dstart = block_exscan256(hc[0], L.sh[0]);
#pragma unroll
        for (int q = 0; q < 7; ++q) hc[q] = hc[q + 1];
        for (int t = 0; t < tiles; ++t) {
            const int wbase = t * kSortTile + wave * (kSortItems * 64);
            unsigned long long key[kSortItems];
            unsigned int val[kSortItems];
#pragma unroll
            for (int r = 0; r < kSortItems; ++r) {
                const int e = wbase + r * 64 + lane;
                key[r] = e < m ? kin[e] : 0ull;
                val[r] = e < m ? vin[e] : 0u;
            }
#pragma unroll
            for (int w = 0; w < kWaves; ++w) L.wcnt[w][tid] = 0;
            // the counts are cleared (and the previous tile's LDS reads done) before ranking
            __syncthreads();
            dstart += scatter_tile(key, val, t, m, 8 * p, dstart, L, kout, vout);
        }
        // this pass's writes are read by the next one (one block: workgroup-scope visibility)
        __syncthreads();
    }
    if (((8 - p0) & 1) != 0) {
        // an odd number of passes ran: the range is sorted in the second buffer
        for (int e = tid; e < m; e += kSortThreads) {
            k0[a + e] = k1[a + e];
            v0[a + e] = v1[a + e];
        }
    }
}

// ---- onesweep -------------------------------------------------------------------------------
constexpr unsigned int kStatAgg = 1u << 30;      // the tile's own digit count is published
constexpr unsigned int kStatPre = 2u << 30;      // the inclusive prefix over tiles is published
constexpr unsigned int kStatCount = (1u << 27) - 1;
constexpr int kOnesweepMaxTiles = 64;
constexpr int kMaxSpins = 1 << 20;   // per thread and look-back: ~1 s at worst, then an error

__global__ void k_sort_zero(unsigned int* __restrict__ digit_hist, int npasses) {
    for (int i = threadIdx.x; i < npasses * 256; i += blockDim.x) digit_hist[i] = 0u;
}

__global__ __launch_bounds__(256) void k_sort_hist(const unsigned long long* __restrict__ k0,
                                                   const unsigned long long* __restrict__ k1, int n, int begin_bit,
                                                   int npasses, unsigned int* __restrict__ digit_hist,
                                                   const DevState* __restrict__ st, int first, int lead) {
    const unsigned long long* __restrict__ keys = start_buffer(st, first, lead) ? k1 : k0;
    __shared__ unsigned int h[8][256];
    for (int p = 0; p < npasses; ++p) h[p][threadIdx.x] = 0u;
    __syncthreads();
    for (int e = blockIdx.x * 256 + threadIdx.x; e < n; e += gridDim.x * 256) {
        const unsigned long long k = keys[e];
        for (int p = 0; p < npasses; ++p) atomicAdd(&h[p][(unsigned)(k >> (begin_bit + 8 * p)) & 255u], 1u);
    }
    __syncthreads();
    for (int p = 0; p < npasses; ++p) {
        const unsigned int c = h[p][threadIdx.x];
        if (c) atomicAdd(&digit_hist[p * 256 + threadIdx.x], c);
    }
}

__global__ __launch_bounds__(kSortThreads) void k_onesweep(unsigned long long* __restrict__ k0,
                                                           unsigned int* __restrict__ v0,
                                                           unsigned long long* __restrict__ k1,
                                                           unsigned int* __restrict__ v1, int n, int shift,
                                                           const unsigned int* __restrict__ digit_hist,
                                                           unsigned int* __restrict__ status,
                                                           unsigned long long* __restrict__ ticket, int tiles,
                                                           int* __restrict__ err, const DevState* __restrict__ st,
                                                           int first, int lead, int pass) {
    if (pass_skipped(st, pass)) return;   // takes no ticket: the epochs of later passes stay distinct
    const int src = pass_src(st, first, lead, pass);
    const unsigned long long* __restrict__ kin = src ? k1 : k0;
    const unsigned int* __restrict__ vin = src ? v1 : v0;
    unsigned long long* __restrict__ kout = src ? k0 : k1;
    unsigned int* __restrict__ vout = src ? v0 : v1;
    __shared__ TileLds L;
    __shared__ unsigned int tcount[256];
    __shared__ int s_tile;
    __shared__ unsigned int s_epoch;
    const int tid = threadIdx.x;
    const int lane = tid & 63;
    const int wave = tid >> 6;
    if (tid == 0) {
        const unsigned long long t = atomicAdd(ticket, 1ull);
        s_tile = (int)(t % (unsigned long long)tiles);
        s_epoch = (unsigned int)((t / (unsigned long long)tiles) & 7ull);
    }
    tcount[tid] = 0u;
#pragma unroll
    for (int w = 0; w < kWaves; ++w) L.wcnt[w][tid] = 0;
    // exclusive scan of this pass's digit totals (its barrier also publishes s_tile and the zeros)
    const unsigned int gstart = block_exscan256(digit_hist[tid], L.sh[0]);
    const int tile = s_tile;
    const unsigned int epoch = s_epoch << 27;
    // the tile's pairs, held in registers (each wave a contiguous quarter), and its digit counts
    const int wbase = tile * kSortTile + wave * (kSortItems * 64);
    unsigned long long key[kSortItems];
    unsigned int val[kSortItems];
#pragma unroll
    for (int r = 0; r < kSortItems; ++r) {
        const int e = wbase + r * 64 + lane;
        key[r] = e < n ? kin[e] : 0ull;
        val[r] = e < n ? vin[e] : 0u;
        if (e < n) atomicAdd(&tcount[(unsigned)(key[r] >> shift) & 255u], 1u);
    }
    __syncthreads();
    // publish, look back, publish the inclusive prefix (one digit per thread)
    unsigned int* my = status + (size_t)tile * 256 + tid;
    const unsigned int cnt = tcount[tid];
    unsigned int excl = 0u;
    if (tile == 0) {
        __hip_atomic_store(my, kStatPre | epoch | cnt, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    } else {
        __hip_atomic_store(my, kStatAgg | epoch | cnt, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        int spins = 0;                       // bounded over the whole look-back
        for (int t = tile - 1; t >= 0; --t) {
            const unsigned int* src = status + (size_t)t * 256 + tid;
            unsigned int v = 0u;
            for (;;) {
                v = __hip_atomic_load(src, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                if ((v >> 30) != 0u && (v & (7u << 27)) == epoch) break;
                if (++spins >= kMaxSpins) break;
                __builtin_amdgcn_s_sleep(1);
            }
            if (spins >= kMaxSpins) {
                atomicOr(err, kErrSortWait);
                break;
            }
            excl += v & kStatCount;
            if ((v >> 30) == 2u) break;
        }
        __hip_atomic_store(my, kStatPre | epoch | ((excl + cnt) & kStatCount), __ATOMIC_RELAXED,
                           __HIP_MEMORY_SCOPE_AGENT);
    }
    scatter_tile(key, val, tile, n, shift, gstart + excl, L, kout, vout);
}

// (key, value) pairs in lexicographic order (values are distinct body ids: a total order).
__device__ __forceinline__ bool pair_less(unsigned long long ka, unsigned int va, unsigned long long kb,
                                          unsigned int vb) {
    return ka < kb || (ka == kb && va < vb);
}

// Heapsort of pairs [0, m) of k/v in place: O(m log m) whatever the input order.
__device__ void heapsort_pairs(unsigned long long* __restrict__ k, unsigned int* __restrict__ v, int m) {
    auto sift = [&](int root, int end) {
        while (2 * root + 1 < end) {
            int c = 2 * root + 1;
            if (c + 1 < end && pair_less(k[c], v[c], k[c + 1], v[c + 1])) ++c;
            if (!pair_less(k[root], v[root], k[c], v[c])) return;
            const unsigned long long tk = k[root];
            const unsigned int tv = v[root];
            k[root] = k[c];
            v[root] = v[c];
            k[c] = tk;
            v[c] = tv;
            root = c;
        }
    };
    for (int r = m / 2 - 1; r >= 0; --r) sift(r, m);
    for (int end = m - 1; end > 0; --end) {
        const unsigned long long tk = k[0];
        const unsigned int tv = v[0];
        k[0] = k[end];
        v[0] = v[end];
        k[end] = tk;
        v[end] = tv;
        sift(0, end);
    }
}

// After a sort that skipped its low passes: the passes that ran left the pairs ordered by
// their top 8 (8 - skip) key bits, in buffer (8 - skip) & 1. Every run of pairs whose keys agree
// in those bits is ordered by (whole key, value) by the thread that owns the run's first
// position (insertion sort up to 32 pairs, heapsort beyond), and the result goes to buffer 0
// (pairs outside runs are copied there when the passes ended in buffer 1). Such runs are the
// bodies of one cell at the sorted depth, few and short when the depth was chosen from the
// previous tree.
__global__ __launch_bounds__(256) void k_sort_fixup(unsigned long long* __restrict__ k0,
                                                    unsigned int* __restrict__ v0,
                                                    const unsigned long long* __restrict__ k1,
                                                    const unsigned int* __restrict__ v1, int n,
                                                    const DevState* __restrict__ st, int first, int lead,
                                                    const int* __restrict__ seg,
                                                    unsigned int* __restrict__ clear_hist) {
    // the passes are done with the digit table: cleared for the next iteration's k_keys
    if (clear_hist && blockIdx.x == 0)
        for (int i = threadIdx.x; i < 8 * 256; i += blockDim.x) clear_hist[i] = 0u;
    const int skip = passes_below(st, first);
    if (skip == 0) return;                       // all 8 passes ran: sorted, ties by value, in buffer 0
    // the passes that ran ended in buffer 1 (unless the keys started where they end in buffer 0)
    const bool moved = !lead && ((8 - skip) & 1) != 0;
    const unsigned long long* __restrict__ ks = moved ? k1 : k0;
    const unsigned int* __restrict__ vs = moved ? v1 : v0;
    const unsigned long long top = ~0ull << (8 * skip);
    // seg (sorted position -> graph, each graph sorted in its own range): runs stop at graphs
    for (int i = blockIdx.x * 256 + threadIdx.x; i < n; i += gridDim.x * 256) {
        const unsigned long long ki = ks[i] & top;
        const int gi = seg ? seg[i] : 0;
        if (i > 0 && (ks[i - 1] & top) == ki && (!seg || seg[i - 1] == gi)) continue;   // not a run's first
        int j = i + 1;
        while (j < n && (ks[j] & top) == ki && (!seg || seg[j] == gi)) ++j;
        if (moved)
            for (int a = i; a < j; ++a) {
                k0[a] = ks[a];
                v0[a] = vs[a];
            }
        if (j - i == 1) continue;
        if (j - i > 32) {
            heapsort_pairs(k0 + i, v0 + i, j - i);
            continue;
        }
        for (int a = i + 1; a < j; ++a) {
            const unsigned long long k = k0[a];
            const unsigned int v = v0[a];
            int b = a - 1;
            while (b >= i && pair_less(k, v, k0[b], v0[b])) {
                k0[b + 1] = k0[b];
                v0[b + 1] = v0[b];
                --b;
            }
            k0[b + 1] = k;
            v0[b + 1] = v;
        }
    }
}

}  // namespace

void launch_sort_fixup(SortBuffers& buf, int n, const DevState* st, hipStream_t s, int first, bool lead,
                       const int* seg, bool clear_hist) {
    if (n < 1) return;
    // one position per thread: a thread that finds a run's first position sorts the run with
    // dependent loads, so positions taken one after another by a thread add up their latencies
    int blocks = (n + 255) / 256;
    if (blocks > 65535) blocks = 65535;
    k_sort_fixup<<<blocks, 256, 0, s>>>(buf.keys[0], buf.vals[0], buf.keys[1], buf.vals[1], n, st, first,
                                        lead ? 8 : 0, seg, clear_hist ? buf.digit_hist : nullptr);
}

bool sort_onesweep(const SortBuffers& buf, int n) {
    return buf.onesweep && (n + kSortTile - 1) / kSortTile <= kOnesweepMaxTiles;
}

void segmented_sort_pairs(SortBuffers& buf, const int* node_off, const int* order, int nseg, hipStream_t s,
                          const DevState* skip) {
    if (nseg > 0)
        k_seg_sort<<<nseg, kSortThreads, 0, s>>>(buf.keys[0], buf.vals[0], buf.keys[1], buf.vals[1], node_off, order,
                                                 skip);
}

void radix_sort_pairs(SortBuffers& buf, int n, int begin_bit, int end_bit, hipStream_t s, const DevState* skip,
                      int first, bool lead_in, bool hist_ready) {
    if (n <= 1) return;
    const int tiles = (n + kSortTile - 1) / kSortTile;
    const int npasses = (end_bit - begin_bit + 7) / 8;
    const int lead = lead_in && skip ? npasses : 0;
    // Every pass picks its buffers on the device (pass_src): with a device skip the passes that
    // run may end in buffer 1, and launch_sort_fixup moves the result to buffer 0.
    unsigned long long *k0 = buf.keys[0], *k1 = buf.keys[1];
    unsigned int *v0 = buf.vals[0], *v1 = buf.vals[1];
    // One launch per pass pays off while the look-back chains stay short; with hundreds of
    // tiles resident at once the chains grow with the tile count and three launches per pass win.
    if (buf.onesweep && tiles <= kOnesweepMaxTiles) {
        if (!hist_ready) {
            k_sort_zero<<<1, 256, 0, s>>>(buf.digit_hist, npasses);
            int hb = (n + 255) / 256;
            if (hb > 256) hb = 256;
            k_sort_hist<<<hb, 256, 0, s>>>(k0, k1, n, begin_bit, npasses, buf.digit_hist, skip, first, lead);
        }
        for (int p = first; p < npasses; ++p)
            k_onesweep<<<tiles, kSortThreads, 0, s>>>(k0, v0, k1, v1, n, begin_bit + 8 * p, buf.digit_hist + 256 * p,
                                                      buf.status, buf.ticket, tiles, buf.err, skip, first, lead, p);
    } else {
        for (int p = first; p < npasses; ++p) {
            const int shift = begin_bit + 8 * p;
            k_upsweep<<<tiles, kSortThreads, 0, s>>>(k0, k1, n, shift, buf.hist, tiles, skip, first, lead, p);
            k_rowscan<<<256, 256, 0, s>>>(buf.hist, tiles, buf.rowtot, skip, p);
            k_scatter<<<tiles, kSortThreads, 0, s>>>(k0, v0, k1, v1, n, shift, buf.hist, tiles, buf.rowtot, skip, first,
                                                     lead, p);
        }
    }
    if (!skip && ((npasses - first) & 1)) {
        hipMemcpyAsync(k0, k1, sizeof(unsigned long long) * n, hipMemcpyDeviceToDevice, s);
        hipMemcpyAsync(v0, v1, sizeof(unsigned int) * n, hipMemcpyDeviceToDevice, s);
    }
}

}  // namespace gfa2
