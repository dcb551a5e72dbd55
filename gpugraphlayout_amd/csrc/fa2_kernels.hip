// ForceAtlas2 force, speed and displacement kernels for gfx950 (FP32 VALU throughout).
#include "engine_internal.h"

#include <algorithm>

namespace gfa2 {
namespace {

// ---------------------------------------------------------------------------------------
// Gravity + attraction, one pass over the symmetric CSR.
// Reference: GravityKernel (src/RPFA2Kernels.cu:35-73) and AttractiveForceKernel
// (:75-109). The reference adds each edge's p_t - p_s to F_s and its negation to F_t with
// four float atomics per edge, in no fixed order; here each node gathers
// sum_j (p_j - p_i) over its own row, which holds exactly the same terms, in a fixed
// order and without atomics. Nodes are pre-sorted into degree classes:
//   light (deg <= 32): one thread per node, sequential sum
//   medium (deg <= 2048): one wave per node, lane-strided sums + butterfly reduction
//   heavy: one 256-thread block per node, thread-strided sums + LDS tree reduction
// ---------------------------------------------------------------------------------------
constexpr int kLightMax = 32;

__device__ __forceinline__ float2 gravity(float2 p, float m, int strong, float k_g) {
    float fg;
    if (strong) fg = k_g * m;
    else if (p.x != 0.0f || p.y != 0.0f) fg = k_g * m * rsqrtf(p.x * p.x + p.y * p.y);
    else fg = 0.0f;
    return make_float2(-p.x * fg, -p.y * fg);
}

// One edge's attraction term on node p from neighbour q. Linear: q - p. With linlog (the
// reference's CPU rule, src/RPCPUForceAtlas2.cpp:63-67): (q - p) * log(1 + d) / d, and 0 at
// d = 0 (the reference multiplies the zero vector by FLT_MAX there). With edge weights:
// times w^delta.
template <bool GEN>
__device__ __forceinline__ void attract_term(float2 p, float2 q, int linlog, const float* __restrict__ ew, int e,
                                             float& ax, float& ay) {
    const float dx = q.x - p.x;
    const float dy = q.y - p.y;
    if constexpr (!GEN) {
        ax += dx;
        ay += dy;
    } else {
        float fa = 1.0f;
        if (linlog) {
            const float d = sqrtf(dx * dx + dy * dy);
            fa = d == 0.0f ? 0.0f : logf(1.0f + d) / d;
        }
        if (ew) fa *= ew[e];
        ax += dx * fa;
        ay += dy * fa;
    }
}

// The terms of CSR entries e, e + S, e + 2S, ... below e1, added in that order. Four entries at
// a time: their index and position loads are all issued before the first term is added (the
// loads past e1 read node 0 and are not added), so a row's tail costs no more round trips to
// memory than a full group of four.
// With HUB, the positions of the kHubCache highest-degree nodes come from the block's LDS copy
// (hub), which spares the L2 the most requested lines.
constexpr int kHubCache = 2048;

template <bool GEN, int S, bool HUB>
__device__ __forceinline__ void gather_terms(float2 p, const float2* __restrict__ pos_rank,
                                             const float2* hub, const int* __restrict__ col, int e, int e1,
                                             int linlog, const float* __restrict__ ew, float& ax, float& ay) {
    for (; e < e1; e += 4 * S) {
        int j[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) j[u] = e + u * S < e1 ? col[e + u * S] : 0;
        float2 q[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            if (HUB && j[u] < kHubCache) q[u] = hub[j[u]];
            else q[u] = pos_rank[j[u]];
        }
#pragma unroll
        for (int u = 0; u < 4; ++u)
            if (e + u * S < e1) attract_term<GEN>(p, q[u], linlog, ew, e + u * S, ax, ay);
    }
}

// col_rank[e] = rank[col[e]] (once per graph).
__global__ __launch_bounds__(256) void k_gather_rank(const int* __restrict__ col, const int* __restrict__ rank,
                                                     long long nnz, int* __restrict__ col_rank) {
    for (long long e = blockIdx.x * 256ll + threadIdx.x; e < nnz; e += (long long)gridDim.x * 256)
        col_rank[e] = rank[col[e]];
}

// pos_rank[rank[v]] = pos[v] (every iteration, before the attraction).
__global__ __launch_bounds__(256) void k_rank_pos(const float2* __restrict__ pos, const int* __restrict__ rank, int n,
                                                  float2* __restrict__ pos_rank) {
    const int v = blockIdx.x * 256 + threadIdx.x;
    if (v < n) pos_rank[rank[v]] = pos[v];
}

// Attraction + gravity, fga[v] = gravity(v) + sum over CSR entries of v (in entry order for the
// light rows; lane-strided then a shuffle tree for medium rows; see below for heavy rows).
// Degree classes: light rows (<= 32 entries) one thread each, entries four at a time;
// medium rows (<= 2,048) one wave each; heavy rows in chunks of kAttrChunk entries, one block
// per chunk writing its partial (thread-strided, then an LDS tree), added per row in chunk order
// by k_attract_heavy. Neighbour positions come from the degree-ranked copy. The work is a list
// of virtual blocks (light, medium, chunks) that the grid strides over: one launch block per
// virtual block by default, or a smaller grid that leaves room on every compute unit for
// kernels running beside it on another stream. Every row's sum is the same whatever the grid.
template <bool GEN, bool HUB>
__global__ __launch_bounds__(256) void k_attract(const float2* __restrict__ pos, const float* __restrict__ mass,
                                                 int n, const float2* __restrict__ pos_rank,
                                                 const int* __restrict__ rowptr, const int* __restrict__ col,
                                                 const int* __restrict__ lists, int nlight, int nmedium,
                                                 const int2* __restrict__ chunks, int nchunks,
                                                 int light_blocks, int medium_blocks,
                                                 int strong, float k_g, float2* __restrict__ fga,
                                                 float2* __restrict__ part, int linlog,
                                                 const float* __restrict__ ew) {
    __shared__ float2 red[256];
    __shared__ float2 hub[HUB ? kHubCache : 1];
    const int tid = threadIdx.x;
    if constexpr (HUB) {
        for (int i = tid; i < kHubCache && i < n; i += 256) hub[i] = pos_rank[i];
        __syncthreads();
    }
    const int total = light_blocks + medium_blocks + nchunks;
    for (int vb = blockIdx.x; vb < total; vb += gridDim.x) {   // uniform per block
        int blk = vb;
        if (blk < light_blocks) {
            const int k = blk * 256 + tid;
            if (k < nlight) {
                const int v = lists[k];
                const float2 p = pos[v];
                float ax = 0.0f, ay = 0.0f;
                const int e1 = rowptr[v + 1];
                gather_terms<GEN, 1, HUB>(p, pos_rank, hub, col, rowptr[v], e1, linlog, ew, ax, ay);
                const float2 g = gravity(p, mass[v], strong, k_g);
                fga[v] = make_float2(g.x + ax, g.y + ay);
            }
            continue;
        }
        blk -= light_blocks;
        if (blk < medium_blocks) {
            const int k = blk * 4 + (tid >> 6);
            if (k < nmedium) {    // uniform per wave
                const int lane = tid & 63;
                const int v = lists[nlight + k];
                const float2 p = pos[v];
                float ax = 0.0f, ay = 0.0f;
                const int e1 = rowptr[v + 1];
                gather_terms<GEN, 64, HUB>(p, pos_rank, hub, col, rowptr[v] + lane, e1, linlog, ew, ax, ay);
#pragma unroll
                for (int o = 32; o > 0; o >>= 1) {
                    ax += __shfl_xor(ax, o, 64);
                    ay += __shfl_xor(ay, o, 64);
                }
                if (lane == 0) {
                    const float2 g = gravity(p, mass[v], strong, k_g);
                    fga[v] = make_float2(g.x + ax, g.y + ay);
                }
            }
            continue;
        }
        blk -= medium_blocks;
        const int2 ch = chunks[blk];
        const int v = lists[nlight + nmedium + ch.x];
        const float2 p = pos[v];
        float ax = 0.0f, ay = 0.0f;
        const int e1 = min(ch.y + kAttrChunk, rowptr[v + 1]);
        gather_terms<GEN, 256, HUB>(p, pos_rank, hub, col, ch.y + tid, e1, linlog, ew, ax, ay);
        red[tid] = make_float2(ax, ay);
        __syncthreads();
        for (int o = 128; o > 0; o >>= 1) {
            if (tid < o) {
                float2 a = red[tid], b = red[tid + o];
                red[tid] = make_float2(a.x + b.x, a.y + b.y);
            }
            __syncthreads();
        }
        if (tid == 0) part[blk] = red[0];
        // red[0] is written again only by thread 0, and every other entry only after the
        // barriers above, so the next virtual block may start at once
    }
}

// Heavy rows: gravity + the row's chunk partials in chunk order.
__global__ __launch_bounds__(256) void k_attract_heavy(const float2* __restrict__ pos, const float* __restrict__ mass,
                                                       const int* __restrict__ heavy, int nheavy,
                                                       const int* __restrict__ c0, const float2* __restrict__ part,
                                                       int strong, float k_g, float2* __restrict__ fga) {
    const int h = blockIdx.x * 256 + threadIdx.x;
    if (h >= nheavy) return;
    const int v = heavy[h];
    float ax = 0.0f, ay = 0.0f;
    for (int c = c0[h]; c < c0[h + 1]; ++c) {
        ax += part[c].x;
        ay += part[c].y;
    }
    const float2 g = gravity(pos[v], mass[v], strong, k_g);
    fga[v] = make_float2(g.x + ax, g.y + ay);
}

// A body's swing and traction terms (SpeedKernel, src/RPFA2Kernels.cu:111-161) from its total
// force F and previous force P: m |F - P| and m |F + P| / 2. The squared norms are written as
// explicit fmaf so that every kernel that forms them (walk, combine, exact) rounds alike.
__device__ __forceinline__ void swing_traction(float mi, float2 F, float2 P, float& sw, float& tr) {
    float dx = F.x - P.x, dy = F.y - P.y;
    sw = mi * sqrtf(fmaf(dx, dx, dy * dy));
    dx = F.x + P.x;
    dy = F.y + P.y;
    tr = mi * (sqrtf(fmaf(dx, dx, dy * dy)) * 0.5f);
}

// The end of a body's force evaluation in the exact kernel: total force
// F = F_ga + F_r, optional repulsion capture, and the group's swing/traction/interaction sums
// (SpeedKernel's per-body terms, src/RPFA2Kernels.cu:111-161), reduced within the G lanes.
template <int G>
__device__ __forceinline__ void group_finish(bool active, bool group_ok, int gl, int gidx, unsigned int id, float mi,
                                             float ax, float ay, unsigned int inter,
                                             const float2* __restrict__ fga, const float2* __restrict__ fprev,
                                             float2* __restrict__ f, float2* __restrict__ rep,
                                             WalkPart* __restrict__ part) {
    float sw = 0.0f, tr = 0.0f;
    if (active) {
        const float2 a = fga[id];
        const float2 F = make_float2(a.x + ax, a.y + ay);
        f[id] = F;
        if (rep) rep[id] = make_float2(ax, ay);
        swing_traction(mi, F, fprev[id], sw, tr);
    }
    // Reduce within the group (xor offsets below G stay inside it).
#pragma unroll
    for (int o = G / 2; o > 0; o >>= 1) {
        sw += __shfl_xor(sw, o, 64);
        tr += __shfl_xor(tr, o, 64);
        inter += __shfl_xor(inter, o, 64);
    }
    if (group_ok && gl == 0) {
        WalkPart wp;
        wp.swing = sw;
        wp.traction = tr;
        wp.interactions = inter;
        wp.pad = 0;
        part[gidx] = wp;
    }
}

// XCD-aware block order: consecutive logical blocks stay on one XCD's L2 (blocks are
// dealt to the 8 XCDs round-robin by hardware block id).
__device__ __forceinline__ int xcd_block() {
    const int nb = gridDim.x;
    const int xcd = blockIdx.x & 7;
    const int per = (nb + 7) >> 3;
    const int full = nb & 7;
    if (full == 0) return xcd * per + (blockIdx.x >> 3);
    return (xcd < full ? xcd * per : full * per + (xcd - full) * (per - 1)) + (blockIdx.x >> 3);
}

// ---------------------------------------------------------------------------------------
// Barnes-Hut repulsion walk. Reference: ForceCalculationKernel, src/RPBHKernels.cu:566-693.
//
// Groups of G consecutive bodies of a graph's sorted (in-order) permutation walk that graph's
// pre-order record array together; groups start at the graph's first sorted position, as the
// reference's warps start at position 0 (a graph in a batch is walked as it is alone). G = 32
// reproduces the reference's 32-lane warp vote; G = 64 votes over the whole wave. One
// wavefront walks one group in two parts:
//
//  traversal  records are tested 64 at a time, one per lane, taken from a stack of pending
//             record indices in LDS. A body record is always accepted. A cell at level L is accepted when every active body b of the group
//             has d2(b, c) >= dq[L-1] (the reference's all-lanes vote, :661), with
//             d2 = fma(dx, dx, fma(dy, dy, eps)) and dq[d] = (2 radius)^2 itolsq 0.25^d + eps
//             (:576-590, formed with ldexp, exact for the power-of-4 scaling). Most cells are
//             decided without the bodies one by one: d2 is monotone in |dx| and |dy|, so the
//             distance from the cell's centroid to the group's bounding box gives a float lower
//             bound of every body's d2 (accept when it clears dq), and the first, middle or
//             last body's own d2 below dq rejects. Only the cells left are voted on, body by
//             body, all lanes at once (lanes = bodies, as in the reference). An accepted record
//             is appended {x, y, k_r m} to an LDS list; an opened cell pushes all its 1-4
//             children at once (their record indices, listed per cell by k_kids when the tree is
//             built), so a cell's children are tested in the same round. The children go on
//             top, so the traversal runs depth-first in blocks of up to 64 records (one record
//             per round near the stack's capacity, where a depth-first walk adds at most three
//             entries per tree level).
//  sums       when the list is nearly full, and at the end, every body adds the list's terms
//             (k_r m_n) d / d2 (k_r m_n formed once per record as it is listed; hardware
//             reciprocal, where the reference divides) and multiplies the sum by its own mass
//             m_i at the end (the reference's k_r d m_i m_n / d2 per term, an ulp or two apart);
//             lanes 0-31 take the even entries and lanes 32-63 the odd ones when G = 32, added
//             together before the multiply.
// The accepted set is the reference's; the order of the terms is the traversal's (fixed for a
// given tree, so every launch shape and a graph alone or in a batch give the same bits).
// The epilogue adds gravity + attraction, writes the total force, and forms the group's
// swing/traction sums (SpeedKernel, src/RPFA2Kernels.cu:111-161).
// ---------------------------------------------------------------------------------------
// LDS per wave: 448 * 4 + (256 + 8) * 12 = 4,960 bytes, so that eight waves fit per SIMD
// (eight 256-thread blocks of 19,840 bytes in the CU's 160 KiB); full 64-entry rounds need
// top <= 128, far above the stacks the walks reach
constexpr int kWalkStack = 448;     // stack entries (record indices) per wave
constexpr int kWalkSlack = 128;     // the depth-first rounds' headroom: at most 3 per tree level
constexpr int kWalkList = 256;      // accepted records buffered per wave
static_assert(kWalkSlack >= 3 * (kMaxLevels + 1) + 4, "a depth-first descent must fit the headroom");
static_assert(kWalkStack - kWalkSlack >= 64 * 3 + 128, "a full round must fit above a stack of 128");

__device__ __forceinline__ float vote_d2(float px, float py, float cx, float cy, float epssq) {
    const float dx = px - cx, dy = py - cy;
    return fmaf(dx, dx, fmaf(dy, dy, epssq));
}

// Orders this wave's LDS writes before its later LDS reads by other lanes.
__device__ __forceinline__ void wave_lds_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

template <int G, int BT>
__global__ __launch_bounds__(BT) void k_walk(const float4* __restrict__ rec, const int4* __restrict__ kids,
                                              const float4* __restrict__ sbody,
                                              const unsigned int* __restrict__ perm, const int2* __restrict__ groups,
                                              int ngroups, const int* __restrict__ node_off,
                                              const int* __restrict__ rec_begin, const float4* __restrict__ seg_root,
                                              const float2* __restrict__ fga, const float2* __restrict__ fprev,
                                              float2* __restrict__ f, float2* __restrict__ rep,
                                              WalkPart* __restrict__ part, float k_r, float itolsq, float epssq,
                                              const DevState* __restrict__ st) {
    static_assert(G == 32 || G == 64, "vote group of 32 or 64 lanes");
    constexpr int W = BT / 64;
    constexpr int kStride = (G == 32) ? 2 : 1;     // list entries per drain step (one per half-wave for G = 32)
    constexpr int kUnroll = 4;                     // drain steps per loop trip
    constexpr int kPad = kStride * kUnroll;
    __shared__ int stk_all[W][kWalkStack];
    __shared__ float3 lst_all[W][kWalkList + kPad];
    if (st->err != 0) return;
    const int lane = threadIdx.x & 63;
    const int wv = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);   // uniform: one group per wave
    int* stk = stk_all[wv];
    float3* lst = lst_all[wv];
    const int gidx = xcd_block() * W + wv;
    if (gidx >= ngroups) return;                  // uniform per wave; no block barriers below
    const int2 grp = groups[gidx];
    const int g = grp.y;
    const int nact = min(G, node_off[g + 1] - grp.x);
    const int bl = (G == 32) ? (lane & 31) : lane;    // this lane's body within the group
    const bool bact = bl < nact;
    const int k = grp.x + min(bl, nact - 1);          // inactive lanes duplicate the last body
    const float4 me = sbody[k];     // = the body's record, read without the record-index hop
    const float px = me.x, py = me.y, mi = me.z;
    const unsigned int id = perm[k];
    // the group's bounding box (inactive lanes duplicate an active body) and its first body
    float mnx = px, mny = py, mxx = px, mxy = py;
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
        mnx = fminf(mnx, __shfl_xor(mnx, o, 64));
        mny = fminf(mny, __shfl_xor(mny, o, 64));
        mxx = fmaxf(mxx, __shfl_xor(mxx, o, 64));
        mxy = fmaxf(mxy, __shfl_xor(mxy, o, 64));
    }
    const float b0x = __int_as_float(__builtin_amdgcn_readfirstlane(__float_as_int(px)));
    const float b0y = __int_as_float(__builtin_amdgcn_readfirstlane(__float_as_int(py)));
    // its middle and last bodies
    const int jm = __builtin_amdgcn_readfirstlane((nact - 1) >> 1), jl = __builtin_amdgcn_readfirstlane(nact - 1);
    const float bmx = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(px), jm));
    const float bmy = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(py), jm));
    const float blx = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(px), jl));
    const float bly = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(py), jl));
    // the lanes whose bodies vote: the group's active bodies (for G = 32 in both half-waves,
    // which hold the same bodies; each half is read as its own 32-bit vote)
    const unsigned long long voters = __builtin_amdgcn_ballot_w64(bact);
    // each half-wave's voters (G = 32), and the byte offset that selects a half-wave's own cell
    const unsigned long long vlo = voters & __builtin_amdgcn_ballot_w64(lane < 32);
    const unsigned long long vhi = voters & __builtin_amdgcn_ballot_w64(lane >= 32);
    const int upper4 = (lane >> 5) << 2;    // 4 in the upper half-wave, else 0 (j2 >= j)
    const float two_r = seg_root[g].z * 2;
    const float base = two_r * two_r * itolsq;
    // the root's children start the walk (a graph of one body has no cell records)
    const int root = rec_begin[g];
    const bool has_root = root + 1 < rec_begin[g + 1];
    int top = 0;                  // stack size (uniform)
    int nlist = 0;                // buffered accepted records (uniform)
    unsigned int naccepted = 0;   // accepted records of the whole walk (uniform)
    if (has_root) {
        const int4 k0 = kids[root];
        if (lane == 0) {
            stk[0] = k0.x;
            stk[1] = k0.y;
            stk[2] = k0.z;
            stk[3] = k0.w;
        }
        top = 1 + (k0.y >= 0) + (k0.z >= 0) + (k0.w >= 0);
    }
    wave_lds_sync();
    float ax = 0.0f, ay = 0.0f;
    const int half = (G == 32) ? (lane >> 5) : 0;
    while (top > 0) {
        // a round takes up to 64 entries; each may push up to 4 children (net +3), so near the
        // capacity the rounds shrink to depth-first steps, whose growth the slack bounds
        int take = top < 64 ? top : 64;
        const int room = (kWalkStack - kWalkSlack - top) / 3;
        if (room < take) take = room > 1 ? room : 1;
        top -= take;
        const bool live = lane < take;
        const unsigned long long livem = take >= 64 ? ~0ull : (1ull << take) - 1ull;
        // lanes past `take` repeat the last entry (in bounds) and take no part below
        const int r = stk[top + min(lane, take - 1)];
        const float4 c = rec[r];
        const int4 kd = kids[r];      // read with the record; used only if the cell is opened
        const unsigned int info = __float_as_uint(c.w);
        const bool body = info == 0xffffffffu;
        const float thr = ldexpf(base, -2 * (int)(info & 31u)) + epssq;
        // distance from the group's box (mnx <= mxx: at most one term is positive)
        const float ex = fmaxf(fmaxf(mnx - c.x, c.x - mxx), 0.0f);
        const float ey = fmaxf(fmaxf(mny - c.y, c.y - mxy), 0.0f);
        // lane masks: bodies; cells every body clears (d2 >= thr); cells the first, middle and
        // last bodies all clear (any of them below thr rejects)
        const unsigned long long bodym = __builtin_amdgcn_ballot_w64(body);
        const unsigned long long clearm = __builtin_amdgcn_ballot_w64(fmaf(ex, ex, fmaf(ey, ey, epssq)) >= thr);
        const unsigned long long p3m = __builtin_amdgcn_ballot_w64(vote_d2(b0x, b0y, c.x, c.y, epssq) >= thr) &
                                       __builtin_amdgcn_ballot_w64(vote_d2(bmx, bmy, c.x, c.y, epssq) >= thr) &
                                       __builtin_amdgcn_ballot_w64(vote_d2(blx, bly, c.x, c.y, epssq) >= thr);
        // the remaining cells, voted on by the group's bodies (lanes = bodies); with G = 32 the
        // two half-waves hold the same bodies and vote on two cells at once
        unsigned long long okm = livem & (bodym | clearm);
        unsigned long long um = livem & p3m & ~(bodym | clearm);
        while (um) {
            const int j = __builtin_ctzll(um);
            um &= um - 1;
            int j2 = j;
            if constexpr (G == 32) {
                if (um) {
                    j2 = __builtin_ctzll(um);
                    um &= um - 1;
                }
            }
            float cx, cy, tj;
            if constexpr (G == 32) {
                // each half-wave fetches its own cell's values (the upper half: the second cell)
                const int src = (j << 2) + upper4 * (j2 - j);
                cx = __int_as_float(__builtin_amdgcn_ds_bpermute(src, __float_as_int(c.x)));
                cy = __int_as_float(__builtin_amdgcn_ds_bpermute(src, __float_as_int(c.y)));
                tj = __int_as_float(__builtin_amdgcn_ds_bpermute(src, __float_as_int(thr)));
            } else {
                cx = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(c.x), j));
                cy = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(c.y), j));
                tj = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(thr), j));
            }
            const unsigned long long fails = __builtin_amdgcn_ballot_w64(!(vote_d2(px, py, cx, cy, epssq) >= tj));
            if constexpr (G == 32) {
                if ((fails & vlo) == 0ull) okm |= 1ull << j;
                if ((fails & vhi) == 0ull) okm |= 1ull << j2;
            } else {
                if ((fails & voters) == 0ull) okm |= 1ull << j;
            }
        }
        // opened cells push their 1-4 children, in four planes (every opened cell's first
        // child, then every second child, ...); okm holds live lanes only
        const unsigned long long am = okm;
        const unsigned long long m0 = livem & ~bodym & ~okm;
        const unsigned long long m1 = m0 & __builtin_amdgcn_ballot_w64(kd.y >= 0);
        const unsigned long long m2 = m0 & __builtin_amdgcn_ballot_w64(kd.z >= 0);
        const unsigned long long m3 = m0 & __builtin_amdgcn_ballot_w64(kd.w >= 0);
        const bool accept = (okm >> lane) & 1ull;
        const bool open = live && !body && !accept;
        const int p1 = top + __popcll(m0), p2 = p1 + __popcll(m1), p3 = p2 + __popcll(m2);
        const int na = __popcll(am);
        if (open) {
            stk[top + __builtin_amdgcn_mbcnt_hi((unsigned)(m0 >> 32), __builtin_amdgcn_mbcnt_lo((unsigned)m0, 0u))] =
                kd.x;
            if (kd.y >= 0)
                stk[p1 + __builtin_amdgcn_mbcnt_hi((unsigned)(m1 >> 32), __builtin_amdgcn_mbcnt_lo((unsigned)m1, 0u))] =
                    kd.y;
            if (kd.z >= 0)
                stk[p2 + __builtin_amdgcn_mbcnt_hi((unsigned)(m2 >> 32), __builtin_amdgcn_mbcnt_lo((unsigned)m2, 0u))] =
                    kd.z;
            if (kd.w >= 0)
                stk[p3 + __builtin_amdgcn_mbcnt_hi((unsigned)(m3 >> 32), __builtin_amdgcn_mbcnt_lo((unsigned)m3, 0u))] =
                    kd.w;
        }
        if (accept) {
            const int ra =
                __builtin_amdgcn_mbcnt_hi((unsigned)(am >> 32), __builtin_amdgcn_mbcnt_lo((unsigned)am, 0u));
            lst[nlist + ra] = make_float3(c.x, c.y, k_r * c.z);
        }
        top = p3 + __popcll(m3);
        nlist += na;
        naccepted += (unsigned int)na;
        if (nlist > kWalkList - 64 || top == 0) {
            // pad to whole loop trips with massless entries (their terms are exactly 0)
            const int padded = (nlist + kPad - 1) / kPad * kPad;
            if (nlist + lane < padded) lst[nlist + lane] = make_float3(0.0f, 0.0f, 0.0f);
            wave_lds_sync();
            for (int q0 = 0; q0 < padded; q0 += kPad) {
#pragma unroll
                for (int u = 0; u < kUnroll; ++u) {
                    const float3 e = lst[q0 + u * kStride + half];
                    const float dx = px - e.x, dy = py - e.y;
                    const float w = e.z * __builtin_amdgcn_rcpf(fmaf(dx, dx, fmaf(dy, dy, epssq)));
                    ax = fmaf(dx, w, ax);
                    ay = fmaf(dy, w, ay);
                }
            }
            nlist = 0;
        }
        wave_lds_sync();
    }
    if constexpr (G == 32) {
        // even entries (lanes 0-31) + odd entries (lanes 32-63)
        ax += __shfl_down(ax, 32, 64);
        ay += __shfl_down(ay, 32, 64);
    }
    ax *= mi;   // the body's own mass, taken out of every term
    ay *= mi;
    const bool own = bact && lane < G;
    if (!fga) {
        // split: the repulsion alone goes to f and the group's interaction count to its
        // partial; k_combine adds F_ga and forms the swing/traction sums once the
        // attraction, run concurrently on another stream, is done
        if (own) f[id] = make_float2(ax, ay);
        if (lane == 0) {
            WalkPart wp;
            wp.swing = 0.0f;
            wp.traction = 0.0f;
            wp.interactions = naccepted * (unsigned int)nact;   // = the sum over the group's bodies
            wp.pad = 0;
            part[gidx] = wp;
        }
        return;
    }
    // epilogue: total force, repulsion capture, the group's swing/traction/interaction sums
    float sw = 0.0f, tr = 0.0f;
    unsigned int inter = own ? naccepted : 0u;
    if (own) {
        const float2 a = fga[id];
        const float2 F = make_float2(a.x + ax, a.y + ay);
        f[id] = F;
        if (rep) rep[id] = make_float2(ax, ay);
        swing_traction(mi, F, fprev[id], sw, tr);
    }
#pragma unroll
    for (int o = G / 2; o > 0; o >>= 1) {
        sw += __shfl_xor(sw, o, 64);
        tr += __shfl_xor(tr, o, 64);
        inter += __shfl_xor(inter, o, 64);
    }
    if (lane == 0) {
        WalkPart wp;
        wp.swing = sw;
        wp.traction = tr;
        wp.interactions = inter;
        wp.pad = 0;
        part[gidx] = wp;
    }
}

// The walk's epilogue for a split walk (k_walk with fga == nullptr): F = F_ga + F_r per body,
// the optional repulsion capture, and the group's swing/traction sums, with the arithmetic and
// the reduction order of the fused epilogue, so both give the same bits. Groups of G = 32 take
// half a wave each (the xor offsets of the reduction stay inside a half).
template <int G>
__global__ __launch_bounds__(256) void k_combine(const float4* __restrict__ sbody,
                                                 const unsigned int* __restrict__ perm, const int2* __restrict__ groups,
                                                 int ngroups, const int* __restrict__ node_off,
                                                 const float2* __restrict__ fga, const float2* __restrict__ fprev,
                                                 float2* __restrict__ f, float2* __restrict__ rep,
                                                 WalkPart* __restrict__ part, const DevState* __restrict__ st) {
    if (st->err != 0) return;
    const int lane = threadIdx.x & 63;
    const int gidx = (blockIdx.x * 4 + (threadIdx.x >> 6)) * (64 / G) + lane / G;
    const int bl = lane % G;
    const bool gok = gidx < ngroups;
    bool own = false;
    unsigned int id = 0;
    float mi = 0.0f;
    if (gok) {
        const int2 grp = groups[gidx];
        own = bl < node_off[grp.y + 1] - grp.x;
        if (own) {
            const int k = grp.x + bl;
            id = perm[k];
            mi = sbody[k].z;
        }
    }
    float sw = 0.0f, tr = 0.0f;
    if (own) {
        const float2 r = f[id];
        const float2 a = fga[id];
        const float2 F = make_float2(a.x + r.x, a.y + r.y);
        f[id] = F;
        if (rep) rep[id] = r;
        swing_traction(mi, F, fprev[id], sw, tr);
    }
#pragma unroll
    for (int o = G / 2; o > 0; o >>= 1) {
        sw += __shfl_xor(sw, o, 64);
        tr += __shfl_xor(tr, o, 64);
    }
    if (gok && bl == 0) {
        part[gidx].swing = sw;
        part[gidx].traction = tr;
    }
}

// ---------------------------------------------------------------------------------------
// Speed update, one block per graph. Reference: SpeedKernel's last block,
// src/RPFA2Kernels.cu:167-219 (Gephi's adaptive speed). The graph's group partials are
// summed in double in a fixed order (thread-strided, then an LDS tree), the same for a
// graph alone or inside a batch; the update itself is the reference's float arithmetic,
// including its double literals. Graph 0 is mirrored into DevState.
// ---------------------------------------------------------------------------------------
constexpr int kSpeedThreads = 1024;

__global__ __launch_bounds__(kSpeedThreads) void k_speed(const WalkPart* __restrict__ part,
                                                         const int* __restrict__ grp_first,
                                                         const int* __restrict__ node_off, SegState* __restrict__ seg,
                                                         DevState* st, const int* __restrict__ kids_part,
                                                         int kids_blocks) {
    __shared__ double rs[kSpeedThreads], rt[kSpeedThreads];
    __shared__ unsigned long long ri[kSpeedThreads];
    __shared__ int rm[8];
    if (st->err != 0) return;
    const int g = blockIdx.x;
    const int tid = threadIdx.x;
    if (g == 0) {
        // the tree's largest cell at level 4, 8, ..., 32, from the child-list kernel's blocks
        if (tid < 8) rm[tid] = 0;
        __syncthreads();
        // thread t reads level slot t & 7 of every 128th block (the stride keeps the slot),
        // eight loads in flight
        int m = 0;
        const int nq = kids_blocks * 8;
        int q = tid;
        for (; q + 7 * kSpeedThreads < nq; q += 8 * kSpeedThreads) {
            int v[8];
#pragma unroll
            for (int u = 0; u < 8; ++u) v[u] = kids_part[q + u * kSpeedThreads];
#pragma unroll
            for (int u = 0; u < 8; ++u) m = max(m, v[u]);
        }
        for (; q < nq; q += kSpeedThreads) m = max(m, kids_part[q]);
        if (m > 0) atomicMax(&rm[tid & 7], m);
        __syncthreads();
    }
    const int g0 = grp_first[g], g1 = grp_first[g + 1];
    double s = 0.0, t = 0.0;
    unsigned long long it = 0;
    // thread-strided sums in group order; 8 partials are loaded before they are added (same
    // order, so the same bits), keeping 8 loads in flight per thread
    constexpr int kAhead = 8;
    int k = g0 + tid;
    for (; k + (kAhead - 1) * kSpeedThreads < g1; k += kAhead * kSpeedThreads) {
        WalkPart p[kAhead];
#pragma unroll
        for (int j = 0; j < kAhead; ++j) p[j] = part[k + j * kSpeedThreads];
#pragma unroll
        for (int j = 0; j < kAhead; ++j) {
            s += (double)p[j].swing;
            t += (double)p[j].traction;
            it += p[j].interactions;
        }
    }
    for (; k < g1; k += kSpeedThreads) {
        const WalkPart p = part[k];
        s += (double)p.swing;
        t += (double)p.traction;
        it += p.interactions;
    }
    rs[tid] = s;
    rt[tid] = t;
    ri[tid] = it;
    __syncthreads();
    for (int o = kSpeedThreads / 2; o > 0; o >>= 1) {
        if (tid < o) {
            rs[tid] += rs[tid + o];
            rt[tid] += rt[tid + o];
            ri[tid] += ri[tid + o];
        }
        __syncthreads();
    }
    if (tid == 0) {
        const int n = node_off[g + 1] - node_off[g];
        SegState& q = seg[g];
        const double S = rs[0], T = rt[0];
        q.swing = S;
        q.traction = T;
        q.interactions = ri[0];
        const float swg = (float)S, etra = (float)T;
        float gs = q.global_speed, se = q.speed_efficiency;
        const float jitter = q.jitter_tolerance, ksmax = q.k_s_max;
        const float eojt = (float)(0.05 * sqrtf((float)n));
        const float minjt = sqrtf(eojt);
        float jt = jitter * fmaxf(minjt, fminf(ksmax, eojt * etra / powf((float)n, 2.0f)));
        const float min_se = (float)0.05;
        if (swg / etra > 2.0) {
            if (se > min_se) se = (float)(se * 0.5);
            jt = fmaxf(jt, jitter);
        }
        const float target = jt * se * etra / swg;
        if (swg > jt * etra) {
            if (se > min_se) se = (float)(se * 0.7);
        } else if (gs < 1000) {
            se = (float)(se * 1.3);
        }
        gs += fminf(target - gs, (float)0.5 * gs);
        q.global_speed = gs;
        q.speed_efficiency = se;
        q.iteration += 1;
        if (g == 0) {
            st->swing = S;
            st->traction = T;
            st->interactions = ri[0];
            st->global_speed = gs;
            st->speed_efficiency = se;
            // the next path-key sort: p 8-bit passes order the bodies by their top 4p levels,
            // and the pairs left tied are the bodies of one 4p-level cell, which the fix-up
            // orders; take the fewest passes whose cells (in this tree) hold at most
            // kSortRunMax bodies each
            int p = 1;
            while (p < 8 && rm[p - 1] > kSortRunMax) ++p;
            if (st->full_sorts > 0) {
                st->full_sorts -= 1;      // still settling after new positions
                p = 8;
            }
            st->sort_skip = 8 - p;
            for (int q = 0; q < 8; ++q) st->runmax[q] = rm[q];
        }
    }
    // per graph (one block each): no word shared by the blocks of a batch
    if (tid == 0) seg[g].interactions_total += ri[0];
}

// ---------------------------------------------------------------------------------------
// Displacement. Reference: DisplacementKernel, src/RPFA2Kernels.cu:223-251:
// factor = gs/(1.0 + sqrtf(gs*swg)) (double, because of the literal), p += F*factor,
// F' = F. A single graph also emits bounding-box partials of the new positions for the
// next tree (a batch recomputes its per-graph boxes instead).
// ---------------------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_displace(float2* __restrict__ pos, const float2* __restrict__ f,
                                                  float2* __restrict__ fprev, int n, const int* __restrict__ seg_of,
                                                  const SegState* __restrict__ seg, float4* __restrict__ part,
                                                  DevState* st) {
    __shared__ float4 red[256];
    const int tid = threadIdx.x;
    const bool live = st->err == 0;
    const float gs0 = seg[0].global_speed;
    float4 acc = make_float4(INFINITY, INFINITY, -INFINITY, -INFINITY);
    for (int i = blockIdx.x * 256 + tid; i < n; i += gridDim.x * 256) {
        float2 p = pos[i];
        if (live) {
            const float gs = seg_of ? seg[seg_of[i]].global_speed : gs0;
            const float2 F = f[i];
            const float2 P = fprev[i];
            const float dx = F.x - P.x, dy = F.y - P.y;
            const float swg = sqrtf(dx * dx + dy * dy);
            const float factor = (float)((double)gs / (1.0 + (double)sqrtf(gs * swg)));
            p.x = fmaf(F.x, factor, p.x);
            p.y = fmaf(F.y, factor, p.y);
            pos[i] = p;
            fprev[i] = F;
        }
        acc.x = fminf(acc.x, p.x);
        acc.y = fminf(acc.y, p.y);
        acc.z = fmaxf(acc.z, p.x);
        acc.w = fmaxf(acc.w, p.y);
    }
    if (part == nullptr) {
        if (blockIdx.x == 0 && tid == 0 && live) st->iteration += 1;
        return;
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
        if (blockIdx.x == 0) {
            st->nbbox = gridDim.x;
            if (live) st->iteration += 1;
        }
    }
}

// ---------------------------------------------------------------------------------------
// Exact repulsion (no counterpart in the reference's GPU path). Each lane is one node; the
// group's G lanes run over all nodes j of their graph in ascending order, reading the packed
// {x, y, mass} of node j (one address per group: a broadcast load), with the walk's term
// k_r m_i m_j (p_i - p_j) / (|p_i - p_j|^2 + eps^2) (node i itself adds 0). Every body
// interacts with every body of its graph, so the interaction count is the sum of n_g^2.
// ---------------------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_pack(const float2* __restrict__ pos, const float* __restrict__ mass, int n,
                                              float4* __restrict__ body) {
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        const float2 p = pos[i];
        body[i] = make_float4(p.x, p.y, mass[i], 0.0f);
    }
}

template <int G, int BT>
__global__ __launch_bounds__(BT) void k_exact(const float4* __restrict__ body, const int2* __restrict__ groups,
                                              int ngroups, const int* __restrict__ node_off,
                                              const float2* __restrict__ fga, const float2* __restrict__ fprev,
                                              float2* __restrict__ f, float2* __restrict__ rep,
                                              WalkPart* __restrict__ part, float k_r, float epssq,
                                              const DevState* __restrict__ st) {
    if (st->err != 0) return;
    const int lb = xcd_block();
    const int gidx = (lb * BT + (int)threadIdx.x) / G;
    const int gl = (lb * BT + (int)threadIdx.x) % G;
    const bool group_ok = gidx < ngroups;
    int2 grp = make_int2(0, 0);
    if (group_ok) grp = groups[gidx];
    const int g = grp.y;
    const int k = grp.x + gl;
    const bool active = group_ok && k < node_off[g + 1];
    float px = 0.0f, py = 0.0f, mi = 0.0f, kmi = 0.0f;
    if (active) {
        const float4 me = body[k];
        px = me.x;
        py = me.y;
        mi = me.z;
        kmi = k_r * mi;
    }
    float ax = 0.0f, ay = 0.0f;
    unsigned int inter = 0;
    if (group_ok) {
        const int j0 = node_off[g], j1 = node_off[g + 1];
        for (int j = j0; j < j1; ++j) {
            const float4 c = body[j];
            const float dx = px - c.x;
            const float dy = py - c.y;
            const float d2 = dx * dx + dy * dy + epssq;
            const float w = kmi * c.z * __builtin_amdgcn_rcpf(d2);
            ax = fmaf(dx, w, ax);
            ay = fmaf(dy, w, ay);
        }
        inter = active ? (unsigned int)(j1 - j0) : 0u;
    }
    group_finish<G>(active, group_ok, gl, gidx, (unsigned int)k, mi, ax, ay, inter, fga, fprev, f, rep, part);
}

// All-pairs repulsion into out, one thread per node, FP32 or FP64 arithmetic; the graph of
// node i is found from node_off by binary search.
template <typename T>
__global__ __launch_bounds__(256) void k_exact_only(const float4* __restrict__ body, int n, int nseg,
                                                    const int* __restrict__ node_off, float k_r, float epssq,
                                                    double2* __restrict__ out) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    int lo = 0, hi = nseg;           // node_off[lo] <= i < node_off[hi]
    while (hi - lo > 1) {
        const int mid = (lo + hi) >> 1;
        if (node_off[mid] <= i) lo = mid;
        else hi = mid;
    }
    const float4 me = body[i];
    const T px = me.x, py = me.y, kmi = (T)k_r * (T)me.z, eps = epssq;
    T ax = 0, ay = 0;
    for (int j = node_off[lo]; j < node_off[lo + 1]; ++j) {
        const float4 c = body[j];
        const T dx = px - (T)c.x;
        const T dy = py - (T)c.y;
        const T d2 = dx * dx + dy * dy + eps;
        const T w = kmi * (T)c.z / d2;
        ax += dx * w;
        ay += dy * w;
    }
    out[i] = make_double2((double)ax, (double)ay);
}

// w^delta per CSR entry (delta = 0 gives 1).
__global__ __launch_bounds__(256) void k_edge_pow(const float* __restrict__ w, long long nnz, float delta,
                                                  float* __restrict__ out) {
    for (long long e = blockIdx.x * (long long)blockDim.x + threadIdx.x; e < nnz;
         e += (long long)gridDim.x * blockDim.x)
        out[e] = delta == 0.0f ? 1.0f : (delta == 1.0f ? w[e] : powf(w[e], delta));
}

}  // namespace

int walk_blocks(int ngroups, int group, int bt) {
    const int per_block = bt / group;
    int b = (ngroups + per_block - 1) / per_block;
    return b < 1 ? 1 : b;
}

namespace {

template <int G, int BT>
void walk_launch(const Bufs& b, const Params& p, bool split, hipStream_t s) {
    constexpr int W = BT / 64;     // one group per wave
    const int blocks = (b.ngroups + W - 1) / W;
    k_walk<G, BT><<<blocks < 1 ? 1 : blocks, BT, 0, s>>>(b.rec, b.kids, b.sbody, b.sort.vals[0], b.groups, b.ngroups,
                                                          b.node_off, b.rec_begin, b.seg_root, split ? nullptr : b.fga,
                                                          b.fprev, b.f, split ? nullptr : b.rep, b.grp_part, p.k_r,
                                                          p.itolsq, p.epssq, b.st);
}

template <int G>
void walk_dispatch(const Bufs& b, const Params& p, bool split, hipStream_t s) {
    switch (p.walk_bt) {
        case 64: walk_launch<G, 64>(b, p, split, s); break;
        case 128: walk_launch<G, 128>(b, p, split, s); break;
        case 512: walk_launch<G, 512>(b, p, split, s); break;
        default: walk_launch<G, 256>(b, p, split, s); break;
    }
}

}  // namespace

void launch_col_rank(const Bufs& b, hipStream_t s) {
    if (b.nnz > 0) k_gather_rank<<<(int)std::min<long long>((b.nnz + 255) / 256, 16384), 256, 0, s>>>(b.col, b.rank, b.nnz, b.col_rank);
}

void launch_attraction(const Bufs& b, const Params& p, hipStream_t s) {
    k_rank_pos<<<(b.n + 255) / 256, 256, 0, s>>>(b.pos, b.rank, b.n, b.pos_rank);
    const int lb = (b.nlight + 255) / 256;
    const int mb = (b.nmedium + 3) / 4;
    int total = lb + mb + b.nattr_chunks;
    if (b.attr_grid > 0 && total > b.attr_grid) total = b.attr_grid;   // the launch grid
    if (total > 0) {
        // a grid smaller than the work (beside other kernels): every block stages the hubs'
        // positions once and uses them for several virtual blocks
        const bool hub = b.attr_grid > 0 && lb + mb + b.nattr_chunks > 2 * total;
        const bool gen = p.linlog || b.ew;
#define GFA2_ATTRACT(G, H)                                                                                   \
    k_attract<G, H><<<total, 256, 0, s>>>(b.pos, b.mass, b.n, b.pos_rank, b.rowptr, b.col_rank, b.node_lists,   \
                                          b.nlight, b.nmedium, b.attr_chunks, b.nattr_chunks, lb, mb,         \
                                          p.strong_gravity, p.k_g, b.fga, b.attr_part, G ? p.linlog : 0,      \
                                          G ? b.ew : nullptr)
        if (gen) {
            if (hub) GFA2_ATTRACT(true, true);
            else GFA2_ATTRACT(true, false);
        } else {
            if (hub) GFA2_ATTRACT(false, true);
            else GFA2_ATTRACT(false, false);
        }
#undef GFA2_ATTRACT
    }
    if (b.nheavy > 0)
        k_attract_heavy<<<(b.nheavy + 255) / 256, 256, 0, s>>>(b.pos, b.mass, b.node_lists + b.nlight + b.nmedium,
                                                               b.nheavy, b.attr_c0, b.attr_part, p.strong_gravity,
                                                               p.k_g, b.fga);
}

void launch_walk(const Bufs& b, const Params& p, bool split, hipStream_t s) {
    if (p.group == 64) walk_dispatch<64>(b, p, split, s);
    else walk_dispatch<32>(b, p, split, s);
}

void launch_combine(const Bufs& b, const Params& p, hipStream_t s) {
    const int per_block = 4 * (64 / p.group);
    const int blocks = (b.ngroups + per_block - 1) / per_block;
    if (p.group == 64)
        k_combine<64><<<blocks < 1 ? 1 : blocks, 256, 0, s>>>(b.sbody, b.sort.vals[0], b.groups, b.ngroups,
                                                              b.node_off, b.fga, b.fprev, b.f, b.rep, b.grp_part, b.st);
    else
        k_combine<32><<<blocks < 1 ? 1 : blocks, 256, 0, s>>>(b.sbody, b.sort.vals[0], b.groups, b.ngroups,
                                                              b.node_off, b.fga, b.fprev, b.f, b.rep, b.grp_part, b.st);
}

void launch_speed(const Bufs& b, hipStream_t s) {
    k_speed<<<b.nseg, kSpeedThreads, 0, s>>>(b.grp_part, b.grp_first, b.node_off, b.seg_st, b.st, b.kids_part,
                                              b.kids_blocks);
}

void launch_displacement(const Bufs& b, hipStream_t s) {
    k_displace<<<displacement_blocks(b), 256, 0, s>>>(b.pos, b.f, b.fprev, b.n, b.nseg > 1 ? b.seg_of : nullptr,
                                                       b.seg_st, b.nseg > 1 ? nullptr : b.bbox_part, b.st);
}

void launch_exact(const Bufs& b, const Params& p, hipStream_t s) {
    int pb = (b.n + 255) / 256;
    if (pb > 4096) pb = 4096;
    k_pack<<<pb, 256, 0, s>>>(b.pos, b.mass, b.n, b.sbody);
    constexpr int BT = 256;
    if (p.group == 64)
        k_exact<64, BT><<<walk_blocks(b.ngroups, 64, BT), BT, 0, s>>>(b.sbody, b.groups, b.ngroups, b.node_off, b.fga,
                                                                      b.fprev, b.f, b.rep, b.grp_part, p.k_r,
                                                                      p.epssq, b.st);
    else
        k_exact<32, BT><<<walk_blocks(b.ngroups, 32, BT), BT, 0, s>>>(b.sbody, b.groups, b.ngroups, b.node_off, b.fga,
                                                                      b.fprev, b.f, b.rep, b.grp_part, p.k_r,
                                                                      p.epssq, b.st);
}

void launch_exact_only(const Bufs& b, const Params& p, bool fp64, double2* out, hipStream_t s) {
    int pb = (b.n + 255) / 256;
    if (pb > 4096) pb = 4096;
    k_pack<<<pb, 256, 0, s>>>(b.pos, b.mass, b.n, b.sbody);
    const int blocks = (b.n + 255) / 256;
    if (fp64)
        k_exact_only<double><<<blocks, 256, 0, s>>>(b.sbody, b.n, b.nseg, b.node_off, p.k_r, p.epssq, out);
    else
        k_exact_only<float><<<blocks, 256, 0, s>>>(b.sbody, b.n, b.nseg, b.node_off, p.k_r, p.epssq, out);
}

void launch_edge_weights(const Bufs& b, float delta, hipStream_t s) {
    if (!b.ew_raw || !b.ew || b.nnz == 0) return;
    long long blocks = (b.nnz + 255) / 256;
    if (blocks > 4096) blocks = 4096;
    k_edge_pow<<<(int)blocks, 256, 0, s>>>(b.ew_raw, b.nnz, delta, b.ew);
}

}  // namespace gfa2
