// Learned initial layout: a small message-passing network whose output positions
// ForceAtlas2 then refines (no counterpart in the reference).
//
// Network (fixed shape): inputs X [V x 16] (9 features, zero padded), hidden 64, output 2.
//   A0 = mean_agg(X)                     mean over CSR neighbours (0 for isolated nodes)
//   H1 = relu(X W1s + A0 W1n + b1)       [V x 64]
//   A1 = mean_agg(H1)
//   H2 = relu(H1 W2s + A1 W2n + b2)      [V x 64]
//   Y  = H2 Wo + bo                      [V x 2]
// Weights are stored in BF16 (round to nearest even from FP32 master copies kept for
// Adam); every product accumulates in FP32. Gradients are taken with respect to the
// BF16-rounded weights and applied to the masters.
//
// Features: column 0 = log(deg+1); columns 1..8 = seeded standard normal fields smoothed by
// `steps` rounds of z <- (z + mean_agg(z)) / 2; every column then standardised (zero mean,
// unit variance over the nodes).
//
// Loss (invariant to translation and rotation/reflection of the target; the target is
// centred and scaled to unit RMS radius): over a symmetric set P of ordered pairs - each
// node's CSR neighbours plus the nodes (i +/- o_k) mod V for R seeded offsets o_k -
//   L = (1/|P|) sum_{(i,j) in P} (|y_i - y_j| - |t_i - t_j|)^2
// Because P is symmetric, dL/dy_i = (4/|P|) sum_{j in P(i)} (|y_ij| - |t_ij|) (y_i-y_j)/|y_ij|
// is a per-node gather: no atomics anywhere, so training is deterministic.
//
// Every reduction over nodes (weight gradients, column statistics, loss) writes per-block
// partials that a second kernel sums in a fixed order.
#include "engine_internal.h"
#include "mpnn_internal.h"

#include <vector>

namespace gfa2 {
namespace {

__device__ __forceinline__ float bf16_to_f32(unsigned short b) { return __uint_as_float((unsigned int)b << 16); }

__device__ __forceinline__ unsigned short f32_to_bf16(float f) {
    unsigned int u = __float_as_uint(f);
    if ((u & 0x7f800000u) == 0x7f800000u) return (unsigned short)(u >> 16);   // inf / nan
    u += 0x7fffu + ((u >> 16) & 1u);
    return (unsigned short)(u >> 16);
}

__device__ __forceinline__ unsigned long long mix64(unsigned long long z) {
    z += 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

// ---- features -----------------------------------------------------------------------------

// Z [V x 8]: standard normal by Box-Muller from two 24-bit uniforms of mix64(seed, i*8+f).
__global__ __launch_bounds__(256) void k_normal_fields(int n, unsigned long long seed, float* __restrict__ Z) {
    const int t = blockIdx.x * 256 + threadIdx.x;
    if (t >= n * 8) return;
    const unsigned long long r = mix64(seed ^ mix64((unsigned long long)t));
    const float u1 = ((float)(r >> 40) + 0.5f) * (1.0f / 16777216.0f);
    const float u2 = (float)((r >> 16) & 0xffffffull) * (1.0f / 16777216.0f);
    Z[t] = sqrtf(-2.0f * logf(u1)) * cosf(6.2831853071795864f * u2);
}

// out[i][f] = (in[i][f] + mean_j in[j][f]) / 2 for 8 columns.
// X (unnormalised) columns 0..8, zero padding 9..15.
__global__ __launch_bounds__(256) void k_assemble(const int* __restrict__ rowptr, int n, const float* __restrict__ Z,
                                                  float* __restrict__ X) {
    const int t = blockIdx.x * 256 + threadIdx.x;
    if (t >= n * kIn) return;
    const int i = t / kIn, f = t % kIn;
    float v = 0.0f;
    if (f == 0) v = logf((float)(rowptr[i + 1] - rowptr[i]) + 1.0f);
    else if (f <= 8) v = Z[i * 8 + f - 1];
    X[t] = v;
}

// Per-block column sums and sums of squares of X (kIn columns).
__global__ __launch_bounds__(256) void k_colstats(const float* __restrict__ X, int n, double* __restrict__ part) {
    __shared__ double s1[256], s2[256];
    const int f = threadIdx.x & (kIn - 1), lane_row = threadIdx.x / kIn;   // 16 rows x 16 columns
    double a = 0.0, b = 0.0;
    for (int i = blockIdx.x * 16 + lane_row; i < n; i += gridDim.x * 16) {
        const double v = X[i * kIn + f];
        a += v;
        b += v * v;
    }
    s1[threadIdx.x] = a;
    s2[threadIdx.x] = b;
    __syncthreads();
    if (threadIdx.x < kIn) {
        double A = 0.0, B = 0.0;
        for (int r = 0; r < 16; ++r) {
            A += s1[r * kIn + threadIdx.x];
            B += s2[r * kIn + threadIdx.x];
        }
        part[(blockIdx.x * kIn + threadIdx.x) * 2] = A;
        part[(blockIdx.x * kIn + threadIdx.x) * 2 + 1] = B;
    }
}

__global__ void k_colnorm_finish(const double* __restrict__ part, int nb, int n, float* __restrict__ mu_sd) {
    const int f = threadIdx.x;
    if (f >= kIn) return;
    double A = 0.0, B = 0.0;
    for (int b = 0; b < nb; ++b) {
        A += part[(b * kIn + f) * 2];
        B += part[(b * kIn + f) * 2 + 1];
    }
    const double mu = A / n;
    const double var = B / n - mu * mu;
    mu_sd[2 * f] = (float)mu;
    mu_sd[2 * f + 1] = (f <= 8 && var > 1e-12) ? (float)(1.0 / sqrt(var)) : 0.0f;
}

__global__ __launch_bounds__(256) void k_colnorm_apply(float* __restrict__ X, int n, const float* __restrict__ mu_sd) {
    const int t = blockIdx.x * 256 + threadIdx.x;
    if (t >= n * kIn) return;
    const int f = t % kIn;
    X[t] = (X[t] - mu_sd[2 * f]) * mu_sd[2 * f + 1];
}

// ---- neighbour aggregation -----------------------------------------------------------------

// out[i][:] = epilogue(sum over CSR neighbours j of in[j][:], times 1/deg(j) for the transposed
// form), F features per node (8, 16 or 64):
//   kAggMean       out = sum / deg(i) (0 for an isolated node)     forward mean
//   kAggSmooth     out = (in[i] + sum / deg(i)) / 2                 feature smoothing
//   kAggTranspose  out = sum of in[j] / deg(j)                      backward of the mean
// Rows of at most kHeavyDeg entries: F lanes per row, entries in CSR order (four loads in
// flight). Longer rows (the hubs of a skewed graph, which would otherwise leave one wave
// looping over 10^5 entries) are split into chunks of kChunk entries: one wave per chunk sums
// its entries in 64/F interleaved slots, combined in slot order, into a partial; then one
// thread per (row, feature) adds the row's partials in chunk order. Every sum has a fixed
// order, so results are deterministic.
enum AggMode { kAggMean = 0, kAggSmooth = 1, kAggTranspose = 2 };

template <int MODE>
__device__ __forceinline__ float agg_add(float s, const int* __restrict__ rowptr, const float* __restrict__ in,
                                         int j, int F, int f) {
    const float v = in[(size_t)j * F + f];
    if (MODE == kAggTranspose) return fmaf(v, 1.0f / (float)(rowptr[j + 1] - rowptr[j]), s);
    return s + v;
}

template <int MODE>
__device__ __forceinline__ float agg_finish(float s, int deg, float self) {
    if (MODE == kAggTranspose) return s;
    const float m = deg > 0 ? s / (float)deg : 0.0f;
    return MODE == kAggSmooth ? 0.5f * (self + m) : m;
}

template <int F, int MODE>
__global__ __launch_bounds__(256) void k_agg_rows(const int* __restrict__ rowptr, const int* __restrict__ col, int n,
                                                  const float* __restrict__ in, float* __restrict__ out) {
    constexpr int kRows = 256 / F;
    const int f = threadIdx.x % F;
    for (long long r0 = (long long)blockIdx.x * kRows; r0 < n; r0 += (long long)gridDim.x * kRows) {
        const long long ii = r0 + threadIdx.x / F;
        if (ii >= n) continue;
        const int i = (int)ii;
        const int e0 = rowptr[i], e1 = rowptr[i + 1];
        if (e1 - e0 > kHeavyDeg) continue;            // the chunk path writes this row
        float s = 0.0f;
        int e = e0;
        for (; e + 4 <= e1; e += 4) {
            const int j0 = col[e], j1 = col[e + 1], j2 = col[e + 2], j3 = col[e + 3];
            s = agg_add<MODE>(s, rowptr, in, j0, F, f);
            s = agg_add<MODE>(s, rowptr, in, j1, F, f);
            s = agg_add<MODE>(s, rowptr, in, j2, F, f);
            s = agg_add<MODE>(s, rowptr, in, j3, F, f);
        }
        for (; e < e1; ++e) s = agg_add<MODE>(s, rowptr, in, col[e], F, f);
        out[(size_t)i * F + f] = agg_finish<MODE>(s, e1 - e0, MODE == kAggSmooth ? in[(size_t)i * F + f] : 0.0f);
    }
}

// One wave per chunk {row, first entry}; part[c][f] = the chunk's sum.
template <int F, int MODE>
__global__ __launch_bounds__(256) void k_agg_chunks(const int* __restrict__ rowptr, const int* __restrict__ col,
                                                    const int2* __restrict__ chunks, int nchunks,
                                                    const float* __restrict__ in, float* __restrict__ part) {
    constexpr int S = 64 / F;
    const int lane = threadIdx.x & 63, slot = lane / F, f = lane % F;
    for (int c = blockIdx.x * 4 + (threadIdx.x >> 6); c < nchunks; c += gridDim.x * 4) {
        const int2 ch = chunks[c];
        const int end = min(ch.y + kChunk, rowptr[ch.x + 1]);
        float s = 0.0f;
        int e = ch.y + slot;
        for (; e + 3 * S < end; e += 4 * S) {
            const int j0 = col[e], j1 = col[e + S], j2 = col[e + 2 * S], j3 = col[e + 3 * S];
            s = agg_add<MODE>(s, rowptr, in, j0, F, f);
            s = agg_add<MODE>(s, rowptr, in, j1, F, f);
            s = agg_add<MODE>(s, rowptr, in, j2, F, f);
            s = agg_add<MODE>(s, rowptr, in, j3, F, f);
        }
        for (; e < end; e += S) s = agg_add<MODE>(s, rowptr, in, col[e], F, f);
        float tot = 0.0f;
#pragma unroll
        for (int q = 0; q < S; ++q) tot += __shfl(s, q * F + f, 64);
        if (slot == 0) part[(size_t)c * F + f] = tot;
    }
}

// One thread per (heavy row, feature): the row's chunk partials in chunk order.
template <int F, int MODE>
__global__ __launch_bounds__(256) void k_agg_heavy(const int* __restrict__ rowptr, const int* __restrict__ heavy_row,
                                                   const int* __restrict__ heavy_c0, int nheavy,
                                                   const float* __restrict__ in, const float* __restrict__ part,
                                                   float* __restrict__ out) {
    const long long t = (long long)blockIdx.x * 256 + threadIdx.x;
    if (t >= (long long)nheavy * F) return;
    const int h = (int)(t / F), f = (int)(t % F);
    const int i = heavy_row[h];
    float s = 0.0f;
    for (int c = heavy_c0[h]; c < heavy_c0[h + 1]; ++c) s += part[(size_t)c * F + f];
    out[(size_t)i * F + f] = agg_finish<MODE>(s, rowptr[i + 1] - rowptr[i],
                                              MODE == kAggSmooth ? in[(size_t)i * F + f] : 0.0f);
}

// H[i][o] = relu(b[o] + sum_k (A[i][k] Ws[k][o] + B[i][k] Wn[k][o])), accumulated in the
// order b, A_0 Ws_0, B_0 Wn_0, A_1 Ws_1, ... (FMAs). Register-tiled: a block holds the BF16
// weights as FP32 in LDS and walks tiles of 64 nodes, staging A and B transposed in LDS; each
// thread owns 4 nodes x 4 outputs (16 accumulators), so one k step is 4 LDS reads of 16 bytes
// for 32 FMAs.
template <int K>
__global__ __launch_bounds__(256) void k_dense(const float* __restrict__ A, const float* __restrict__ B, int n,
                                               const unsigned short* __restrict__ Ws,
                                               const unsigned short* __restrict__ Wn,
                                               const unsigned short* __restrict__ bias, float* __restrict__ H) {
    __shared__ float4 ws[K * kHid / 4], wn[K * kHid / 4];     // [k][o/4]
    __shared__ float4 ta[K * 16], tb[K * 16];                 // [k][node/4], 64 nodes
    for (int q = threadIdx.x; q < K * kHid; q += 256) {
        reinterpret_cast<float*>(ws)[q] = bf16_to_f32(Ws[q]);
        reinterpret_cast<float*>(wn)[q] = bf16_to_f32(Wn[q]);
    }
    const int oq = threadIdx.x & 15, rq = threadIdx.x >> 4;   // outputs 4*oq.., nodes 4*rq..
    const float4 bo = make_float4(bf16_to_f32(bias[4 * oq]), bf16_to_f32(bias[4 * oq + 1]),
                                  bf16_to_f32(bias[4 * oq + 2]), bf16_to_f32(bias[4 * oq + 3]));
    const int tiles = (n + 63) / 64;
    for (int tile = blockIdx.x; tile < tiles; tile += gridDim.x) {
        const long long i0 = (long long)tile * 64;
        __syncthreads();                                      // weights written / last tile read
        for (int q = threadIdx.x; q < 64 * (K / 4); q += 256) {
            const int r = q & 63, k4 = q >> 6;                // node r, features 4*k4..
            float4 a = make_float4(0.f, 0.f, 0.f, 0.f), b = a;
            if (i0 + r < n) {
                a = reinterpret_cast<const float4*>(A + (i0 + r) * K)[k4];
                b = reinterpret_cast<const float4*>(B + (i0 + r) * K)[k4];
            }
            float* fa = reinterpret_cast<float*>(ta);
            float* fb = reinterpret_cast<float*>(tb);
            fa[(4 * k4 + 0) * 64 + r] = a.x;
            fa[(4 * k4 + 1) * 64 + r] = a.y;
            fa[(4 * k4 + 2) * 64 + r] = a.z;
            fa[(4 * k4 + 3) * 64 + r] = a.w;
            fb[(4 * k4 + 0) * 64 + r] = b.x;
            fb[(4 * k4 + 1) * 64 + r] = b.y;
            fb[(4 * k4 + 2) * 64 + r] = b.z;
            fb[(4 * k4 + 3) * 64 + r] = b.w;
        }
        __syncthreads();
        float z[4][4];
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            z[r][0] = bo.x;
            z[r][1] = bo.y;
            z[r][2] = bo.z;
            z[r][3] = bo.w;
        }
#pragma unroll 4
        for (int k = 0; k < K; ++k) {
            const float4 a = ta[k * 16 + rq], b = tb[k * 16 + rq];
            const float4 u = ws[k * 16 + oq], v = wn[k * 16 + oq];
            const float av[4] = {a.x, a.y, a.z, a.w}, bv[4] = {b.x, b.y, b.z, b.w};
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                z[r][0] = fmaf(bv[r], v.x, fmaf(av[r], u.x, z[r][0]));
                z[r][1] = fmaf(bv[r], v.y, fmaf(av[r], u.y, z[r][1]));
                z[r][2] = fmaf(bv[r], v.z, fmaf(av[r], u.z, z[r][2]));
                z[r][3] = fmaf(bv[r], v.w, fmaf(av[r], u.w, z[r][3]));
            }
        }
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            const long long i = i0 + 4 * rq + r;
            if (i < n)
                reinterpret_cast<float4*>(H + i * kHid)[oq] =
                    make_float4(fmaxf(z[r][0], 0.f), fmaxf(z[r][1], 0.f), fmaxf(z[r][2], 0.f), fmaxf(z[r][3], 0.f));
        }
    }
}

// Y[i][c] = sum_k H[i][k] Wo[k][c] + bo[c]; thread per node.
__global__ __launch_bounds__(256) void k_head(const float* __restrict__ H, int n, const unsigned short* __restrict__ Wo,
                                              const unsigned short* __restrict__ bo, float* __restrict__ Y) {
    __shared__ float w[kHid * 2];
    if (threadIdx.x < kHid * 2) w[threadIdx.x] = bf16_to_f32(Wo[threadIdx.x]);
    __syncthreads();
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    float y0 = bf16_to_f32(bo[0]), y1 = bf16_to_f32(bo[1]);
    for (int k = 0; k < kHid; ++k) {
        const float h = H[i * kHid + k];
        y0 = fmaf(h, w[2 * k], y0);
        y1 = fmaf(h, w[2 * k + 1], y1);
    }
    Y[2 * i] = y0;
    Y[2 * i + 1] = y1;
}

// ---- loss ----------------------------------------------------------------------------------

// Per node: loss terms over P(i) and dY[i]; per-block loss partials.
__global__ __launch_bounds__(256) void k_loss(const int* __restrict__ rowptr, const int* __restrict__ col, int n,
                                              const float* __restrict__ Y, const float* __restrict__ T,
                                              const int* __restrict__ offs, int R, double inv_pairs,
                                              float* __restrict__ dY, double* __restrict__ part) {
    __shared__ double red[256];
    const int i = blockIdx.x * 256 + threadIdx.x;
    double l = 0.0;
    if (i < n) {
        const float yx = Y[2 * i], yy = Y[2 * i + 1], tx = T[2 * i], ty = T[2 * i + 1];
        float gx = 0.0f, gy = 0.0f;
        const int e0 = rowptr[i], e1 = rowptr[i + 1];
        const int cnt = (e1 - e0) + 2 * R;
        for (int q = 0; q < cnt; ++q) {
            int j;
            if (q < e1 - e0) j = col[e0 + q];
            else {
                const int k = q - (e1 - e0);
                const int o = offs[k >> 1];
                j = (k & 1) ? (int)(((long long)i + n - o) % n) : (int)(((long long)i + o) % n);
            }
            const float dx = yx - Y[2 * j], dy = yy - Y[2 * j + 1];
            const float ex = tx - T[2 * j], ey = ty - T[2 * j + 1];
            const float d = sqrtf(dx * dx + dy * dy) + 1e-12f;
            const float t = sqrtf(ex * ex + ey * ey);
            const float r = d - t;
            l += (double)r * r;
            const float s = r / d;
            gx = fmaf(s, dx, gx);
            gy = fmaf(s, dy, gy);
        }
        const float c = (float)(4.0 * inv_pairs);
        dY[2 * i] = c * gx;
        dY[2 * i + 1] = c * gy;
    }
    red[threadIdx.x] = l;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if (threadIdx.x < o) red[threadIdx.x] += red[threadIdx.x + o];
        __syncthreads();
    }
    if (threadIdx.x == 0) part[blockIdx.x] = red[0] * inv_pairs;
}

// ---- backward ------------------------------------------------------------------------------

// dH2[i][k] = sum_c dY[i][c] Wo[k][c], masked by H2 > 0 (relu): thread per (i, k).
__global__ __launch_bounds__(256) void k_head_back(const float* __restrict__ dY, const float* __restrict__ H2, int n,
                                                   const unsigned short* __restrict__ Wo, float* __restrict__ dZ2) {
    const int t = blockIdx.x * 256 + threadIdx.x;
    if (t >= n * kHid) return;
    const int i = t / kHid, k = t % kHid;
    const float g = fmaf(dY[2 * i], bf16_to_f32(Wo[2 * k]), dY[2 * i + 1] * bf16_to_f32(Wo[2 * k + 1]));
    dZ2[t] = H2[t] > 0.0f ? g : 0.0f;
}

// D2[i][k] = sum_o D[i][o] W[k][o] (dense input gradient, K = 64), one wave per node.
__global__ __launch_bounds__(256) void k_dense_back_x(const float* __restrict__ D, int n,
                                                      const unsigned short* __restrict__ W, float* __restrict__ out) {
    __shared__ float w[kHid * kHid];
    for (int k = threadIdx.x; k < kHid * kHid; k += 256) w[k] = bf16_to_f32(W[k]);
    __syncthreads();
    const int lane = threadIdx.x & 63;
    for (int i = blockIdx.x * 4 + (threadIdx.x >> 6); i < n; i += gridDim.x * 4) {
        const float d = D[i * kHid + lane];
        float s = 0.0f;
#pragma unroll 16
        for (int o = 0; o < kHid; ++o) s = fmaf(__shfl(d, o, 64), w[lane * kHid + o], s);
        out[i * kHid + lane] = s;
    }
}

// dH1 = P + Q masked by H1 > 0.
__global__ __launch_bounds__(256) void k_add_mask(const float* __restrict__ P, const float* __restrict__ Q,
                                                  const float* __restrict__ H, int total, float* __restrict__ out) {
    const int t = blockIdx.x * 256 + threadIdx.x;
    if (t >= total) return;
    out[t] = H[t] > 0.0f ? P[t] + Q[t] : 0.0f;
}

// Weight-gradient partials: for a chunk of 64 nodes, G[k][o] = sum_i A[i][k] D[i][o]
// (A: K columns; column K stands for the bias, A = 1). Thread per (k, o) pairs.
template <int K>
__global__ __launch_bounds__(256) void k_wgrad(const float* __restrict__ A, const float* __restrict__ D, int n,
                                               float* __restrict__ part) {
    __shared__ float sa[64 * K];
    __shared__ float sd[64 * kHid];
    const int i0 = blockIdx.x * 64;
    for (int t = threadIdx.x; t < 64 * K; t += 256) {
        const int i = i0 + t / K;
        sa[t] = i < n ? A[i * K + t % K] : 0.0f;
    }
    for (int t = threadIdx.x; t < 64 * kHid; t += 256) {
        const int i = i0 + t / kHid;
        sd[t] = i < n ? D[i * kHid + t % kHid] : 0.0f;
    }
    __syncthreads();
    float* out = part + (size_t)blockIdx.x * (K + 1) * kHid;
    for (int t = threadIdx.x; t < (K + 1) * kHid; t += 256) {
        const int k = t / kHid, o = t % kHid;
        float s = 0.0f;
        for (int r = 0; r < 64; ++r) s = fmaf(k < K ? sa[r * K + k] : 1.0f, sd[r * kHid + o], s);
        out[t] = s;
    }
}

// Head gradient partials: chunk of 64 nodes, G[k][c] = sum_i H2[i][k] dY[i][c], k = 64 -> bias.
__global__ __launch_bounds__(256) void k_head_wgrad(const float* __restrict__ H, const float* __restrict__ dY, int n,
                                                    float* __restrict__ part) {
    const int i0 = blockIdx.x * 64;
    float* out = part + (size_t)blockIdx.x * (kHid + 1) * 2;
    for (int t = threadIdx.x; t < (kHid + 1) * 2; t += 256) {
        const int k = t >> 1, c = t & 1;
        float s = 0.0f;
        for (int r = 0; r < 64 && i0 + r < n; ++r) {
            const int i = i0 + r;
            s = fmaf(k < kHid ? H[i * kHid + k] : 1.0f, dY[2 * i + c], s);
        }
        out[t] = s;
    }
}

// Fixed-order sum of nb partial blocks of `len` floats into grad (accumulated in double).
__global__ __launch_bounds__(256) void k_sum_parts(const float* __restrict__ part, int nb, int len,
                                                   float* __restrict__ grad) {
    const int t = blockIdx.x * 256 + threadIdx.x;
    if (t >= len) return;
    double s = 0.0;
    for (int b = 0; b < nb; ++b) s += part[(size_t)b * len + t];
    grad[t] = (float)s;
}

// Adam on the FP32 masters, then the BF16 copies.
__global__ __launch_bounds__(256) void k_adam(float* __restrict__ w, const float* __restrict__ g, float* __restrict__ m,
                                              float* __restrict__ v, unsigned short* __restrict__ wb, int len, float lr,
                                              float b1, float b2, float c1, float c2, float eps) {
    const int t = blockIdx.x * 256 + threadIdx.x;
    if (t >= len) return;
    const float gt = g[t];
    const float mt = b1 * m[t] + (1.0f - b1) * gt;
    const float vt = b2 * v[t] + (1.0f - b2) * gt * gt;
    m[t] = mt;
    v[t] = vt;
    const float wt = w[t] - lr * (mt / c1) / (sqrtf(vt / c2) + eps);
    w[t] = wt;
    wb[t] = f32_to_bf16(wt);
}

__global__ __launch_bounds__(256) void k_to_bf16(const float* __restrict__ w, int len, unsigned short* __restrict__ wb) {
    const int t = blockIdx.x * 256 + threadIdx.x;
    if (t < len) wb[t] = f32_to_bf16(w[t]);
}

// Output scaled to the layout size: xy = s * (Y - mean Y) / rms, with mean / rms from
// per-block partials summed in a fixed order.
__global__ __launch_bounds__(256) void k_ystats(const float* __restrict__ Y, int n, double* __restrict__ part) {
    __shared__ double a[256], b[256], c[256];
    double sx = 0.0, sy = 0.0, sq = 0.0;
    for (int i = blockIdx.x * 256 + threadIdx.x; i < n; i += gridDim.x * 256) {
        const double x = Y[2 * i], y = Y[2 * i + 1];
        sx += x;
        sy += y;
        sq += x * x + y * y;
    }
    a[threadIdx.x] = sx;
    b[threadIdx.x] = sy;
    c[threadIdx.x] = sq;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if (threadIdx.x < o) {
            a[threadIdx.x] += a[threadIdx.x + o];
            b[threadIdx.x] += b[threadIdx.x + o];
            c[threadIdx.x] += c[threadIdx.x + o];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        part[3 * blockIdx.x] = a[0];
        part[3 * blockIdx.x + 1] = b[0];
        part[3 * blockIdx.x + 2] = c[0];
    }
}

__global__ __launch_bounds__(256) void k_place(const float* __restrict__ Y, int n, const double* __restrict__ part,
                                               int nb, float scale, float2* __restrict__ xy) {
    __shared__ float coef[3];
    if (threadIdx.x == 0) {
        double sx = 0.0, sy = 0.0, sq = 0.0;
        for (int b = 0; b < nb; ++b) {
            sx += part[3 * b];
            sy += part[3 * b + 1];
            sq += part[3 * b + 2];
        }
        const double mx = sx / n, my = sy / n;
        const double var = sq / n - mx * mx - my * my;
        coef[0] = (float)mx;
        coef[1] = (float)my;
        coef[2] = var > 0 ? (float)(scale / sqrt(var)) : scale;
    }
    __syncthreads();
    for (int i = blockIdx.x * 256 + threadIdx.x; i < n; i += gridDim.x * 256)
        xy[i] = make_float2((Y[2 * i] - coef[0]) * coef[2], (Y[2 * i + 1] - coef[1]) * coef[2]);
}

int blocks_for(long long work) {
    long long b = (work + 255) / 256;
    return (int)(b < 1 ? 1 : b);
}

template <int F, int MODE>
void aggregate(const MpnnGraph& g, const float* in, float* out, hipStream_t st) {
    const long long rb = ((long long)g.n * F + 255) / 256;
    k_agg_rows<F, MODE><<<(int)(rb < 1 ? 1 : (rb > 65536 ? 65536 : rb)), 256, 0, st>>>(g.rowptr, g.col, g.n, in, out);
    if (g.nchunks > 0) {
        const int cb = (g.nchunks + 3) / 4;
        k_agg_chunks<F, MODE><<<cb < 8192 ? cb : 8192, 256, 0, st>>>(g.rowptr, g.col, g.chunks, g.nchunks, in, g.part);
        k_agg_heavy<F, MODE><<<blocks_for((long long)g.nheavy * F), 256, 0, st>>>(g.rowptr, g.heavy_row, g.heavy_c0,
                                                                                 g.nheavy, in, g.part, out);
    }
}

}  // namespace

void mpnn_agg_plan(const int* rowptr, int n, std::vector<int>& heavy_row, std::vector<int>& heavy_c0,
                   std::vector<int2>& chunks) {
    heavy_row.clear();
    heavy_c0.assign(1, 0);
    chunks.clear();
    for (int i = 0; i < n; ++i) {
        const int e0 = rowptr[i], e1 = rowptr[i + 1];
        if (e1 - e0 <= kHeavyDeg) continue;
        heavy_row.push_back(i);
        for (int e = e0; e < e1; e += kChunk) chunks.push_back(make_int2(i, e));
        heavy_c0.push_back((int)chunks.size());
    }
}

void mpnn_features(const MpnnGraph& g, unsigned long long seed, int steps, MpnnScratch& s, hipStream_t st) {
    const int n = g.n;
    k_normal_fields<<<blocks_for((long long)n * 8), 256, 0, st>>>(n, seed, s.z0);
    float* a = s.z0;
    float* b = s.z1;
    for (int k = 0; k < steps; ++k) {
        aggregate<8, kAggSmooth>(g, a, b, st);
        float* t = a;
        a = b;
        b = t;
    }
    k_assemble<<<blocks_for((long long)n * kIn), 256, 0, st>>>(g.rowptr, n, a, g.X);
    const int nb = 256;
    k_colstats<<<nb, 256, 0, st>>>(g.X, n, s.dpart);
    k_colnorm_finish<<<1, 64, 0, st>>>(s.dpart, nb, n, s.musd);
    k_colnorm_apply<<<blocks_for((long long)n * kIn), 256, 0, st>>>(g.X, n, s.musd);
    aggregate<kIn, kAggMean>(g, g.X, g.A0, st);
}

void mpnn_forward(const MpnnGraph& g, const MpnnWeights& w, MpnnScratch& s, hipStream_t st) {
    const int n = g.n;
    const int db = (n + 63) / 64 < 1024 ? (n + 63) / 64 : 1024;
    k_dense<kIn><<<db, 256, 0, st>>>(g.X, g.A0, n, w.bf + kOffW1s, w.bf + kOffW1n, w.bf + kOffB1, s.H1);
    aggregate<kHid, kAggMean>(g, s.H1, s.A1, st);
    k_dense<kHid><<<db, 256, 0, st>>>(s.H1, s.A1, n, w.bf + kOffW2s, w.bf + kOffW2n, w.bf + kOffB2, s.H2);
    k_head<<<blocks_for(n), 256, 0, st>>>(s.H2, n, w.bf + kOffWo, w.bf + kOffBo, s.Y);
}

int mpnn_loss_backward(const MpnnGraph& g, const MpnnWeights& w, MpnnScratch& s, const int* d_offs, int R,
                       double* loss, hipStream_t st) {
    const int n = g.n;
    const long long pairs = (long long)g.nnz + 2ll * R * n;
    const double inv = 1.0 / (double)pairs;
    const int lb = blocks_for(n);
    k_loss<<<lb, 256, 0, st>>>(g.rowptr, g.col, n, s.Y, g.T, d_offs, R, inv, s.dY, s.dpart);
    // head
    const int chunks = (n + 63) / 64;
    k_head_wgrad<<<chunks, 256, 0, st>>>(s.H2, s.dY, n, s.wpart);
    k_sum_parts<<<blocks_for((kHid + 1) * 2), 256, 0, st>>>(s.wpart, chunks, (kHid + 1) * 2, w.grad + kOffWo);
    k_head_back<<<blocks_for((long long)n * kHid), 256, 0, st>>>(s.dY, s.H2, n, w.bf + kOffWo, s.dZ2);
    // layer 2: grads of W2s (+b2) from H1, of W2n from A1
    k_wgrad<kHid><<<chunks, 256, 0, st>>>(s.H1, s.dZ2, n, s.wpart);
    k_sum_parts<<<blocks_for((kHid + 1) * kHid), 256, 0, st>>>(s.wpart, chunks, (kHid + 1) * kHid, s.gtmp);
    // gtmp rows 0..63 -> W2s, row 64 -> b2
    hipMemcpyAsync(w.grad + kOffW2s, s.gtmp, sizeof(float) * kHid * kHid, hipMemcpyDeviceToDevice, st);
    hipMemcpyAsync(w.grad + kOffB2, s.gtmp + kHid * kHid, sizeof(float) * kHid, hipMemcpyDeviceToDevice, st);
    k_wgrad<kHid><<<chunks, 256, 0, st>>>(s.A1, s.dZ2, n, s.wpart);
    k_sum_parts<<<blocks_for(kHid * kHid), 256, 0, st>>>(s.wpart, chunks, (kHid + 1) * kHid, s.gtmp);
    hipMemcpyAsync(w.grad + kOffW2n, s.gtmp, sizeof(float) * kHid * kHid, hipMemcpyDeviceToDevice, st);
    // dH1 = dZ2 W2s^T + agg^T(dZ2 W2n^T), masked
    const int db = (n + 3) / 4 < 2048 ? (n + 3) / 4 : 2048;
    k_dense_back_x<<<db, 256, 0, st>>>(s.dZ2, n, w.bf + kOffW2s, s.P);
    k_dense_back_x<<<db, 256, 0, st>>>(s.dZ2, n, w.bf + kOffW2n, s.Q);
    aggregate<kHid, kAggTranspose>(g, s.Q, s.A1, st);
    k_add_mask<<<blocks_for((long long)n * kHid), 256, 0, st>>>(s.P, s.A1, s.H1, n * kHid, s.dZ1);
    // layer 1
    k_wgrad<kIn><<<chunks, 256, 0, st>>>(g.X, s.dZ1, n, s.wpart);
    k_sum_parts<<<blocks_for((kIn + 1) * kHid), 256, 0, st>>>(s.wpart, chunks, (kIn + 1) * kHid, s.gtmp);
    hipMemcpyAsync(w.grad + kOffW1s, s.gtmp, sizeof(float) * kIn * kHid, hipMemcpyDeviceToDevice, st);
    hipMemcpyAsync(w.grad + kOffB1, s.gtmp + kIn * kHid, sizeof(float) * kHid, hipMemcpyDeviceToDevice, st);
    k_wgrad<kIn><<<chunks, 256, 0, st>>>(g.A0, s.dZ1, n, s.wpart);
    k_sum_parts<<<blocks_for(kIn * kHid), 256, 0, st>>>(s.wpart, chunks, (kIn + 1) * kHid, s.gtmp);
    hipMemcpyAsync(w.grad + kOffW1n, s.gtmp, sizeof(float) * kIn * kHid, hipMemcpyDeviceToDevice, st);
    if (loss) {
        std::vector<double> h((size_t)lb);
        if (hipMemcpyAsync(h.data(), s.dpart, sizeof(double) * lb, hipMemcpyDeviceToHost, st) != hipSuccess) return -1;
        if (hipStreamSynchronize(st) != hipSuccess) return -1;
        double L = 0.0;
        for (double v : h) L += v;
        *loss = L;
    }
    return hipGetLastError() == hipSuccess ? 0 : -1;
}

void mpnn_adam(MpnnWeights& w, float lr, int step, hipStream_t st) {
    const float b1 = 0.9f, b2 = 0.999f;
    const float c1 = 1.0f - powf(b1, (float)step), c2 = 1.0f - powf(b2, (float)step);
    k_adam<<<blocks_for(kParams), 256, 0, st>>>(w.master, w.grad, w.m, w.v, w.bf, kParams, lr, b1, b2, c1, c2, 1e-8f);
}

void mpnn_sync_bf16(MpnnWeights& w, hipStream_t st) { k_to_bf16<<<blocks_for(kParams), 256, 0, st>>>(w.master, kParams, w.bf); }

void mpnn_place(const MpnnGraph& g, MpnnScratch& s, float scale, float2* xy, hipStream_t st) {
    const int nb = 128;
    k_ystats<<<nb, 256, 0, st>>>(s.Y, g.n, s.dpart);
    k_place<<<blocks_for(g.n) < 1024 ? blocks_for(g.n) : 1024, 256, 0, st>>>(s.Y, g.n, s.dpart, nb, scale, xy);
}

}  // namespace gfa2
