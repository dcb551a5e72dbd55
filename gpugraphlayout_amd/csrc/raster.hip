// Rasteriser of a layout on the device: the picture GraphLayout::writeToPNG draws
// (src/RPGraphLayout.cpp:170-207: bounding box fitted to the image, white background,
// black nodes as discs of radius 3 with opacity 10000/V, black edges as lines with opacity
// 100000/E). Each node and each edge (drawn once, from its smaller to its larger id) adds
// to per-pixel hit counters with integer atomics; a pixel's brightness is then
// (1 - a_node)^nodes * (1 - a_edge)^edges, the result of blending that many layers in any
// order. The host layer's CPU rasteriser uses the same pixel walk and the same shading
// arithmetic (powers by repeated squaring, IEEE operations), so both produce the same bytes;
// this file is compiled with -ffp-contract=off, since the *_rn intrinsics alone do not stop
// a multiply and an add from being fused.
#include "engine_internal.h"
#include "../../include/gfa2.h"

#include <mutex>

// Fixed scratch of gfa2_render_noalloc (one copy per device, loaded with the code object and
// held until the process ends): the node and edge hit counters, 2 x kNoAllocMaxPixels words,
// then the picture, kNoAllocMaxPixels bytes; 9 MiB in all. Host code never passes this name
// to a kernel or a copy: it uses the device address hipGetSymbolAddress gives for it.
namespace gfa2 {
constexpr long long kNoAllocMaxPixels = 1ll << 20;
constexpr size_t kScratchWords = (size_t)(2 * kNoAllocMaxPixels + kNoAllocMaxPixels / 4);
}  // namespace gfa2
static __device__ __align__(256) unsigned int scratch[gfa2::kScratchWords];

namespace gfa2 {
namespace {

constexpr int kScratchMaxDevices = 64;
std::mutex g_scratch_mu;                               // also serialises the calls of one process
unsigned int* g_scratch[kScratchMaxDevices] = {};      // device address of `scratch`, per device

__global__ __launch_bounds__(256) void k_fill_zero(unsigned int* __restrict__ p, long long n) {
    for (long long i = blockIdx.x * 256ll + threadIdx.x; i < n; i += (long long)gridDim.x * 256) p[i] = 0u;
}

struct Frame {
    float x0, y0, sx, sy;
    int w, h;
};

__device__ __forceinline__ void hit(unsigned int* __restrict__ cnt, const Frame& f, int x, int y) {
    if (x < 0 || y < 0 || x >= f.w || y >= f.h) return;
    atomicAdd(&cnt[(size_t)(f.h - 1 - y) * f.w + x], 1u);   // y up
}

__device__ __forceinline__ float to_px(float v, float v0, float s) { return __fmul_rn(__fsub_rn(v, v0), s); }

__global__ __launch_bounds__(256) void k_minmax(const float2* __restrict__ pos, int n, float4* __restrict__ part) {
    __shared__ float4 red[256];
    float4 m = make_float4(3.402823466e38f, -3.402823466e38f, 3.402823466e38f, -3.402823466e38f);
    for (int i = blockIdx.x * 256 + threadIdx.x; i < n; i += gridDim.x * 256) {
        const float2 p = pos[i];
        m.x = fminf(m.x, p.x);
        m.y = fmaxf(m.y, p.x);
        m.z = fminf(m.z, p.y);
        m.w = fmaxf(m.w, p.y);
    }
    red[threadIdx.x] = m;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if ((int)threadIdx.x < o) {
            const float4 a = red[threadIdx.x], b = red[threadIdx.x + o];
            red[threadIdx.x] = make_float4(fminf(a.x, b.x), fmaxf(a.y, b.y), fminf(a.z, b.z), fmaxf(a.w, b.w));
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) part[blockIdx.x] = red[0];
}

__global__ void k_minmax_final(const float4* __restrict__ part, int nb, float4* __restrict__ out) {
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    float4 m = part[0];
    for (int b = 1; b < nb; ++b) {
        const float4 a = part[b];
        m = make_float4(fminf(m.x, a.x), fmaxf(m.y, a.y), fminf(m.z, a.z), fmaxf(m.w, a.w));
    }
    *out = m;
}

// Discs of radius 3 around the rounded node centre.
__global__ __launch_bounds__(256) void k_raster_nodes(const float2* __restrict__ pos, int n, Frame f,
                                                      unsigned int* __restrict__ cnt) {
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    const float2 p = pos[i];
    const int cx = (int)lroundf(to_px(p.x, f.x0, f.sx)), cy = (int)lroundf(to_px(p.y, f.y0, f.sy));
    for (int dy = -3; dy <= 3; ++dy)
        for (int dx = -3; dx <= 3; ++dx)
            if (dx * dx + dy * dy <= 9) hit(cnt, f, cx + dx, cy + dy);
}

// One thread per CSR entry (i, j) with i < j: the line from node i to node j, sampled at
// ceil(max(|dx|, |dy|)) + 1 evenly spaced points (positions formed in double), each rounded
// to the nearest pixel.
__global__ __launch_bounds__(256) void k_raster_edges(const float2* __restrict__ pos, const int* __restrict__ rowptr,
                                                      const int* __restrict__ col, int n, long long nnz, Frame f,
                                                      unsigned int* __restrict__ cnt) {
    for (long long e = blockIdx.x * 256ll + threadIdx.x; e < nnz; e += (long long)gridDim.x * 256) {
        const int j = col[e];
        int lo = 0, hi = n;              // row: rowptr[lo] <= e < rowptr[lo + 1]
        while (hi - lo > 1) {
            const int mid = (lo + hi) >> 1;
            if ((long long)rowptr[mid] <= e) lo = mid;
            else hi = mid;
        }
        const int i = lo;
        if (!(i < j)) continue;
        const float2 a = pos[i], b = pos[j];
        const float x0 = to_px(a.x, f.x0, f.sx), y0 = to_px(a.y, f.y0, f.sy);
        const float x1 = to_px(b.x, f.x0, f.sx), y1 = to_px(b.y, f.y0, f.sy);
        const float ddx = __fsub_rn(x1, x0), ddy = __fsub_rn(y1, y0);
        const float len = fmaxf(fabsf(ddx), fabsf(ddy));
        const int steps = (int)ceilf(len);
        if (steps == 0) {
            hit(cnt, f, (int)lroundf(x0), (int)lroundf(y0));
            continue;
        }
        // sample positions in double (IEEE-exact divide on both processors), then float
        for (int k = 0; k <= steps; ++k) {
            const double t = __ddiv_rn((double)k, (double)steps);
            const float sx = (float)__dadd_rn((double)x0, __dmul_rn(t, (double)ddx));
            const float sy = (float)__dadd_rn((double)y0, __dmul_rn(t, (double)ddy));
            hit(cnt, f, (int)lroundf(sx), (int)lroundf(sy));
        }
    }
}

__device__ __forceinline__ float powi(float b, unsigned int k) {
    float r = 1.0f;
    while (k) {
        if (k & 1u) r = __fmul_rn(r, b);
        b = __fmul_rn(b, b);
        k >>= 1;
    }
    return r;
}

__global__ __launch_bounds__(256) void k_shade(const unsigned int* __restrict__ cn, const unsigned int* __restrict__ ce,
                                               long long npix, float keep_n, float keep_e,
                                               unsigned char* __restrict__ out) {
    for (long long p = blockIdx.x * 256ll + threadIdx.x; p < npix; p += (long long)gridDim.x * 256) {
        const float keep = __fmul_rn(powi(keep_n, cn[p]), powi(keep_e, ce[p]));
        out[p] = (unsigned char)lroundf(__fmul_rn(255.0f, keep));
    }
}

}  // namespace
}  // namespace gfa2

using namespace gfa2;

extern "C" {

int gfa2_get_bbox(gfa2_handle h, float* out) {
    const Bufs* b = nullptr;
    int device = 0;
    hipStream_t s = nullptr;
    if (!h || !out) return GFA2_ERR_INVALID_ARG;
    if (!engine_graph(h, &b, &device, &s)) return GFA2_ERR_NO_GRAPH;
    if (hipSetDevice(device) != hipSuccess) return GFA2_ERR_HIP;
    const int nb = 256;
    float4* part = nullptr;
    if (hipMalloc(&part, sizeof(float4) * (nb + 1)) != hipSuccess) return GFA2_ERR_OOM;
    k_minmax<<<nb, 256, 0, s>>>(b->pos, b->n, part);
    k_minmax_final<<<1, 64, 0, s>>>(part, nb, part + nb);
    float4 m;
    hipError_t e = hipGetLastError();
    if (e == hipSuccess) e = hipMemcpyAsync(&m, part + nb, sizeof(float4), hipMemcpyDeviceToHost, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    (void)hipFree(part);
    if (e != hipSuccess) return GFA2_ERR_HIP;
    out[0] = m.x;
    out[1] = m.y;
    out[2] = m.z;
    out[3] = m.w;
    return GFA2_OK;
}

int gfa2_render(gfa2_handle h, const gfa2_raster* r, uint8_t* gray) {
    const Bufs* b = nullptr;
    int device = 0;
    hipStream_t s = nullptr;
    if (!h || !r || !gray || r->width <= 0 || r->height <= 0) return GFA2_ERR_INVALID_ARG;
    if (!engine_graph(h, &b, &device, &s)) return GFA2_ERR_NO_GRAPH;
    if (b->nseg > 1) return GFA2_ERR_UNSUPPORTED;
    if (hipSetDevice(device) != hipSuccess) return GFA2_ERR_HIP;
    const long long npix = (long long)r->width * r->height;
    unsigned int* cnt = nullptr;
    unsigned char* img = nullptr;
    if (hipMalloc(&cnt, sizeof(unsigned int) * 2 * (size_t)npix) != hipSuccess) return GFA2_ERR_OOM;
    if (hipMalloc(&img, (size_t)npix) != hipSuccess) {
        (void)hipFree(cnt);
        return GFA2_ERR_OOM;
    }
    const Frame f{r->x0, r->y0, r->sx, r->sy, r->width, r->height};
    hipError_t e = hipMemsetAsync(cnt, 0, sizeof(unsigned int) * 2 * (size_t)npix, s);
    if (e == hipSuccess) {
        k_raster_nodes<<<(b->n + 255) / 256, 256, 0, s>>>(b->pos, b->n, f, cnt);
        long long eb = (b->nnz + 255) / 256;
        if (eb > 8192) eb = 8192;
        if (eb > 0) k_raster_edges<<<(int)eb, 256, 0, s>>>(b->pos, b->rowptr, b->col, b->n, b->nnz, f, cnt + npix);
        long long pb = (npix + 255) / 256;
        if (pb > 8192) pb = 8192;
        k_shade<<<(int)pb, 256, 0, s>>>(cnt, cnt + npix, npix, 1.0f - r->node_alpha, 1.0f - r->edge_alpha, img);
        e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipMemcpyAsync(gray, img, (size_t)npix, hipMemcpyDeviceToHost, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    (void)hipFree(cnt);
    (void)hipFree(img);
    return e == hipSuccess ? GFA2_OK : GFA2_ERR_HIP;
}

int gfa2_render_noalloc(gfa2_handle h, const gfa2_raster* r, uint8_t* gray) {
    const Bufs* b = nullptr;
    int device = 0;
    hipStream_t s = nullptr;
    if (!h || !r || !gray || r->width <= 0 || r->height <= 0) return GFA2_ERR_INVALID_ARG;
    const long long npix = (long long)r->width * r->height;
    if (npix > kNoAllocMaxPixels) return GFA2_ERR_SCRATCH_LIMIT;
    if (!engine_graph(h, &b, &device, &s)) return GFA2_ERR_NO_GRAPH;
    if (b->nseg > 1) return GFA2_ERR_UNSUPPORTED;
    if (device < 0 || device >= kScratchMaxDevices) return GFA2_ERR_UNSUPPORTED;
    if (hipSetDevice(device) != hipSuccess) return GFA2_ERR_HIP;
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch[device]) {
        // once per device, with that device current: this device's copy of the array
        void* p = nullptr;
        if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) != hipSuccess || !p) return GFA2_ERR_HIP;
        g_scratch[device] = static_cast<unsigned int*>(p);
    }
    unsigned int* cnt = g_scratch[device];                       // node counters [0, npix), edges [npix, 2 npix)
    unsigned char* img = reinterpret_cast<unsigned char*>(cnt + 2 * kNoAllocMaxPixels);
    const Frame f{r->x0, r->y0, r->sx, r->sy, r->width, r->height};
    // the counters keep the previous call's hits: cleared by a kernel (no memset node)
    long long zb = (2 * npix + 255) / 256;
    if (zb > 8192) zb = 8192;
    k_fill_zero<<<(int)zb, 256, 0, s>>>(cnt, 2 * npix);
    k_raster_nodes<<<(b->n + 255) / 256, 256, 0, s>>>(b->pos, b->n, f, cnt);
    long long eb = (b->nnz + 255) / 256;
    if (eb > 8192) eb = 8192;
    if (eb > 0) k_raster_edges<<<(int)eb, 256, 0, s>>>(b->pos, b->rowptr, b->col, b->n, b->nnz, f, cnt + npix);
    long long pb = (npix + 255) / 256;
    if (pb > 8192) pb = 8192;
    k_shade<<<(int)pb, 256, 0, s>>>(cnt, cnt + npix, npix, 1.0f - r->node_alpha, 1.0f - r->edge_alpha, img);
    hipError_t e = hipGetLastError();
    if (e == hipSuccess) e = hipMemcpyAsync(gray, img, (size_t)npix, hipMemcpyDeviceToHost, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    return e == hipSuccess ? GFA2_OK : GFA2_ERR_HIP;
}

}  // extern "C"
