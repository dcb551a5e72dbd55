// C ABI of the learned initialiser (include/gfa2.h, gfa2_mpnn_*).
#include "engine_internal.h"
#include "mpnn_internal.h"
#include "../../include/gfa2.h"

#include <cmath>
#include <new>
#include <vector>

using namespace gfa2;

struct gfa2_mpnn {
    int device = 0;
    hipStream_t stream = nullptr;
    MpnnWeights w;
    MpnnScratch s;
    std::vector<MpnnGraph> graphs;
    std::vector<void*> allocs;          // weights and scratch
    std::vector<std::vector<void*>> graph_allocs;
    int adam_step = 0;
    int* offs = nullptr;                // training offsets, steps x R
    long long offs_cap = 0;
};

namespace {

int st_of(hipError_t e) { return e == hipSuccess ? GFA2_OK : (e == hipErrorOutOfMemory ? GFA2_ERR_OOM : GFA2_ERR_HIP); }

#define MP_TRY(expr)                           \
    do {                                       \
        hipError_t _e = (expr);                \
        if (_e != hipSuccess) return st_of(_e); \
    } while (0)

template <typename T>
int mp_alloc(std::vector<void*>& list, T** p, size_t count) {
    void* q = nullptr;
    hipError_t e = hipMalloc(&q, sizeof(T) * (count ? count : 1));
    if (e != hipSuccess) return st_of(e);
    list.push_back(q);
    *p = static_cast<T*>(q);
    return GFA2_OK;
}

struct SplitMix {
    uint64_t x;
    uint64_t next() {
        uint64_t z = (x += 0x9E3779B97F4A7C15ull);
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
        return z ^ (z >> 31);
    }
    double uniform() { return (double)(next() >> 11) * (1.0 / 9007199254740992.0); }
};

// Scratch sized for n nodes (grows, never shrinks).
int ensure_scratch(gfa2_mpnn* m, int n) {
    MpnnScratch& s = m->s;
    if (n <= s.cap) return GFA2_OK;
    for (void* p : {(void*)s.z0, (void*)s.z1, (void*)s.H1, (void*)s.A1, (void*)s.H2, (void*)s.Y, (void*)s.dY,
                    (void*)s.dZ2, (void*)s.P, (void*)s.Q, (void*)s.dZ1, (void*)s.wpart})
        if (p) (void)hipFree(p);
    const size_t nh = (size_t)n * kHid;
    const size_t chunks = ((size_t)n + 63) / 64;
    MP_TRY(hipMalloc(&s.z0, sizeof(float) * (size_t)n * 8));
    MP_TRY(hipMalloc(&s.z1, sizeof(float) * (size_t)n * 8));
    MP_TRY(hipMalloc(&s.H1, sizeof(float) * nh));
    MP_TRY(hipMalloc(&s.A1, sizeof(float) * nh));
    MP_TRY(hipMalloc(&s.H2, sizeof(float) * nh));
    MP_TRY(hipMalloc(&s.Y, sizeof(float) * (size_t)n * 2));
    MP_TRY(hipMalloc(&s.dY, sizeof(float) * (size_t)n * 2));
    MP_TRY(hipMalloc(&s.dZ2, sizeof(float) * nh));
    MP_TRY(hipMalloc(&s.P, sizeof(float) * nh));
    MP_TRY(hipMalloc(&s.Q, sizeof(float) * nh));
    MP_TRY(hipMalloc(&s.dZ1, sizeof(float) * nh));
    MP_TRY(hipMalloc(&s.wpart, sizeof(float) * chunks * (kHid + 1) * kHid));
    s.cap = n;
    return GFA2_OK;
}

// Heavy rows and chunks of a graph's aggregation, from its row offsets on the host.
int upload_agg_plan(gfa2_mpnn* m, const int* rp, MpnnGraph& g, std::vector<void*>& mine) {
    std::vector<int> heavy_row, heavy_c0;
    std::vector<int2> chunks;
    mpnn_agg_plan(rp, g.n, heavy_row, heavy_c0, chunks);
    g.nheavy = (int)heavy_row.size();
    g.nchunks = (int)chunks.size();
    if (g.nchunks == 0) return GFA2_OK;
    int st;
    if ((st = mp_alloc(mine, &g.heavy_row, heavy_row.size())) != GFA2_OK ||
        (st = mp_alloc(mine, &g.heavy_c0, heavy_c0.size())) != GFA2_OK ||
        (st = mp_alloc(mine, &g.chunks, chunks.size())) != GFA2_OK ||
        (st = mp_alloc(mine, &g.part, chunks.size() * (size_t)kHid)) != GFA2_OK)
        return st;
    MP_TRY(hipMemcpyAsync(g.heavy_row, heavy_row.data(), sizeof(int) * heavy_row.size(), hipMemcpyHostToDevice, m->stream));
    MP_TRY(hipMemcpyAsync(g.heavy_c0, heavy_c0.data(), sizeof(int) * heavy_c0.size(), hipMemcpyHostToDevice, m->stream));
    MP_TRY(hipMemcpyAsync(g.chunks, chunks.data(), sizeof(int2) * chunks.size(), hipMemcpyHostToDevice, m->stream));
    MP_TRY(hipStreamSynchronize(m->stream));   // the host vectors go out of scope
    return GFA2_OK;
}

int check_graph(const gfa2_mpnn* m, int id) {
    if (!m || id < 0 || id >= (int)m->graphs.size() || m->graphs[id].n == 0) return GFA2_ERR_INVALID_ARG;
    return GFA2_OK;
}

}  // namespace

extern "C" {

int32_t gfa2_mpnn_num_params(void) { return kParams; }

int gfa2_mpnn_create(int device, uint64_t seed, gfa2_mpnn_t* out) {
    if (!out) return GFA2_ERR_INVALID_ARG;
    *out = nullptr;
    int count = 0;
    if (hipGetDeviceCount(&count) != hipSuccess || count == 0) return GFA2_ERR_NO_DEVICE;
    if (device < 0 || device >= count) return GFA2_ERR_INVALID_ARG;
    MP_TRY(hipSetDevice(device));
    gfa2_mpnn* m = new (std::nothrow) gfa2_mpnn();
    if (!m) return GFA2_ERR_OOM;
    m->device = device;
    int st = GFA2_OK;
    if (hipStreamCreateWithFlags(&m->stream, hipStreamNonBlocking) != hipSuccess ||
        (st = mp_alloc(m->allocs, &m->w.master, kParams)) != GFA2_OK ||
        (st = mp_alloc(m->allocs, &m->w.bf, kParams)) != GFA2_OK ||
        (st = mp_alloc(m->allocs, &m->w.grad, kParams)) != GFA2_OK ||
        (st = mp_alloc(m->allocs, &m->w.m, kParams)) != GFA2_OK ||
        (st = mp_alloc(m->allocs, &m->w.v, kParams)) != GFA2_OK ||
        (st = mp_alloc(m->allocs, &m->s.dpart, 8192)) != GFA2_OK ||
        (st = mp_alloc(m->allocs, &m->s.musd, 2 * kIn)) != GFA2_OK ||
        (st = mp_alloc(m->allocs, &m->s.gtmp, (kHid + 1) * kHid)) != GFA2_OK) {
        gfa2_mpnn_destroy(m);
        return st != GFA2_OK ? st : GFA2_ERR_HIP;
    }
    // Glorot-uniform weights (fan_in + fan_out of each matrix), zero biases.
    std::vector<float> w(kParams, 0.0f);
    SplitMix rng{seed};
    auto fill = [&](int off, int fan_in, int fan_out, int rows_used) {
        const double a = std::sqrt(6.0 / (fan_in + fan_out));
        for (int k = 0; k < rows_used; ++k)
            for (int o = 0; o < fan_out; ++o) w[off + k * fan_out + o] = (float)((2.0 * rng.uniform() - 1.0) * a);
    };
    fill(kOffW1s, 9, kHid, 9);
    fill(kOffW1n, 9, kHid, 9);
    fill(kOffW2s, kHid, kHid, kHid);
    fill(kOffW2n, kHid, kHid, kHid);
    fill(kOffWo, kHid, 2, kHid);
    st = gfa2_mpnn_set_params(m, w.data());
    if (st != GFA2_OK) {
        gfa2_mpnn_destroy(m);
        return st;
    }
    *out = m;
    return GFA2_OK;
}

int gfa2_mpnn_set_params(gfa2_mpnn_t m, const float* params) {
    if (!m || !params) return GFA2_ERR_INVALID_ARG;
    MP_TRY(hipSetDevice(m->device));
    MP_TRY(hipMemcpyAsync(m->w.master, params, sizeof(float) * kParams, hipMemcpyHostToDevice, m->stream));
    MP_TRY(hipMemsetAsync(m->w.m, 0, sizeof(float) * kParams, m->stream));
    MP_TRY(hipMemsetAsync(m->w.v, 0, sizeof(float) * kParams, m->stream));
    mpnn_sync_bf16(m->w, m->stream);
    MP_TRY(hipStreamSynchronize(m->stream));
    m->adam_step = 0;
    return GFA2_OK;
}

int gfa2_mpnn_get_params(gfa2_mpnn_t m, float* params) {
    if (!m || !params) return GFA2_ERR_INVALID_ARG;
    MP_TRY(hipSetDevice(m->device));
    MP_TRY(hipMemcpyAsync(params, m->w.master, sizeof(float) * kParams, hipMemcpyDeviceToHost, m->stream));
    MP_TRY(hipStreamSynchronize(m->stream));
    return GFA2_OK;
}

int gfa2_mpnn_add_graph(gfa2_mpnn_t m, int32_t n, const int64_t* rowptr, const int32_t* col, const float* target_xy,
                        uint64_t feature_seed, int32_t smoothing_steps, int32_t* graph_id) {
    if (!m || !rowptr || n < 1 || smoothing_steps < 0) return GFA2_ERR_INVALID_ARG;
    const int64_t nnz = rowptr[n];
    if (rowptr[0] != 0 || nnz < 0 || nnz >= (int64_t)1 << 31 || (nnz > 0 && !col)) return GFA2_ERR_INVALID_GRAPH;
    std::vector<int> rp((size_t)n + 1);
    for (int v = 0; v <= n; ++v) {
        if (v < n && (rowptr[v + 1] < rowptr[v])) return GFA2_ERR_INVALID_GRAPH;
        rp[v] = (int)rowptr[v];
    }
    for (int64_t e = 0; e < nnz; ++e)
        if (col[e] < 0 || col[e] >= n) return GFA2_ERR_INVALID_GRAPH;
    MP_TRY(hipSetDevice(m->device));
    int st = ensure_scratch(m, n);
    if (st != GFA2_OK) return st;
    MpnnGraph g;
    g.n = n;
    g.nnz = nnz;
    std::vector<void*> mine;
    if ((st = mp_alloc(mine, &g.rowptr, (size_t)n + 1)) != GFA2_OK || (st = mp_alloc(mine, &g.col, (size_t)nnz)) != GFA2_OK ||
        (st = mp_alloc(mine, &g.X, (size_t)n * kIn)) != GFA2_OK || (st = mp_alloc(mine, &g.A0, (size_t)n * kIn)) != GFA2_OK) {
        for (void* p : mine) (void)hipFree(p);
        return st;
    }
    MP_TRY(hipMemcpyAsync(g.rowptr, rp.data(), sizeof(int) * (n + 1), hipMemcpyHostToDevice, m->stream));
    if (nnz) MP_TRY(hipMemcpyAsync(g.col, col, sizeof(int) * nnz, hipMemcpyHostToDevice, m->stream));
    if ((st = upload_agg_plan(m, rp.data(), g, mine)) != GFA2_OK) {
        for (void* p : mine) (void)hipFree(p);
        return st;
    }
    if (target_xy) {
        // centre and scale to unit RMS radius
        double cx = 0.0, cy = 0.0;
        for (int i = 0; i < n; ++i) {
            cx += target_xy[2 * i];
            cy += target_xy[2 * i + 1];
        }
        cx /= n;
        cy /= n;
        double r2 = 0.0;
        for (int i = 0; i < n; ++i) {
            const double dx = target_xy[2 * i] - cx, dy = target_xy[2 * i + 1] - cy;
            r2 += dx * dx + dy * dy;
        }
        const double rms = std::sqrt(r2 / n);
        std::vector<float> t((size_t)2 * n);
        for (int i = 0; i < n; ++i) {
            t[2 * i] = (float)((target_xy[2 * i] - cx) / (rms > 0 ? rms : 1.0));
            t[2 * i + 1] = (float)((target_xy[2 * i + 1] - cy) / (rms > 0 ? rms : 1.0));
        }
        if ((st = mp_alloc(mine, &g.T, (size_t)2 * n)) != GFA2_OK) {
            for (void* p : mine) (void)hipFree(p);
            return st;
        }
        g.target_scale = (float)rms;
        MP_TRY(hipMemcpyAsync(g.T, t.data(), sizeof(float) * 2 * n, hipMemcpyHostToDevice, m->stream));
    }
    mpnn_features(g, feature_seed, smoothing_steps, m->s, m->stream);
    MP_TRY(hipGetLastError());
    MP_TRY(hipStreamSynchronize(m->stream));
    m->graphs.push_back(g);
    m->graph_allocs.push_back(mine);
    if (graph_id) *graph_id = (int32_t)m->graphs.size() - 1;
    return GFA2_OK;
}

int gfa2_mpnn_add_engine_graph(gfa2_mpnn_t m, gfa2_handle h, uint64_t feature_seed, int32_t smoothing_steps,
                               int32_t* graph_id) {
    if (!m || !h || smoothing_steps < 0) return GFA2_ERR_INVALID_ARG;
    EngineView v;
    if (!engine_view(h, &v)) return GFA2_ERR_NO_GRAPH;
    if (v.nseg > 1) return GFA2_ERR_UNSUPPORTED;
    if (v.device != m->device) return GFA2_ERR_INVALID_ARG;
    MP_TRY(hipSetDevice(m->device));
    const int n = v.n;
    int st = ensure_scratch(m, n);
    if (st != GFA2_OK) return st;
    MpnnGraph g;
    g.n = n;
    g.nnz = v.nnz;
    std::vector<void*> mine;
    if ((st = mp_alloc(mine, &g.rowptr, (size_t)n + 1)) != GFA2_OK || (st = mp_alloc(mine, &g.col, (size_t)g.nnz)) != GFA2_OK ||
        (st = mp_alloc(mine, &g.X, (size_t)n * kIn)) != GFA2_OK || (st = mp_alloc(mine, &g.A0, (size_t)n * kIn)) != GFA2_OK) {
        for (void* p : mine) (void)hipFree(p);
        return st;
    }
    // the engine's CSR was validated when it was set; copy it device to device after the
    // engine's queued work
    MP_TRY(hipStreamSynchronize(v.stream));
    MP_TRY(hipMemcpyAsync(g.rowptr, v.rowptr, sizeof(int) * ((size_t)n + 1), hipMemcpyDeviceToDevice, m->stream));
    if (g.nnz) MP_TRY(hipMemcpyAsync(g.col, v.col, sizeof(int) * (size_t)g.nnz, hipMemcpyDeviceToDevice, m->stream));
    std::vector<int> rp((size_t)n + 1);
    MP_TRY(hipMemcpyAsync(rp.data(), v.rowptr, sizeof(int) * ((size_t)n + 1), hipMemcpyDeviceToHost, m->stream));
    MP_TRY(hipStreamSynchronize(m->stream));
    if ((st = upload_agg_plan(m, rp.data(), g, mine)) != GFA2_OK) {
        for (void* p : mine) (void)hipFree(p);
        return st;
    }
    mpnn_features(g, feature_seed, smoothing_steps, m->s, m->stream);
    MP_TRY(hipGetLastError());
    MP_TRY(hipStreamSynchronize(m->stream));
    m->graphs.push_back(g);
    m->graph_allocs.push_back(mine);
    if (graph_id) *graph_id = (int32_t)m->graphs.size() - 1;
    return GFA2_OK;
}

int gfa2_mpnn_release_graph(gfa2_mpnn_t m, int32_t graph_id) {
    int st = check_graph(m, graph_id);
    if (st != GFA2_OK) return st;
    MP_TRY(hipSetDevice(m->device));
    MP_TRY(hipStreamSynchronize(m->stream));
    for (void* p : m->graph_allocs[graph_id]) (void)hipFree(p);
    m->graph_allocs[graph_id].clear();
    m->graphs[graph_id] = MpnnGraph{};     // n = 0, no target: skipped by training, rejected by id
    return GFA2_OK;
}

int gfa2_mpnn_features(gfa2_mpnn_t m, int32_t graph_id, float* X) {
    int st = check_graph(m, graph_id);
    if (st != GFA2_OK || !X) return st != GFA2_OK ? st : GFA2_ERR_INVALID_ARG;
    const MpnnGraph& g = m->graphs[graph_id];
    MP_TRY(hipSetDevice(m->device));
    MP_TRY(hipMemcpy(X, g.X, sizeof(float) * (size_t)g.n * kIn, hipMemcpyDeviceToHost));
    return GFA2_OK;
}

int gfa2_mpnn_loss_grad(gfa2_mpnn_t m, int32_t graph_id, const int32_t* offsets, int32_t R, double* loss, float* grad) {
    int st = check_graph(m, graph_id);
    if (st != GFA2_OK) return st;
    const MpnnGraph& g = m->graphs[graph_id];
    if (!g.T || R < 0 || (R > 0 && !offsets)) return GFA2_ERR_INVALID_ARG;
    for (int k = 0; k < R; ++k)
        if (offsets[k] < 1 || offsets[k] >= g.n) return GFA2_ERR_INVALID_ARG;
    MP_TRY(hipSetDevice(m->device));
    if (R > m->offs_cap) {
        if (m->offs) (void)hipFree(m->offs);
        MP_TRY(hipMalloc(&m->offs, sizeof(int) * R));
        m->offs_cap = R;
    }
    if (R) MP_TRY(hipMemcpyAsync(m->offs, offsets, sizeof(int) * R, hipMemcpyHostToDevice, m->stream));
    mpnn_forward(g, m->w, m->s, m->stream);
    if (mpnn_loss_backward(g, m->w, m->s, m->offs, R, loss, m->stream) != 0) return GFA2_ERR_HIP;
    if (grad) {
        MP_TRY(hipMemcpyAsync(grad, m->w.grad, sizeof(float) * kParams, hipMemcpyDeviceToHost, m->stream));
        MP_TRY(hipStreamSynchronize(m->stream));
    }
    return GFA2_OK;
}

int gfa2_mpnn_train(gfa2_mpnn_t m, int32_t steps, float lr, int32_t R, uint64_t seed, double* last_loss) {
    if (!m || steps < 0 || R < 0) return GFA2_ERR_INVALID_ARG;
    std::vector<int> ids;
    for (size_t k = 0; k < m->graphs.size(); ++k)
        if (m->graphs[k].T) ids.push_back((int)k);
    if (ids.empty()) return GFA2_ERR_NO_GRAPH;
    MP_TRY(hipSetDevice(m->device));
    // all steps' offsets drawn up front (seeded) and uploaded once
    SplitMix rng{seed};
    std::vector<int> offs((size_t)steps * (R > 0 ? R : 1), 1);
    for (int s = 0; s < steps; ++s) {
        const int n = m->graphs[ids[s % ids.size()]].n;
        for (int k = 0; k < R; ++k) offs[(size_t)s * R + k] = n > 1 ? 1 + (int)(rng.next() % (uint64_t)(n - 1)) : 1;
    }
    if ((long long)offs.size() > m->offs_cap) {
        if (m->offs) (void)hipFree(m->offs);
        MP_TRY(hipMalloc(&m->offs, sizeof(int) * offs.size()));
        m->offs_cap = (long long)offs.size();
    }
    MP_TRY(hipMemcpyAsync(m->offs, offs.data(), sizeof(int) * offs.size(), hipMemcpyHostToDevice, m->stream));
    for (int s = 0; s < steps; ++s) {
        const MpnnGraph& g = m->graphs[ids[s % ids.size()]];
        mpnn_forward(g, m->w, m->s, m->stream);
        double* lp = (s == steps - 1) ? last_loss : nullptr;
        if (mpnn_loss_backward(g, m->w, m->s, m->offs + (size_t)s * R, R, lp, m->stream) != 0) return GFA2_ERR_HIP;
        mpnn_adam(m->w, lr, ++m->adam_step, m->stream);
    }
    MP_TRY(hipGetLastError());
    MP_TRY(hipStreamSynchronize(m->stream));
    return GFA2_OK;
}

float gfa2_mpnn_target_scale(gfa2_mpnn_t m) {
    if (!m) return 0.0f;
    double s = 0.0;
    int c = 0;
    for (const MpnnGraph& g : m->graphs)
        if (g.T) {
            s += g.target_scale;
            ++c;
        }
    return c ? (float)(s / c) : 0.0f;
}

int gfa2_mpnn_infer(gfa2_mpnn_t m, int32_t graph_id, float scale, float* xy) {
    int st = check_graph(m, graph_id);
    if (st != GFA2_OK || !xy) return st != GFA2_OK ? st : GFA2_ERR_INVALID_ARG;
    const MpnnGraph& g = m->graphs[graph_id];
    MP_TRY(hipSetDevice(m->device));
    mpnn_forward(g, m->w, m->s, m->stream);
    mpnn_place(g, m->s, scale, reinterpret_cast<float2*>(m->s.dY), m->stream);
    MP_TRY(hipMemcpyAsync(xy, m->s.dY, sizeof(float) * 2 * g.n, hipMemcpyDeviceToHost, m->stream));
    MP_TRY(hipStreamSynchronize(m->stream));
    return GFA2_OK;
}

int gfa2_mpnn_infer_to_engine(gfa2_mpnn_t m, int32_t graph_id, float scale, gfa2_handle h) {
    int st = check_graph(m, graph_id);
    if (st != GFA2_OK) return st;
    EngineView v;
    if (!engine_view(h, &v)) return GFA2_ERR_NO_GRAPH;
    const MpnnGraph& g = m->graphs[graph_id];
    if (v.n != g.n || v.device != m->device) return GFA2_ERR_INVALID_ARG;
    MP_TRY(hipSetDevice(m->device));
    MP_TRY(hipStreamSynchronize(v.stream));
    mpnn_forward(g, m->w, m->s, m->stream);
    mpnn_place(g, m->s, scale, v.pos, m->stream);
    MP_TRY(hipGetLastError());
    MP_TRY(hipStreamSynchronize(m->stream));
    return engine_positions_written(h);
}

int gfa2_mpnn_destroy(gfa2_mpnn_t m) {
    if (!m) return GFA2_ERR_INVALID_ARG;
    (void)hipSetDevice(m->device);
    if (m->stream) (void)hipStreamSynchronize(m->stream);
    for (auto& list : m->graph_allocs)
        for (void* p : list) (void)hipFree(p);
    for (void* p : m->allocs) (void)hipFree(p);
    MpnnScratch& s = m->s;
    for (void* p : {(void*)s.z0, (void*)s.z1, (void*)s.H1, (void*)s.A1, (void*)s.H2, (void*)s.Y, (void*)s.dY,
                    (void*)s.dZ2, (void*)s.P, (void*)s.Q, (void*)s.dZ1, (void*)s.wpart})
        if (p) (void)hipFree(p);
    if (m->offs) (void)hipFree(m->offs);
    if (m->stream) (void)hipStreamDestroy(m->stream);
    delete m;
    return GFA2_OK;
}

}  // extern "C"
