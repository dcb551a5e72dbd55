// Shapes, parameter layout and buffers of the learned-initialisation network.
#pragma once
#include <hip/hip_runtime.h>

#include <vector>

namespace gfa2 {

constexpr int kIn = 16;      // input features (9 used, zero padded)
constexpr int kHid = 64;     // hidden width
// Parameter layout (FP32 masters and BF16 copies share it): W[k][o] row-major.
constexpr int kOffW1s = 0;
constexpr int kOffW1n = kOffW1s + kIn * kHid;
constexpr int kOffB1 = kOffW1n + kIn * kHid;
constexpr int kOffW2s = kOffB1 + kHid;
constexpr int kOffW2n = kOffW2s + kHid * kHid;
constexpr int kOffB2 = kOffW2n + kHid * kHid;
constexpr int kOffWo = kOffB2 + kHid;
constexpr int kOffBo = kOffWo + kHid * 2;
constexpr int kParams = kOffBo + 2;

// Rows longer than kHeavyDeg entries are aggregated in chunks of kChunk entries.
constexpr int kHeavyDeg = 512;
constexpr int kChunk = 512;

struct MpnnGraph {
    int n = 0;
    long long nnz = 0;
    int* rowptr = nullptr;
    int* col = nullptr;
    int nheavy = 0;             // rows with more than kHeavyDeg entries
    int nchunks = 0;
    int* heavy_row = nullptr;   // [nheavy]
    int* heavy_c0 = nullptr;    // [nheavy + 1] first chunk of each heavy row
    int2* chunks = nullptr;     // [nchunks] {row, first entry}
    float* part = nullptr;      // [nchunks x kHid] chunk partial sums
    float* X = nullptr;     // [n x kIn] features
    float* A0 = nullptr;    // [n x kIn] mean_agg(X)
    float* T = nullptr;     // [n x 2] normalised target, or null
    float target_scale = 0.0f;
};

struct MpnnWeights {
    float* master = nullptr;
    unsigned short* bf = nullptr;
    float* grad = nullptr;
    float* m = nullptr;
    float* v = nullptr;
};

struct MpnnScratch {
    int cap = 0;            // nodes the buffers hold
    float *z0 = nullptr, *z1 = nullptr;
    double* dpart = nullptr;    // >= 4096 doubles
    float* musd = nullptr;      // 2 * kIn
    float *H1 = nullptr, *A1 = nullptr, *H2 = nullptr, *Y = nullptr, *dY = nullptr;
    float *dZ2 = nullptr, *P = nullptr, *Q = nullptr, *dZ1 = nullptr;
    float* wpart = nullptr;     // chunks * (kHid+1) * kHid
    float* gtmp = nullptr;      // (kHid+1) * kHid
};

// Host side: the heavy rows and chunk list of a CSR given by its row offsets.
void mpnn_agg_plan(const int* rowptr, int n, std::vector<int>& heavy_row, std::vector<int>& heavy_c0,
                   std::vector<int2>& chunks);
void mpnn_features(const MpnnGraph& g, unsigned long long seed, int steps, MpnnScratch& s, hipStream_t st);
void mpnn_forward(const MpnnGraph& g, const MpnnWeights& w, MpnnScratch& s, hipStream_t st);
int mpnn_loss_backward(const MpnnGraph& g, const MpnnWeights& w, MpnnScratch& s, const int* d_offs, int R,
                       double* loss, hipStream_t st);
void mpnn_adam(MpnnWeights& w, float lr, int step, hipStream_t st);
void mpnn_sync_bf16(MpnnWeights& w, hipStream_t st);
void mpnn_place(const MpnnGraph& g, MpnnScratch& s, float scale, float2* xy, hipStream_t st);

}  // namespace gfa2
