// Autotuner and learned selector support.
//
// No reference counterpart: the reference fixes its tiers at compile time
// (kernels/mod.rs:63-106 thresholds, kernels/gemm.rs:290-312). Here every operation family has a
// small table of launch configurations; the tuner times all of them per shape (exhaustive
// search), and a 6 -> 64 -> 64 -> K MLP cost model, trained on those timings by the kernels below,
// predicts log2(time) of every configuration so that the fastest one can be picked without
// timing. Families: 0 = f64 GEMM, 1 = f32 GEMM (MFMA tile sizes), 2 = bf16 -> f32 NN GEMM
// (tile sizes), 3 = f32 row sums (persistent-grid sizes; result bits do not depend on it).
#include <algorithm>
#include <cmath>
#include <map>
#include <mutex>
#include <tuple>
#include <vector>

#include "ops.hpp"

namespace nrk {

constexpr int MLP_IN = 6, MLP_H = 64, MLP_MAXK = 8;
constexpr int kRowSumBlocks[4] = {1, 2, 4, 8};

static int config_count(int dtype) {
    switch (dtype) {
        case NRK_DT_F64:
        case NRK_DT_F32:
        case NRK_DT_BF16: return 5;
        case NRK_DT_ROWSUM_F32: return 4;
        case NRK_DT_F32_BF16X3:
        case NRK_DT_F32_BF16X6:
        case NRK_DT_F64_OZAKI: return 5;
        default: return 0;
    }
}

static nrk_status run_config(int dtype, int config, int d0, int d1, int d2, const void *a, const void *b, void *c,
                             hipStream_t s) {
    switch (dtype) {
        case NRK_DT_F64: return gemm_config<double>(config, d0, d2, d1, (const double *)a, (const double *)b, (double *)c, s);
        case NRK_DT_F32: return gemm_config<float>(config, d0, d2, d1, (const float *)a, (const float *)b, (float *)c, s);
        case NRK_DT_BF16: {
            nrk_gemm_desc d{d0, d1, d2, 0, a, b, (float *)c};
            return gemm_bf16_config(config, d, s);
        }
        case NRK_DT_ROWSUM_F32: {
            if (config < 0 || config > 3) return fail(NRK_ERR_INVALID_ARGUMENT, "row-sum config %d outside [0,3]", config);
            const int saved = reduce_blocks_per_cu();
            set_reduce_blocks_per_cu(kRowSumBlocks[config]);
            Chain ch;
            NRK_TRY(make_chain(nullptr, 0, &ch));
            nrk_reduce_outputs o{};
            o.sum = c;
            nrk_status st = seg_reduce<float>(ch, (const float *)a, (size_t)d0, (size_t)d1, 0.0, &o, nullptr, s);
            set_reduce_blocks_per_cu(saved);
            return st;
        }
        case NRK_DT_F32_BF16X3:
            if (config < 0 || config > 4) return fail(NRK_ERR_INVALID_ARGUMENT, "config %d outside [0,4]", config);
            return gemm_f32_bf16x(d0, d2, d1, (const float *)a, (const float *)b, (float *)c, 3, config, s);
        case NRK_DT_F32_BF16X6:
            if (config < 0 || config > 4) return fail(NRK_ERR_INVALID_ARGUMENT, "config %d outside [0,4]", config);
            return gemm_f32_bf16x(d0, d2, d1, (const float *)a, (const float *)b, (float *)c, 6, config, s);
        case NRK_DT_F64_OZAKI:
            if (config < 0 || config >= config_count(dtype)) return fail(NRK_ERR_INVALID_ARGUMENT, "config %d outside [0,%d)", config, config_count(dtype));
            return gemm_f64_ozaki(d0, d2, d1, (const double *)a, (const double *)b, (double *)c, 16, config, s);
        default: return fail(NRK_ERR_INVALID_ARGUMENT, "unknown operation family %d", dtype);
    }
}

__global__ void fill_pattern_kernel(unsigned int *p, size_t words, unsigned int salt, int bf16) {
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < words; i += (size_t)gridDim.x * blockDim.x) {
        unsigned int h = (unsigned int)(i * 2654435761u) ^ salt;
        h ^= h >> 13;
        // small finite values: exponent of 1.0f-ish floats, or pairs of bf16 near 1
        p[i] = bf16 ? (0x3f803f80u ^ (h & 0x007f007fu)) : (0x3f800000u | (h & 0x007fffffu));
    }
}

// ---------------- MLP kernels ----------------
// Parameter layout (fp32, contiguous): W1[H][IN] b1[H] W2[H][H] b2[H] W3[K][H] b3[K].
struct MlpView {
    const float *W1, *b1, *W2, *b2, *W3, *b3;
};
__host__ __device__ inline size_t mlp_param_count(int K) {
    return (size_t)MLP_H * MLP_IN + MLP_H + MLP_H * MLP_H + MLP_H + (size_t)K * MLP_H + K;
}
__device__ __forceinline__ MlpView mlp_view(const float *p, int K) {
    MlpView v;
    v.W1 = p;
    v.b1 = v.W1 + MLP_H * MLP_IN;
    v.W2 = v.b1 + MLP_H;
    v.b2 = v.W2 + MLP_H * MLP_H;
    v.W3 = v.b2 + MLP_H;
    v.b3 = v.W3 + K * MLP_H;
    return v;
}

// One thread per sample; parameters staged in LDS. Optionally saves the hidden activations.
__global__ __launch_bounds__(256) void mlp_forward_kernel(const float *__restrict__ params, int K,
                                                          const float *__restrict__ x, int N, float *__restrict__ y,
                                                          float *__restrict__ h1_out, float *__restrict__ h2_out) {
    __shared__ float sp[MLP_H * MLP_IN + MLP_H + MLP_H * MLP_H + MLP_H + MLP_MAXK * MLP_H + MLP_MAXK];
    const int total = (int)mlp_param_count(K);
    for (int i = threadIdx.x; i < total; i += blockDim.x) sp[i] = params[i];
    __syncthreads();
    const int n = blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= N) return;
    const MlpView w = mlp_view(sp, K);
    float in[MLP_IN];
#pragma unroll
    for (int i = 0; i < MLP_IN; ++i) in[i] = x[(size_t)n * MLP_IN + i];
    float h1[MLP_H];
#pragma unroll
    for (int j = 0; j < MLP_H; ++j) {
        float s = w.b1[j];
#pragma unroll
        for (int i = 0; i < MLP_IN; ++i) s = fmaf(w.W1[j * MLP_IN + i], in[i], s);
        h1[j] = s > 0.f ? s : 0.f;
        if (h1_out) h1_out[(size_t)n * MLP_H + j] = h1[j];
    }
    float h2[MLP_H];
    for (int j = 0; j < MLP_H; ++j) {
        float s = w.b2[j];
#pragma unroll
        for (int i = 0; i < MLP_H; ++i) s = fmaf(w.W2[j * MLP_H + i], h1[i], s);
        h2[j] = s > 0.f ? s : 0.f;
        if (h2_out) h2_out[(size_t)n * MLP_H + j] = h2[j];
    }
    for (int o = 0; o < K; ++o) {
        float s = w.b3[o];
#pragma unroll
        for (int i = 0; i < MLP_H; ++i) s = fmaf(w.W3[o * MLP_H + i], h2[i], s);
        y[(size_t)n * K + o] = s;
    }
}

// Per-sample backward: dy = 2 mask (y - t) / count; dh2 = (W3^T dy) * [h2 > 0];
// dh1 = (W2^T dh2) * [h1 > 0]. Also writes the per-sample loss sum(mask (y - t)^2).
__global__ __launch_bounds__(256) void mlp_backward_sample_kernel(const float *__restrict__ params, int K, int N,
                                                                  const float *__restrict__ y, const float *__restrict__ t,
                                                                  const float *__restrict__ mask, float inv_count,
                                                                  const float *__restrict__ h1, const float *__restrict__ h2,
                                                                  float *__restrict__ dy, float *__restrict__ dh2,
                                                                  float *__restrict__ dh1, float *__restrict__ loss) {
    __shared__ float sW2[MLP_H * MLP_H];
    __shared__ float sW3[MLP_MAXK * MLP_H];
    const float *W2 = params + MLP_H * MLP_IN + MLP_H;
    const float *W3 = W2 + MLP_H * MLP_H + MLP_H;
    for (int i = threadIdx.x; i < MLP_H * MLP_H; i += blockDim.x) sW2[i] = W2[i];
    for (int i = threadIdx.x; i < K * MLP_H; i += blockDim.x) sW3[i] = W3[i];
    __syncthreads();
    const int n = blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= N) return;
    float g[MLP_MAXK];
    float l = 0.f;
    for (int o = 0; o < K; ++o) {
        const size_t idx = (size_t)n * K + o;
        const float d = (y[idx] - t[idx]) * mask[idx];
        l = fmaf(d, d, l);
        g[o] = 2.f * d * inv_count;
        dy[idx] = g[o];
    }
    loss[n] = l;
    float d2[MLP_H];
    for (int i = 0; i < MLP_H; ++i) {
        float s = 0.f;
        for (int o = 0; o < K; ++o) s = fmaf(sW3[o * MLP_H + i], g[o], s);
        d2[i] = h2[(size_t)n * MLP_H + i] > 0.f ? s : 0.f;
        dh2[(size_t)n * MLP_H + i] = d2[i];
    }
    for (int i = 0; i < MLP_H; ++i) {
        float s = 0.f;
#pragma unroll 8
        for (int j = 0; j < MLP_H; ++j) s = fmaf(sW2[j * MLP_H + i], d2[j], s);
        dh1[(size_t)n * MLP_H + i] = h1[(size_t)n * MLP_H + i] > 0.f ? s : 0.f;
    }
}

// Weight gradients: one thread per parameter, summing over the samples in a fixed order.
// grad[W_l[j][i]] = sum_n delta_l[n][j] * in_l[n][i];  grad[b_l[j]] = sum_n delta_l[n][j].
__global__ __launch_bounds__(256) void mlp_weight_grad_kernel(int K, int N, const float *__restrict__ x,
                                                              const float *__restrict__ h1, const float *__restrict__ h2,
                                                              const float *__restrict__ dy, const float *__restrict__ dh2,
                                                              const float *__restrict__ dh1, float *__restrict__ grad) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    const int n1 = MLP_H * MLP_IN, n2 = n1 + MLP_H, n3 = n2 + MLP_H * MLP_H, n4 = n3 + MLP_H, n5 = n4 + K * MLP_H,
              n6 = n5 + K;
    if (p >= n6) return;
    const float *delta, *in;
    int dw, iw, j, i;
    if (p < n1) { delta = dh1; dw = MLP_H; j = p / MLP_IN; i = p % MLP_IN; in = x; iw = MLP_IN; }
    else if (p < n2) { delta = dh1; dw = MLP_H; j = p - n1; i = -1; in = nullptr; iw = 0; }
    else if (p < n3) { delta = dh2; dw = MLP_H; j = (p - n2) / MLP_H; i = (p - n2) % MLP_H; in = h1; iw = MLP_H; }
    else if (p < n4) { delta = dh2; dw = MLP_H; j = p - n3; i = -1; in = nullptr; iw = 0; }
    else if (p < n5) { delta = dy; dw = K; j = (p - n4) / MLP_H; i = (p - n4) % MLP_H; in = h2; iw = MLP_H; }
    else { delta = dy; dw = K; j = p - n5; i = -1; in = nullptr; iw = 0; }
    float s = 0.f, c = 0.f;  // compensated
    for (int n = 0; n < N; ++n) {
        const float v = delta[(size_t)n * dw + j] * (i >= 0 ? in[(size_t)n * iw + i] : 1.f);
        const float yk = v - c;
        const float tt = s + yk;
        c = (tt - s) - yk;
        s = tt;
    }
    grad[p] = s;
}

// Adam (Kingma & Ba 2015) with bias correction.
__global__ void adam_kernel(float *__restrict__ params, const float *__restrict__ grad, float *__restrict__ m,
                            float *__restrict__ v, int count, float lr, float b1, float b2, float eps, float c1, float c2) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= count) return;
    const float g = grad[i];
    const float mi = b1 * m[i] + (1.f - b1) * g;
    const float vi = b2 * v[i] + (1.f - b2) * g * g;
    m[i] = mi;
    v[i] = vi;
    params[i] -= lr * (mi / c1) / (sqrtf(vi / c2) + eps);
}

__global__ void sum_kernel(const float *__restrict__ x, int n, float *__restrict__ out) {
    __shared__ float s[256];
    float a = 0.f;
    for (int i = threadIdx.x; i < n; i += 256) a += x[i];
    s[threadIdx.x] = a;
    __syncthreads();
    for (int k = 128; k > 0; k >>= 1) {
        if (threadIdx.x < k) s[threadIdx.x] += s[threadIdx.x + k];
        __syncthreads();
    }
    if (threadIdx.x == 0) *out = s[0];
}


// ---------------- the selector on the dispatch path ----------------
// Per operation family: an optional cost model (the MLP above, evaluated on the host, a few
// thousand multiply-adds) and a decision table per shape. A dispatch with no forced tile asks
// select_config: a recorded decision wins; else, inside the shapes the model was trained on, the
// model's fastest predicted configuration, or, when its two best predictions are closer than the
// margin, the configuration the exhaustive timing finds (not while the stream is being captured
// into a graph, where the model's choice stands); else -1, and the caller keeps its fixed rule.
// Every configuration of a family gives the same result bits, so the choice only moves time.
struct SelModel {
    bool set = false;
    int K = 0;
    float dtype_feature = 0.f;
    std::vector<float> params;
    float x_mean[MLP_IN], x_scale[MLP_IN];
    float lo[3], hi[3];  // log2 d0, d1, d2 range of the training shapes
    float y_offset = 0.f, margin = 0.f;
};
enum { SEL_RULE = 0, SEL_MODEL = 1, SEL_TUNED = 2, SEL_TABLE = 3 };
struct SelDecision {
    int config, source;
};
struct SelStats {
    long long n[4] = {0, 0, 0, 0};  // decisions by source, counted per dispatch
};
constexpr int SEL_FAMILIES = 7;
static std::mutex g_sel_mutex;
static SelModel g_sel_model[SEL_FAMILIES];
static SelStats g_sel_stats[SEL_FAMILIES];
static std::map<std::tuple<int, int, int, int>, SelDecision> g_sel_table;

static void host_mlp(const SelModel &m, const float (&f)[MLP_IN], float *y) {
    const float *W1 = m.params.data(), *b1 = W1 + MLP_H * MLP_IN, *W2 = b1 + MLP_H, *b2 = W2 + MLP_H * MLP_H;
    const float *W3 = b2 + MLP_H, *b3 = W3 + m.K * MLP_H;
    float x[MLP_IN], h1[MLP_H], h2[MLP_H];
    for (int i = 0; i < MLP_IN; ++i) x[i] = (f[i] - m.x_mean[i]) / m.x_scale[i];
    for (int j = 0; j < MLP_H; ++j) {
        float s = b1[j];
        for (int i = 0; i < MLP_IN; ++i) s = std::fma(W1[j * MLP_IN + i], x[i], s);
        h1[j] = s > 0.f ? s : 0.f;
    }
    for (int j = 0; j < MLP_H; ++j) {
        float s = b2[j];
        for (int i = 0; i < MLP_H; ++i) s = std::fma(W2[j * MLP_H + i], h1[i], s);
        h2[j] = s > 0.f ? s : 0.f;
    }
    for (int o = 0; o < m.K; ++o) {
        float s = b3[o];
        for (int i = 0; i < MLP_H; ++i) s = std::fma(W3[o * MLP_H + i], h2[i], s);
        y[o] = s + m.y_offset;
    }
}

// The decision for one shape (source: SEL_*), without the table; the caller holds no lock.
static SelDecision decide(int family, int d0, int d1, int d2, hipStream_t stream) {
    SelModel m;
    {
        std::lock_guard<std::mutex> lock(g_sel_mutex);
        if (!g_sel_model[family].set) return {-1, SEL_RULE};
        m = g_sel_model[family];
    }
    const float f[MLP_IN] = {std::log2((float)d0), std::log2((float)d1), std::log2((float)std::max(d2, 1)), m.dtype_feature, 0.f,
                             0.f};
    for (int i = 0; i < 3; ++i)
        if (f[i] < m.lo[i] - 0.5f || f[i] > m.hi[i] + 0.5f) return {-1, SEL_RULE};  // outside the trained shapes
    float y[MLP_MAXK];
    host_mlp(m, f, y);
    int best = 0, second = -1;
    for (int o = 1; o < m.K; ++o) {
        if (y[o] < y[best]) {
            second = best;
            best = o;
        } else if (second < 0 || y[o] < y[second]) {
            second = o;
        }
    }
    if (second >= 0 && y[second] - y[best] < m.margin) {
        hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
        if (hipStreamIsCapturing(stream, &cs) == hipSuccess && cs == hipStreamCaptureStatusNone) {
            std::vector<float> t(m.K);
            const int32_t a0 = d0, a1 = d1, a2 = std::max(d2, 1);
            if (nrk_tune(family, 1, &a0, &a1, &a2, 3, t.data()) == NRK_OK)
                return {(int)(std::min_element(t.begin(), t.end()) - t.begin()), SEL_TUNED};
        }
    }
    return {best, SEL_MODEL};
}

int select_config(int family, size_t d0, size_t d1, size_t d2, hipStream_t stream) {
    if (family < 0 || family >= SEL_FAMILIES || d0 > (1u << 30) || d1 > (1u << 30) || d2 > (1u << 30)) return -1;
    const auto key = std::make_tuple(family, (int)d0, (int)d1, (int)d2);
    {
        std::lock_guard<std::mutex> lock(g_sel_mutex);
        auto it = g_sel_table.find(key);
        if (it != g_sel_table.end()) {
            ++g_sel_stats[family].n[it->second.source == SEL_TUNED ? SEL_TABLE : it->second.source];
            return it->second.config;
        }
        if (!g_sel_model[family].set) {
            ++g_sel_stats[family].n[SEL_RULE];
            return -1;
        }
    }
    const SelDecision d = decide(family, (int)d0, (int)d1, (int)d2, stream);
    std::lock_guard<std::mutex> lock(g_sel_mutex);
    ++g_sel_stats[family].n[d.source];
    if (d.config >= 0) g_sel_table[key] = d;
    return d.config;
}

}  // namespace nrk

using namespace nrk;

extern "C" {

nrk_status nrk_config_count(int32_t family, int32_t *count) {
    if (!count) return fail(NRK_ERR_INVALID_ARGUMENT, "count is NULL");
    *count = config_count(family);
    return *count ? NRK_OK : fail(NRK_ERR_INVALID_ARGUMENT, "unknown operation family %d", family);
}

nrk_status nrk_dev_run_config(int32_t family, int32_t config, int32_t d0, int32_t d1, int32_t d2, const void *a,
                              const void *b, void *c, nrk_stream stream) {
    NRK_TRY(ensure_device());
    return run_config(family, config, d0, d1, d2, a, b, c, (hipStream_t)stream);
}

nrk_status nrk_tune(int32_t family, int32_t count, const int32_t *d0, const int32_t *d1, const int32_t *d2, int32_t reps,
                    float *times_ns) {
    NRK_TRY(ensure_device());
    const int K = config_count(family);
    if (!K) return fail(NRK_ERR_INVALID_ARGUMENT, "unknown operation family %d", family);
    if (count <= 0) return NRK_OK;
    if (!d0 || !d1 || (family != NRK_DT_ROWSUM_F32 && !d2) || !times_ns) return fail(NRK_ERR_INVALID_ARGUMENT, "NULL array");
    if (reps < 1) reps = 1;
    size_t amax = 16, bmax = 16, cmax = 16;
    for (int i = 0; i < count; ++i) {
        if (d0[i] <= 0 || d1[i] <= 0 || (family != NRK_DT_ROWSUM_F32 && d2[i] <= 0))
            return fail(NRK_ERR_INVALID_ARGUMENT, "shape %d has a non-positive dimension", i);
        if (family == NRK_DT_ROWSUM_F32) {
            amax = std::max(amax, (size_t)d0[i] * d1[i]);
            cmax = std::max(cmax, (size_t)d0[i]);
        } else {
            amax = std::max(amax, (size_t)d0[i] * d2[i]);
            bmax = std::max(bmax, (size_t)d2[i] * d1[i]);
            cmax = std::max(cmax, (size_t)d0[i] * d1[i]);
        }
    }
    const bool f64 = family == NRK_DT_F64 || family == NRK_DT_F64_OZAKI;
    const size_t es = f64 ? 8 : (family == NRK_DT_BF16 ? 2 : 4);
    const size_t cs = f64 ? 8 : 4;
    void *a = nullptr, *b = nullptr, *c = nullptr;
    hipStream_t s = nullptr;
    NRK_HIP(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
    std::vector<hipEvent_t> ev(2 * K);
    nrk_status st = NRK_OK;
    auto cleanup = [&]() {
        for (auto e : ev)
            if (e) (void)hipEventDestroy(e);
        (void)hipFree(a);
        (void)hipFree(b);
        (void)hipFree(c);
        (void)hipStreamDestroy(s);
    };
    for (auto &e : ev) e = nullptr;
    if (hipMalloc(&a, amax * es) != hipSuccess || hipMalloc(&b, bmax * es) != hipSuccess || hipMalloc(&c, cmax * cs) != hipSuccess) {
        cleanup();
        return fail(NRK_ERR_OUT_OF_MEMORY, "tuner operand allocation failed");
    }
    for (auto &e : ev) (void)hipEventCreate(&e);
    hipLaunchKernelGGL(fill_pattern_kernel, dim3(1024), dim3(256), 0, s, (unsigned *)a, amax * es / 4, 0x1234u, (int)(family == NRK_DT_BF16));
    hipLaunchKernelGGL(fill_pattern_kernel, dim3(1024), dim3(256), 0, s, (unsigned *)b, bmax * es / 4, 0x9876u, (int)(family == NRK_DT_BF16));
    for (int i = 0; i < count && st == NRK_OK; ++i) {
        const int x0 = d0[i], x1 = d1[i], x2 = family == NRK_DT_ROWSUM_F32 ? 0 : d2[i];
        for (int cfg = 0; cfg < K && st == NRK_OK; ++cfg) {
            st = run_config(family, cfg, x0, x1, x2, a, b, c, s);  // warm-up
            if (st != NRK_OK) break;
            (void)hipEventRecord(ev[2 * cfg], s);
            for (int r = 0; r < reps && st == NRK_OK; ++r) st = run_config(family, cfg, x0, x1, x2, a, b, c, s);
            (void)hipEventRecord(ev[2 * cfg + 1], s);
        }
        if (st != NRK_OK) break;
        if (hipStreamSynchronize(s) != hipSuccess) {
            st = fail(NRK_ERR_HIP, "tuner stream failed at shape %d", i);
            break;
        }
        for (int cfg = 0; cfg < K; ++cfg) {
            float ms = 0.f;
            (void)hipEventElapsedTime(&ms, ev[2 * cfg], ev[2 * cfg + 1]);
            times_ns[(size_t)i * K + cfg] = ms * 1e6f / (float)reps;
        }
    }
    cleanup();
    return st;
}

size_t nrk_mlp_param_count(int32_t K) { return (K < 1 || K > MLP_MAXK) ? 0 : mlp_param_count(K); }

nrk_status nrk_mlp_forward(const float *params, int32_t K, const float *x, int32_t N, float *y, nrk_stream stream) {
    NRK_TRY(ensure_device());
    if (K < 1 || K > MLP_MAXK) return fail(NRK_ERR_INVALID_ARGUMENT, "K = %d outside [1,%d]", K, MLP_MAXK);
    if (N <= 0) return NRK_OK;
    if (!params || !x || !y) return fail(NRK_ERR_INVALID_ARGUMENT, "NULL pointer");
    hipLaunchKernelGGL(mlp_forward_kernel, dim3((N + 255) / 256), dim3(256), 0, (hipStream_t)stream, params, K, x, N, y,
                       (float *)nullptr, (float *)nullptr);
    NRK_HIP(hipGetLastError());
    return NRK_OK;
}

nrk_status nrk_mlp_train_step(float *params, float *adam_m, float *adam_v, int32_t step, int32_t K, const float *x,
                              const float *target, const float *mask, int32_t N, float lr, float *loss_out,
                              float *grad_out, nrk_stream stream) {
    NRK_TRY(ensure_device());
    if (K < 1 || K > MLP_MAXK) return fail(NRK_ERR_INVALID_ARGUMENT, "K = %d outside [1,%d]", K, MLP_MAXK);
    if (N <= 0) return fail(NRK_ERR_INVALID_ARGUMENT, "empty batch");
    if (!params || !x || !target || !mask) return fail(NRK_ERR_INVALID_ARGUMENT, "NULL pointer");
    if (lr != 0.f && (!adam_m || !adam_v || step < 1)) return fail(NRK_ERR_INVALID_ARGUMENT, "Adam state missing or step < 1");
    hipStream_t s = (hipStream_t)stream;
    const size_t P = mlp_param_count(K);
    const size_t floats = (size_t)N * (K + MLP_H + MLP_H + K + MLP_H + MLP_H + 1) + P + 1;
    void *ws = nullptr;
    NRK_TRY(get_workspace(s, floats * sizeof(float), &ws));
    float *y = (float *)ws, *h1 = y + (size_t)N * K, *h2 = h1 + (size_t)N * MLP_H, *dy = h2 + (size_t)N * MLP_H;
    float *dh2 = dy + (size_t)N * K, *dh1 = dh2 + (size_t)N * MLP_H, *loss = dh1 + (size_t)N * MLP_H;
    float *grad = grad_out ? grad_out : loss + N;
    // count of valid targets for the mean (host side: the caller's mask is device memory)
    std::vector<float> hmask((size_t)N * K);
    NRK_HIP(hipMemcpyAsync(hmask.data(), mask, hmask.size() * sizeof(float), hipMemcpyDeviceToHost, s));
    NRK_HIP(hipStreamSynchronize(s));
    double cnt = 0;
    for (float v : hmask) cnt += v;
    const float inv_count = cnt > 0 ? (float)(1.0 / cnt) : 0.f;
    const int nb = (N + 255) / 256;
    hipLaunchKernelGGL(mlp_forward_kernel, dim3(nb), dim3(256), 0, s, params, K, x, N, y, h1, h2);
    hipLaunchKernelGGL(mlp_backward_sample_kernel, dim3(nb), dim3(256), 0, s, params, K, N, y, target, mask, inv_count, h1,
                       h2, dy, dh2, dh1, loss);
    hipLaunchKernelGGL(mlp_weight_grad_kernel, dim3((unsigned)((P + 255) / 256)), dim3(256), 0, s, K, N, x, h1, h2, dy, dh2,
                       dh1, grad);
    if (loss_out) hipLaunchKernelGGL(sum_kernel, dim3(1), dim3(256), 0, s, loss, N, loss_out);
    if (lr != 0.f) {
        const float b1 = 0.9f, b2 = 0.999f;
        const float c1 = 1.f - powf(b1, (float)step), c2 = 1.f - powf(b2, (float)step);
        hipLaunchKernelGGL(adam_kernel, dim3((unsigned)((P + 255) / 256)), dim3(256), 0, s, params, grad, adam_m, adam_v,
                           (int)P, lr, b1, b2, 1e-8f, c1, c2);
    }
    NRK_HIP(hipGetLastError());
    return NRK_OK;
}


nrk_status nrk_selector_set_model(int32_t family, float dtype_feature, const float *params, int32_t K, const float *x_mean,
                                  const float *x_scale, const float *domain_lo, const float *domain_hi, float y_offset,
                                  float margin) {
    if (family < 0 || family >= SEL_FAMILIES || config_count(family) == 0)
        return fail(NRK_ERR_INVALID_ARGUMENT, "unknown operation family %d", family);
    if (K != config_count(family)) return fail(NRK_ERR_INVALID_ARGUMENT, "family %d has %d configurations, the model %d", family, config_count(family), K);
    if (!params || !x_mean || !x_scale || !domain_lo || !domain_hi) return fail(NRK_ERR_INVALID_ARGUMENT, "NULL pointer");
    SelModel m;
    m.set = true;
    m.K = K;
    m.dtype_feature = dtype_feature;
    m.params.assign(params, params + mlp_param_count(K));
    for (int i = 0; i < MLP_IN; ++i) {
        if (!(x_scale[i] > 0.f)) return fail(NRK_ERR_INVALID_ARGUMENT, "x_scale[%d] must be positive", i);
        m.x_mean[i] = x_mean[i];
        m.x_scale[i] = x_scale[i];
    }
    for (int i = 0; i < 3; ++i) {
        m.lo[i] = domain_lo[i];
        m.hi[i] = domain_hi[i];
    }
    m.y_offset = y_offset;
    m.margin = margin;
    std::lock_guard<std::mutex> lock(g_sel_mutex);
    g_sel_model[family] = std::move(m);
    for (auto it = g_sel_table.begin(); it != g_sel_table.end();)  // the new model's choices replace the old ones'
        it = (std::get<0>(it->first) == family && it->second.source == SEL_MODEL) ? g_sel_table.erase(it) : std::next(it);
    return NRK_OK;
}

nrk_status nrk_selector_clear(int32_t family) {
    if (family < -1 || family >= SEL_FAMILIES) return fail(NRK_ERR_INVALID_ARGUMENT, "unknown operation family %d", family);
    std::lock_guard<std::mutex> lock(g_sel_mutex);
    for (int f = 0; f < SEL_FAMILIES; ++f) {
        if (family >= 0 && f != family) continue;
        g_sel_model[f] = SelModel();
        g_sel_stats[f] = SelStats();
    }
    for (auto it = g_sel_table.begin(); it != g_sel_table.end();)
        it = (family < 0 || std::get<0>(it->first) == family) ? g_sel_table.erase(it) : std::next(it);
    return NRK_OK;
}

nrk_status nrk_selector_table_put(int32_t family, int32_t d0, int32_t d1, int32_t d2, int32_t config) {
    if (family < 0 || family >= SEL_FAMILIES || config_count(family) == 0)
        return fail(NRK_ERR_INVALID_ARGUMENT, "unknown operation family %d", family);
    if (config < 0 || config >= config_count(family)) return fail(NRK_ERR_INVALID_ARGUMENT, "config %d outside [0,%d)", config, config_count(family));
    if (d0 <= 0 || d1 <= 0 || d2 < 0) return fail(NRK_ERR_INVALID_ARGUMENT, "non-positive dimension");
    std::lock_guard<std::mutex> lock(g_sel_mutex);
    g_sel_table[std::make_tuple((int)family, (int)d0, (int)d1, (int)d2)] = SelDecision{config, SEL_TUNED};
    return NRK_OK;
}

nrk_status nrk_selector_table_size(int32_t *count) {
    if (!count) return fail(NRK_ERR_INVALID_ARGUMENT, "count is NULL");
    std::lock_guard<std::mutex> lock(g_sel_mutex);
    int32_t n = 0;
    for (const auto &e : g_sel_table) n += e.second.source == SEL_TUNED;
    *count = n;
    return NRK_OK;
}

nrk_status nrk_selector_table_get(int32_t index, int32_t *family, int32_t *d0, int32_t *d1, int32_t *d2, int32_t *config) {
    if (!family || !d0 || !d1 || !d2 || !config) return fail(NRK_ERR_INVALID_ARGUMENT, "NULL pointer");
    std::lock_guard<std::mutex> lock(g_sel_mutex);
    int32_t i = 0;
    for (const auto &e : g_sel_table) {
        if (e.second.source != SEL_TUNED) continue;
        if (i++ == index) {
            std::tie(*family, *d0, *d1, *d2) = e.first;
            *config = e.second.config;
            return NRK_OK;
        }
    }
    return fail(NRK_ERR_INVALID_ARGUMENT, "table index %d out of range", index);
}

nrk_status nrk_selector_choose(int32_t family, int32_t d0, int32_t d1, int32_t d2, int32_t *config) {
    NRK_TRY(ensure_device());
    if (!config) return fail(NRK_ERR_INVALID_ARGUMENT, "config is NULL");
    if (family < 0 || family >= SEL_FAMILIES || config_count(family) == 0)
        return fail(NRK_ERR_INVALID_ARGUMENT, "unknown operation family %d", family);
    if (d0 <= 0 || d1 <= 0 || d2 < 0) return fail(NRK_ERR_INVALID_ARGUMENT, "non-positive dimension");
    *config = select_config(family, (size_t)d0, (size_t)d1, (size_t)d2, nullptr);
    return NRK_OK;
}

nrk_status nrk_selector_stats(int32_t family, int64_t *by_rule, int64_t *by_model, int64_t *by_timing, int64_t *by_table) {
    if (family < 0 || family >= SEL_FAMILIES) return fail(NRK_ERR_INVALID_ARGUMENT, "unknown operation family %d", family);
    std::lock_guard<std::mutex> lock(g_sel_mutex);
    const SelStats &st = g_sel_stats[family];
    if (by_rule) *by_rule = st.n[SEL_RULE];
    if (by_model) *by_model = st.n[SEL_MODEL];
    if (by_timing) *by_timing = st.n[SEL_TUNED];
    if (by_table) *by_table = st.n[SEL_TABLE];
    return NRK_OK;
}

}  // extern "C"
