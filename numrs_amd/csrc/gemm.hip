// Plain GEMM on gfx950 matrix cores: C(m,n) := A(m,k) * B(k,n), row-major, overwrite.
//
// Reference: kernels/gemm.rs:246-312 (gemm_2d -> gemm_float_tier). Contract kept:
//   m == 0 or n == 0 -> return without writing; k == 0 -> C zero-filled (NaN garbage
//   overwritten, gemm.rs:295-303); beta = 0 always. K is never split across workgroups, so each
//   C element is produced by exactly one wavefront in a fixed k order: results are bit-identical
//   across launches, like the reference's M-only row split (gemm.rs:364-381).
//
// Kernel: workgroup tile BM x BN, K-tile BK, (BM/WM)*(BN/WN) wavefronts each owning a WM x WN
// sub-tile of 16x16 MFMA accumulators. Operands are staged global -> registers -> LDS with two
// LDS buffers, so the next K-tile's global loads are in flight while the current one feeds the
// MFMAs. A is stored K-major in LDS (As[k][m]) so that the 16 lanes of an MFMA operand column
// read consecutive words. Workgroup ids are remapped so that the tiles one XCD executes are
// neighbours in the tile grid (each XCD has its own L2), then grouped GROUP_M tile-rows at a
// time for A/B reuse.
//
// Batches (array/linalg.rs:56-136, batched_gemm / par_batch_f64/f32): blockIdx.y is the panel
// t, whose operands start at A + t sa, B + t sb and C + t sc (element strides). A stride of 0
// repeats one panel for every t, which is matmul's batch broadcast (linalg.rs:212-225) without
// the broadcast_to copy. Every panel runs the same code as a single gemm_2d, so panel t of a
// batch is bit-identical to gemm_2d on that panel.
//
// MFMA operand layout (16x16x4, f64 and f32): lane l supplies A[l%16][l/16] and B[l/16][l%16];
// accumulator element r of lane l is D[4*(l/16) + r][l%16] for the f32 instruction and
// D[l/16 + 4*r][l%16] for the f64 one (Mfma<T>::row below).
#include "ops.hpp"

namespace nrk {

template <typename T>
struct Mfma;
template <>
struct Mfma<double> {
    typedef double acc_t __attribute__((ext_vector_type(4)));
    static __device__ __forceinline__ acc_t mma(double a, double b, acc_t c) {
        return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
    }
    static __device__ __forceinline__ int row(int lane, int r) { return (lane >> 4) + 4 * r; }
};
template <>
struct Mfma<float> {
    typedef float acc_t __attribute__((ext_vector_type(4)));
    static __device__ __forceinline__ acc_t mma(float a, float b, acc_t c) {
        return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
    }
    static __device__ __forceinline__ int row(int lane, int r) { return 4 * (lane >> 4) + r; }
};

constexpr int XCDS = 8;

// Map the linear workgroup id to a (tile_m, tile_n) so that consecutive ids on one XCD cover
// a compact block of the tile grid.
__device__ __forceinline__ void tile_of(int nbm, int nbn, int group_m, int &tm, int &tn) {
    const int total = nbm * nbn;
    int g = blockIdx.x;
    if (total % XCDS == 0) {
        const int per = total / XCDS;
        g = (g % XCDS) * per + g / XCDS;  // hardware places workgroup g on XCD g % 8
    }
    const int in_group = group_m * nbn;
    const int grp = g / in_group;
    const int first_m = grp * group_m;
    const int gm = min(nbm - first_m, group_m);
    const int r = g % in_group;
    tm = first_m + r % gm;
    tn = r / gm;
}

template <typename T, int BM, int BN, int BK, int WM, int WN>
__global__ __launch_bounds__((BM / WM) * (BN / WN) * 64) void gemm_kernel(int M, int N, int K, const T *__restrict__ A,
                                                                         const T *__restrict__ B, T *__restrict__ C,
                                                                         int group_m, size_t sa, size_t sb, size_t sc) {
    constexpr int WGN = BN / WN;
    constexpr int NT = (BM / WM) * WGN * 64;
    constexpr int TM = WM / 16, TN = WN / 16;
    constexpr int PAD = 4;
    constexpr int LA = (BM * BK) / NT;  // A elements per thread per K-tile
    constexpr int LB = (BK * BN) / NT;
    static_assert((BM * BK) % NT == 0 && (BK * BN) % NT == 0, "tile not divisible by threads");
    static_assert(BK % 4 == 0, "BK multiple of the MFMA K");
    typedef typename Mfma<T>::acc_t acc_t;

    __shared__ T As[2][BK][BM + PAD];
    __shared__ T Bs[2][BK][BN + PAD];
    A += blockIdx.y * sa;
    B += blockIdx.y * sb;
    C += blockIdx.y * sc;

    const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
    const int wm = wave / WGN, wn = wave % WGN;
    int tm, tn;
    tile_of((M + BM - 1) / BM, (N + BN - 1) / BN, group_m, tm, tn);
    const int bm0 = tm * BM, bn0 = tn * BN;

    acc_t acc[TM][TN];
#pragma unroll
    for (int i = 0; i < TM; ++i)
#pragma unroll
        for (int j = 0; j < TN; ++j) acc[i][j] = acc_t{0, 0, 0, 0};

    T ra[LA], rb[LB];
    auto load_tile = [&](int k0) {
#pragma unroll
        for (int s = 0; s < LA; ++s) {
            const int q = tid + s * NT;
            const int row = q / BK, col = q % BK;
            const int gr = bm0 + row, gc = k0 + col;
            ra[s] = (gr < M && gc < K) ? A[(size_t)gr * K + gc] : (T)0;
        }
#pragma unroll
        for (int s = 0; s < LB; ++s) {
            const int q = tid + s * NT;
            const int row = q / BN, col = q % BN;
            const int gr = k0 + row, gc = bn0 + col;
            rb[s] = (gr < K && gc < N) ? B[(size_t)gr * N + gc] : (T)0;
        }
    };
    auto store_tile = [&](int buf) {
#pragma unroll
        for (int s = 0; s < LA; ++s) {
            const int q = tid + s * NT;
            As[buf][q % BK][q / BK] = ra[s];
        }
#pragma unroll
        for (int s = 0; s < LB; ++s) {
            const int q = tid + s * NT;
            Bs[buf][q / BN][q % BN] = rb[s];
        }
    };

    const int nk = (K + BK - 1) / BK;
    load_tile(0);
    store_tile(0);
    __syncthreads();
    for (int kt = 0; kt < nk; ++kt) {
        const int cur = kt & 1;
        if (kt + 1 < nk) load_tile((kt + 1) * BK);
#pragma unroll
        for (int kk = 0; kk < BK; kk += 4) {
            T a[TM], b[TN];
#pragma unroll
            for (int i = 0; i < TM; ++i) a[i] = As[cur][kk + (lane >> 4)][wm * WM + i * 16 + (lane & 15)];
#pragma unroll
            for (int j = 0; j < TN; ++j) b[j] = Bs[cur][kk + (lane >> 4)][wn * WN + j * 16 + (lane & 15)];
#pragma unroll
            for (int i = 0; i < TM; ++i)
#pragma unroll
                for (int j = 0; j < TN; ++j) acc[i][j] = Mfma<T>::mma(a[i], b[j], acc[i][j]);
        }
        if (kt + 1 < nk) store_tile(cur ^ 1);
        __syncthreads();
    }
#pragma unroll
    for (int i = 0; i < TM; ++i) {
#pragma unroll
        for (int j = 0; j < TN; ++j) {
            const int col = bn0 + wn * WN + j * 16 + (lane & 15);
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int row = bm0 + wm * WM + i * 16 + Mfma<T>::row(lane, r);
                if (row < M && col < N) C[(size_t)row * N + col] = acc[i][j][r];
            }
        }
    }
}

template <typename T>
__global__ void zero_fill_kernel(T *c, size_t n) {
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) c[i] = (T)0;
}

// `batch` panels (grid y), at most 65535 per launch: larger batches go out in several launches.
template <typename T, int BM, int BN, int BK, int WM, int WN>
static nrk_status launch_gemm(int m, int n, int k, const T *a, const T *b, T *c, hipStream_t stream, size_t batch = 1,
                              size_t sa = 0, size_t sb = 0, size_t sc = 0) {
    const int nbm = (m + BM - 1) / BM, nbn = (n + BN - 1) / BN;
    const int threads = (BM / WM) * (BN / WN) * 64;
    const int group_m = 8;
    for (size_t t0 = 0; t0 < batch; t0 += 65535) {
        const unsigned nb = (unsigned)std::min<size_t>(65535, batch - t0);
        hipLaunchKernelGGL((gemm_kernel<T, BM, BN, BK, WM, WN>), dim3(nbm * nbn, nb), dim3(threads), 0, stream, m, n, k,
                           a + t0 * sa, b + t0 * sb, c + t0 * sc, group_m, sa, sb, sc);
        NRK_HIP(hipGetLastError());
    }
    return NRK_OK;
}

// Default tile: the largest of 128 x 128, 64 x 64 and 32 x 32 that still gives every CU a
// workgroup over the whole batch (measured at 512^3 f64: 17.8 us with 32 x 32 against 35.3 us
// with 64 x 64; at 2048^3 128 x 128 is best). The tile never changes a result bit.
template <typename T>
static nrk_status launch_default(size_t batch, size_t m, size_t k, size_t n, const T *a, size_t sa, const T *b, size_t sb,
                                 T *c, size_t sc, hipStream_t stream) {
    const size_t cus = (size_t)cu_count();
    const int M = (int)m, N = (int)n, K = (int)k;
    if (batch * ceil_div(m, 128) * ceil_div(n, 128) >= cus)
        return launch_gemm<T, 128, 128, 16, 64, 64>(M, N, K, a, b, c, stream, batch, sa, sb, sc);
    if (batch * ceil_div(m, 64) * ceil_div(n, 64) > cus)
        return launch_gemm<T, 64, 64, 16, 32, 32>(M, N, K, a, b, c, stream, batch, sa, sb, sc);
    return launch_gemm<T, 32, 32, 16, 16, 16>(M, N, K, a, b, c, stream, batch, sa, sb, sc);
}

template <typename T>
nrk_status gemm(size_t m, size_t k, size_t n, const T *a, const T *b, T *c, hipStream_t stream) {
    NRK_TRY(ensure_device());
    if (m == 0 || n == 0) return NRK_OK;
    if (!c) return fail(NRK_ERR_INVALID_ARGUMENT, "C is NULL");
    if (k == 0) {
        const size_t total = m * n;
        int grid = (int)std::min<size_t>(ceil_div(total, 256), (size_t)cu_count() * 8);
        hipLaunchKernelGGL(zero_fill_kernel<T>, dim3(grid), dim3(256), 0, stream, c, total);
        NRK_HIP(hipGetLastError());
        return NRK_OK;
    }
    if (!a || !b) return fail(NRK_ERR_INVALID_ARGUMENT, "A or B is NULL");
    if (m > (size_t)INT32_MAX / 2 || n > (size_t)INT32_MAX / 2 || k > (size_t)INT32_MAX / 2)
        return fail(NRK_ERR_INVALID_ARGUMENT, "dimension exceeds 2^30");
    const int cfg = select_config(sizeof(T) == 8 ? NRK_DT_F64 : NRK_DT_F32, m, n, k, stream);
    if (cfg >= 0) return gemm_config<T>(cfg, m, k, n, a, b, c, stream);
    return launch_default<T>(1, m, k, n, a, 0, b, 0, c, 0, stream);
}

template nrk_status gemm<double>(size_t, size_t, size_t, const double *, const double *, double *, hipStream_t);
template nrk_status gemm<float>(size_t, size_t, size_t, const float *, const float *, float *, hipStream_t);

// C[t] := A[t] * B[t] for t < batch; C panels are contiguous (sc = m n), A and B panels sit
// `sa` / `sb` elements apart (0: one panel shared by every t). The tile is chosen as for one
// gemm_2d but counting the tiles of the whole batch.
template <typename T>
nrk_status batched_gemm(size_t batch, size_t m, size_t k, size_t n, const T *a, size_t sa, const T *b, size_t sb, T *c,
                        hipStream_t stream) {
    NRK_TRY(ensure_device());
    if (batch == 0 || m == 0 || n == 0) return NRK_OK;
    if (!c) return fail(NRK_ERR_INVALID_ARGUMENT, "C is NULL");
    if (k == 0) return gemm<T>(batch * m, 0, n, a, b, c, stream);  // every panel zero-filled
    if (!a || !b) return fail(NRK_ERR_INVALID_ARGUMENT, "A or B is NULL");
    if (m > (size_t)INT32_MAX / 2 || n > (size_t)INT32_MAX / 2 || k > (size_t)INT32_MAX / 2)
        return fail(NRK_ERR_INVALID_ARGUMENT, "dimension exceeds 2^30");
    return launch_default<T>(batch, m, k, n, a, sa, b, sb, c, m * n, stream);
}

// Tile configurations the tuner times and the selector chooses from (all 4 wavefronts).
template <typename T>
nrk_status gemm_config(int config, size_t m, size_t k, size_t n, const T *a, const T *b, T *c, hipStream_t stream) {
    NRK_TRY(ensure_device());
    if (m == 0 || n == 0 || k == 0) return gemm<T>(m, k, n, a, b, c, stream);
    if (!a || !b || !c) return fail(NRK_ERR_INVALID_ARGUMENT, "A, B or C is NULL");
    const int M = (int)m, N = (int)n, K = (int)k;
    switch (config) {
        case 0: return launch_gemm<T, 32, 32, 16, 16, 16>(M, N, K, a, b, c, stream);
        case 1: return launch_gemm<T, 64, 64, 16, 32, 32>(M, N, K, a, b, c, stream);
        case 2: return launch_gemm<T, 128, 64, 16, 64, 32>(M, N, K, a, b, c, stream);
        case 3: return launch_gemm<T, 64, 128, 16, 32, 64>(M, N, K, a, b, c, stream);
        case 4: return launch_gemm<T, 128, 128, 16, 64, 64>(M, N, K, a, b, c, stream);
        default: return fail(NRK_ERR_INVALID_ARGUMENT, "GEMM config %d outside [0,4]", config);
    }
}
template nrk_status gemm_config<double>(int, size_t, size_t, size_t, const double *, const double *, double *, hipStream_t);
template nrk_status gemm_config<float>(int, size_t, size_t, size_t, const float *, const float *, float *, hipStream_t);

}  // namespace nrk

using namespace nrk;

namespace {
template <typename T>
nrk_status host_gemm(size_t m, size_t k, size_t n, const T *a, size_t a_len, const T *b, size_t b_len, T *c, size_t c_len) {
    NRK_TRY(ensure_device());
    if (c_len > 0 && !c) return fail(NRK_ERR_INVALID_ARGUMENT, "c is NULL");
    const bool lens_ok = a_len == m * k && b_len == k * n && c_len == m * n;
    // gemm.rs:250-253: any other length goes to gemm_generic (gemm.rs:408-445), which zeroes C and
    // then, if every slice is long enough, multiplies the leading parts in its own k order
    if (!lens_ok) return host_gemm_generic<T>(m, k, n, a, a_len, b, b_len, c, c_len);
    if (m == 0 || n == 0) return NRK_OK;
    void *da = nullptr, *db = nullptr, *dc = nullptr;
    NRK_TRY(staging(0, a_len * sizeof(T) + 16, &da));
    NRK_TRY(staging(2, b_len * sizeof(T) + 16, &db));
    NRK_TRY(staging(1, c_len * sizeof(T) + 16, &dc));
    hipStream_t s = nullptr;
    if (a_len) NRK_TRY(copy_h2d(da, a, a_len * sizeof(T), s));
    if (b_len) NRK_TRY(copy_h2d(db, b, b_len * sizeof(T), s));
    NRK_TRY(gemm<T>(m, k, n, (const T *)da, (const T *)db, (T *)dc, s));
    NRK_TRY(copy_d2h(c, dc, c_len * sizeof(T), s));
    NRK_HIP(hipStreamSynchronize(s));
    return NRK_OK;
}

// array/linalg.rs:56-136 batched_gemm on host slices: batch, m or n = 0, or slice lengths other
// than batch m k / batch k n / batch m n, return without writing C (linalg.rs:60-73); otherwise
// every panel is gemm_2d's (so k = 0 zero-fills C).
template <typename T>
nrk_status host_batched_gemm(size_t batch, size_t m, size_t k, size_t n, const T *a, size_t a_len, const T *b, size_t b_len,
                             T *c, size_t c_len) {
    NRK_TRY(ensure_device());
    if (batch == 0 || m == 0 || n == 0) return NRK_OK;
    if (a_len != batch * m * k || b_len != batch * k * n || c_len != batch * m * n) return NRK_OK;
    if (!c) return fail(NRK_ERR_INVALID_ARGUMENT, "c is NULL");
    if (k > 0 && (!a || !b)) return fail(NRK_ERR_INVALID_ARGUMENT, "a or b is NULL");
    void *da = nullptr, *db = nullptr, *dc = nullptr;
    NRK_TRY(staging(0, a_len * sizeof(T) + 16, &da));
    NRK_TRY(staging(2, b_len * sizeof(T) + 16, &db));
    NRK_TRY(staging(1, c_len * sizeof(T) + 16, &dc));
    hipStream_t s = nullptr;
    if (a_len) NRK_TRY(copy_h2d(da, a, a_len * sizeof(T), s));
    if (b_len) NRK_TRY(copy_h2d(db, b, b_len * sizeof(T), s));
    NRK_TRY(batched_gemm<T>(batch, m, k, n, (const T *)da, m * k, (const T *)db, k * n, (T *)dc, s));
    NRK_TRY(copy_d2h(c, dc, c_len * sizeof(T), s));
    NRK_HIP(hipStreamSynchronize(s));
    return NRK_OK;
}
}  // namespace

extern "C" {

nrk_status nrk_gemm_2d_f64(size_t m, size_t k, size_t n, const double *a, size_t a_len, const double *b, size_t b_len,
                           double *c, size_t c_len) {
    return host_gemm<double>(m, k, n, a, a_len, b, b_len, c, c_len);
}
nrk_status nrk_gemm_2d_f32(size_t m, size_t k, size_t n, const float *a, size_t a_len, const float *b, size_t b_len,
                           float *c, size_t c_len) {
    return host_gemm<float>(m, k, n, a, a_len, b, b_len, c, c_len);
}
nrk_status nrk_dev_gemm_f64(size_t m, size_t k, size_t n, const double *a, const double *b, double *c, nrk_stream stream) {
    return gemm<double>(m, k, n, a, b, c, (hipStream_t)stream);
}
nrk_status nrk_dev_gemm_f32(size_t m, size_t k, size_t n, const float *a, const float *b, float *c, nrk_stream stream) {
    return gemm<float>(m, k, n, a, b, c, (hipStream_t)stream);
}
nrk_status nrk_batched_gemm_f64(size_t batch, size_t m, size_t k, size_t n, const double *a, size_t a_len, const double *b,
                                size_t b_len, double *c, size_t c_len) {
    return host_batched_gemm<double>(batch, m, k, n, a, a_len, b, b_len, c, c_len);
}
nrk_status nrk_batched_gemm_f32(size_t batch, size_t m, size_t k, size_t n, const float *a, size_t a_len, const float *b,
                                size_t b_len, float *c, size_t c_len) {
    return host_batched_gemm<float>(batch, m, k, n, a, a_len, b, b_len, c, c_len);
}
nrk_status nrk_dev_batched_gemm_f64(size_t batch, size_t m, size_t k, size_t n, const double *a, size_t stride_a,
                                    const double *b, size_t stride_b, double *c, nrk_stream stream) {
    return batched_gemm<double>(batch, m, k, n, a, stride_a, b, stride_b, c, (hipStream_t)stream);
}
nrk_status nrk_dev_batched_gemm_f32(size_t batch, size_t m, size_t k, size_t n, const float *a, size_t stride_a,
                                    const float *b, size_t stride_b, float *c, nrk_stream stream) {
    return batched_gemm<float>(batch, m, k, n, a, stride_a, b, stride_b, c, (hipStream_t)stream);
}

}  // extern "C"
