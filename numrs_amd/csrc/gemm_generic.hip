// gemm_generic (kernels/gemm.rs:408-445) on the GPU: C := A * B for any element type, every
// output accumulated in T in ascending k from T::zero(), one `acc + a_ik * b_kj` per step.
//
// This is the reference's path for every T that is not f32/f64 (integers: Array<i32/i64>::matmul,
// tests/test_matmul_dispatch.rs:205-219), and for f32/f64 when gemm_2d's slice lengths do not
// match m*k, k*n, m*n exactly (gemm.rs:250-253). Integers wrap, as Rust's release arithmetic does:
// the products and sums are formed in the unsigned type of the same width, so the result is the
// exact product modulo 2^32 / 2^64 (the same bits for the signed and unsigned types, and for
// every summation order). Floats keep the reference's order to the bit: ascending k, a product
// rounded, then a sum rounded (never an FMA).
//
// Tiling: a workgroup of 256 threads owns a 64 x 64 block of C; K moves through LDS in steps of
// 32 (A as [k][row], B as [k][col]); thread (ty, tx) owns rows ty + 16 i and columns tx + 16 j,
// i, j < 4. The k loop of a step stops at K, so no padding term is ever added.
#include "ops.hpp"

namespace nrk {
namespace {

constexpr int GB = 64, GK = 32, GT = 256;

template <typename T>
struct Arith {  // floats: one IEEE rounding per operation
    typedef T acc_t;
    __device__ static acc_t load(T v) { return v; }
    __device__ static acc_t madd(acc_t acc, acc_t a, acc_t b) { return rn_add(acc, rn_mul(a, b)); }
    __device__ static T store(acc_t v) { return v; }
};
template <>
struct Arith<int32_t> {  // wrapping 32-bit
    typedef uint32_t acc_t;
    __device__ static acc_t load(int32_t v) { return (uint32_t)v; }
    __device__ static acc_t madd(acc_t acc, acc_t a, acc_t b) { return acc + a * b; }
    __device__ static int32_t store(acc_t v) { return (int32_t)v; }
};
template <>
struct Arith<int64_t> {  // wrapping 64-bit
    typedef uint64_t acc_t;
    __device__ static acc_t load(int64_t v) { return (uint64_t)v; }
    __device__ static acc_t madd(acc_t acc, acc_t a, acc_t b) { return acc + a * b; }
    __device__ static int64_t store(acc_t v) { return (int64_t)v; }
};

// Panel z of the batch: C_z (m x n) = A_z (m x k) * B_z (k x n), panel strides in elements.
template <typename T>
__global__ __launch_bounds__(GT) void gemm_generic_kernel(int m, int n, int k, const T *__restrict__ A, size_t sa,
                                                          const T *__restrict__ B, size_t sb, T *__restrict__ Cm,
                                                          size_t sc, size_t panel0) {
    typedef typename Arith<T>::acc_t U;
    __shared__ U As[GK][GB + 1];
    __shared__ U Bs[GK][GB + 1];
    const size_t z = panel0 + blockIdx.z;
    const T *a = A + z * sa;
    const T *b = B + z * sb;
    T *c = Cm + z * sc;
    const int t = threadIdx.x, tx = t & 15, ty = t >> 4;
    const int r0 = blockIdx.y * GB, c0 = blockIdx.x * GB;
    U acc[4][4];
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < 4; ++j) acc[i][j] = (U)0;
    for (int k0 = 0; k0 < k; k0 += GK) {
        const int kl = min(GK, k - k0);
        // A block: 64 rows x 32 k (k fastest across threads); B block: 32 k x 64 cols
        for (int e = t; e < GB * GK; e += GT) {
            const int rr = e / GK, kk = e % GK;
            const int r = r0 + rr, kx = k0 + kk;
            As[kk][rr] = (r < m && kk < kl) ? Arith<T>::load(a[(size_t)r * k + kx]) : (U)0;
        }
        for (int e = t; e < GB * GK; e += GT) {
            const int kk = e / GB, cc = e % GB;
            const int col = c0 + cc, kx = k0 + kk;
            Bs[kk][cc] = (col < n && kk < kl) ? Arith<T>::load(b[(size_t)kx * n + col]) : (U)0;
        }
        __syncthreads();
        for (int kk = 0; kk < kl; ++kk) {  // kl is uniform: ascending k, no padding term
            U av[4], bv[4];
#pragma unroll
            for (int i = 0; i < 4; ++i) av[i] = As[kk][ty + 16 * i];
#pragma unroll
            for (int j = 0; j < 4; ++j) bv[j] = Bs[kk][tx + 16 * j];
#pragma unroll
            for (int i = 0; i < 4; ++i)
#pragma unroll
                for (int j = 0; j < 4; ++j) acc[i][j] = Arith<T>::madd(acc[i][j], av[i], bv[j]);
        }
        __syncthreads();
    }
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        const int r = r0 + ty + 16 * i;
        if (r >= m) continue;
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const int col = c0 + tx + 16 * j;
            if (col < n) c[(size_t)r * n + col] = Arith<T>::store(acc[i][j]);
        }
    }
}

template <typename T>
__global__ void fill_zero_kernel(T *p, size_t count) {
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < count; i += (size_t)gridDim.x * blockDim.x) p[i] = (T)0;
}

}  // namespace

// batch panels (strides in elements; a stride of 0 broadcasts that operand); k = 0 gives zeros.
template <typename T>
nrk_status gemm_generic(size_t batch, size_t m, size_t k, size_t n, const T *a, size_t sa, const T *b, size_t sb, T *c,
                        hipStream_t stream) {
    NRK_TRY(ensure_device());
    if (batch == 0 || m == 0 || n == 0) return NRK_OK;
    if (!c) return fail(NRK_ERR_INVALID_ARGUMENT, "C is NULL");
    if (k > 0 && (!a || !b)) return fail(NRK_ERR_INVALID_ARGUMENT, "A or B is NULL");
    if (m > (1u << 30) || n > (1u << 30) || k > (1u << 30)) return fail(NRK_ERR_INVALID_ARGUMENT, "dimension too large");
    if (k == 0) {
        const size_t total = batch * m * n;
        hipLaunchKernelGGL(fill_zero_kernel<T>, dim3((unsigned)std::min<size_t>(ceil_div(total, 256), 65535)), dim3(256), 0,
                           stream, c, total);
        NRK_HIP(hipGetLastError());
        return NRK_OK;
    }
    const unsigned gx = (unsigned)ceil_div(n, (size_t)GB), gy = (unsigned)ceil_div(m, (size_t)GB);
    if (gy > 65535) return fail(NRK_ERR_INVALID_ARGUMENT, "m too large for the generic GEMM grid");
    for (size_t z0 = 0; z0 < batch; z0 += 65535) {
        const unsigned gz = (unsigned)std::min<size_t>(65535, batch - z0);
        hipLaunchKernelGGL(gemm_generic_kernel<T>, dim3(gx, gy, gz), dim3(GT), 0, stream, (int)m, (int)n, (int)k, a, sa, b, sb,
                           c, m * n, z0);
        NRK_HIP(hipGetLastError());
    }
    return NRK_OK;
}
template nrk_status gemm_generic<float>(size_t, size_t, size_t, size_t, const float *, size_t, const float *, size_t, float *,
                                        hipStream_t);
template nrk_status gemm_generic<double>(size_t, size_t, size_t, size_t, const double *, size_t, const double *, size_t,
                                         double *, hipStream_t);
template nrk_status gemm_generic<int32_t>(size_t, size_t, size_t, size_t, const int32_t *, size_t, const int32_t *, size_t,
                                          int32_t *, hipStream_t);
template nrk_status gemm_generic<int64_t>(size_t, size_t, size_t, size_t, const int64_t *, size_t, const int64_t *, size_t,
                                          int64_t *, hipStream_t);

// The host-slice contract of gemm_generic (gemm.rs:408-423): C is zeroed over its own length;
// with any slice shorter than its matrix nothing else happens; otherwise the leading m*k, k*n
// elements are multiplied into the leading m*n of C.
template <typename T>
nrk_status host_gemm_generic(size_t m, size_t k, size_t n, const T *a, size_t a_len, const T *b, size_t b_len, T *c,
                             size_t c_len) {
    NRK_TRY(ensure_device());
    if (c_len > 0 && !c) return fail(NRK_ERR_INVALID_ARGUMENT, "c is NULL");
    for (size_t i = 0; i < c_len; ++i) c[i] = (T)0;
    if (a_len < m * k || b_len < k * n || c_len < m * n) return NRK_OK;
    if (m == 0 || n == 0 || k == 0) return NRK_OK;
    if ((!a && m * k) || (!b && k * n)) return fail(NRK_ERR_INVALID_ARGUMENT, "a or b is NULL");
    void *da = nullptr, *db = nullptr, *dc = nullptr;
    NRK_TRY(staging(0, m * k * sizeof(T) + 16, &da));
    NRK_TRY(staging(2, k * n * sizeof(T) + 16, &db));
    NRK_TRY(staging(1, m * n * sizeof(T) + 16, &dc));
    hipStream_t s = nullptr;
    NRK_TRY(copy_h2d(da, a, m * k * sizeof(T), s));
    NRK_TRY(copy_h2d(db, b, k * n * sizeof(T), s));
    NRK_TRY(gemm_generic<T>(1, m, k, n, (const T *)da, 0, (const T *)db, 0, (T *)dc, s));
    NRK_TRY(copy_d2h(c, dc, m * n * sizeof(T), s));
    NRK_HIP(hipStreamSynchronize(s));
    return NRK_OK;
}
template nrk_status host_gemm_generic<float>(size_t, size_t, size_t, const float *, size_t, const float *, size_t, float *,
                                             size_t);
template nrk_status host_gemm_generic<double>(size_t, size_t, size_t, const double *, size_t, const double *, size_t,
                                              double *, size_t);

}  // namespace nrk

using namespace nrk;

extern "C" {

nrk_status nrk_gemm_generic_i64(size_t m, size_t k, size_t n, const int64_t *a, size_t a_len, const int64_t *b, size_t b_len,
                                int64_t *c, size_t c_len) {
    return host_gemm_generic<int64_t>(m, k, n, a, a_len, b, b_len, c, c_len);
}
nrk_status nrk_gemm_generic_i32(size_t m, size_t k, size_t n, const int32_t *a, size_t a_len, const int32_t *b, size_t b_len,
                                int32_t *c, size_t c_len) {
    return host_gemm_generic<int32_t>(m, k, n, a, a_len, b, b_len, c, c_len);
}
nrk_status nrk_dev_gemm_generic_i64(size_t batch, size_t m, size_t k, size_t n, const int64_t *a, size_t stride_a,
                                    const int64_t *b, size_t stride_b, int64_t *c, nrk_stream stream) {
    return gemm_generic<int64_t>(batch, m, k, n, a, stride_a, b, stride_b, c, (hipStream_t)stream);
}
nrk_status nrk_dev_gemm_generic_i32(size_t batch, size_t m, size_t k, size_t n, const int32_t *a, size_t stride_a,
                                    const int32_t *b, size_t stride_b, int32_t *c, nrk_stream stream) {
    return gemm_generic<int32_t>(batch, m, k, n, a, stride_a, b, stride_b, c, (hipStream_t)stream);
}

}  // extern "C"
