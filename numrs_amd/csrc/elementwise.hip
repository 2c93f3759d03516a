// Elementwise map (scalar chain) and zip (binary op) kernels for gfx950.
//
// Reference: kernels/elementwise.rs:51-258 applies a Rust closure per element; on the GPU the
// closure becomes an op code (a closure cannot be shipped to the device). Every op is one IEEE
// rounding: `x * k + c` is two roundings, never an FMA (array/operations.rs:448-474,
// expr/fused_eval.rs:15-21).
//
// Both kernels stream 16-byte vectors with non-temporal loads, grid-stride over the array, and
// fall back to scalar accesses for a misaligned base or the tail.
#include "common.hpp"

namespace nrk {

constexpr int EB = 256;  // threads per workgroup
constexpr int EUNROLL = 4;

template <typename T>
__device__ __forceinline__ T apply_bop(int op, T a, T b) {
    switch (op) {
        case NRK_BOP_ADD: return rn_add(a, b);
        case NRK_BOP_SUB: return rn_sub(a, b);
        case NRK_BOP_MUL: return rn_mul(a, b);
        case NRK_BOP_DIV: return rn_div(a, b);
        case NRK_BOP_REM: return __builtin_fmod(a, b);  // exact, truncated like Rust's %
        case NRK_BOP_MAXIMUM: return a > b ? a : b;      // ufuncs.rs:169
        case NRK_BOP_MINIMUM: return a < b ? a : b;      // ufuncs.rs:173
        default: return a;
    }
}

template <typename T>
__device__ __forceinline__ T apply_chain(const Chain &c, T v) {
    for (int k = 0; k < c.len; ++k) v = apply_sop<T>(c.op[k], (T)c.s[k], v);
    return v;
}

template <typename T>
__global__ __launch_bounds__(EB) void map_kernel(Chain chain, const T *x, T *y, size_t n, int vec_ok) {
    typedef typename Vec16<T>::type V;
    constexpr int VN = Vec16<T>::N;
    const size_t tid = blockIdx.x * (size_t)EB + threadIdx.x;
    const size_t nthreads = (size_t)gridDim.x * EB;
    size_t done = 0;
    if (vec_ok) {
        const size_t nv = n / VN;
        const V *xv = reinterpret_cast<const V *>(x);
        V *yv = reinterpret_cast<V *>(y);
        size_t i = tid;
        for (; i + (EUNROLL - 1) * nthreads < nv; i += EUNROLL * nthreads) {
            V v[EUNROLL];
#pragma unroll
            for (int u = 0; u < EUNROLL; ++u) v[u] = __builtin_nontemporal_load(xv + i + u * nthreads);
#pragma unroll
            for (int u = 0; u < EUNROLL; ++u) {
#pragma unroll
                for (int e = 0; e < VN; ++e) v[u][e] = apply_chain(chain, v[u][e]);
                __builtin_nontemporal_store(v[u], yv + i + u * nthreads);
            }
        }
        for (; i < nv; i += nthreads) {
            V v = __builtin_nontemporal_load(xv + i);
#pragma unroll
            for (int e = 0; e < VN; ++e) v[e] = apply_chain(chain, v[e]);
            __builtin_nontemporal_store(v, yv + i);
        }
        done = nv * VN;
    }
    for (size_t i = done + tid; i < n; i += nthreads) y[i] = apply_chain(chain, x[i]);
}

template <typename T>
__global__ __launch_bounds__(EB) void zip_kernel(int op, const T *a, const T *b, T *y, size_t n, int vec_ok) {
    typedef typename Vec16<T>::type V;
    constexpr int VN = Vec16<T>::N;
    const size_t tid = blockIdx.x * (size_t)EB + threadIdx.x;
    const size_t nthreads = (size_t)gridDim.x * EB;
    size_t done = 0;
    if (vec_ok) {
        const size_t nv = n / VN;
        const V *av = reinterpret_cast<const V *>(a);
        const V *bv = reinterpret_cast<const V *>(b);
        V *yv = reinterpret_cast<V *>(y);
        size_t i = tid;
        for (; i + (EUNROLL - 1) * nthreads < nv; i += EUNROLL * nthreads) {
            V va[EUNROLL], vb[EUNROLL];
#pragma unroll
            for (int u = 0; u < EUNROLL; ++u) {
                va[u] = __builtin_nontemporal_load(av + i + u * nthreads);
                vb[u] = __builtin_nontemporal_load(bv + i + u * nthreads);
            }
#pragma unroll
            for (int u = 0; u < EUNROLL; ++u) {
#pragma unroll
                for (int e = 0; e < VN; ++e) va[u][e] = apply_bop(op, va[u][e], vb[u][e]);
                __builtin_nontemporal_store(va[u], yv + i + u * nthreads);
            }
        }
        for (; i < nv; i += nthreads) {
            V va = __builtin_nontemporal_load(av + i);
            V vb = __builtin_nontemporal_load(bv + i);
#pragma unroll
            for (int e = 0; e < VN; ++e) va[e] = apply_bop(op, va[e], vb[e]);
            __builtin_nontemporal_store(va, yv + i);
        }
        done = nv * VN;
    }
    for (size_t i = done + tid; i < n; i += nthreads) y[i] = apply_bop(op, a[i], b[i]);
}

static int stream_grid(size_t n_vec) {
    size_t want = ceil_div(n_vec, (size_t)EB * EUNROLL);
    size_t cap = (size_t)cu_count() * 16;
    if (want < 1) want = 1;
    return (int)(want < cap ? want : cap);
}

template <typename T>
nrk_status map(const Chain &chain, const T *x, T *y, size_t n, hipStream_t stream) {
    NRK_TRY(ensure_device());
    if (n == 0) return NRK_OK;
    if (!x || !y) return fail(NRK_ERR_INVALID_ARGUMENT, "x or y is NULL");
    const int vec_ok = (reinterpret_cast<uintptr_t>(x) % 16 == 0) && (reinterpret_cast<uintptr_t>(y) % 16 == 0);
    hipLaunchKernelGGL(map_kernel<T>, dim3(stream_grid(n / Vec16<T>::N + 1)), dim3(EB), 0, stream, chain, x, y, n, vec_ok);
    NRK_HIP(hipGetLastError());
    return NRK_OK;
}

template <typename T>
nrk_status zip(int op, const T *a, const T *b, T *y, size_t n, hipStream_t stream) {
    NRK_TRY(ensure_device());
    if (op < 0 || op >= NRK_BOP_COUNT) return fail(NRK_ERR_INVALID_ARGUMENT, "op %d is not an NRK_BOP_* code", op);
    if (n == 0) return NRK_OK;
    if (!a || !b || !y) return fail(NRK_ERR_INVALID_ARGUMENT, "a, b or y is NULL");
    const int vec_ok = (reinterpret_cast<uintptr_t>(a) % 16 == 0) && (reinterpret_cast<uintptr_t>(b) % 16 == 0) &&
                       (reinterpret_cast<uintptr_t>(y) % 16 == 0);
    hipLaunchKernelGGL(zip_kernel<T>, dim3(stream_grid(n / Vec16<T>::N + 1)), dim3(EB), 0, stream, op, a, b, y, n, vec_ok);
    NRK_HIP(hipGetLastError());
    return NRK_OK;
}

template nrk_status map<float>(const Chain &, const float *, float *, size_t, hipStream_t);
template nrk_status map<double>(const Chain &, const double *, double *, size_t, hipStream_t);
template nrk_status zip<float>(int, const float *, const float *, float *, size_t, hipStream_t);
template nrk_status zip<double>(int, const double *, const double *, double *, size_t, hipStream_t);

}  // namespace nrk
