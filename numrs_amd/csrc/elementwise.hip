// Elementwise map (scalar chain) and zip (binary op) kernels for gfx950.
//
// Reference: kernels/elementwise.rs:51-258 applies a Rust closure per element; on the GPU the
// closure becomes an op code (a closure cannot be shipped to the device). Every op is one IEEE
// rounding: `x * k + c` is two roundings, never an FMA (array/operations.rs:448-474,
// expr/fused_eval.rs:15-21).
//
// Both kernels stream 16-byte vectors with non-temporal loads, grid-stride over the array, and
// fall back to scalar accesses for a misaligned base or the tail. The op dispatch is hoisted out
// of the element loops: the binary op is a template parameter, and a scalar chain is applied
// one step at a time to a whole batch of registers (one uniform branch per step).
#include "common.hpp"

namespace nrk {

constexpr int EB = 256;  // threads per workgroup
constexpr int EUNROLL = 2;  // 16-byte vectors in flight per thread and pass

template <typename T, int OP>
__device__ __forceinline__ T apply_bop(T a, T b) {
    if constexpr (OP == NRK_BOP_ADD) return rn_add(a, b);
    else if constexpr (OP == NRK_BOP_SUB) return rn_sub(a, b);
    else if constexpr (OP == NRK_BOP_MUL) return rn_mul(a, b);
    else if constexpr (OP == NRK_BOP_DIV) return rn_div(a, b);
    else if constexpr (OP == NRK_BOP_REM) return __builtin_fmod(a, b);  // exact, truncated like Rust's %
    else if constexpr (OP == NRK_BOP_MAXIMUM) return a > b ? a : b;     // ufuncs.rs:169
    else if constexpr (OP == NRK_BOP_MINIMUM) return a < b ? a : b;     // ufuncs.rs:173
    else if constexpr (OP == NRK_BOP_COPYSIGN) return dev_copysign(a, b);
    else if constexpr (OP == NRK_BOP_HYPOT) return (T)fm::hypot((double)a, (double)b);
    else if constexpr (OP == NRK_BOP_ATAN2) return (T)fm::atan2((double)a, (double)b);
    else if constexpr (OP == NRK_BOP_PRELU) return a > (T)0 ? a : rn_mul(b, a);  // ufuncs.rs:1394
    else if constexpr (OP == NRK_BOP_POWER) return fm::pow_t(a, b);  // ufuncs.rs:166
    else return (T)fm::logaddexp((double)a, (double)b);                // LOGADDEXP
}

// The map applies the NRK_SOP_* steps of its class (MATH_NONE / LIGHT / ALL), or, with OP >= 0,
// the chain whose single elementary-function step OP sits at index `at`. The host launches the
// one-op kernel when the chain has exactly one such step (every ufunc call), else the smallest
// class holding every step. One kernel over all cases held ~200 VGPRs (2 waves per SIMD): the
// loop over steps kept every case's constants live; each op alone needs at most 64.
template <typename T, int CLS, int OP, int N>
__device__ __forceinline__ void map_chain(const Chain &c, int at, T (&v)[N]) {
    if constexpr (OP >= 0) apply_chain_one_n<T, N, OP>(c, at, v);
    else apply_chain_n<T, N, CLS>(c, v);
}

template <typename T, int CLS, int OP = -1>
__global__ __launch_bounds__(EB) void map_kernel(Chain chain, int at, const T *x, T *y, size_t n, int vec_ok) {
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
            T r[EUNROLL * VN];
#pragma unroll
            for (int u = 0; u < EUNROLL; ++u) {
                const V v = __builtin_nontemporal_load(xv + i + u * nthreads);
#pragma unroll
                for (int e = 0; e < VN; ++e) r[u * VN + e] = v[e];
            }
            map_chain<T, CLS, OP>(chain, at, r);
#pragma unroll
            for (int u = 0; u < EUNROLL; ++u) {
                V v;
#pragma unroll
                for (int e = 0; e < VN; ++e) v[e] = r[u * VN + e];
                __builtin_nontemporal_store(v, yv + i + u * nthreads);
            }
        }
        for (; i < nv; i += nthreads) {
            const V v = __builtin_nontemporal_load(xv + i);
            T r[VN];
#pragma unroll
            for (int e = 0; e < VN; ++e) r[e] = v[e];
            map_chain<T, CLS, OP>(chain, at, r);
            V w;
#pragma unroll
            for (int e = 0; e < VN; ++e) w[e] = r[e];
            __builtin_nontemporal_store(w, yv + i);
        }
        done = nv * VN;
    }
    for (size_t i = done + tid; i < n; i += nthreads) {
        T r[1] = {x[i]};
        map_chain<T, CLS, OP>(chain, at, r);
        y[i] = r[0];
    }
}

template <typename T, int OP>
__global__ __launch_bounds__(EB) void zip_kernel(const T *a, const T *b, T *y, size_t n, int vec_ok) {
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
                for (int e = 0; e < VN; ++e) va[u][e] = apply_bop<T, OP>(va[u][e], vb[u][e]);
                __builtin_nontemporal_store(va[u], yv + i + u * nthreads);
            }
        }
        for (; i < nv; i += nthreads) {
            V va = __builtin_nontemporal_load(av + i);
            V vb = __builtin_nontemporal_load(bv + i);
#pragma unroll
            for (int e = 0; e < VN; ++e) va[e] = apply_bop<T, OP>(va[e], vb[e]);
            __builtin_nontemporal_store(va, yv + i);
        }
        done = nv * VN;
    }
    for (size_t i = done + tid; i < n; i += nthreads) y[i] = apply_bop<T, OP>(a[i], b[i]);
}

// Up to 64 workgroups per CU: many short-lived workgroups keep more loads in flight than a
// grid-stride loop over 16 per CU (measured, 8192 x 8192: copy 4.7 -> 5.6 TB/s f64 and 4.6 -> 6.0
// TB/s f32; exp f64 4.4 -> 5.4 TB/s).
static int stream_grid(size_t n_vec) {
    size_t want = ceil_div(n_vec, (size_t)EB * EUNROLL);
    size_t cap = (size_t)cu_count() * 64;
    if (want < 1) want = 1;
    return (int)(want < cap ? want : cap);
}

template <typename T>
nrk_status map(const Chain &chain, const T *x, T *y, size_t n, hipStream_t stream) {
    NRK_TRY(ensure_device());
    if (n == 0) return NRK_OK;
    if (!x || !y) return fail(NRK_ERR_INVALID_ARGUMENT, "x or y is NULL");
    const int vec_ok = (reinterpret_cast<uintptr_t>(x) % 16 == 0) && (reinterpret_cast<uintptr_t>(y) % 16 == 0);
    int cls = MATH_NONE, n_math = 0, at = -1;
    for (int k = 0; k < chain.len; ++k) {
        const int c = sop_math_class(chain.op[k]);
        cls = std::max(cls, c);
        if (c != MATH_NONE) {
            ++n_math;
            at = k;
        }
    }
    const dim3 grid(stream_grid(n / Vec16<T>::N + 1));
    static_assert(NRK_SOP_EXP == 18 && NRK_SOP_COUNT == 58, "the one-op kernels below list codes 18 .. 57");
    if (n_math == 1) {
        switch (chain.op[at]) {
#define NRK_MAP_ONE(OP) \
    case OP: hipLaunchKernelGGL((map_kernel<T, MATH_ALL, OP>), grid, dim3(EB), 0, stream, chain, at, x, y, n, vec_ok); break;
            NRK_MAP_ONE(18) NRK_MAP_ONE(19) NRK_MAP_ONE(20) NRK_MAP_ONE(21) NRK_MAP_ONE(22) NRK_MAP_ONE(23) NRK_MAP_ONE(24)
            NRK_MAP_ONE(25) NRK_MAP_ONE(26) NRK_MAP_ONE(27) NRK_MAP_ONE(28) NRK_MAP_ONE(29) NRK_MAP_ONE(30) NRK_MAP_ONE(31)
            NRK_MAP_ONE(32) NRK_MAP_ONE(33) NRK_MAP_ONE(34) NRK_MAP_ONE(35) NRK_MAP_ONE(36) NRK_MAP_ONE(37) NRK_MAP_ONE(38)
            NRK_MAP_ONE(39) NRK_MAP_ONE(40) NRK_MAP_ONE(41) NRK_MAP_ONE(42) NRK_MAP_ONE(43) NRK_MAP_ONE(44) NRK_MAP_ONE(45)
            NRK_MAP_ONE(46) NRK_MAP_ONE(47) NRK_MAP_ONE(48) NRK_MAP_ONE(49) NRK_MAP_ONE(50) NRK_MAP_ONE(51) NRK_MAP_ONE(52)
            NRK_MAP_ONE(53) NRK_MAP_ONE(54) NRK_MAP_ONE(55) NRK_MAP_ONE(56) NRK_MAP_ONE(57)
#undef NRK_MAP_ONE
            default: return fail(NRK_ERR_INVALID_ARGUMENT, "op %d has no map kernel", chain.op[at]);
        }
    } else if (cls == MATH_NONE) {
        hipLaunchKernelGGL((map_kernel<T, MATH_NONE>), grid, dim3(EB), 0, stream, chain, -1, x, y, n, vec_ok);
    } else if (cls == MATH_LIGHT) {
        hipLaunchKernelGGL((map_kernel<T, MATH_LIGHT>), grid, dim3(EB), 0, stream, chain, -1, x, y, n, vec_ok);
    } else {
        hipLaunchKernelGGL((map_kernel<T, MATH_ALL>), grid, dim3(EB), 0, stream, chain, -1, x, y, n, vec_ok);
    }
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
    const dim3 grid(stream_grid(n / Vec16<T>::N + 1));
    switch (op) {
#define NRK_ZIP_CASE(OP) \
    case OP: hipLaunchKernelGGL((zip_kernel<T, OP>), grid, dim3(EB), 0, stream, a, b, y, n, vec_ok); break;
        NRK_ZIP_CASE(NRK_BOP_ADD)
        NRK_ZIP_CASE(NRK_BOP_SUB)
        NRK_ZIP_CASE(NRK_BOP_MUL)
        NRK_ZIP_CASE(NRK_BOP_DIV)
        NRK_ZIP_CASE(NRK_BOP_REM)
        NRK_ZIP_CASE(NRK_BOP_MAXIMUM)
        NRK_ZIP_CASE(NRK_BOP_MINIMUM)
        NRK_ZIP_CASE(NRK_BOP_COPYSIGN)
        NRK_ZIP_CASE(NRK_BOP_HYPOT)
        NRK_ZIP_CASE(NRK_BOP_ATAN2)
        NRK_ZIP_CASE(NRK_BOP_LOGADDEXP)
        NRK_ZIP_CASE(NRK_BOP_PRELU)
        NRK_ZIP_CASE(NRK_BOP_POWER)
#undef NRK_ZIP_CASE
    }
    NRK_HIP(hipGetLastError());
    return NRK_OK;
}

// ---- three-operand streams: fma (ufuncs.rs:864, a.mul_add(b, c), one rounding) and lerp
// (ufuncs.rs:893, a + t * (b - a), three roundings). c is unused by lerp.
enum { TRI_FMA = 0, TRI_LERP = 1 };

template <typename T, int OP>
__device__ __forceinline__ T apply_tri(T a, T b, T c, T t) {
    if constexpr (OP == TRI_FMA) {
        if constexpr (sizeof(T) == 4) return __builtin_fmaf(a, b, c);  // one binary32 rounding
        else return __builtin_fma(a, b, c);
    }
    else return rn_add(a, rn_mul(t, rn_sub(b, a)));
}

template <typename T, int OP>
__global__ __launch_bounds__(EB) void tri_kernel(const T *a, const T *b, const T *c, T t, T *y, size_t n, int vec_ok) {
    typedef typename Vec16<T>::type V;
    constexpr int VN = Vec16<T>::N;
    const size_t tid = blockIdx.x * (size_t)EB + threadIdx.x;
    const size_t nthreads = (size_t)gridDim.x * EB;
    size_t done = 0;
    if (vec_ok) {
        const size_t nv = n / VN;
        const V *av = reinterpret_cast<const V *>(a);
        const V *bv = reinterpret_cast<const V *>(b);
        const V *cv = reinterpret_cast<const V *>(c);
        V *yv = reinterpret_cast<V *>(y);
        for (size_t i = tid; i < nv; i += nthreads) {
            V va = __builtin_nontemporal_load(av + i);
            const V vb = __builtin_nontemporal_load(bv + i);
            V vc = va;
            if constexpr (OP == TRI_FMA) vc = __builtin_nontemporal_load(cv + i);
#pragma unroll
            for (int e = 0; e < VN; ++e) va[e] = apply_tri<T, OP>(va[e], vb[e], vc[e], t);
            __builtin_nontemporal_store(va, yv + i);
        }
        done = nv * VN;
    }
    for (size_t i = done + tid; i < n; i += nthreads)
        y[i] = apply_tri<T, OP>(a[i], b[i], OP == TRI_FMA ? c[i] : (T)0, t);
}

template <typename T>
nrk_status tri(int op, const T *a, const T *b, const T *c, T t, T *y, size_t n, hipStream_t stream) {
    NRK_TRY(ensure_device());
    if (n == 0) return NRK_OK;
    if (!a || !b || !y || (op == TRI_FMA && !c)) return fail(NRK_ERR_INVALID_ARGUMENT, "an input or y is NULL");
    auto al = [](const void *p) { return reinterpret_cast<uintptr_t>(p) % 16 == 0; };
    const int vec_ok = al(a) && al(b) && al(y) && (op != TRI_FMA || al(c));
    const dim3 grid(stream_grid(n / Vec16<T>::N + 1));
    if (op == TRI_FMA)
        hipLaunchKernelGGL((tri_kernel<T, TRI_FMA>), grid, dim3(EB), 0, stream, a, b, c, t, y, n, vec_ok);
    else
        hipLaunchKernelGGL((tri_kernel<T, TRI_LERP>), grid, dim3(EB), 0, stream, a, b, c, t, y, n, vec_ok);
    NRK_HIP(hipGetLastError());
    return NRK_OK;
}
template nrk_status tri<float>(int, const float *, const float *, const float *, float, float *, size_t, hipStream_t);
template nrk_status tri<double>(int, const double *, const double *, const double *, double, double *, size_t, hipStream_t);

// ---- ufunc_ops.rs:583 ufunc_outer and :841 ufunc_where with UfuncOp::apply (:156): add, subtract,
// multiply, divide, and maximum / minimum that return NaN when either operand is NaN.
template <typename T, int OP>
__device__ __forceinline__ T ufop_apply(T a, T b) {
    if constexpr (OP == NRK_BOP_ADD) return rn_add(a, b);
    else if constexpr (OP == NRK_BOP_SUB) return rn_sub(a, b);
    else if constexpr (OP == NRK_BOP_MUL) return rn_mul(a, b);
    else if constexpr (OP == NRK_BOP_DIV) return rn_div(a, b);
    else if constexpr (OP == NRK_BOP_MAXIMUM) return (a != a || b != b) ? a + b : (a > b ? a : b);
    else return (a != a || b != b) ? a + b : (a < b ? a : b);
}

// y[i * nb + j] = a[i] op b[j]: row i per blockIdx.y step, columns coalesced across the x blocks
template <typename T, int OP>
__global__ __launch_bounds__(EB) void outer_kernel(const T *a, size_t na, const T *b, size_t nb, T *y) {
    for (size_t i = blockIdx.y; i < na; i += gridDim.y) {
        const T ai = a[i];
        for (size_t j = blockIdx.x * (size_t)EB + threadIdx.x; j < nb; j += (size_t)gridDim.x * EB)
            y[i * nb + j] = ufop_apply<T, OP>(ai, b[j]);
    }
}

template <typename T, int OP>
__global__ __launch_bounds__(EB) void where_kernel(const T *a, const T *b, const uint8_t *mask, T *y, size_t n) {
    const size_t nthreads = (size_t)gridDim.x * EB;
    for (size_t i = blockIdx.x * (size_t)EB + threadIdx.x; i < n; i += nthreads) {
        const T av = a[i];
        y[i] = mask[i] ? ufop_apply<T, OP>(av, b[i]) : av;
    }
}

static bool ufunc_op_ok(int op) {
    return op == NRK_BOP_ADD || op == NRK_BOP_SUB || op == NRK_BOP_MUL || op == NRK_BOP_DIV || op == NRK_BOP_MAXIMUM ||
           op == NRK_BOP_MINIMUM;
}

template <typename T>
nrk_status outer(int op, const T *a, size_t na, const T *b, size_t nb, T *y, hipStream_t stream) {
    NRK_TRY(ensure_device());
    if (!ufunc_op_ok(op)) return fail(NRK_ERR_INVALID_ARGUMENT, "outer: op %d is not ADD, SUB, MUL, DIV, MAXIMUM or MINIMUM", op);
    if (na == 0 || nb == 0) return NRK_OK;
    if (!a || !b || !y) return fail(NRK_ERR_INVALID_ARGUMENT, "a, b or y is NULL");
    const dim3 grid((unsigned)std::min<size_t>(ceil_div(nb, (size_t)EB), 64), (unsigned)std::min<size_t>(na, 8192));
    switch (op) {
#define NRK_OUTER_CASE(OP) \
    case OP: hipLaunchKernelGGL((outer_kernel<T, OP>), grid, dim3(EB), 0, stream, a, na, b, nb, y); break;
        NRK_OUTER_CASE(NRK_BOP_ADD) NRK_OUTER_CASE(NRK_BOP_SUB) NRK_OUTER_CASE(NRK_BOP_MUL) NRK_OUTER_CASE(NRK_BOP_DIV)
        NRK_OUTER_CASE(NRK_BOP_MAXIMUM) NRK_OUTER_CASE(NRK_BOP_MINIMUM)
#undef NRK_OUTER_CASE
    }
    NRK_HIP(hipGetLastError());
    return NRK_OK;
}
template nrk_status outer<float>(int, const float *, size_t, const float *, size_t, float *, hipStream_t);
template nrk_status outer<double>(int, const double *, size_t, const double *, size_t, double *, hipStream_t);

template <typename T>
nrk_status where_op(int op, const T *a, const T *b, const uint8_t *mask, T *y, size_t n, hipStream_t stream) {
    NRK_TRY(ensure_device());
    if (!ufunc_op_ok(op)) return fail(NRK_ERR_INVALID_ARGUMENT, "where: op %d is not ADD, SUB, MUL, DIV, MAXIMUM or MINIMUM", op);
    if (n == 0) return NRK_OK;
    if (!a || !b || !mask || !y) return fail(NRK_ERR_INVALID_ARGUMENT, "a, b, mask or y is NULL");
    const dim3 grid((unsigned)std::min<size_t>(ceil_div(n, (size_t)EB), (size_t)cu_count() * 32));
    switch (op) {
#define NRK_WHERE_CASE(OP) \
    case OP: hipLaunchKernelGGL((where_kernel<T, OP>), grid, dim3(EB), 0, stream, a, b, mask, y, n); break;
        NRK_WHERE_CASE(NRK_BOP_ADD) NRK_WHERE_CASE(NRK_BOP_SUB) NRK_WHERE_CASE(NRK_BOP_MUL) NRK_WHERE_CASE(NRK_BOP_DIV)
        NRK_WHERE_CASE(NRK_BOP_MAXIMUM) NRK_WHERE_CASE(NRK_BOP_MINIMUM)
#undef NRK_WHERE_CASE
    }
    NRK_HIP(hipGetLastError());
    return NRK_OK;
}
template nrk_status where_op<float>(int, const float *, const float *, const uint8_t *, float *, size_t, hipStream_t);
template nrk_status where_op<double>(int, const double *, const double *, const uint8_t *, double *, size_t, hipStream_t);

// ---- softmax / log_softmax passes (ufuncs.rs:1411-1457). ms = {max, sum} on the device.
// mode 0: y = exp(x - max); mode 1: y = y / sum; mode 2: y = (x - max) - ln(sum);
// mode 3: y = exp(x - max) / sum (the reference's per-element expression, one pass).
// inv_s: 1 / sum in binary64 (binary32 only): e^(x - max) / sum evaluated unrounded in binary64
// and rounded once, instead of two binary32 roundings and a binary64 division per element
template <typename T, int MODE>
__device__ __forceinline__ T softmax_elem(T xv, T yv, T m, T s, T lse, double inv_s) {
    if constexpr (MODE == 0) return fm::exp_t(rn_sub(xv, m));
    else if constexpr (MODE == 1) return rn_div(yv, s);
    else if constexpr (MODE == 3) {
        if constexpr (sizeof(T) == 4) return (T)(fm::f32::exp_d(rn_sub(xv, m)) * inv_s);
        else return rn_div(fm::exp_t(rn_sub(xv, m)), s);
    } else return rn_sub(rn_sub(xv, m), lse);
}

template <typename T, int MODE>
__global__ __launch_bounds__(EB) void softmax_pass_kernel(const T *x, T *y, const T *ms, size_t n, int vec_ok) {
    typedef typename Vec16<T>::type V;
    constexpr int VN = Vec16<T>::N;
    const T m = ms[0];
    const T s = ms[1];
    const T lse = MODE == 2 ? (T)fm::log((double)s) : (T)0;
    const double inv_s = 1.0 / (double)s;
    const size_t tid = blockIdx.x * (size_t)EB + threadIdx.x;
    const size_t nthreads = (size_t)gridDim.x * EB;
    size_t done = 0;
    if (vec_ok) {  // x and y 16-byte aligned: EUNROLL vectors per thread in flight
        const size_t nv = n / VN;
        const V *src = reinterpret_cast<const V *>(MODE == 1 ? y : x);
        V *yv = reinterpret_cast<V *>(y);
        size_t i = tid;
        for (; i + (EUNROLL - 1) * nthreads < nv; i += EUNROLL * nthreads) {
            V v[EUNROLL];
#pragma unroll
            for (int u = 0; u < EUNROLL; ++u) v[u] = src[i + u * nthreads];  // cached: x was just read
#pragma unroll
            for (int u = 0; u < EUNROLL; ++u) {
                V w;
#pragma unroll
                for (int e = 0; e < VN; ++e) w[e] = softmax_elem<T, MODE>(v[u][e], v[u][e], m, s, lse, inv_s);
                __builtin_nontemporal_store(w, yv + i + u * nthreads);
            }
        }
        for (; i < nv; i += nthreads) {
            const V v = src[i];
            V w;
#pragma unroll
            for (int e = 0; e < VN; ++e) w[e] = softmax_elem<T, MODE>(v[e], v[e], m, s, lse, inv_s);
            __builtin_nontemporal_store(w, yv + i);
        }
        done = nv * VN;
    }
    for (size_t i = done + tid; i < n; i += nthreads) y[i] = softmax_elem<T, MODE>(x[i], MODE == 1 ? y[i] : (T)0, m, s, lse, inv_s);
}

template <typename T>
nrk_status softmax_pass(int mode, const T *x, T *y, const T *ms, size_t n, hipStream_t stream) {
    if (n == 0) return NRK_OK;
    const int vec_ok = ((reinterpret_cast<uintptr_t>(x) | reinterpret_cast<uintptr_t>(y)) & 15) == 0;
    const dim3 grid(stream_grid(n / Vec16<T>::N + 1));
    if (mode == 0) hipLaunchKernelGGL((softmax_pass_kernel<T, 0>), grid, dim3(EB), 0, stream, x, y, ms, n, vec_ok);
    else if (mode == 1) hipLaunchKernelGGL((softmax_pass_kernel<T, 1>), grid, dim3(EB), 0, stream, x, y, ms, n, vec_ok);
    else if (mode == 2) hipLaunchKernelGGL((softmax_pass_kernel<T, 2>), grid, dim3(EB), 0, stream, x, y, ms, n, vec_ok);
    else hipLaunchKernelGGL((softmax_pass_kernel<T, 3>), grid, dim3(EB), 0, stream, x, y, ms, n, vec_ok);
    NRK_HIP(hipGetLastError());
    return NRK_OK;
}
template nrk_status softmax_pass<float>(int, const float *, float *, const float *, size_t, hipStream_t);
template nrk_status softmax_pass<double>(int, const double *, double *, const double *, size_t, hipStream_t);

template nrk_status map<float>(const Chain &, const float *, float *, size_t, hipStream_t);
template nrk_status map<double>(const Chain &, const double *, double *, size_t, hipStream_t);
template nrk_status zip<float>(int, const float *, const float *, float *, size_t, hipStream_t);
template nrk_status zip<double>(int, const double *, const double *, double *, size_t, hipStream_t);

}  // namespace nrk
