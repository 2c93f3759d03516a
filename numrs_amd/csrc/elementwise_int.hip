// Generic-T elementwise work beyond f32/f64: the integer zip and map (kernels/elementwise.rs:
// 217-258 binary_dispatch / unary_dispatch for T = i32, i64, u32, u64, with the integer operators
// of array/arithmetic.rs and the bitwise ops of bitwise_ops.rs), and one N-D strided zip for every
// dtype that serves broadcasting (operations.rs:370-440 add/subtract/multiply/divide_broadcast),
// non-contiguous operands (the views borrow::operand would copy) and binary_serial's heterogeneous
// operand types (bitwise_ops.rs:190-243: a shift amount of another integer type).
//
// Integer arithmetic is Rust's release arithmetic: +, -, * and negation wrap; a shift amount is
// masked to the width (b & (bits - 1)); / and % truncate toward zero. Rust panics on x / 0, x % 0
// and MIN / -1: the host entry points check the divisor first and return NRK_ERR_INVALID_ARGUMENT
// without writing; the device forms define x / 0 = x % 0 = 0, MIN / -1 = MIN and MIN % -1 = 0.
#include "ops.hpp"

#include <limits>
#include <type_traits>

namespace nrk {
namespace {

constexpr int IB = 256;

template <typename T>
struct Vec16I;
template <>
struct Vec16I<int64_t> {
    typedef long long type __attribute__((ext_vector_type(2)));
    static constexpr int N = 2;
};
template <>
struct Vec16I<uint64_t> {
    typedef unsigned long long type __attribute__((ext_vector_type(2)));
    static constexpr int N = 2;
};
template <>
struct Vec16I<int32_t> {
    typedef int type __attribute__((ext_vector_type(4)));
    static constexpr int N = 4;
};
template <>
struct Vec16I<uint32_t> {
    typedef unsigned int type __attribute__((ext_vector_type(4)));
    static constexpr int N = 4;
};

template <typename T>
using U_of = typename std::make_unsigned<T>::type;

template <typename T>
__device__ __forceinline__ T wadd(T a, T b) { return (T)((U_of<T>)a + (U_of<T>)b); }
template <typename T>
__device__ __forceinline__ T wsub(T a, T b) { return (T)((U_of<T>)a - (U_of<T>)b); }
template <typename T>
__device__ __forceinline__ T wmul(T a, T b) { return (T)((U_of<T>)a * (U_of<T>)b); }
template <typename T>
__device__ __forceinline__ T idiv(T a, T b) {
    if (b == 0) return 0;
    if (std::is_signed<T>::value && b == (T)-1) return wsub<T>((T)0, a);  // MIN / -1 wraps to MIN
    return a / b;
}
template <typename T>
__device__ __forceinline__ T irem(T a, T b) {
    if (b == 0) return 0;
    if (std::is_signed<T>::value && b == (T)-1) return 0;
    return a % b;
}
template <typename T>
__device__ __forceinline__ T ishl(T a, uint64_t s) { return (T)((U_of<T>)a << (s & (sizeof(T) * 8 - 1))); }
template <typename T>
__device__ __forceinline__ T ishr(T a, uint64_t s) { return (T)(a >> (s & (sizeof(T) * 8 - 1))); }  // arithmetic for signed

template <typename T>
__device__ __forceinline__ T int_bop(int op, T a, T b) {
    switch (op) {
        case NRK_BOP_ADD: return wadd(a, b);
        case NRK_BOP_SUB: return wsub(a, b);
        case NRK_BOP_MUL: return wmul(a, b);
        case NRK_BOP_DIV: return idiv(a, b);
        case NRK_BOP_REM: return irem(a, b);
        case NRK_BOP_MAXIMUM: return a > b ? a : b;
        case NRK_BOP_MINIMUM: return a < b ? a : b;
        case NRK_BOP_BITAND: return a & b;
        case NRK_BOP_BITOR: return a | b;
        case NRK_BOP_BITXOR: return a ^ b;
        case NRK_BOP_SHL: return ishl(a, (uint64_t)b);
        default: return ishr(a, (uint64_t)b);  // NRK_BOP_SHR
    }
}

// One chain step on a batch of registers (the op is wavefront-uniform: one branch per step).
template <typename T, int N>
__device__ __forceinline__ void int_step_n(int op, T s, T (&v)[N]) {
    switch (op) {
#define NRK_ISTEP(OP, EXPR)                             \
    case OP:                                            \
        _Pragma("unroll") for (int i = 0; i < N; ++i) { \
            const T x = v[i];                           \
            v[i] = (EXPR);                              \
        }                                               \
        break;
        NRK_ISTEP(NRK_SOP_ADD, wadd(x, s))
        NRK_ISTEP(NRK_SOP_SUB, wsub(x, s))
        NRK_ISTEP(NRK_SOP_MUL, wmul(x, s))
        NRK_ISTEP(NRK_SOP_DIV, idiv(x, s))
        NRK_ISTEP(NRK_SOP_RSUB, wsub(s, x))
        NRK_ISTEP(NRK_SOP_RDIV, idiv(s, x))
        NRK_ISTEP(NRK_SOP_NEG, wsub((T)0, x))
        NRK_ISTEP(NRK_SOP_ABS, x < (T)0 ? wsub((T)0, x) : x)
        NRK_ISTEP(NRK_SOP_SQUARE, wmul(x, x))
        NRK_ISTEP(NRK_SOP_SIGN, (T)(x > (T)0 ? 1 : (x == (T)0 ? 0 : -1)))
        NRK_ISTEP(NRK_SOP_CLAMP_LO, x < s ? s : x)
        NRK_ISTEP(NRK_SOP_CLAMP_HI, x > s ? s : x)
#undef NRK_ISTEP
        default: break;
    }
}

struct IntChain {
    int len;
    int op[NRK_MAX_CHAIN];
    int64_t s[NRK_MAX_CHAIN];
};

template <typename T>
__global__ __launch_bounds__(IB) void int_map_kernel(IntChain ch, const T *x, T *y, size_t n, int vec_ok) {
    typedef typename Vec16I<T>::type V;
    constexpr int VN = Vec16I<T>::N;
    const size_t tid = blockIdx.x * (size_t)IB + threadIdx.x, nth = (size_t)gridDim.x * IB;
    size_t done = 0;
    if (vec_ok) {
        const size_t nv = n / VN;
        for (size_t i = tid; i < nv; i += nth) {
            V v = __builtin_nontemporal_load(reinterpret_cast<const V *>(x) + i);
            T r[VN];
#pragma unroll
            for (int e = 0; e < VN; ++e) r[e] = (T)v[e];
            for (int k = 0; k < ch.len; ++k) int_step_n<T, VN>(ch.op[k], (T)ch.s[k], r);
#pragma unroll
            for (int e = 0; e < VN; ++e) v[e] = r[e];
            __builtin_nontemporal_store(v, reinterpret_cast<V *>(y) + i);
        }
        done = nv * VN;
    }
    for (size_t i = done + tid; i < n; i += nth) {
        T r[1] = {x[i]};
        for (int k = 0; k < ch.len; ++k) int_step_n<T, 1>(ch.op[k], (T)ch.s[k], r);
        y[i] = r[0];
    }
}

template <typename T>
__global__ __launch_bounds__(IB) void int_zip_kernel(int op, const T *a, const T *b, T *y, size_t n, int vec_ok) {
    typedef typename Vec16I<T>::type V;
    constexpr int VN = Vec16I<T>::N;
    const size_t tid = blockIdx.x * (size_t)IB + threadIdx.x, nth = (size_t)gridDim.x * IB;
    size_t done = 0;
    if (vec_ok) {
        const size_t nv = n / VN;
        for (size_t i = tid; i < nv; i += nth) {
            V va = __builtin_nontemporal_load(reinterpret_cast<const V *>(a) + i);
            const V vb = __builtin_nontemporal_load(reinterpret_cast<const V *>(b) + i);
#pragma unroll
            for (int e = 0; e < VN; ++e) va[e] = int_bop<T>(op, (T)va[e], (T)vb[e]);
            __builtin_nontemporal_store(va, reinterpret_cast<V *>(y) + i);
        }
        done = nv * VN;
    }
    for (size_t i = done + tid; i < n; i += nth) y[i] = int_bop<T>(op, a[i], b[i]);
}

// ---- N-D strided zip -------------------------------------------------------------------
constexpr int ND_MAX = NRK_MAX_NDIM;
struct NdShape {
    int ndim;                      // >= 1 after collapsing; the last axis is the inner loop
    int64_t shape[ND_MAX];
    int64_t sa[ND_MAX], sb[ND_MAX];  // element strides of a and b (0: broadcast)
    size_t rows;                   // product of shape[0 .. ndim-2]
};

// b of another integer dtype (a shift amount), read as a 64-bit value
__device__ __forceinline__ uint64_t load_shift(const void *p, int dt, int64_t idx) {
    switch (dt) {
        case NRK_DTYPE_I64: return (uint64_t)((const int64_t *)p)[idx];
        case NRK_DTYPE_I32: return (uint64_t)(int64_t)((const int32_t *)p)[idx];
        case NRK_DTYPE_U64: return ((const uint64_t *)p)[idx];
        default: return (uint64_t)((const uint32_t *)p)[idx];
    }
}

template <typename T>
__device__ __forceinline__ T float_bop(int op, T a, T b) {
    switch (op) {
        case NRK_BOP_ADD: return rn_add(a, b);
        case NRK_BOP_SUB: return rn_sub(a, b);
        case NRK_BOP_MUL: return rn_mul(a, b);
        case NRK_BOP_DIV: return rn_div(a, b);
        case NRK_BOP_REM: return __builtin_fmod(a, b);
        case NRK_BOP_MAXIMUM: return a > b ? a : b;
        case NRK_BOP_MINIMUM: return a < b ? a : b;
        case NRK_BOP_COPYSIGN: return dev_copysign(a, b);
        default: return a > (T)0 ? a : rn_mul(b, a);  // NRK_BOP_PRELU
    }
}

// Tile = (row r, a run of IB * 4 consecutive positions of the inner axis). The row's offsets
// are unravelled once per tile; lanes then walk the inner axis with the operands' inner strides.
template <typename T>
__global__ __launch_bounds__(IB) void nd_zip_kernel(int op, NdShape s, const T *a, const void *b, int b_dt, T *y) {
    const int64_t inner = s.shape[s.ndim - 1];
    const size_t per_row = (size_t)((inner + IB * 4 - 1) / (IB * 4));
    const size_t tiles = s.rows * per_row;
    const int64_t ia = s.sa[s.ndim - 1], ib = s.sb[s.ndim - 1];
    for (size_t t = blockIdx.x; t < tiles; t += gridDim.x) {
        const size_t r = t / per_row;
        const int64_t c0 = (int64_t)(t - r * per_row) * IB * 4;
        int64_t oa = 0, ob = 0;
        size_t rem = r;
        for (int d = s.ndim - 2; d >= 0; --d) {
            const int64_t q = (int64_t)(rem % (size_t)s.shape[d]);
            rem /= (size_t)s.shape[d];
            oa += q * s.sa[d];
            ob += q * s.sb[d];
        }
        T *yr = y + r * (size_t)inner;
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            const int64_t c = c0 + u * IB + threadIdx.x;
            if (c >= inner) break;
            const T av = a[oa + c * ia];
            T out;
            if constexpr (std::is_floating_point<T>::value) {
                out = float_bop<T>(op, av, ((const T *)b)[ob + c * ib]);
            } else {
                if (op == NRK_BOP_SHL || op == NRK_BOP_SHR) {
                    const uint64_t sh = load_shift(b, b_dt, ob + c * ib);
                    out = op == NRK_BOP_SHL ? ishl<T>(av, sh) : ishr<T>(av, sh);
                } else {
                    out = int_bop<T>(op, av, ((const T *)b)[ob + c * ib]);
                }
            }
            yr[c] = out;
        }
    }
}

int grid_for(size_t items, int per_cu) {
    const size_t cap = (size_t)cu_count() * per_cu;
    return (int)std::max<size_t>(1, std::min(items, cap));
}

}  // namespace

static bool int_dtype(int dt) { return dt == NRK_DTYPE_I64 || dt == NRK_DTYPE_I32 || dt == NRK_DTYPE_U64 || dt == NRK_DTYPE_U32; }
static size_t dtype_size(int dt) { return (dt == NRK_DTYPE_F64 || dt == NRK_DTYPE_I64 || dt == NRK_DTYPE_U64) ? 8 : 4; }

template <typename T>
nrk_status int_zip(int op, const T *a, const T *b, T *y, size_t n, hipStream_t stream) {
    NRK_TRY(ensure_device());
    const bool ok_op = (op >= NRK_BOP_ADD && op <= NRK_BOP_MINIMUM) || (op >= NRK_BOP_BITAND && op <= NRK_BOP_SHR);
    if (!ok_op) return fail(NRK_ERR_INVALID_ARGUMENT, "op %d is not an integer NRK_BOP_* code", op);
    if (n == 0) return NRK_OK;
    if (!a || !b || !y) return fail(NRK_ERR_INVALID_ARGUMENT, "a, b or y is NULL");
    const int vec_ok = ((uintptr_t)a % 16 == 0) && ((uintptr_t)b % 16 == 0) && ((uintptr_t)y % 16 == 0);
    hipLaunchKernelGGL(int_zip_kernel<T>, dim3(grid_for(ceil_div(n, (size_t)IB * 4), 32)), dim3(IB), 0, stream, op, a, b, y, n,
                       vec_ok);
    NRK_HIP(hipGetLastError());
    return NRK_OK;
}

template <typename T>
nrk_status int_map(const nrk_int_op *chain, int32_t chain_len, const T *x, T *y, size_t n, hipStream_t stream) {
    NRK_TRY(ensure_device());
    if (chain_len < 0 || chain_len > NRK_MAX_CHAIN) return fail(NRK_ERR_INVALID_ARGUMENT, "chain_len %d outside [0,%d]", chain_len, NRK_MAX_CHAIN);
    if (chain_len > 0 && !chain) return fail(NRK_ERR_INVALID_ARGUMENT, "chain is NULL");
    IntChain ch{};
    ch.len = chain_len;
    for (int k = 0; k < chain_len; ++k) {
        const int o = chain[k].op;
        const bool ok = (o >= NRK_SOP_ADD && o <= NRK_SOP_SQUARE) || o == NRK_SOP_SIGN || o == NRK_SOP_CLAMP_LO || o == NRK_SOP_CLAMP_HI;
        if (!ok) return fail(NRK_ERR_INVALID_ARGUMENT, "chain[%d].op = %d has no integer form", k, o);
        if (!std::is_signed<T>::value && (o == NRK_SOP_NEG || o == NRK_SOP_ABS || o == NRK_SOP_SIGN))
            return fail(NRK_ERR_INVALID_ARGUMENT, "chain[%d].op = %d is not defined for unsigned integers", k, o);
        ch.op[k] = o;
        ch.s[k] = chain[k].scalar;
    }
    if (n == 0) return NRK_OK;
    if (!x || !y) return fail(NRK_ERR_INVALID_ARGUMENT, "x or y is NULL");
    const int vec_ok = ((uintptr_t)x % 16 == 0) && ((uintptr_t)y % 16 == 0);
    hipLaunchKernelGGL(int_map_kernel<T>, dim3(grid_for(ceil_div(n, (size_t)IB * 4), 32)), dim3(IB), 0, stream, ch, x, y, n, vec_ok);
    NRK_HIP(hipGetLastError());
    return NRK_OK;
}

// Drops axes of extent 1 and merges neighbours that are contiguous for both operands.
static nrk_status collapse(int ndim, const size_t *shape, const int64_t *sa, const int64_t *sb, NdShape *out) {
    if (ndim < 0 || ndim > ND_MAX) return fail(NRK_ERR_INVALID_ARGUMENT, "ndim %d outside [0,%d]", ndim, ND_MAX);
    NdShape s{};
    s.ndim = 0;
    for (int d = 0; d < ndim; ++d) {
        if (shape[d] == 1) continue;
        if (s.ndim > 0) {
            const int p = s.ndim - 1;
            if (s.sa[p] == sa[d] * (int64_t)shape[d] && s.sb[p] == sb[d] * (int64_t)shape[d]) {
                s.shape[p] *= (int64_t)shape[d];
                s.sa[p] = sa[d];
                s.sb[p] = sb[d];
                continue;
            }
        }
        s.shape[s.ndim] = (int64_t)shape[d];
        s.sa[s.ndim] = sa[d];
        s.sb[s.ndim] = sb[d];
        ++s.ndim;
    }
    if (s.ndim == 0) {
        s.ndim = 1;
        s.shape[0] = 1;
        s.sa[0] = s.sb[0] = 0;
    }
    s.rows = 1;
    for (int d = 0; d < s.ndim - 1; ++d) s.rows *= (size_t)s.shape[d];
    *out = s;
    return NRK_OK;
}

nrk_status zip_strided(int dtype, int op, int ndim, const size_t *shape, const void *a, const int64_t *sa, const void *b,
                       int b_dtype, const int64_t *sb, void *y, hipStream_t stream) {
    NRK_TRY(ensure_device());
    if (dtype < NRK_DTYPE_F64 || dtype > NRK_DTYPE_U32) return fail(NRK_ERR_INVALID_ARGUMENT, "dtype %d is not an NRK_DTYPE_* code", dtype);
    const bool fl = !int_dtype(dtype);
    if (fl) {
        if (!(op >= NRK_BOP_ADD && op <= NRK_BOP_COPYSIGN) && op != NRK_BOP_PRELU)
            return fail(NRK_ERR_INVALID_ARGUMENT, "op %d has no strided floating-point form", op);
        if (b_dtype != dtype) return fail(NRK_ERR_INVALID_ARGUMENT, "b's dtype must equal a's for a floating-point op");
    } else {
        const bool ok_op = (op >= NRK_BOP_ADD && op <= NRK_BOP_MINIMUM) || (op >= NRK_BOP_BITAND && op <= NRK_BOP_SHR);
        if (!ok_op) return fail(NRK_ERR_INVALID_ARGUMENT, "op %d is not an integer NRK_BOP_* code", op);
        const bool shift = op == NRK_BOP_SHL || op == NRK_BOP_SHR;
        if (shift ? !int_dtype(b_dtype) : b_dtype != dtype)
            return fail(NRK_ERR_INVALID_ARGUMENT, "b's dtype %d does not fit op %d on dtype %d", b_dtype, op, dtype);
    }
    if (ndim > 0 && (!shape || !sa || !sb)) return fail(NRK_ERR_INVALID_ARGUMENT, "shape or strides is NULL");
    size_t total = 1;
    for (int d = 0; d < ndim; ++d) total *= shape[d];
    if (total == 0) return NRK_OK;
    if (!a || !b || !y) return fail(NRK_ERR_INVALID_ARGUMENT, "a, b or y is NULL");
    NdShape s;
    NRK_TRY(collapse(ndim, shape, sa, sb, &s));
    const size_t tiles = s.rows * ceil_div((size_t)s.shape[s.ndim - 1], (size_t)IB * 4);
    const dim3 grid(grid_for(tiles, 32));
    switch (dtype) {
        case NRK_DTYPE_F64: hipLaunchKernelGGL(nd_zip_kernel<double>, grid, dim3(IB), 0, stream, op, s, (const double *)a, b, b_dtype, (double *)y); break;
        case NRK_DTYPE_F32: hipLaunchKernelGGL(nd_zip_kernel<float>, grid, dim3(IB), 0, stream, op, s, (const float *)a, b, b_dtype, (float *)y); break;
        case NRK_DTYPE_I64: hipLaunchKernelGGL(nd_zip_kernel<int64_t>, grid, dim3(IB), 0, stream, op, s, (const int64_t *)a, b, b_dtype, (int64_t *)y); break;
        case NRK_DTYPE_I32: hipLaunchKernelGGL(nd_zip_kernel<int32_t>, grid, dim3(IB), 0, stream, op, s, (const int32_t *)a, b, b_dtype, (int32_t *)y); break;
        case NRK_DTYPE_U64: hipLaunchKernelGGL(nd_zip_kernel<uint64_t>, grid, dim3(IB), 0, stream, op, s, (const uint64_t *)a, b, b_dtype, (uint64_t *)y); break;
        default: hipLaunchKernelGGL(nd_zip_kernel<uint32_t>, grid, dim3(IB), 0, stream, op, s, (const uint32_t *)a, b, b_dtype, (uint32_t *)y); break;
    }
    NRK_HIP(hipGetLastError());
    return NRK_OK;
}

// ---- host forms ----
template <typename T>
static nrk_status check_divisors(int op, const T *a, const T *b, size_t n, int64_t sa_step, int64_t sb_step) {
    if (op != NRK_BOP_DIV && op != NRK_BOP_REM) return NRK_OK;
    for (size_t i = 0; i < n; ++i) {
        const T d = b[(int64_t)i * sb_step];
        if (d == 0) return fail(NRK_ERR_INVALID_ARGUMENT, "attempt to %s by zero", op == NRK_BOP_DIV ? "divide" : "calculate the remainder with a divisor of");
        if (std::is_signed<T>::value && d == (T)-1 && a[(int64_t)i * sa_step] == std::numeric_limits<T>::min())
            return fail(NRK_ERR_INVALID_ARGUMENT, "attempt to %s with overflow", op == NRK_BOP_DIV ? "divide" : "calculate the remainder");
    }
    return NRK_OK;
}

template <typename T>
static nrk_status host_int_zip(int op, const T *a, const T *b, T *y, size_t n) {
    NRK_TRY(ensure_device());
    if (n == 0) return NRK_OK;
    if (!a || !b || !y) return fail(NRK_ERR_INVALID_ARGUMENT, "a, b or y is NULL");
    NRK_TRY(check_divisors<T>(op, a, b, n, 1, 1));
    void *da, *db, *dy;
    NRK_TRY(staging(0, n * sizeof(T) + 16, &da));
    NRK_TRY(staging(2, n * sizeof(T) + 16, &db));
    NRK_TRY(staging(1, n * sizeof(T) + 16, &dy));
    hipStream_t s = nullptr;
    NRK_TRY(copy_h2d(da, a, n * sizeof(T), s));
    NRK_TRY(copy_h2d(db, b, n * sizeof(T), s));
    NRK_TRY(int_zip<T>(op, (const T *)da, (const T *)db, (T *)dy, n, s));
    NRK_TRY(copy_d2h(y, dy, n * sizeof(T), s));
    NRK_HIP(hipStreamSynchronize(s));
    return NRK_OK;
}

template <typename T>
static nrk_status host_int_map(const nrk_int_op *chain, int32_t chain_len, const T *x, T *y, size_t n) {
    NRK_TRY(ensure_device());
    if (n == 0) return int_map<T>(chain, chain_len, x, y, 0, nullptr);  // validates the chain
    if (!x || !y) return fail(NRK_ERR_INVALID_ARGUMENT, "x or y is NULL");
    for (int k = 0; chain && k < chain_len; ++k) {
        const int o = chain[k].op;
        if ((o == NRK_SOP_DIV) && chain[k].scalar == 0) return fail(NRK_ERR_INVALID_ARGUMENT, "attempt to divide by zero");
        if (o == NRK_SOP_DIV && std::is_signed<T>::value && chain[k].scalar == -1)
            return fail(NRK_ERR_UNSUPPORTED, "integer division by -1 may overflow; use NEG");
        if (o == NRK_SOP_RDIV) return fail(NRK_ERR_UNSUPPORTED, "scalar / x on integers: the divisors are data; use the zip");
    }
    void *dx, *dy;
    NRK_TRY(staging(0, n * sizeof(T) + 16, &dx));
    NRK_TRY(staging(1, n * sizeof(T) + 16, &dy));
    hipStream_t s = nullptr;
    NRK_TRY(copy_h2d(dx, x, n * sizeof(T), s));
    NRK_TRY(int_map<T>(chain, chain_len, (const T *)dx, (T *)dy, n, s));
    NRK_TRY(copy_d2h(y, dy, n * sizeof(T), s));
    NRK_HIP(hipStreamSynchronize(s));
    return NRK_OK;
}

// numpy broadcasting of two shapes (right-aligned; extents equal or 1), as Array::broadcast_shape
static nrk_status broadcast_shapes(int an, const size_t *ash, int bn, const size_t *bsh, int *on, size_t *osh, int64_t *sa,
                                   int64_t *sb) {
    const int n = an > bn ? an : bn;
    if (n > ND_MAX) return fail(NRK_ERR_INVALID_ARGUMENT, "more than %d axes", ND_MAX);
    int64_t ca = 1, cb = 1;
    for (int d = n - 1; d >= 0; --d) {
        const int ia = d - (n - an), ibx = d - (n - bn);
        const size_t ea = ia >= 0 ? ash[ia] : 1, eb = ibx >= 0 ? bsh[ibx] : 1;
        if (ea != eb && ea != 1 && eb != 1)
            return fail(NRK_ERR_INVALID_ARGUMENT, "shapes do not broadcast: extent %zu against %zu on axis %d", ea, eb, d);
        osh[d] = ea == 1 ? eb : ea;
        sa[d] = ea == 1 ? 0 : ca;
        sb[d] = eb == 1 ? 0 : cb;
        ca *= (int64_t)ea;
        cb *= (int64_t)eb;
    }
    *on = n;
    return NRK_OK;
}

static nrk_status host_zip_broadcast(int dtype, int op, const void *a, int a_ndim, const size_t *a_shape, const void *b,
                                     int b_ndim, const size_t *b_shape, void *y) {
    NRK_TRY(ensure_device());
    if (a_ndim < 0 || b_ndim < 0) return fail(NRK_ERR_INVALID_ARGUMENT, "negative ndim");
    if ((a_ndim && !a_shape) || (b_ndim && !b_shape)) return fail(NRK_ERR_INVALID_ARGUMENT, "shape is NULL");
    if (dtype < NRK_DTYPE_F64 || dtype > NRK_DTYPE_U32) return fail(NRK_ERR_INVALID_ARGUMENT, "dtype %d is not an NRK_DTYPE_* code", dtype);
    int on = 0;
    size_t osh[ND_MAX];
    int64_t sa[ND_MAX], sb[ND_MAX];
    NRK_TRY(broadcast_shapes(a_ndim, a_shape, b_ndim, b_shape, &on, osh, sa, sb));
    size_t na = 1, nb = 1, no = 1;
    for (int d = 0; d < a_ndim; ++d) na *= a_shape[d];
    for (int d = 0; d < b_ndim; ++d) nb *= b_shape[d];
    for (int d = 0; d < on; ++d) no *= osh[d];
    if (no == 0) return NRK_OK;
    if (!a || !b || !y) return fail(NRK_ERR_INVALID_ARGUMENT, "a, b or y is NULL");
    const size_t es = dtype_size(dtype);
    if (int_dtype(dtype) && (op == NRK_BOP_DIV || op == NRK_BOP_REM)) {  // Rust panics: refuse before writing
        for (size_t i = 0; i < nb; ++i) {
            bool zero = es == 8 ? ((const int64_t *)b)[i] == 0 : ((const int32_t *)b)[i] == 0;
            if (zero) return fail(NRK_ERR_INVALID_ARGUMENT, "attempt to divide by zero");
        }
        if (dtype == NRK_DTYPE_I64 || dtype == NRK_DTYPE_I32) {  // MIN / -1 anywhere in the broadcast
            bool has_m1 = false, has_min = false;
            for (size_t i = 0; i < nb; ++i) has_m1 |= es == 8 ? ((const int64_t *)b)[i] == -1 : ((const int32_t *)b)[i] == -1;
            for (size_t i = 0; i < na; ++i)
                has_min |= es == 8 ? ((const int64_t *)a)[i] == INT64_MIN : ((const int32_t *)a)[i] == INT32_MIN;
            if (has_m1 && has_min) return fail(NRK_ERR_UNSUPPORTED, "an integer division of MIN by -1 may occur (Rust panics on it)");
        }
    }
    void *da, *db, *dy;
    NRK_TRY(staging(0, na * es + 16, &da));
    NRK_TRY(staging(2, nb * es + 16, &db));
    NRK_TRY(staging(1, no * es + 16, &dy));
    hipStream_t s = nullptr;
    NRK_TRY(copy_h2d(da, a, na * es, s));
    NRK_TRY(copy_h2d(db, b, nb * es, s));
    NRK_TRY(zip_strided(dtype, op, on, osh, da, sa, db, dtype, sb, dy, s));
    NRK_TRY(copy_d2h(y, dy, no * es, s));
    NRK_HIP(hipStreamSynchronize(s));
    return NRK_OK;
}

}  // namespace nrk

using namespace nrk;

extern "C" {

#define NRK_INT_ENTRY(SUF, T)                                                                                        \
    nrk_status nrk_zip_##SUF(int32_t op, const T *a, const T *b, T *y, size_t n) { return host_int_zip<T>(op, a, b, y, n); } \
    nrk_status nrk_dev_zip_##SUF(int32_t op, const T *a, const T *b, T *y, size_t n, nrk_stream stream) {                \
        return int_zip<T>(op, a, b, y, n, (hipStream_t)stream);                                                      \
    }                                                                                                                \
    nrk_status nrk_map_##SUF(const nrk_int_op *chain, int32_t chain_len, const T *x, T *y, size_t n) {              \
        return host_int_map<T>(chain, chain_len, x, y, n);                                                           \
    }                                                                                                                \
    nrk_status nrk_dev_map_##SUF(const nrk_int_op *chain, int32_t chain_len, const T *x, T *y, size_t n, nrk_stream stream) { \
        return int_map<T>(chain, chain_len, x, y, n, (hipStream_t)stream);                                           \
    }
NRK_INT_ENTRY(i64, int64_t)
NRK_INT_ENTRY(i32, int32_t)
NRK_INT_ENTRY(u64, uint64_t)
NRK_INT_ENTRY(u32, uint32_t)
#undef NRK_INT_ENTRY

nrk_status nrk_zip_broadcast(int32_t dtype, int32_t op, const void *a, int32_t a_ndim, const size_t *a_shape, const void *b,
                             int32_t b_ndim, const size_t *b_shape, void *y) {
    return host_zip_broadcast(dtype, op, a, a_ndim, a_shape, b, b_ndim, b_shape, y);
}

nrk_status nrk_dev_zip_strided(int32_t dtype, int32_t op, int32_t ndim, const size_t *shape, const void *a,
                               const int64_t *a_strides, const void *b, int32_t b_dtype, const int64_t *b_strides, void *y,
                               nrk_stream stream) {
    return zip_strided(dtype, op, ndim, shape, a, a_strides, b, b_dtype, b_strides, y, (hipStream_t)stream);
}

}  // extern "C"
