// Internal helpers shared by every translation unit of libnumrs_amd.so.
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>

#include "numrs_amd.h"
#include "mathfn.hpp"

namespace nrk {

// ---- error reporting (thread-local last error text) ----
void set_error(const char *fmt, ...) __attribute__((format(printf, 1, 2)));
nrk_status fail(nrk_status code, const char *fmt, ...) __attribute__((format(printf, 2, 3)));
nrk_status hip_fail(hipError_t e, const char *what, const char *file, int line);

#define NRK_HIP(call)                                                                   \
    do {                                                                                \
        hipError_t nrk_e_ = (call);                                                     \
        if (nrk_e_ != hipSuccess) return ::nrk::hip_fail(nrk_e_, #call, __FILE__, __LINE__); \
    } while (0)

#define NRK_TRY(expr)                         \
    do {                                      \
        nrk_status nrk_s_ = (expr);           \
        if (nrk_s_ != NRK_OK) return nrk_s_;  \
    } while (0)

// Make sure the calling thread has a current gfx950 device (nrk_init(0) on first use).
nrk_status ensure_device();
// Number of compute units of the current device (cached).
int cu_count();

// ---- per-stream scratch ----
// Every stream gets its own scratch: a zero-invariant counter area (kernels that use a counter
// reset it to zero before they exit, so it is zero at every launch) and a partials area whose
// contents are undefined at launch.
struct Scratch {
    unsigned int *counters;  // zero at every launch boundary
    size_t n_counters;
    void *partials;
    size_t partial_bytes;
};
nrk_status get_scratch(hipStream_t stream, size_t n_counters, size_t partial_bytes, Scratch *out);

// Caller-provided device scratch (device address and size): the reducer then allocates nothing.
// Its first 256 bytes hold the finish counter, cleared by a kernel at every call.
struct ExtScratch {
    void *base;
    size_t bytes;
};

// Per-stream workspace for multi-pass algorithms (split operands, residues). Contents undefined
// at entry; grows on demand (growing synchronises the stream first).
nrk_status get_workspace(hipStream_t stream, size_t bytes, void **out);

// Host-pointer staging: device copies of host operands, cached per thread and grown on demand.
struct Staging {
    void *ptr = nullptr;
    size_t bytes = 0;
};
nrk_status staging(int slot, size_t bytes, void **ptr);

// Copies between a host pointer (pageable) and device memory for the host-pointer entry points,
// through pinned buffers and a small host-copy pool. copy_h2d returns once the source has been
// read (the DMAs are queued on s); copy_d2h returns with the data in dst.
nrk_status copy_h2d(void *dst, const void *src, size_t bytes, hipStream_t s);
nrk_status copy_d2h(void *dst, const void *src, size_t bytes, hipStream_t s);

inline size_t ceil_div(size_t a, size_t b) { return (a + b - 1) / b; }
inline size_t align_up(size_t a, size_t b) { return ceil_div(a, b) * b; }

// ---- device-side helpers ----
template <typename T>
struct Vec16;  // 16-byte vector of T
template <>
struct Vec16<float> {
    typedef float type __attribute__((ext_vector_type(4)));
    static constexpr int N = 4;
};
template <>
struct Vec16<double> {
    typedef double type __attribute__((ext_vector_type(2)));
    static constexpr int N = 2;
};

template <typename T>
__device__ __forceinline__ T dev_nan();
template <>
__device__ __forceinline__ float dev_nan<float>() { return __builtin_nanf(""); }
template <>
__device__ __forceinline__ double dev_nan<double>() { return __builtin_nan(""); }

template <typename T>
__device__ __forceinline__ T dev_inf();
template <>
__device__ __forceinline__ float dev_inf<float>() { return __builtin_inff(); }
template <>
__device__ __forceinline__ double dev_inf<double>() { return __builtin_inf(); }

// One rounding per operation, never contracted.
__device__ __forceinline__ float rn_add(float a, float b) { return __fadd_rn(a, b); }
__device__ __forceinline__ double rn_add(double a, double b) { return __dadd_rn(a, b); }
__device__ __forceinline__ float rn_sub(float a, float b) { return __fsub_rn(a, b); }
__device__ __forceinline__ double rn_sub(double a, double b) { return __dsub_rn(a, b); }
__device__ __forceinline__ float rn_mul(float a, float b) { return __fmul_rn(a, b); }
__device__ __forceinline__ double rn_mul(double a, double b) { return __dmul_rn(a, b); }
// binary32 division and square root are evaluated in binary64 and rounded once: since
// 53 >= 2*24 + 2 the double rounding is innocuous (Figueroa 1995), so the result is the
// correctly rounded binary32 value that the reference's CPU code produces.
__device__ __forceinline__ float rn_div(float a, float b) { return (float)__ddiv_rn((double)a, (double)b); }
__device__ __forceinline__ double rn_div(double a, double b) { return __ddiv_rn(a, b); }
__device__ __forceinline__ float rn_sqrt(float a) { return (float)__dsqrt_rn((double)a); }
__device__ __forceinline__ double rn_sqrt(double a) { return __dsqrt_rn(a); }

__device__ __forceinline__ float dev_fabs(float a) { return __builtin_fabsf(a); }
__device__ __forceinline__ double dev_fabs(double a) { return __builtin_fabs(a); }
// fmax/fmin return the non-NaN operand when exactly one operand is NaN.
__device__ __forceinline__ float dev_fmax(float a, float b) { return __builtin_fmaxf(a, b); }
__device__ __forceinline__ double dev_fmax(double a, double b) { return __builtin_fmax(a, b); }
__device__ __forceinline__ float dev_fmin(float a, float b) { return __builtin_fminf(a, b); }
__device__ __forceinline__ double dev_fmin(double a, double b) { return __builtin_fmin(a, b); }

// Scalar chain passed by value to kernels.
struct Chain {
    int len;
    int op[NRK_MAX_CHAIN];
    double s[NRK_MAX_CHAIN];
};

__device__ __forceinline__ float dev_floor(float a) { return __builtin_floorf(a); }
__device__ __forceinline__ double dev_floor(double a) { return __builtin_floor(a); }
__device__ __forceinline__ float dev_ceil(float a) { return __builtin_ceilf(a); }
__device__ __forceinline__ double dev_ceil(double a) { return __builtin_ceil(a); }
__device__ __forceinline__ float dev_trunc(float a) { return __builtin_truncf(a); }
__device__ __forceinline__ double dev_trunc(double a) { return __builtin_trunc(a); }
__device__ __forceinline__ float dev_copysign(float a, float b) { return __builtin_copysignf(a, b); }
__device__ __forceinline__ double dev_copysign(double a, double b) { return __builtin_copysign(a, b); }

// Rust's f64::round: half-way cases away from zero (x - trunc(x) is exact).
template <typename T>
__device__ __forceinline__ T round_half_away(T x) {
    const T t = dev_trunc(x);
    return dev_fabs(rn_sub(x, t)) >= (T)0.5 ? rn_add(t, dev_copysign((T)1, x)) : t;
}
// ufuncs.rs:756: 0 for +-0, 1 above zero, -1 otherwise (NaN included).
template <typename T>
__device__ __forceinline__ T sign_ref(T x) {
    return x == (T)0 ? (T)0 : (x > (T)0 ? (T)1 : (T)-1);
}
// Elementary functions: FP64 evaluation, binary32 rounded once.
template <typename T, typename F>
__device__ __forceinline__ T via_f64(T x, F f) {
    return (T)f((double)x);
}

// Which steps a kernel instantiation compiles: MATH_NONE the IEEE steps (ADD .. CLAMP_HI),
// MATH_LIGHT also the elementary functions, activations and special functions, MATH_ALL also
// pow, gamma and lgamma (a table-driven double-double log inside; alone they take the map kernel
// from ~100 to over 200 VGPRs). The register budget of a kernel is that of its largest case.
enum { MATH_NONE = 0, MATH_LIGHT = 1, MATH_ALL = 2 };
__host__ __device__ inline int sop_math_class(int op) {
    if (op == NRK_SOP_POW || op == NRK_SOP_GAMMA || op == NRK_SOP_LGAMMA) return MATH_ALL;
    return op >= NRK_SOP_EXP ? MATH_LIGHT : MATH_NONE;
}

// Apply the whole chain to N register values, each step (NRK_SOP_*) with one IEEE rounding: one
// (wavefront-uniform) branch per step, then N straight-line operations, so the op dispatch is
// not repeated per element. MATH (above) selects the steps compiled in; kernels built with
// MATH_NONE (the reductions) keep their register budget, and their chains are validated to hold
// no other step.
template <typename T, int N, int MATH = MATH_NONE>
__device__ __forceinline__ void apply_step_n(int op, double sd, T (&v)[N]) {
    {
        const T s = (T)sd;
        switch (op) {
#define NRK_CHAIN_CASE(OP, EXPR)                      \
    case OP:                                          \
        _Pragma("unroll") for (int i = 0; i < N; ++i) { \
            const T x = v[i];                         \
            v[i] = (EXPR);                            \
        }                                             \
        break;
            NRK_CHAIN_CASE(NRK_SOP_ADD, rn_add(x, s))
            NRK_CHAIN_CASE(NRK_SOP_SUB, rn_sub(x, s))
            NRK_CHAIN_CASE(NRK_SOP_MUL, rn_mul(x, s))
            NRK_CHAIN_CASE(NRK_SOP_DIV, rn_div(x, s))
            NRK_CHAIN_CASE(NRK_SOP_RSUB, rn_sub(s, x))
            NRK_CHAIN_CASE(NRK_SOP_RDIV, rn_div(s, x))
            NRK_CHAIN_CASE(NRK_SOP_NEG, -x)
            NRK_CHAIN_CASE(NRK_SOP_ABS, dev_fabs(x))
            NRK_CHAIN_CASE(NRK_SOP_SQUARE, rn_mul(x, x))
            NRK_CHAIN_CASE(NRK_SOP_SQRT, rn_sqrt(x))
            NRK_CHAIN_CASE(NRK_SOP_FLOOR, dev_floor(x))
            NRK_CHAIN_CASE(NRK_SOP_CEIL, dev_ceil(x))
            NRK_CHAIN_CASE(NRK_SOP_TRUNC, dev_trunc(x))
            NRK_CHAIN_CASE(NRK_SOP_ROUND, round_half_away(x))
            NRK_CHAIN_CASE(NRK_SOP_FRACT, rn_sub(x, dev_trunc(x)))
            NRK_CHAIN_CASE(NRK_SOP_SIGN, sign_ref(x))
            NRK_CHAIN_CASE(NRK_SOP_CLAMP_LO, x < s ? s : x)
            NRK_CHAIN_CASE(NRK_SOP_CLAMP_HI, x > s ? s : x)
            default:
                if constexpr (MATH >= MATH_LIGHT) {
                    switch (op) {
                        NRK_CHAIN_CASE(NRK_SOP_EXP, fm::exp_t(x))
                        NRK_CHAIN_CASE(NRK_SOP_EXP2, via_f64(x, [](double d) { return fm::exp2(d); }))
                        NRK_CHAIN_CASE(NRK_SOP_EXPM1, via_f64(x, [](double d) { return fm::expm1(d); }))
                        NRK_CHAIN_CASE(NRK_SOP_LOG, fm::log_t(x))
                        NRK_CHAIN_CASE(NRK_SOP_LOG2, fm::log2_t(x))
                        NRK_CHAIN_CASE(NRK_SOP_LOG10, fm::log10_t(x))
                        NRK_CHAIN_CASE(NRK_SOP_LOG1P, fm::log1p_t(x))
                        NRK_CHAIN_CASE(NRK_SOP_TANH, fm::tanh_t(x))
                        NRK_CHAIN_CASE(NRK_SOP_SINH, via_f64(x, [](double d) { return fm::sinh(d); }))
                        NRK_CHAIN_CASE(NRK_SOP_COSH, via_f64(x, [](double d) { return fm::cosh(d); }))
                        NRK_CHAIN_CASE(NRK_SOP_ASINH, via_f64(x, [](double d) { return fm::asinh(d); }))
                        NRK_CHAIN_CASE(NRK_SOP_ACOSH, via_f64(x, [](double d) { return fm::acosh(d); }))
                        NRK_CHAIN_CASE(NRK_SOP_ATANH, via_f64(x, [](double d) { return fm::atanh(d); }))
                        NRK_CHAIN_CASE(NRK_SOP_ATAN, via_f64(x, [](double d) { return fm::atan(d); }))
                        NRK_CHAIN_CASE(NRK_SOP_ASIN, via_f64(x, [](double d) { return fm::asin(d); }))
                        NRK_CHAIN_CASE(NRK_SOP_ACOS, via_f64(x, [](double d) { return fm::acos(d); }))
                        NRK_CHAIN_CASE(NRK_SOP_SIN, fm::sin_t(x))
                        NRK_CHAIN_CASE(NRK_SOP_COS, fm::cos_t(x))
                        NRK_CHAIN_CASE(NRK_SOP_TAN, via_f64(x, [](double d) { return fm::tan(d); }))
                        NRK_CHAIN_CASE(NRK_SOP_RELU, via_f64(x, [](double d) { return fm::relu(d); }))
                        NRK_CHAIN_CASE(NRK_SOP_LEAKY_RELU, (T)((double)x > 0.0 ? (double)x : sd * (double)x))
                        NRK_CHAIN_CASE(NRK_SOP_ELU, (T)fm::elu((double)x, sd))
                        NRK_CHAIN_CASE(NRK_SOP_SELU, via_f64(x, [](double d) { return fm::selu(d); }))
                        NRK_CHAIN_CASE(NRK_SOP_HARDSIGMOID, via_f64(x, [](double d) { return fm::hardsigmoid(d); }))
                        NRK_CHAIN_CASE(NRK_SOP_HARDSWISH, via_f64(x, [](double d) { return fm::hardswish(d); }))
                        NRK_CHAIN_CASE(NRK_SOP_SOFTPLUS, via_f64(x, [](double d) { return fm::softplus(d); }))
                        NRK_CHAIN_CASE(NRK_SOP_MISH, via_f64(x, [](double d) { return fm::mish(d); }))
                        NRK_CHAIN_CASE(NRK_SOP_SILU, via_f64(x, [](double d) { return fm::silu(d); }))
                        NRK_CHAIN_CASE(NRK_SOP_GELU, via_f64(x, [](double d) { return fm::gelu(d); }))
                        NRK_CHAIN_CASE(NRK_SOP_LOGIT, via_f64(x, [](double d) { return fm::logit(d); }))
                        NRK_CHAIN_CASE(NRK_SOP_SINC, via_f64(x, [](double d) { return fm::sinc(d); }))
                        NRK_CHAIN_CASE(NRK_SOP_SMOOTHSTEP, via_f64(x, [](double d) { return fm::smoothstep_poly(d); }))
                        NRK_CHAIN_CASE(NRK_SOP_SMOOTHERSTEP, via_f64(x, [](double d) { return fm::smootherstep_poly(d); }))
                        NRK_CHAIN_CASE(NRK_SOP_ERF, via_f64(x, [](double d) { return fm::erf_ref(d); }))
                        NRK_CHAIN_CASE(NRK_SOP_ERFC, via_f64(x, [](double d) { return 1.0 - fm::erf_ref(d); }))
                        NRK_CHAIN_CASE(NRK_SOP_DIGAMMA, via_f64(x, [](double d) { return fm::digamma(d); }))
                        NRK_CHAIN_CASE(NRK_SOP_CBRT, via_f64(x, [](double d) { return fm::cbrt(d); }))
                        default:
                            if constexpr (MATH >= MATH_ALL) {
                                switch (op) {
                                    NRK_CHAIN_CASE(NRK_SOP_POW, fm::pow_t(x, s))
                                    NRK_CHAIN_CASE(NRK_SOP_GAMMA, via_f64(x, [](double d) { return fm::gamma_ref(d); }))
                                    NRK_CHAIN_CASE(NRK_SOP_LGAMMA, via_f64(x, [](double d) { return fm::lgamma_ref(d); }))
                                    default: break;
                                }
                            }
                            break;
                    }
                }
                break;
#undef NRK_CHAIN_CASE
        }
    }
}

template <typename T, int N, int MATH = MATH_NONE>
__device__ __forceinline__ void apply_chain_n(const Chain &c, T (&v)[N]) {
    for (int k = 0; k < c.len; ++k) apply_step_n<T, N, MATH>(c.op[k], c.s[k], v);
}

// The chain with its one elementary-function step fixed at compile time (OP at index `at`), the
// other steps from the IEEE switch: a kernel built this way holds only OP's registers.
template <typename T, int N, int OP>
__device__ __forceinline__ void apply_chain_one_n(const Chain &c, int at, T (&v)[N]) {
    for (int k = 0; k < c.len; ++k) {
        if (k == at) apply_step_n<T, N, MATH_ALL>(OP, c.s[k], v);
        else apply_step_n<T, N, MATH_NONE>(c.op[k], c.s[k], v);
    }
}
// Validates and copies a chain. Without allow_math, elementary-function steps are refused (the
// fused reductions are built without them).
nrk_status make_chain(const nrk_scalar_op *chain, int32_t chain_len, Chain *out, bool allow_math = false);

}  // namespace nrk
