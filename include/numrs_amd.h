/*
 * numrs_amd.h -- C ABI of the MI355X (gfx950) kernels layer behind NumRS2's Array<T>.
 *
 * Every entry point replaces one function of the reference's crate-private `src/kernels`
 * dispatch layer (or a public entry point that bypasses it: argmax/argmin, axis folds,
 * the i64 sum). The citation above each declaration names the reference function
 * (file:line under the reference's `src/`) whose argument meaning and results it keeps.
 *
 * Conventions (all entry points):
 *   - Return an nrk_status; 0 (NRK_OK) on success. nrk_last_error() describes the last failure
 *     of the calling thread.
 *   - Matrices are dense, row-major, in logical order (the reference's `borrow::operand` layout).
 *   - GEMM output is caller-allocated and overwritten (beta = 0).
 *   - Entry points named nrk_<fn> take HOST pointers, exactly like the Rust slices they replace:
 *     they stage the operands through the device and block until the result is on the host.
 *   - Entry points named nrk_dev_<fn> take DEVICE pointers (from nrk_device_malloc) and an
 *     nrk_stream; they enqueue work and return without waiting. Scalar results are written to
 *     device memory.
 *   - Reductions are deterministic: the result bits do not depend on the grid size, the number
 *     of CUs or the order in which workgroups finish (fixed segment tree, fixed-order finish).
 *   - Empty input: sum/mean/min/max/var return 0 and prod returns 1, as the reference's
 *     kernels do (kernels/reduce.rs:207-421). argmax/argmin of an empty input is an error
 *     (NRK_ERR_INVALID_ARGUMENT), as math/statistics/mod.rs:83-87 errors.
 *
 * Concurrency: calls on different streams may run concurrently; each stream has its own
 * scratch workspace. Calls on one stream are serialised by that stream.
 */
#ifndef NUMRS_AMD_H
#define NUMRS_AMD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef int32_t nrk_status;
enum {
    NRK_OK = 0,
    NRK_ERR_INVALID_ARGUMENT = 1, /* shape/length/argument contract violated            */
    NRK_ERR_NO_DEVICE = 2,        /* no gfx950 device, or the device index is out of range */
    NRK_ERR_OUT_OF_MEMORY = 3,    /* hipMalloc of an operand or of scratch failed          */
    NRK_ERR_HIP = 4,              /* any other HIP runtime error (see nrk_last_error)       */
    NRK_ERR_UNSUPPORTED = 5       /* a combination the library does not implement          */
};

/* A hipStream_t. NULL means the null stream of the current device. */
typedef void *nrk_stream;

/* ------------------------------------------------------------------------------------------ */
/* Runtime                                                                                     */
/* ------------------------------------------------------------------------------------------ */

/* Select the device used by subsequent calls from this thread (default 0) and check that it
 * is a gfx950 part. Idempotent. */
nrk_status nrk_init(int device);
const char *nrk_status_string(nrk_status status);
/* Text of the last error raised on the calling thread ("" if none). */
const char *nrk_last_error(void);
/* Library version as "MAJOR.MINOR.PATCH". */
const char *nrk_version(void);
/* Compute units of the calling thread's device (the built-in tile heuristics depend on it). No
 * reference counterpart: the reference sizes its CPU tiers by rayon's thread count instead. */
nrk_status nrk_device_cu_count(int32_t *count);

nrk_status nrk_device_malloc(void **ptr, size_t bytes);
/* Test hook for the out-of-memory paths (no reference counterpart): the library's allocation
 * number `after` from now on (0: the next one) -- scratch, workspace, staging or
 * nrk_device_malloc -- fails with NRK_ERR_OUT_OF_MEMORY, as a real one does; the hook then
 * disarms. after < 0 disarms it. */
nrk_status nrk_debug_fail_allocations(int32_t after);
nrk_status nrk_device_free(void *ptr);
nrk_status nrk_memcpy_htod(void *dst, const void *src, size_t bytes, nrk_stream stream);
nrk_status nrk_memcpy_dtoh(void *dst, const void *src, size_t bytes, nrk_stream stream);
nrk_status nrk_memcpy_dtod(void *dst, const void *src, size_t bytes, nrk_stream stream);
nrk_status nrk_memset_zero(void *dst, size_t bytes, nrk_stream stream);
nrk_status nrk_stream_create(nrk_stream *stream);
nrk_status nrk_stream_destroy(nrk_stream stream);
nrk_status nrk_stream_synchronize(nrk_stream stream);
nrk_status nrk_device_synchronize(void);

/* HIP events for timing work on a stream. */
nrk_status nrk_event_create(void **event);
nrk_status nrk_event_destroy(void *event);
nrk_status nrk_event_record(void *event, nrk_stream stream);
nrk_status nrk_stream_wait_event(nrk_stream stream, void *event); /* later work on stream waits for event */
nrk_status nrk_event_elapsed_ms(float *ms, void *start, void *stop);

/* ------------------------------------------------------------------------------------------ */
/* Elementwise: the reference's closures become op codes                                       */
/* ------------------------------------------------------------------------------------------ */

/* One step of a scalar chain; every step is one IEEE rounding (never contracted to an FMA),
 * the contract of array/operations.rs:448-474 and expr/fused_eval.rs:15-21. */
enum {
    NRK_SOP_ADD = 0,    /* y = x + s  (Array::add_scalar, operations.rs:448)      */
    NRK_SOP_SUB = 1,    /* y = x - s  (Array::subtract_scalar, operations.rs:455) */
    NRK_SOP_MUL = 2,    /* y = x * s  (Array::multiply_scalar, operations.rs:462) */
    NRK_SOP_DIV = 3,    /* y = x / s  (Array::divide_scalar, operations.rs:469)   */
    NRK_SOP_RSUB = 4,   /* y = s - x  (scalar - &Array, arithmetic.rs:342-431)    */
    NRK_SOP_RDIV = 5,   /* y = s / x                                              */
    NRK_SOP_NEG = 6,    /* y = -x     (Neg for &Array, arithmetic.rs:448-468)     */
    NRK_SOP_ABS = 7,    /* y = |x|                                                */
    NRK_SOP_SQUARE = 8, /* y = x * x                                              */
    NRK_SOP_SQRT = 9,   /* y = sqrt(x), correctly rounded                         */
    /* ufuncs.rs:702-780, exact: Rust's f64 floor / ceil / trunc / round (half away from
     * zero) / fract (x - trunc(x)); sign as ufuncs.rs:756 (0 for +-0, 1 for x > 0, else -1,
     * so NaN gives -1); clamp (ufuncs.rs:774, x.clamp(lo, hi)) as CLAMP_LO s = lo then
     * CLAMP_HI s = hi (NaN passes through). */
    NRK_SOP_FLOOR = 10,
    NRK_SOP_CEIL = 11,
    NRK_SOP_TRUNC = 12,
    NRK_SOP_ROUND = 13,
    NRK_SOP_FRACT = 14,
    NRK_SOP_SIGN = 15,
    NRK_SOP_CLAMP_LO = 16, /* y = x < s ? s : x */
    NRK_SOP_CLAMP_HI = 17, /* y = x > s ? s : x */
    /* ufuncs.rs:398-475, 639: exp, exp2, expm1, ln, log2, log10, ln_1p, tanh: FP64 evaluation
     * within 1 ulp of the correctly rounded result (tanh 2 ulp); binary32 inputs are evaluated
     * in FP64 and rounded once. Elementwise map only (nrk_map_* / nrk_dev_map_*): a fused
     * map -> reduce chain holding one is rejected with NRK_ERR_UNSUPPORTED. */
    NRK_SOP_EXP = 18,
    NRK_SOP_EXP2 = 19,
    NRK_SOP_EXPM1 = 20,
    NRK_SOP_LOG = 21,
    NRK_SOP_LOG2 = 22,
    NRK_SOP_LOG10 = 23,
    NRK_SOP_LOG1P = 24,
    NRK_SOP_TANH = 25,
    /* ufuncs.rs:619-700: sinh, cosh, asinh, acosh, atanh (within 2 ulp, acosh 3; elementwise
     * map only) */
    NRK_SOP_SINH = 26,
    NRK_SOP_COSH = 27,
    NRK_SOP_ASINH = 28,
    NRK_SOP_ACOSH = 29,
    NRK_SOP_ATANH = 30,
    /* ufuncs.rs:542-590: atan (within 1.5 ulp), asin, acos (within 2 ulp); elementwise map only */
    NRK_SOP_ATAN = 31,
    NRK_SOP_ASIN = 32,
    NRK_SOP_ACOS = 33,
    /* ufuncs.rs:477-510: sin, cos (within 1 ulp), tan (2.5 ulp) for every finite argument
     * (Payne-Hanek reduction above 2^20); elementwise map only */
    NRK_SOP_SIN = 34,
    NRK_SOP_COS = 35,
    NRK_SOP_TAN = 36,
    /* ufuncs.rs:1243-1480 activations, special functions and steps, as the reference's f64
     * expressions (one IEEE rounding per operation; its exp / ln / tanh / sin within the ulp
     * bounds above). s is the scalar where named. Elementwise map only. */
    NRK_SOP_RELU = 37,        /* :1380 x.max(0.0) (NaN -> 0)                                  */
    NRK_SOP_LEAKY_RELU = 38,  /* :1390 x > 0 ? x : s * x                                      */
    NRK_SOP_ELU = 39,         /* :1322 x > 0 ? x : s * (exp(x) - 1)                           */
    NRK_SOP_SELU = 40,        /* :1334 lambda * (x > 0 ? x : alpha * (exp(x) - 1))            */
    NRK_SOP_HARDSIGMOID = 41, /* :1344 0 / 1 outside [-3, 3], else x / 6 + 0.5                */
    NRK_SOP_HARDSWISH = 42,   /* :1362 0 / x outside [-3, 3], else x * (x + 3) / 6            */
    NRK_SOP_SOFTPLUS = 43,    /* :1291 x > 20 ? x : x < -20 ? exp(x) : ln(1 + exp(x))         */
    NRK_SOP_MISH = 44,        /* :1309 x * tanh(x > 20 ? x : ln(1 + exp(x)))                  */
    NRK_SOP_SILU = 45,        /* :1275 x / (1 + exp(-x))                                      */
    NRK_SOP_GELU = 46,        /* :1262 0.5 x (1 + tanh(sqrt(2/pi) (x + 0.044715 x^3)))        */
    NRK_SOP_LOGIT = 47,       /* :1251 ln(x / (1 - x))                                        */
    NRK_SOP_SINC = 48,        /* :527  x == 0 ? 1 : sin(pi x) / (pi x)                        */
    NRK_SOP_SMOOTHSTEP = 49,  /* :1467 t t (3 - 2 t), after SUB e0, DIV (e1 - e0), clamp 0..1 */
    NRK_SOP_SMOOTHERSTEP = 50, /* :1480 t t t (t (6 t - 15) + 10), the same prefix            */
    NRK_SOP_ERF = 51,         /* :1081 erf_scalar (Chebyshev fit of erfc, |error| < 1.2e-7)   */
    NRK_SOP_ERFC = 52,        /* :1118 1 - erf_scalar(x)                                      */
    NRK_SOP_DIGAMMA = 53,     /* :1202 recurrence to x >= 6, then the asymptotic series       */
    NRK_SOP_POW = 54,         /* :305 power_scalar, x.powf(s), within 1 ulp, C's special cases */
    NRK_SOP_GAMMA = 55,       /* :1122 Lanczos g = 7 (reflection below 0.5), with its powf     */
    NRK_SOP_LGAMMA = 56,      /* :1183 ln |gamma_single(x)| (no reflection, as the reference)  */
    NRK_SOP_CBRT = 57,        /* :367 x.cbrt(), within 1 ulp                                   */
    NRK_SOP_COUNT = 58
};
typedef struct nrk_scalar_op {
    int32_t op;     /* NRK_SOP_* */
    int32_t reserved;
    double scalar;  /* rounded to the element type before use */
} nrk_scalar_op;
#define NRK_MAX_CHAIN 8

/* Binary op codes for zip: out[i] = a[i] (op) b[i], one rounding. */
enum {
    NRK_BOP_ADD = 0, /* arithmetic.rs:35-100  Add            */
    NRK_BOP_SUB = 1, /* arithmetic.rs:100-160 Sub            */
    NRK_BOP_MUL = 2, /* arithmetic.rs:160-220 Mul            */
    NRK_BOP_DIV = 3, /* arithmetic.rs:220-270 Div            */
    NRK_BOP_REM = 4, /* arithmetic.rs:273-320 Rem (truncated, fmod) */
    NRK_BOP_MAXIMUM = 5, /* ufuncs.rs:169  if a > b {a} else {b} */
    NRK_BOP_MINIMUM = 6, /* ufuncs.rs:173  if a < b {a} else {b} */
    NRK_BOP_COPYSIGN = 7,  /* ufuncs.rs:784  a.copysign(b), exact                       */
    NRK_BOP_HYPOT = 8,     /* ufuncs.rs:841  a.hypot(b), within 1 ulp                   */
    NRK_BOP_ATAN2 = 9,     /* ufuncs.rs:587  a.atan2(b) (a = y, b = x), within 1.5 ulp  */
    NRK_BOP_LOGADDEXP = 10, /* ufuncs.rs:916 m + ln(e^-|a-m| + e^-|b-m|), within 2 ulp */
    NRK_BOP_PRELU = 11,    /* ufuncs.rs:1394 a > 0 ? a : b * a (b = alpha)             */
    NRK_BOP_POWER = 12,    /* ufuncs.rs:166 a.powf(b), within 1 ulp                      */
    NRK_BOP_COUNT = 13,    /* the floating-point ops end here; the rest are integer-only  */
    /* integer-only (bitwise_ops.rs:30-243): &, |, ^, and shifts whose amount is masked to the
     * width (Rust's release arithmetic); >> is arithmetic for signed types, logical otherwise */
    NRK_BOP_BITAND = 13,
    NRK_BOP_BITOR = 14,
    NRK_BOP_BITXOR = 15,
    NRK_BOP_SHL = 16,
    NRK_BOP_SHR = 17
};

/* kernels/elementwise.rs:237-258 unary_serial / unary_dispatch with an op-coded closure:
 * y[i] = chain(x[i]). In-place (y == x) is allowed. */
nrk_status nrk_map_f64(const nrk_scalar_op *chain, int32_t chain_len, const double *x, double *y, size_t n);
nrk_status nrk_map_f32(const nrk_scalar_op *chain, int32_t chain_len, const float *x, float *y, size_t n);
nrk_status nrk_dev_map_f64(const nrk_scalar_op *chain, int32_t chain_len, const double *x, double *y, size_t n, nrk_stream stream);
nrk_status nrk_dev_map_f32(const nrk_scalar_op *chain, int32_t chain_len, const float *x, float *y, size_t n, nrk_stream stream);

/* kernels/elementwise.rs:70-109 binary_f64/binary_f32 (and :217 binary_dispatch) with an
 * op-coded closure: y[i] = a[i] (op) b[i]. */
nrk_status nrk_zip_f64(int32_t op, const double *a, const double *b, double *y, size_t n);
nrk_status nrk_zip_f32(int32_t op, const float *a, const float *b, float *y, size_t n);
nrk_status nrk_dev_zip_f64(int32_t op, const double *a, const double *b, double *y, size_t n, nrk_stream stream);
nrk_status nrk_dev_zip_f32(int32_t op, const float *a, const float *b, float *y, size_t n, nrk_stream stream);

/* kernels/elementwise.rs:217-258 binary_dispatch / unary_dispatch for the integer T (i64, i32,
 * u64, u32) with array/arithmetic.rs's operators and bitwise_ops.rs's. Rust's release arithmetic:
 * + - * and negation wrap; / and % truncate; shift amounts are masked to the width. Rust panics on
 * x / 0, x % 0 and MIN / -1: the host forms check first and return NRK_ERR_INVALID_ARGUMENT (or
 * NRK_ERR_UNSUPPORTED where the check would need the data of a scalar-divisor chain) without
 * writing; the device forms give x / 0 = x % 0 = 0, MIN / -1 = MIN, MIN % -1 = 0.
 * zip ops: NRK_BOP_ADD SUB MUL DIV REM MAXIMUM MINIMUM BITAND BITOR BITXOR SHL SHR.
 * map steps: NRK_SOP_ADD SUB MUL DIV RSUB RDIV NEG ABS SQUARE SIGN CLAMP_LO CLAMP_HI (NEG, ABS
 * and SIGN only for the signed types), with an integer scalar. */
typedef struct nrk_int_op {
    int32_t op;       /* NRK_SOP_* */
    int32_t reserved; /* 0 */
    int64_t scalar;   /* the step's integer operand (converted to the array's type) */
} nrk_int_op;
nrk_status nrk_zip_i64(int32_t op, const int64_t *a, const int64_t *b, int64_t *y, size_t n);
nrk_status nrk_zip_i32(int32_t op, const int32_t *a, const int32_t *b, int32_t *y, size_t n);
nrk_status nrk_zip_u64(int32_t op, const uint64_t *a, const uint64_t *b, uint64_t *y, size_t n);
nrk_status nrk_zip_u32(int32_t op, const uint32_t *a, const uint32_t *b, uint32_t *y, size_t n);
nrk_status nrk_dev_zip_i64(int32_t op, const int64_t *a, const int64_t *b, int64_t *y, size_t n, nrk_stream stream);
nrk_status nrk_dev_zip_i32(int32_t op, const int32_t *a, const int32_t *b, int32_t *y, size_t n, nrk_stream stream);
nrk_status nrk_dev_zip_u64(int32_t op, const uint64_t *a, const uint64_t *b, uint64_t *y, size_t n, nrk_stream stream);
nrk_status nrk_dev_zip_u32(int32_t op, const uint32_t *a, const uint32_t *b, uint32_t *y, size_t n, nrk_stream stream);
nrk_status nrk_map_i64(const nrk_int_op *chain, int32_t chain_len, const int64_t *x, int64_t *y, size_t n);
nrk_status nrk_map_i32(const nrk_int_op *chain, int32_t chain_len, const int32_t *x, int32_t *y, size_t n);
nrk_status nrk_map_u64(const nrk_int_op *chain, int32_t chain_len, const uint64_t *x, uint64_t *y, size_t n);
nrk_status nrk_map_u32(const nrk_int_op *chain, int32_t chain_len, const uint32_t *x, uint32_t *y, size_t n);
nrk_status nrk_dev_map_i64(const nrk_int_op *chain, int32_t chain_len, const int64_t *x, int64_t *y, size_t n,
                           nrk_stream stream);
nrk_status nrk_dev_map_i32(const nrk_int_op *chain, int32_t chain_len, const int32_t *x, int32_t *y, size_t n,
                           nrk_stream stream);
nrk_status nrk_dev_map_u64(const nrk_int_op *chain, int32_t chain_len, const uint64_t *x, uint64_t *y, size_t n,
                           nrk_stream stream);
nrk_status nrk_dev_map_u32(const nrk_int_op *chain, int32_t chain_len, const uint32_t *x, uint32_t *y, size_t n,
                           nrk_stream stream);

/* Element types of the dtype-generic entry points below. */
enum { NRK_DTYPE_F64 = 0, NRK_DTYPE_F32 = 1, NRK_DTYPE_I64 = 2, NRK_DTYPE_I32 = 3, NRK_DTYPE_U64 = 4, NRK_DTYPE_U32 = 5 };
#define NRK_MAX_NDIM 8

/* operations.rs:370-440 add/subtract/multiply/divide_broadcast (Array::broadcast_op + binary_serial):
 * y = a op b over numpy's broadcast of a's and b's shapes (right-aligned; extents equal or 1; at
 * most NRK_MAX_NDIM axes), without materialising either operand at the common shape. y is
 * contiguous at the broadcast shape; a, b and y are host pointers; shapes that do not broadcast
 * are NRK_ERR_INVALID_ARGUMENT (the reference's ShapeMismatch). Floating-point ops: NRK_BOP_ADD ..
 * NRK_BOP_COPYSIGN and PRELU; integer ops as for nrk_zip_i64 (divisors checked as there). */
nrk_status nrk_zip_broadcast(int32_t dtype, int32_t op, const void *a, int32_t a_ndim, const size_t *a_shape, const void *b,
                             int32_t b_ndim, const size_t *b_shape, void *y);
/* The device form, over any strides: y (contiguous, `shape`) = a op b with a[i] at
 * sum_d i_d a_strides[d] (elements; 0 broadcasts an axis, negative strides are allowed), likewise
 * b. This serves broadcasting, non-contiguous views (no borrow::operand copy) and binary_serial's
 * heterogeneous operand types (kernels/elementwise.rs:51-66): for NRK_BOP_SHL / SHR the shift
 * amounts b may have any integer dtype b_dtype (bitwise_ops.rs:190-243 left_shift<T, U>); every
 * other op needs b_dtype == dtype. */
nrk_status nrk_dev_zip_strided(int32_t dtype, int32_t op, int32_t ndim, const size_t *shape, const void *a,
                               const int64_t *a_strides, const void *b, int32_t b_dtype, const int64_t *b_strides, void *y,
                               nrk_stream stream);

/* ------------------------------------------------------------------------------------------ */
/* Reductions over a flat slice                                                                */
/* ------------------------------------------------------------------------------------------ */

/* kernels/reduce.rs:207  sum_{f64,f32}(a) -> T                       (0 for empty)          */
nrk_status nrk_sum_f64(const double *a, size_t n, double *out);
nrk_status nrk_sum_f32(const float *a, size_t n, float *out);
/* kernels/reduce.rs:227  mean_{f64,f32}(a) -> T  = sum / n            (0 for empty)          */
nrk_status nrk_mean_f64(const double *a, size_t n, double *out);
nrk_status nrk_mean_f32(const float *a, size_t n, float *out);
/* kernels/reduce.rs:239  sum_sq_dev_{f64,f32}(a, mean) -> sum((a-mean)^2)                    */
nrk_status nrk_sum_sq_dev_f64(const double *a, size_t n, double mean, double *out);
nrk_status nrk_sum_sq_dev_f32(const float *a, size_t n, float mean, float *out);
/* kernels/reduce.rs:284  var_{f64,f32}(a, ddof) -> T, divisor (n - ddof) taken in floating
 * point, so ddof >= n gives inf/NaN as in the reference                 (0 for empty)          */
nrk_status nrk_var_f64(const double *a, size_t n, size_t ddof, double *out);
nrk_status nrk_var_f32(const float *a, size_t n, size_t ddof, float *out);
/* kernels/reduce.rs:323,376  min/max_{f64,f32}(a): any NaN -> NaN      (0 for empty)          */
nrk_status nrk_min_f64(const double *a, size_t n, double *out);
nrk_status nrk_min_f32(const float *a, size_t n, float *out);
nrk_status nrk_max_f64(const double *a, size_t n, double *out);
nrk_status nrk_max_f32(const float *a, size_t n, float *out);
/* kernels/reduce.rs:419  prod_{f64,f32}(a)                            (1 for empty)          */
nrk_status nrk_prod_f64(const double *a, size_t n, double *out);
nrk_status nrk_prod_f32(const float *a, size_t n, float *out);
/* array/operations.rs:303-318  Array::<i64>::sum: wrapping integer fold, bit-exact           */
nrk_status nrk_sum_i64(const int64_t *a, size_t n, int64_t *out);
/* math/statistics/mod.rs:79-103 / :218-242  argmax/argmin(None): index of the first strict
 * maximum (minimum); NaN elements are skipped, except that a NaN at index 0 yields 0.         */
nrk_status nrk_argmax_f64(const double *a, size_t n, uint64_t *out);
nrk_status nrk_argmax_f32(const float *a, size_t n, uint64_t *out);
nrk_status nrk_argmin_f64(const double *a, size_t n, uint64_t *out);
nrk_status nrk_argmin_f32(const float *a, size_t n, uint64_t *out);

/* Device-pointer forms of the above (out is a device pointer to one element). */
nrk_status nrk_dev_sum_f64(const double *a, size_t n, double *out, nrk_stream stream);
nrk_status nrk_dev_sum_f32(const float *a, size_t n, float *out, nrk_stream stream);
nrk_status nrk_dev_mean_f64(const double *a, size_t n, double *out, nrk_stream stream);
nrk_status nrk_dev_mean_f32(const float *a, size_t n, float *out, nrk_stream stream);
nrk_status nrk_dev_var_f64(const double *a, size_t n, size_t ddof, double *out, nrk_stream stream);
nrk_status nrk_dev_var_f32(const float *a, size_t n, size_t ddof, float *out, nrk_stream stream);
nrk_status nrk_dev_min_f64(const double *a, size_t n, double *out, nrk_stream stream);
nrk_status nrk_dev_min_f32(const float *a, size_t n, float *out, nrk_stream stream);
nrk_status nrk_dev_max_f64(const double *a, size_t n, double *out, nrk_stream stream);
nrk_status nrk_dev_max_f32(const float *a, size_t n, float *out, nrk_stream stream);
nrk_status nrk_dev_prod_f64(const double *a, size_t n, double *out, nrk_stream stream);
nrk_status nrk_dev_prod_f32(const float *a, size_t n, float *out, nrk_stream stream);
nrk_status nrk_dev_sum_i64(const int64_t *a, size_t n, int64_t *out, nrk_stream stream);
nrk_status nrk_dev_argmax_f64(const double *a, size_t n, uint64_t *out, nrk_stream stream);
nrk_status nrk_dev_argmax_f32(const float *a, size_t n, uint64_t *out, nrk_stream stream);
nrk_status nrk_dev_argmin_f64(const double *a, size_t n, uint64_t *out, nrk_stream stream);
nrk_status nrk_dev_argmin_f32(const float *a, size_t n, uint64_t *out, nrk_stream stream);

/* ------------------------------------------------------------------------------------------ */
/* Axis reductions over an (outer, len, inner) view of a row-major array                       */
/* ------------------------------------------------------------------------------------------ */
/* The reduced axis has `len` elements with stride `inner`; there are outer*inner outputs, in
 * row-major order of the remaining axes. A 2-D (R, C) array reduced along axis 1 is
 * (outer=R, len=C, inner=1); along axis 0 it is (outer=1, len=R, inner=C).
 * The reference folds each output left to right (math/aggregation.rs:548-561 for sum,
 * :244-261 for max/min with NaN short-circuit; math/statistics/mod.rs:1064-1071 mean,
 * :1403-1494 var = mean(axis) then sum((x-mean)^2)/(len-ddof); :105-183 argmax). len == 0
 * is an error for max/min/argmax/argmin/var (the reference errors on empty input) and gives
 * zeros for sum/mean. For var, len <= ddof is an error (statistics/mod.rs:1423-1428).
 * Here: inner == 1 runs the row kernels (compensated sums and variances, within the bounds the
 * fused reduction states); inner > 1 with at least 2^18 outputs folds each output left to right
 * as the reference does (bit-exact); with fewer, every op reads column tiles over ranges of the
 * axis once (binary64 partials for sums and means, Chan records for variances, strict winners for
 * extrema), then merges each output's ranges. Extrema and their indices are bit-exact on every
 * path. */
enum {
    NRK_AXIS_SUM = 0, NRK_AXIS_MEAN = 1, NRK_AXIS_VAR = 2, NRK_AXIS_MIN = 3,
    NRK_AXIS_MAX = 4, NRK_AXIS_ARGMIN = 5, NRK_AXIS_ARGMAX = 6
};
/* out is T* for SUM/MEAN/VAR/MIN/MAX and uint64_t* for ARGMIN/ARGMAX. */
nrk_status nrk_reduce_axis_f64(int32_t op, const double *x, size_t outer, size_t len, size_t inner, size_t ddof, void *out);
nrk_status nrk_reduce_axis_f32(int32_t op, const float *x, size_t outer, size_t len, size_t inner, size_t ddof, void *out);
nrk_status nrk_dev_reduce_axis_f64(int32_t op, const double *x, size_t outer, size_t len, size_t inner, size_t ddof, void *out, nrk_stream stream);
nrk_status nrk_dev_reduce_axis_f32(int32_t op, const float *x, size_t outer, size_t len, size_t inner, size_t ddof, void *out, nrk_stream stream);

/* ufunc_ops.rs:583 ufunc_outer: y[i * nb + j] = a[i] op b[j] (shape a.shape + b.shape), and
 * ufunc_ops.rs:841 ufunc_where: y[i] = mask[i] ? a[i] op b[i] : a[i] (mask bytes 0 / nonzero).
 * op: NRK_BOP_ADD, SUB, MUL, DIV, MAXIMUM, MINIMUM, with UfuncOp::apply's maximum / minimum
 * (NaN if either operand is NaN, ufunc_ops.rs:156). One rounding per element. */
nrk_status nrk_outer_f64(int32_t op, const double *a, size_t na, const double *b, size_t nb, double *y);
nrk_status nrk_outer_f32(int32_t op, const float *a, size_t na, const float *b, size_t nb, float *y);
nrk_status nrk_dev_outer_f64(int32_t op, const double *a, size_t na, const double *b, size_t nb, double *y, nrk_stream stream);
nrk_status nrk_dev_outer_f32(int32_t op, const float *a, size_t na, const float *b, size_t nb, float *y, nrk_stream stream);
nrk_status nrk_where_f64(int32_t op, const double *a, const double *b, const uint8_t *mask, double *y, size_t n);
nrk_status nrk_where_f32(int32_t op, const float *a, const float *b, const uint8_t *mask, float *y, size_t n);
nrk_status nrk_dev_where_f64(int32_t op, const double *a, const double *b, const uint8_t *mask, double *y, size_t n,
                             nrk_stream stream);
nrk_status nrk_dev_where_f32(int32_t op, const float *a, const float *b, const uint8_t *mask, float *y, size_t n,
                             nrk_stream stream);

/* ufunc_ops.rs:772-838 ufunc_at, in place: for k = 0 .. m-1 in order, row indices[k] of a
 * (n0 rows of row_len) becomes a[indices[k]][:] op b[k][:], op as for accumulate below. Repeated
 * indices fold in k order, the reference's sequence of roundings exactly. Host form: an index
 * >= n0 is NRK_ERR_INVALID_ARGUMENT (IndexOutOfBounds) and a is untouched; device form: indices
 * >= n0 are skipped (their updates are dropped; the valid ones are applied as above). */
nrk_status nrk_at_f64(int32_t op, double *a, size_t n0, size_t row_len, const uint64_t *indices, size_t m, const double *b);
nrk_status nrk_at_f32(int32_t op, float *a, size_t n0, size_t row_len, const uint64_t *indices, size_t m, const float *b);
nrk_status nrk_dev_at_f64(int32_t op, double *a, size_t n0, size_t row_len, const uint64_t *indices, size_t m,
                          const double *b, nrk_stream stream);
nrk_status nrk_dev_at_f32(int32_t op, float *a, size_t n0, size_t row_len, const uint64_t *indices, size_t m,
                          const float *b, nrk_stream stream);

/* ufunc_ops.rs:619-733 ufunc_reduceat along contiguous rows: y[r][k] reduces x[r][j .. end) with
 * j = indices[k] and end = indices[k + 1] (len for the last k), or is x[r][j] when end <= j. op as
 * for accumulate below. Host form: an index >= len is NRK_ERR_INVALID_ARGUMENT (the reference's
 * IndexOutOfBounds); device form (indices on the device): such an output is NaN. */
nrk_status nrk_reduceat_f64(int32_t op, const double *x, size_t rows, size_t len, const uint64_t *indices, size_t nidx, double *y);
nrk_status nrk_reduceat_f32(int32_t op, const float *x, size_t rows, size_t len, const uint64_t *indices, size_t nidx, float *y);
nrk_status nrk_dev_reduceat_f64(int32_t op, const double *x, size_t rows, size_t len, const uint64_t *indices, size_t nidx,
                                double *y, nrk_stream stream);
nrk_status nrk_dev_reduceat_f32(int32_t op, const float *x, size_t rows, size_t len, const uint64_t *indices, size_t nidx,
                                float *y, nrk_stream stream);

/* ufunc_ops.rs:535-580 ufunc_accumulate along contiguous rows: y[r][j] = x[r][0] op ... op x[r][j]
 * for `rows` rows of `len` (axis = None is one row; another axis is moved last by the caller).
 * op: NRK_BOP_ADD (cumsum, statistics/mod.rs:584, running sum from +0), NRK_BOP_MUL (cumprod),
 * NRK_BOP_MAXIMUM / NRK_BOP_MINIMUM (NaN-propagating, ufunc_ops.rs:156; exact, so the same bits
 * as the sequential fold), NRK_BOP_SUB (x0 - x1 - ...), NRK_BOP_DIV (x0 / x1 / ..., as a product of
 * reciprocals). Sums and products run in a fixed tree set by the shape. */
nrk_status nrk_accumulate_f64(int32_t op, const double *x, size_t rows, size_t len, double *y);
nrk_status nrk_accumulate_f32(int32_t op, const float *x, size_t rows, size_t len, float *y);
nrk_status nrk_dev_accumulate_f64(int32_t op, const double *x, size_t rows, size_t len, double *y, nrk_stream stream);
nrk_status nrk_dev_accumulate_f32(int32_t op, const float *x, size_t rows, size_t len, float *y, nrk_stream stream);

/* ufuncs.rs:1411-1457 softmax and log_softmax over all n elements (the reference flattens):
 * m = max, e = exp(x - m), S = sum e (fixed tree), then y = e / S or y = (x - m) - ln S. A NaN
 * anywhere makes every output NaN, as in the reference. y may alias x only for n = 0. */
nrk_status nrk_softmax_f64(const double *x, double *y, size_t n);
nrk_status nrk_softmax_f32(const float *x, float *y, size_t n);
nrk_status nrk_log_softmax_f64(const double *x, double *y, size_t n);
nrk_status nrk_log_softmax_f32(const float *x, float *y, size_t n);
nrk_status nrk_dev_softmax_f64(const double *x, double *y, size_t n, nrk_stream stream);
nrk_status nrk_dev_softmax_f32(const float *x, float *y, size_t n, nrk_stream stream);
nrk_status nrk_dev_log_softmax_f64(const double *x, double *y, size_t n, nrk_stream stream);
nrk_status nrk_dev_log_softmax_f32(const float *x, float *y, size_t n, nrk_stream stream);

/* ufuncs.rs:864 fma: y = a.mul_add(b, c), one rounding; ufuncs.rs:893 lerp: y = a + t * (b - a),
 * three roundings. Equal lengths n; the reference's ShapeMismatch is the caller's check. */
nrk_status nrk_fma_f64(const double *a, const double *b, const double *c, double *y, size_t n);
nrk_status nrk_fma_f32(const float *a, const float *b, const float *c, float *y, size_t n);
nrk_status nrk_dev_fma_f64(const double *a, const double *b, const double *c, double *y, size_t n, nrk_stream stream);
nrk_status nrk_dev_fma_f32(const float *a, const float *b, const float *c, float *y, size_t n, nrk_stream stream);
nrk_status nrk_lerp_f64(const double *a, const double *b, double t, double *y, size_t n);
nrk_status nrk_lerp_f32(const float *a, const float *b, float t, float *y, size_t n);
nrk_status nrk_dev_lerp_f64(const double *a, const double *b, double t, double *y, size_t n, nrk_stream stream);
nrk_status nrk_dev_lerp_f32(const float *a, const float *b, float t, float *y, size_t n, nrk_stream stream);

/* ufuncs.rs:946-979, 1060: dot (sum of the rounded products a[i] * b[i]; 0 for n = 0),
 * norm_l1 (sum |x|), norm_l2 (sqrt of sum x * x) and std (sqrt of var with divisor n - ddof;
 * the reference's std is ddof = 0). Sums run in the fixed tree of sum above. */
nrk_status nrk_dot_f64(const double *a, const double *b, size_t n, double *out);
nrk_status nrk_dot_f32(const float *a, const float *b, size_t n, float *out);
nrk_status nrk_dev_dot_f64(const double *a, const double *b, size_t n, double *out, nrk_stream stream);
nrk_status nrk_dev_dot_f32(const float *a, const float *b, size_t n, float *out, nrk_stream stream);
nrk_status nrk_norm_l1_f64(const double *a, size_t n, double *out);
nrk_status nrk_norm_l1_f32(const float *a, size_t n, float *out);
nrk_status nrk_norm_l2_f64(const double *a, size_t n, double *out);
nrk_status nrk_norm_l2_f32(const float *a, size_t n, float *out);
nrk_status nrk_dev_norm_l1_f64(const double *a, size_t n, double *out, nrk_stream stream);
nrk_status nrk_dev_norm_l1_f32(const float *a, size_t n, float *out, nrk_stream stream);
nrk_status nrk_dev_norm_l2_f64(const double *a, size_t n, double *out, nrk_stream stream);
nrk_status nrk_dev_norm_l2_f32(const float *a, size_t n, float *out, nrk_stream stream);
nrk_status nrk_std_f64(const double *a, size_t n, size_t ddof, double *out);
nrk_status nrk_std_f32(const float *a, size_t n, size_t ddof, float *out);
nrk_status nrk_dev_std_f64(const double *a, size_t n, size_t ddof, double *out, nrk_stream stream);
nrk_status nrk_dev_std_f32(const float *a, size_t n, size_t ddof, float *out, nrk_stream stream);

/* ------------------------------------------------------------------------------------------ */
/* Fused elementwise -> reduction (one HBM pass)                                               */
/* ------------------------------------------------------------------------------------------ */
/* No reference counterpart as one call: it fuses `(x * a) + b` (two roundings,
 * array/operations.rs:448-474) with the reductions above, which the reference runs as separate
 * passes over a materialised intermediate. The input is read once.
 *
 * x is (rows, cols) row-major; y = chain(x) elementwise. Outputs are device pointers; any may
 * be NULL to skip it. row_* outputs have `rows` elements (reduction along axis 1); full_*
 * outputs have one element (reduction over all rows*cols elements, argmax as a flat index).
 * var uses divisor (count - ddof). */
typedef struct nrk_reduce_outputs {
    void *sum;      /* T        */
    void *mean;     /* T        */
    void *var;      /* T        */
    void *min;      /* T        */
    void *max;      /* T        */
    uint64_t *argmin;
    uint64_t *argmax;
} nrk_reduce_outputs;
nrk_status nrk_dev_map_reduce_f64(const nrk_scalar_op *chain, int32_t chain_len, const double *x, size_t rows, size_t cols,
                                  size_t ddof, const nrk_reduce_outputs *row_out, const nrk_reduce_outputs *full_out,
                                  nrk_stream stream);
nrk_status nrk_dev_map_reduce_f32(const nrk_scalar_op *chain, int32_t chain_len, const float *x, size_t rows, size_t cols,
                                  size_t ddof, const nrk_reduce_outputs *row_out, const nrk_reduce_outputs *full_out,
                                  nrk_stream stream);

/* The same call with every sum, mean and variance computed in the reference's own order of
 * rounded additions, so each is bit-identical to the reference's CPU result on y:
 *   full sum/mean: kernels/reduce.rs:207-232 (simd_sum below 10000 elements; else simd_sum of
 *     each 8192-element chunk, folded left to right), mean = sum / n;
 *   full var: kernels/reduce.rs:284-314 (one tier decision at 40000 elements for both passes);
 *   row sum/mean/var: math/aggregation.rs:548-561, math/statistics/mod.rs:1064-1071 and
 *     :1403-1494 (left-to-right folds along the row; var = sum((y - mean)^2) / (cols - ddof)).
 * min/max/argmin/argmax do not depend on an order of additions and come from the fused kernel.
 * Several passes over x (two per requested sum-and-variance); the compensated call above is the
 * faster and more accurate one. */
nrk_status nrk_dev_map_reduce_plain_f64(const nrk_scalar_op *chain, int32_t chain_len, const double *x, size_t rows,
                                        size_t cols, size_t ddof, const nrk_reduce_outputs *row_out,
                                        const nrk_reduce_outputs *full_out, nrk_stream stream);
nrk_status nrk_dev_map_reduce_plain_f32(const nrk_scalar_op *chain, int32_t chain_len, const float *x, size_t rows,
                                        size_t cols, size_t ddof, const nrk_reduce_outputs *row_out,
                                        const nrk_reduce_outputs *full_out, nrk_stream stream);

/* No counterpart in the reference. For callers that cannot allocate device memory (code under
 * hipGraph capture, tight serving loops, harnesses that forbid hipMalloc): nrk_dev_map_reduce_f64
 * with the same arguments and results, whose scratch (chunk records, record-tree buffers, the
 * finish counter) is a fixed per-device array of NRK_NOALLOC_SCRATCH_BYTES, set up by nrk_init.
 * It calls no hipMalloc / hipFree. Whole-array results alone fit for every shape (about 8192 chunk
 * records); row results need about 272 B of records per row, so up to ~61 000 rows. A shape whose
 * records exceed the array is NRK_ERR_INVALID_ARGUMENT, and nothing is launched.
 * Every call on one device shares that device's array: calls that overlap on one device (e.g. on
 * two streams) race and must be serialised by the caller. */
#define NRK_NOALLOC_SCRATCH_BYTES ((size_t)16 << 20)
nrk_status nrk_dev_map_reduce_f64_noalloc(const nrk_scalar_op *chain, int32_t chain_len, const double *x, size_t rows,
                                          size_t cols, size_t ddof, const nrk_reduce_outputs *row_out,
                                          const nrk_reduce_outputs *full_out, nrk_stream stream);

/* expr/fused_eval.rs:421-697 expression fusion: an ExprNode tree (owned.rs:204-223) over up to
 * NRK_MAX_LEAVES equally long operands, as a postfix program (children before their parent):
 *   NRK_EXPR_LEAF   op = operand index                  (ExprNode::Leaf)
 *   NRK_EXPR_STEP   op = NRK_SOP_ADD/SUB/MUL/DIV with `scalar` (ScalarRhs: x op s), RSUB / RDIV
 *                   (ScalarLhs s - x, s / x; s + x and s * x are ADD / MUL), NEG, ABS, SQRT, EXP,
 *                   LOG (UnaryOp::Neg/Abs/Sqrt/Exp/Ln)        -- applied to the top of the stack
 *   NRK_EXPR_BINARY op = NRK_BOP_ADD/SUB/MUL/DIV (BinOp), also REM, MAXIMUM, MINIMUM, COPYSIGN,
 *                   PRELU: pops r then l, pushes l op r    (ExprNode::Binary)
 *   NRK_EXPR_FMA    pops c, b, a, pushes (a * b) + c with two roundings (ExprNode::Fma)
 * Every node rounds once and nothing is contracted, so the result is eval_eager's to the bit;
 * the whole tree is one pass (each operand read once, y written once). Limits: NRK_MAX_EXPR nodes,
 * a stack of 4. */
enum { NRK_EXPR_LEAF = 0, NRK_EXPR_STEP = 1, NRK_EXPR_BINARY = 2, NRK_EXPR_FMA = 3 };
#define NRK_MAX_EXPR 16
#define NRK_MAX_LEAVES 4
typedef struct nrk_expr_node {
    int32_t kind;  /* NRK_EXPR_* */
    int32_t op;    /* see above */
    double scalar; /* NRK_EXPR_STEP's operand */
} nrk_expr_node;
nrk_status nrk_expr_eval_f64(const nrk_expr_node *nodes, int32_t len, const double *const *leaves, int32_t n_leaves,
                             double *y, size_t n);
nrk_status nrk_expr_eval_f32(const nrk_expr_node *nodes, int32_t len, const float *const *leaves, int32_t n_leaves,
                             float *y, size_t n);
nrk_status nrk_dev_expr_eval_f64(const nrk_expr_node *nodes, int32_t len, const double *const *leaves, int32_t n_leaves,
                                 double *y, size_t n, nrk_stream stream);
nrk_status nrk_dev_expr_eval_f32(const nrk_expr_node *nodes, int32_t len, const float *const *leaves, int32_t n_leaves,
                                 float *y, size_t n, nrk_stream stream);
/* The expression feeding the fused reduction (as nrk_dev_map_reduce_*, leaves are (rows, cols)
 * row-major): one pass over the operands when the tree is one operand with scalar steps, or two
 * operands (each with scalar steps) combined by one binary op, followed by scalar steps; any other
 * tree is evaluated into a workspace first. EXP / LOG steps are not allowed here
 * (NRK_ERR_UNSUPPORTED), as for the map -> reduce chains. */
nrk_status nrk_dev_expr_reduce_f64(const nrk_expr_node *nodes, int32_t len, const double *const *leaves, int32_t n_leaves,
                                   size_t rows, size_t cols, size_t ddof, const nrk_reduce_outputs *row_out,
                                   const nrk_reduce_outputs *full_out, nrk_stream stream);
nrk_status nrk_dev_expr_reduce_f32(const nrk_expr_node *nodes, int32_t len, const float *const *leaves, int32_t n_leaves,
                                   size_t rows, size_t cols, size_t ddof, const nrk_reduce_outputs *row_out,
                                   const nrk_reduce_outputs *full_out, nrk_stream stream);

/* ------------------------------------------------------------------------------------------ */
/* GEMM: C := A * B, row-major, overwrite                                                      */
/* ------------------------------------------------------------------------------------------ */
/* kernels/gemm.rs:246-268 gemm_2d<T>(m, k, n, a, b, c) for T = f64 / f32. The slice lengths
 * are passed so that the reference's length-mismatch rule holds: if a_len != m*k, b_len != k*n
 * or c_len != m*n, C is zero-filled where it exists and nothing is read (gemm_generic's
 * defensive branch, gemm.rs:408-423). m == 0 or n == 0 returns without writing; k == 0 writes
 * zeros (gemm.rs:295-303). */
nrk_status nrk_gemm_2d_f64(size_t m, size_t k, size_t n, const double *a, size_t a_len, const double *b, size_t b_len,
                           double *c, size_t c_len);
nrk_status nrk_gemm_2d_f32(size_t m, size_t k, size_t n, const float *a, size_t a_len, const float *b, size_t b_len,
                           float *c, size_t c_len);
/* Device forms: plain FP64 / FP32 on MFMA (v_mfma_f64_16x16x4 / v_mfma_f32_16x16x4). */
nrk_status nrk_dev_gemm_f64(size_t m, size_t k, size_t n, const double *a, const double *b, double *c, nrk_stream stream);
/* kernels/gemm.rs:408-445 gemm_generic<T> for the integer types (Array<i32/i64>::matmul): every
 * output accumulated from 0 in ascending k as acc + a * b, wrapping like Rust's release
 * arithmetic, so the result is exact modulo 2^32 / 2^64 (the same bits for u32 / u64 data).
 * Host form: the reference's slice contract -- C is zeroed over c_len; a, b or c shorter than
 * m*k, k*n, m*n leaves it zeroed; otherwise the leading parts are multiplied. Device form: `batch`
 * panels (C panels of m*n; a panel stride of 0 broadcasts that operand, as matmul's batch axes). */
nrk_status nrk_gemm_generic_i64(size_t m, size_t k, size_t n, const int64_t *a, size_t a_len, const int64_t *b,
                                size_t b_len, int64_t *c, size_t c_len);
nrk_status nrk_gemm_generic_i32(size_t m, size_t k, size_t n, const int32_t *a, size_t a_len, const int32_t *b,
                                size_t b_len, int32_t *c, size_t c_len);
nrk_status nrk_dev_gemm_generic_i64(size_t batch, size_t m, size_t k, size_t n, const int64_t *a, size_t stride_a,
                                    const int64_t *b, size_t stride_b, int64_t *c, nrk_stream stream);
nrk_status nrk_dev_gemm_generic_i32(size_t batch, size_t m, size_t k, size_t n, const int32_t *a, size_t stride_a,
                                    const int32_t *b, size_t stride_b, int32_t *c, nrk_stream stream);
nrk_status nrk_dev_gemm_f32(size_t m, size_t k, size_t n, const float *a, const float *b, float *c, nrk_stream stream);

/* array/linalg.rs:56-136 batched_gemm<T>(batch, m, k, n, a, b, c) for T = f64 / f32 (the
 * par_batch_f64 / par_batch_f32 tier): C[t] := A[t] * B[t] over contiguous panels
 * ([batch, m, k] x [batch, k, n] -> [batch, m, n]). batch, m or n == 0, or slice lengths other
 * than batch*m*k, batch*k*n, batch*m*n, return without writing C (linalg.rs:60-73); otherwise
 * panel t is exactly gemm_2d on that panel (k == 0 zero-fills it), bit for bit. */
nrk_status nrk_batched_gemm_f64(size_t batch, size_t m, size_t k, size_t n, const double *a, size_t a_len, const double *b,
                                size_t b_len, double *c, size_t c_len);
nrk_status nrk_batched_gemm_f32(size_t batch, size_t m, size_t k, size_t n, const float *a, size_t a_len, const float *b,
                                size_t b_len, float *c, size_t c_len);
/* Device forms with panel strides in elements: A[t] = a + t*stride_a, B[t] = b + t*stride_b,
 * C[t] = c + t*m*n. A stride of 0 gives every panel the same operand: Array::matmul's batch
 * broadcast (array/linalg.rs:212-225) without materialising the broadcast_to copy. */
nrk_status nrk_dev_batched_gemm_f64(size_t batch, size_t m, size_t k, size_t n, const double *a, size_t stride_a,
                                    const double *b, size_t stride_b, double *c, nrk_stream stream);
nrk_status nrk_dev_batched_gemm_f32(size_t batch, size_t m, size_t k, size_t n, const float *a, size_t stride_a,
                                    const float *b, size_t stride_b, float *c, nrk_stream stream);

/* Split-precision GEMMs. No reference counterpart: alternative engines for gemm_2d's f32 / f64
 * tier (kernels/gemm.rs:290-312) with the same C := A*B overwrite contract.
 *
 * FP32 as BF16 split products on v_mfma_f32_16x16x32_bf16 with FP32 accumulation:
 * products = 3 (BF16x3: a1b0 + a0b1 + a0b0), 6 (BF16x6, adds a2b0 + a1b1 + a0b2) or 1 (inputs
 * rounded to bf16). Inputs must be finite and below 2^127 in magnitude. */
nrk_status nrk_gemm_f32_bf16x(size_t m, size_t k, size_t n, const float *a, const float *b, float *c, int32_t products);
nrk_status nrk_dev_gemm_f32_bf16x(size_t m, size_t k, size_t n, const float *a, const float *b, float *c, int32_t products,
                                  nrk_stream stream);
/* FP64 by the Ozaki scheme II on v_mfma_i32_16x16x64_i8: `moduli` int8 residue products
 * (2..20; 0 selects 16) with exact INT32 accumulation and an exact CRT recombination. The only
 * rounding is of A's rows and B's columns to integers after scaling each by a power of two to a
 * 2-norm of at most X, X^2 < prod(moduli) / 2 (Cauchy-Schwarz keeps every product entry in the
 * CRT range; X <= 2^62). k <= 130000. A row/column holding Inf/NaN gives NaN outputs. */
nrk_status nrk_gemm_f64_ozaki(size_t m, size_t k, size_t n, const double *a, const double *b, double *c, int32_t moduli);
nrk_status nrk_dev_gemm_f64_ozaki(size_t m, size_t k, size_t n, const double *a, const double *b, double *c, int32_t moduli,
                                  nrk_stream stream);

/* ------------------------------------------------------------------------------------------ */
/* Ragged batches: bf16 GEMMs (fp32 accumulate/output) + f32 row sums in one persistent launch  */
/* ------------------------------------------------------------------------------------------ */
/* Generalises array/linalg.rs:56-136 (batched_gemm / par_batch_f64, uniform panels) to ragged
 * shapes. One problem: C (m x n, f32, row-major) := A (m x k, bf16 bits, row-major) *
 * B (k x n, bf16 bits, row-major). All pointers are device pointers. */
typedef struct nrk_gemm_desc {
    int32_t m, n, k, reserved;
    const void *a;  /* uint16 bf16 bits */
    const void *b;  /* uint16 bf16 bits */
    float *c;
} nrk_gemm_desc;
typedef struct nrk_grouped_plan nrk_grouped_plan;

/* Builds the work queue for `gemm_count` GEMMs (host array of device-pointer descriptors) and
 * `row_count` ragged row sums: rows_out[r] = sum(rows_x[row_offsets[r] .. row_offsets[r+1]))
 * (row_offsets: HOST array of row_count+1 offsets; rows_x, rows_out: device pointers; sums are
 * compensated and deterministic). item_cost_ns (host, gemm_count entries, may be NULL) is the
 * per-tile cost estimate used to order the queue largest-first; NULL uses a flop/byte model.
 * The plan keeps device copies of the descriptors and offsets. */
nrk_status nrk_grouped_plan_create(int32_t gemm_count, const nrk_gemm_desc *descs, int32_t row_count, const float *rows_x,
                                   const int64_t *row_offsets, float *rows_out, const float *item_cost_ns,
                                   nrk_grouped_plan **plan);
/* One persistent launch over the whole queue. Replayable (also from a captured graph): the
 * kernel resets its own queue counters. Runs of one plan must not overlap. */
nrk_status nrk_grouped_plan_run(const nrk_grouped_plan *plan, nrk_stream stream);
nrk_status nrk_grouped_plan_info(const nrk_grouped_plan *plan, int32_t *n_items, int32_t *n_gemm_tiles, int32_t *grid);
nrk_status nrk_grouped_plan_destroy(nrk_grouped_plan *plan);
/* One bf16 GEMM per launch (the unbatched per-call path the grouped launch is compared to). */
nrk_status nrk_dev_gemm_bf16(const nrk_gemm_desc *desc, nrk_stream stream);
/* The same batch without a plan: one nrk_dev_gemm_bf16 launch per GEMM, then one nrk_dev_sum_f32
 * launch per row, issued in a loop on `stream` (descs and row_offsets are HOST arrays). The
 * per-call baseline a caller without the grouped launch runs (no reference counterpart as one
 * call; the reference runs one task per panel, array/linalg.rs:56-136). */
nrk_status nrk_dev_ragged_per_call(int32_t gemm_count, const nrk_gemm_desc *descs, int32_t row_count, const float *rows_x,
                                   const int64_t *row_offsets, float *rows_out, nrk_stream stream);

/* ------------------------------------------------------------------------------------------ */
/* Autotuner and learned selector                                                             */
/* ------------------------------------------------------------------------------------------ */
/* No reference counterpart: the reference hard-codes its tiers (kernels/mod.rs:63-106,
 * kernels/gemm.rs:290-312). Operation families and their launch-configuration tables:
 *   NRK_DT_F64 / NRK_DT_F32 GEMM (d0=m, d1=n, d2=k): configs 0..4 = MFMA tiles 32x32, 64x64,
 *   128x64, 64x128, 128x128;  NRK_DT_BF16 NN GEMM (bf16 bits in, f32 out), same tiles;
 *   NRK_DT_ROWSUM_F32 row sums of a (d0 rows, d1 cols) matrix: configs 0..3 = 1, 2, 4, 8
 *   workgroups per CU (results identical across configs);
 *   NRK_DT_F32_BF16X3 FP32 GEMM as 3 BF16 products (d0=m, d1=n, d2=k): configs 0..4 = tiles
 *   (workgroup / wave) 64x64 / 32x32, 128x128 / 64x64, 128x64 / 64x32, 256x128 / 64x64,
 *   128x256 / 64x64 (results identical across configs);
 *   NRK_DT_F32_BF16X6 FP32 GEMM as 6 BF16 products: the same tile table (3 and 4 do not fit the
 *   LDS with six terms and run as 1);
 *   NRK_DT_F64_OZAKI FP64 GEMM by the Ozaki scheme II, 16 moduli: configs 0..4 = INT8 tiles
 *   (workgroup / wave) 64x64 / 32x32, 128x128 / 64x64, 128x256 / 64x64, 256x256 / 128x64,
 *   256x256 / 128x128 (results identical across configs). */
enum { NRK_DT_F64 = 0, NRK_DT_F32 = 1, NRK_DT_BF16 = 2, NRK_DT_ROWSUM_F32 = 3, NRK_DT_F32_BF16X3 = 4, NRK_DT_F64_OZAKI = 5,
       NRK_DT_F32_BF16X6 = 6 };
nrk_status nrk_config_count(int32_t family, int32_t *count);
/* Run one configuration (device pointers; c receives the result). */
nrk_status nrk_dev_run_config(int32_t family, int32_t config, int32_t d0, int32_t d1, int32_t d2, const void *a,
                              const void *b, void *c, nrk_stream stream);
/* Exhaustive search: times every configuration on every shape (one warm-up, then the mean of
 * `reps` launches, HIP events) on library-owned operands; times_ns is count x n_configs. */
nrk_status nrk_tune(int32_t family, int32_t count, const int32_t *d0, const int32_t *d1, const int32_t *d2, int32_t reps,
                    float *times_ns);

/* The selector on the dispatch path (no reference counterpart: it replaces the compile-time tiers
 * of kernels/mod.rs:63-106 and kernels/gemm.rs:290-312). Every call that runs a GEMM of these
 * families without a forced tile -- nrk_dev_gemm_f64/f32 (NRK_DT_F64/F32), nrk_dev_gemm_bf16
 * (NRK_DT_BF16), nrk_dev_gemm_f32_bf16x (NRK_DT_F32_BF16X3/X6), nrk_dev_gemm_f64_ozaki
 * (NRK_DT_F64_OZAKI) -- takes its tile from, in order: a recorded decision for the shape; the
 * family's cost model (the MLP above, evaluated on the host) when the shape lies within half a
 * binade of the shapes it was trained on, with the exhaustive timing of nrk_tune when its two best
 * predictions are closer than `margin` (log2 units; not while the stream is captured into a
 * graph); else the built-in rule. Decisions are recorded per (family, d0, d1, d2). Every
 * configuration of a family gives the same result bits.
 *
 * set_model: host arrays; params as nrk_mlp_param_count(K) floats (K = nrk_config_count), inputs
 * normalised as (feature - x_mean) / x_scale over the features log2 d0, log2 d1, log2 d2,
 * dtype_feature, 0, 0; predictions are log2(ns) - y_offset; domain_lo/hi[3] bound log2 d0..d2.
 * clear: the model, the recorded decisions and the counters of one family (-1: all).
 * table_put / table_size / table_get: the timed decisions, to persist and restore them.
 * choose: the configuration a dispatch of this shape takes now (-1: the built-in rule).
 * stats: dispatch decisions so far by source. */
nrk_status nrk_selector_set_model(int32_t family, float dtype_feature, const float *params, int32_t K, const float *x_mean,
                                  const float *x_scale, const float *domain_lo, const float *domain_hi, float y_offset,
                                  float margin);
nrk_status nrk_selector_clear(int32_t family);
nrk_status nrk_selector_table_put(int32_t family, int32_t d0, int32_t d1, int32_t d2, int32_t config);
nrk_status nrk_selector_table_size(int32_t *count);
nrk_status nrk_selector_table_get(int32_t index, int32_t *family, int32_t *d0, int32_t *d1, int32_t *d2, int32_t *config);
nrk_status nrk_selector_choose(int32_t family, int32_t d0, int32_t d1, int32_t d2, int32_t *config);
nrk_status nrk_selector_stats(int32_t family, int64_t *by_rule, int64_t *by_model, int64_t *by_timing, int64_t *by_table);
/* MLP cost model 6 -> 64 -> 64 -> K (ReLU, fp32), K <= 8. Parameters (device, contiguous):
 * W1[64][6] b1[64] W2[64][64] b2[64] W3[K][64] b3[K]; nrk_mlp_param_count(K) floats. */
size_t nrk_mlp_param_count(int32_t K);
/* y (N x K) = MLP(x (N x 6)); device pointers. */
nrk_status nrk_mlp_forward(const float *params, int32_t K, const float *x, int32_t N, float *y, nrk_stream stream);
/* One full-batch step on loss = sum(mask * (y - target)^2) / sum(mask): forward, backward,
 * weight gradients (written to grad_out if not NULL), then Adam (lr = 0 skips the update;
 * step >= 1 is the Adam step count). loss_out (device, may be NULL) receives
 * sum(mask * (y - target)^2). All pointers are device pointers except none. */
nrk_status nrk_mlp_train_step(float *params, float *adam_m, float *adam_v, int32_t step, int32_t K, const float *x,
                              const float *target, const float *mask, int32_t N, float lr, float *loss_out,
                              float *grad_out, nrk_stream stream);

/* hipGraph capture of work enqueued on a created (non-NULL) stream. */
nrk_status nrk_graph_begin_capture(nrk_stream stream);
nrk_status nrk_graph_end_capture(nrk_stream stream, void **graph_exec);
nrk_status nrk_graph_launch(void *graph_exec, nrk_stream stream);
nrk_status nrk_graph_destroy(void *graph_exec);

#ifdef __cplusplus
}
#endif
#endif /* NUMRS_AMD_H */
