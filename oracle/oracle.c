/*
 * oracle.c -- TEST INFRASTRUCTURE ONLY. CPU oracle for the numrs_amd parity tests and the CPU
 * baseline leg of bench.py. Nothing in the product (numrs_amd/) links, loads or calls this.
 *
 * Two kinds of functions live here:
 *
 *  1. ref_*   A plain-C restatement of the reference's CPU algorithms for the kernels path,
 *             keeping its fold orders, thresholds and fixed 8192-element chunking, so that the
 *             restatement reproduces the reference's rounding wherever the reference's own
 *             source fixes it. Each function cites the reference file:line it restates.
 *             Built with -ffp-contract=off: every a*b+c is two roundings, as in Rust.
 *
 *  2. dd_* / exact_*   High-precision references: double-double (TwoSum/TwoProd) accumulation and
 *             __int128 integer sums. These are the "FP64 oracle" every floating-point result is
 *             measured against (max relative error), and the exact answer for integers.
 *
 * Parity pinning (what checks each function; see tests/test_oracle.py):
 *   ref_sum/mean/var/std f64+f32: the reference's pinned bit patterns
 *       (tests/test_reduction_determinism.rs:126-197) and NumPy JSON fixtures
 *       (tests/py/array_operations_reference_data.json), plus closed forms.
 *       The bits of ref_simd_sum on generic data depend on scirs2-core 0.6.5's
 *       SimdUnifiedOps::simd_sum (not vendored): its lane order is "parity unpinned"; this file
 *       uses a documented 4-accumulator stand-in (ref_simd_sum below).
 *   ref_min/max: closed form (order-independent) + reduce.rs:843-925 NaN/inf cases.
 *   ref_argmax/argmin: closed form (first strict extremum) + statistics/mod.rs:65-77 examples.
 *   ref_axis_*: NumPy JSON *_axis0/1 fixtures + brute force.
 *   ref_sum_i64: exact_sum_i64 (__int128).
 *   ref_gemm_generic: integer GEMM bit-exact vs brute force; f64 3x3 known answer
 *       (tests/test_linalg_reference.rs:163-196).
 *   dd_*: closed forms (sum of an arithmetic series, Faulhaber) and exact small cases.
 */
#include <math.h>
#include <pthread.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#define PARALLEL_MIN_LEN 10000      /* kernels/mod.rs:74  */
#define PARALLEL_CHUNK 8192         /* kernels/mod.rs:92  */
#define MINMAX_PARALLEL_MIN_LEN 80000 /* kernels/reduce.rs:200 */
#define VAR_PARALLEL_MIN_LEN 40000  /* kernels/reduce.rs:202 */

/* ------------------------------------------------------------------------------------------ */
/* thread pool for the chunked tiers: chunk boundaries and the fold order are fixed, only the */
/* assignment of chunks to threads varies, exactly as in the reference (kernels/mod.rs:76-92)  */
/* ------------------------------------------------------------------------------------------ */
static int g_threads = 1;
void oracle_set_threads(int n) { g_threads = n < 1 ? 1 : n; }
int oracle_get_threads(void) { return g_threads; }

typedef void (*chunk_fn)(const void *ctx, size_t chunk, void *out);
typedef struct {
    const void *ctx;
    size_t nchunks, first, step;
    chunk_fn fn;
    char *outs;
    size_t out_size;
} worker_arg;

static void *worker(void *p) {
    worker_arg *w = (worker_arg *)p;
    for (size_t c = w->first; c < w->nchunks; c += w->step) w->fn(w->ctx, c, w->outs + c * w->out_size);
    return NULL;
}

static void par_chunks(const void *ctx, size_t nchunks, chunk_fn fn, void *outs, size_t out_size) {
    int nt = g_threads;
    if (nt <= 1 || nchunks < 2) {
        for (size_t c = 0; c < nchunks; ++c) fn(ctx, c, (char *)outs + c * out_size);
        return;
    }
    if ((size_t)nt > nchunks) nt = (int)nchunks;
    pthread_t th[256];
    worker_arg args[256];
    if (nt > 256) nt = 256;
    for (int i = 0; i < nt; ++i) {
        args[i] = (worker_arg){ctx, nchunks, (size_t)i, (size_t)nt, fn, (char *)outs, out_size};
        pthread_create(&th[i], NULL, worker, &args[i]);
    }
    for (int i = 0; i < nt; ++i) pthread_join(th[i], NULL);
}

/* ------------------------------------------------------------------------------------------ */
/* simd_sum stand-in                                                                           */
/* ------------------------------------------------------------------------------------------ */
/* kernels/reduce.rs:212,219 call scirs2_core's SimdUnifiedOps::simd_sum, whose lane order is
 * not in the reference tree. Stand-in: 4 interleaved accumulators (element i goes to
 * accumulator i % 4) over the largest multiple of 4, combined ((a0 + a1) + (a2 + a3)), then the
 * remaining 0..3 elements added left to right. Inputs whose partial sums are exact (the pinned
 * determinism data) give the same bits under any order. */
#define DEFINE_SIMD_SUM(T, NAME)                                     \
    static T NAME(const T *a, size_t n) {                            \
        T acc[4] = {0, 0, 0, 0};                                     \
        size_t n4 = n & ~(size_t)3;                                  \
        for (size_t i = 0; i < n4; i += 4) {                         \
            acc[0] = acc[0] + a[i];                                  \
            acc[1] = acc[1] + a[i + 1];                              \
            acc[2] = acc[2] + a[i + 2];                              \
            acc[3] = acc[3] + a[i + 3];                              \
        }                                                            \
        T s = (acc[0] + acc[1]) + (acc[2] + acc[3]);                 \
        for (size_t i = n4; i < n; ++i) s = s + a[i];                \
        return s;                                                    \
    }
DEFINE_SIMD_SUM(double, simd_sum_f64)
DEFINE_SIMD_SUM(float, simd_sum_f32)

/* ------------------------------------------------------------------------------------------ */
/* reduction restatements: kernels/reduce.rs:204-444 (macro define_reduce_kernels!)            */
/* ------------------------------------------------------------------------------------------ */
#define DEFINE_REDUCE(T, SUF, INF)                                                            \
    typedef struct { const T *a; size_t n; T mean; } ctx_##SUF;                              \
    static void chunk_sum_##SUF(const void *c, size_t k, void *out) {                        \
        const ctx_##SUF *x = (const ctx_##SUF *)c;                                            \
        size_t lo = k * PARALLEL_CHUNK, hi = lo + PARALLEL_CHUNK;                             \
        if (hi > x->n) hi = x->n;                                                             \
        *(T *)out = simd_sum_##SUF(x->a + lo, hi - lo);                                       \
    }                                                                                         \
    static void chunk_ssd_##SUF(const void *c, size_t k, void *out) {                        \
        const ctx_##SUF *x = (const ctx_##SUF *)c;                                            \
        size_t lo = k * PARALLEL_CHUNK, hi = lo + PARALLEL_CHUNK;                             \
        if (hi > x->n) hi = x->n;                                                             \
        T s = 0;                                                                              \
        for (size_t i = lo; i < hi; ++i) {                                                    \
            T d = x->a[i] - x->mean;                                                          \
            s = s + d * d;                                                                    \
        }                                                                                     \
        *(T *)out = s;                                                                        \
    }                                                                                         \
    /* reduce.rs:207-224 */                                                                   \
    T ref_sum_##SUF(const T *a, size_t n) {                                                   \
        if (n == 0) return 0;                                                                 \
        if (n < PARALLEL_MIN_LEN) return simd_sum_##SUF(a, n);                                \
        size_t nc = (n + PARALLEL_CHUNK - 1) / PARALLEL_CHUNK;                                \
        T *parts = (T *)malloc(nc * sizeof(T));                                               \
        ctx_##SUF x = {a, n, 0};                                                              \
        par_chunks(&x, nc, chunk_sum_##SUF, parts, sizeof(T));                                \
        T s = 0;                                                                              \
        for (size_t k = 0; k < nc; ++k) s = s + parts[k];                                     \
        free(parts);                                                                          \
        return s;                                                                             \
    }                                                                                         \
    /* reduce.rs:227-232 */                                                                   \
    T ref_mean_##SUF(const T *a, size_t n) {                                                  \
        if (n == 0) return 0;                                                                 \
        return ref_sum_##SUF(a, n) / (T)n;                                                    \
    }                                                                                         \
    /* reduce.rs:239-260 */                                                                   \
    T ref_sum_sq_dev_##SUF(const T *a, size_t n, T mean) {                                    \
        ctx_##SUF x = {a, n, mean};                                                           \
        if (n < PARALLEL_MIN_LEN) {                                                           \
            T s;                                                                              \
            chunk_ssd_##SUF(&(ctx_##SUF){a, n, mean}, 0, &s);                                 \
            return n == 0 ? 0 : s;                                                            \
        }                                                                                     \
        size_t nc = (n + PARALLEL_CHUNK - 1) / PARALLEL_CHUNK;                                \
        T *parts = (T *)malloc(nc * sizeof(T));                                               \
        par_chunks(&x, nc, chunk_ssd_##SUF, parts, sizeof(T));                                \
        T s = 0;                                                                              \
        for (size_t k = 0; k < nc; ++k) s = s + parts[k];                                     \
        free(parts);                                                                          \
        return s;                                                                             \
    }                                                                                         \
    /* reduce.rs:284-314: one tier decision at VAR_PARALLEL_MIN_LEN for both passes */        \
    T ref_var_##SUF(const T *a, size_t n, size_t ddof) {                                      \
        if (n == 0) return 0;                                                                 \
        T divisor = (T)n - (T)ddof;                                                           \
        if (n < VAR_PARALLEL_MIN_LEN) {                                                       \
            T mean = simd_sum_##SUF(a, n) / (T)n;                                             \
            T s = 0;                                                                          \
            for (size_t i = 0; i < n; ++i) {                                                  \
                T d = a[i] - mean;                                                            \
                s = s + d * d;                                                                \
            }                                                                                 \
            return s / divisor;                                                               \
        }                                                                                     \
        size_t nc = (n + PARALLEL_CHUNK - 1) / PARALLEL_CHUNK;                                \
        T *parts = (T *)malloc(nc * sizeof(T));                                               \
        ctx_##SUF x = {a, n, 0};                                                              \
        par_chunks(&x, nc, chunk_sum_##SUF, parts, sizeof(T));                                \
        T s = 0;                                                                              \
        for (size_t k = 0; k < nc; ++k) s = s + parts[k];                                     \
        x.mean = s / (T)n;                                                                    \
        par_chunks(&x, nc, chunk_ssd_##SUF, parts, sizeof(T));                                \
        T q = 0;                                                                              \
        for (size_t k = 0; k < nc; ++k) q = q + parts[k];                                     \
        free(parts);                                                                          \
        return q / divisor;                                                                   \
    }                                                                                         \
    typedef struct { T acc; int nan; } mm_##SUF;                                              \
    static void chunk_max_##SUF(const void *c, size_t k, void *out) {                        \
        const ctx_##SUF *x = (const ctx_##SUF *)c;                                            \
        size_t lo = k * PARALLEL_CHUNK, hi = lo + PARALLEL_CHUNK;                             \
        if (hi > x->n) hi = x->n;                                                             \
        mm_##SUF r = {-INF, 0};                                                               \
        for (size_t i = lo; i < hi; ++i) {                                                    \
            r.nan |= isnan(x->a[i]);                                                          \
            if (x->a[i] > r.acc) r.acc = x->a[i];                                             \
        }                                                                                     \
        *(mm_##SUF *)out = r;                                                                 \
    }                                                                                         \
    static void chunk_min_##SUF(const void *c, size_t k, void *out) {                        \
        const ctx_##SUF *x = (const ctx_##SUF *)c;                                            \
        size_t lo = k * PARALLEL_CHUNK, hi = lo + PARALLEL_CHUNK;                             \
        if (hi > x->n) hi = x->n;                                                             \
        mm_##SUF r = {INF, 0};                                                                \
        for (size_t i = lo; i < hi; ++i) {                                                    \
            r.nan |= isnan(x->a[i]);                                                          \
            if (x->a[i] < r.acc) r.acc = x->a[i];                                             \
        }                                                                                     \
        *(mm_##SUF *)out = r;                                                                 \
    }                                                                                         \
    /* reduce.rs:376-412 */                                                                   \
    T ref_max_##SUF(const T *a, size_t n) {                                                   \
        if (n == 0) return 0;                                                                 \
        ctx_##SUF x = {a, n, 0};                                                              \
        mm_##SUF r;                                                                           \
        if (n < MINMAX_PARALLEL_MIN_LEN) {                                                    \
            ctx_##SUF one = {a, n, 0};                                                        \
            r.acc = -INF; r.nan = 0;                                                          \
            for (size_t i = 0; i < n; ++i) {                                                  \
                r.nan |= isnan(one.a[i]);                                                     \
                if (one.a[i] > r.acc) r.acc = one.a[i];                                       \
            }                                                                                 \
        } else {                                                                              \
            size_t nc = (n + PARALLEL_CHUNK - 1) / PARALLEL_CHUNK;                            \
            mm_##SUF *parts = (mm_##SUF *)malloc(nc * sizeof(mm_##SUF));                      \
            par_chunks(&x, nc, chunk_max_##SUF, parts, sizeof(mm_##SUF));                     \
            r.acc = -INF; r.nan = 0;                                                          \
            for (size_t k = 0; k < nc; ++k) {                                                 \
                if (parts[k].acc > r.acc) r.acc = parts[k].acc;                               \
                r.nan |= parts[k].nan;                                                        \
            }                                                                                 \
            free(parts);                                                                      \
        }                                                                                     \
        return r.nan ? (T)NAN : r.acc;                                                        \
    }                                                                                         \
    /* reduce.rs:323-372 */                                                                   \
    T ref_min_##SUF(const T *a, size_t n) {                                                   \
        if (n == 0) return 0;                                                                 \
        ctx_##SUF x = {a, n, 0};                                                              \
        mm_##SUF r;                                                                           \
        if (n < MINMAX_PARALLEL_MIN_LEN) {                                                    \
            r.acc = INF; r.nan = 0;                                                           \
            for (size_t i = 0; i < n; ++i) {                                                  \
                r.nan |= isnan(a[i]);                                                         \
                if (a[i] < r.acc) r.acc = a[i];                                               \
            }                                                                                 \
        } else {                                                                              \
            size_t nc = (n + PARALLEL_CHUNK - 1) / PARALLEL_CHUNK;                            \
            mm_##SUF *parts = (mm_##SUF *)malloc(nc * sizeof(mm_##SUF));                      \
            par_chunks(&x, nc, chunk_min_##SUF, parts, sizeof(mm_##SUF));                     \
            r.acc = INF; r.nan = 0;                                                           \
            for (size_t k = 0; k < nc; ++k) {                                                 \
                if (parts[k].acc < r.acc) r.acc = parts[k].acc;                               \
                r.nan |= parts[k].nan;                                                        \
            }                                                                                 \
            free(parts);                                                                      \
        }                                                                                     \
        return r.nan ? (T)NAN : r.acc;                                                        \
    }                                                                                         \
    /* reduce.rs:419-421: strictly left to right */                                           \
    T ref_prod_##SUF(const T *a, size_t n) {                                                  \
        T p = 1;                                                                              \
        for (size_t i = 0; i < n; ++i) p = p * a[i];                                          \
        return p;                                                                             \
    }                                                                                         \
    /* math/statistics/mod.rs:90-103: first strict maximum; NaN never wins; NaN at 0 pins 0 */ \
    uint64_t ref_argmax_##SUF(const T *a, size_t n) {                                         \
        uint64_t idx = 0;                                                                     \
        T best = a[0];                                                                        \
        for (size_t i = 1; i < n; ++i)                                                        \
            if (a[i] > best) { best = a[i]; idx = i; }                                        \
        return idx;                                                                           \
    }                                                                                         \
    /* math/statistics/mod.rs:229-242 */                                                      \
    uint64_t ref_argmin_##SUF(const T *a, size_t n) {                                         \
        uint64_t idx = 0;                                                                     \
        T best = a[0];                                                                        \
        for (size_t i = 1; i < n; ++i)                                                        \
            if (a[i] < best) { best = a[i]; idx = i; }                                        \
        return idx;                                                                           \
    }                                                                                         \
    /* Axis folds over the (outer, len, inner) view, output o*inner + i.                      \
     * op 0 sum  math/aggregation.rs:548-561  left-to-right from zero                         \
     * op 1 mean math/statistics/mod.rs:1064-1071  left-to-right sum / len                    \
     * op 2 var  math/statistics/mod.rs:1403-1494  mean(axis), then sum((x-mean)^2)/(len-ddof) \
     * op 3 min / op 4 max  math/aggregation.rs:244-261  NaN short-circuits                   \
     * op 5 argmin / op 6 argmax  math/statistics/mod.rs:164-175 / :303-314                   \
     * out is T* for ops 0-4 and uint64_t* for ops 5-6 */                                     \
    void ref_axis_##SUF(int op, const T *x, size_t outer, size_t len, size_t inner, size_t ddof, \
                        void *out) {                                                          \
        for (size_t o = 0; o < outer; ++o) {                                                  \
            for (size_t i = 0; i < inner; ++i) {                                              \
                const T *p = x + o * len * inner + i;                                         \
                size_t at = o * inner + i;                                                    \
                if (op <= 2) {                                                                \
                    T s = 0;                                                                  \
                    for (size_t j = 0; j < len; ++j) s = s + p[j * inner];                    \
                    if (op == 0) { ((T *)out)[at] = s; continue; }                            \
                    T mean = s / (T)len;                                                      \
                    if (op == 1) { ((T *)out)[at] = mean; continue; }                         \
                    T q = 0;                                                                  \
                    for (size_t j = 0; j < len; ++j) {                                        \
                        T d = p[j * inner] - mean;                                            \
                        q = q + d * d;                                                        \
                    }                                                                         \
                    ((T *)out)[at] = q / (T)(len - ddof);                                     \
                } else if (op == 3 || op == 4) {                                              \
                    T acc = p[0];                                                             \
                    if (!isnan(acc)) {                                                        \
                        for (size_t j = 1; j < len; ++j) {                                    \
                            T v = p[j * inner];                                               \
                            if (isnan(v)) { acc = v; break; }                                 \
                            if (op == 4 ? (v > acc) : (v < acc)) acc = v;                     \
                        }                                                                     \
                    }                                                                         \
                    ((T *)out)[at] = acc;                                                     \
                } else {                                                                      \
                    T best = p[0];                                                            \
                    uint64_t bi = 0;                                                          \
                    for (size_t j = 1; j < len; ++j) {                                        \
                        T v = p[j * inner];                                                   \
                        if (op == 6 ? (v > best) : (v < best)) { best = v; bi = j; }          \
                    }                                                                         \
                    ((uint64_t *)out)[at] = bi;                                               \
                }                                                                             \
            }                                                                                 \
        }                                                                                     \
    }

DEFINE_REDUCE(double, f64, INFINITY)
DEFINE_REDUCE(float, f32, INFINITY)

/* array/operations.rs:303-318: Array::<i64>::sum is op.iter().fold(0, +); release builds wrap
 * (Cargo.toml:244 overflow-checks off), i.e. arithmetic modulo 2^64. */
int64_t ref_sum_i64(const int64_t *a, size_t n) {
    uint64_t s = 0;
    for (size_t i = 0; i < n; ++i) s += (uint64_t)a[i];
    return (int64_t)s;
}

/* ------------------------------------------------------------------------------------------ */
/* elementwise restatements: op codes match include/numrs_amd.h NRK_SOP_* / NRK_BOP_*          */
/* ------------------------------------------------------------------------------------------ */
#define DEFINE_EW(T, SUF, SQRT, FMOD, FABS, COPYSIGN)                                               \
    static T sop_##SUF(int op, T s, T x) {                                                  \
        switch (op) {                                                                       \
            case 0: return x + s;  /* Array::add_scalar, operations.rs:448 */              \
            case 1: return x - s;  /* operations.rs:455 */                                 \
            case 2: return x * s;  /* operations.rs:462 */                                 \
            case 3: return x / s;  /* operations.rs:469 */                                 \
            case 4: return s - x;                                                           \
            case 5: return s / x;                                                           \
            case 6: return -x;     /* arithmetic.rs:448-468 */                              \
            case 7: return FABS(x);                                                         \
            case 8: return x * x;                                                           \
            case 9: return SQRT(x);                                                         \
            default: return x;                                                              \
        }                                                                                   \
    }                                                                                       \
    /* each chain step is one separate map pass (operations.rs:182-200): one rounding each */ \
    void ref_map_##SUF(const int32_t *ops, const double *scalars, int len, const T *x, T *y, \
                       size_t n) {                                                          \
        for (size_t i = 0; i < n; ++i) {                                                    \
            T v = x[i];                                                                     \
            for (int k = 0; k < len; ++k) v = sop_##SUF(ops[k], (T)scalars[k], v);         \
            y[i] = v;                                                                       \
        }                                                                                   \
    }                                                                                       \
    /* elementwise.rs:51-66 binary_serial with arithmetic.rs / ufuncs.rs:169-173 closures */ \
    void ref_zip_##SUF(int op, const T *a, const T *b, T *y, size_t n) {                   \
        for (size_t i = 0; i < n; ++i) {                                                    \
            T p = a[i], q = b[i];                                                           \
            switch (op) {                                                                   \
                case 0: y[i] = p + q; break;                                                \
                case 1: y[i] = p - q; break;                                                \
                case 2: y[i] = p * q; break;                                                \
                case 3: y[i] = p / q; break;                                                \
                case 4: y[i] = FMOD(p, q); break;                                           \
                case 5: y[i] = p > q ? p : q; break;                                        \
                case 6: y[i] = p < q ? p : q; break;                                        \
                case 7: y[i] = COPYSIGN(p, q); break; /* ufuncs.rs:801 m.copysign(s) */      \
                case 11: y[i] = p > 0 ? p : q * p; break; /* ufuncs.rs:1405 prelu */          \
                default: y[i] = p;                                                          \
            }                                                                               \
        }                                                                                   \
    }

/* Integer restatement of binary_serial / unary_dispatch with arithmetic.rs's and bitwise_ops.rs's
 * operators under Rust's release arithmetic: + - * and negation wrap (formed in the unsigned type
 * of the same width); / and % truncate toward zero (C99 does too); a shift amount is masked to
 * the width; >> is arithmetic for the signed types. The callers never pass a zero divisor or
 * MIN / -1 (Rust panics on both). */
#define DEFINE_INT_ELEM(T, U, SUF, BITS)                                                    \
    static T iop_##SUF(int op, T p, T q) {                                                  \
        switch (op) {                                                                       \
            case 0: return (T)((U)p + (U)q);                                                \
            case 1: return (T)((U)p - (U)q);                                                \
            case 2: return (T)((U)p * (U)q);                                                \
            case 3: return p / q;                                                           \
            case 4: return p % q;                                                           \
            case 5: return p > q ? p : q;                                                   \
            case 6: return p < q ? p : q;                                                   \
            case 13: return p & q;                                                          \
            case 14: return p | q;                                                          \
            case 15: return p ^ q;                                                          \
            case 16: return (T)((U)p << ((U)q & (BITS - 1)));                               \
            case 17: return (T)(p >> ((U)q & (BITS - 1)));                                  \
            default: return p;                                                              \
        }                                                                                   \
    }                                                                                       \
    void ref_zip_int_##SUF(int op, const T *a, const T *b, T *y, size_t n) {               \
        for (size_t i = 0; i < n; ++i) y[i] = iop_##SUF(op, a[i], b[i]);                    \
    }                                                                                       \
    void ref_map_int_##SUF(const int32_t *ops, const int64_t *scalars, int len, const T *x, T *y, \
                           size_t n) {                                                      \
        for (size_t i = 0; i < n; ++i) {                                                    \
            T v = x[i];                                                                     \
            for (int k = 0; k < len; ++k) {                                                 \
                const T s = (T)scalars[k];                                                  \
                switch (ops[k]) {                                                           \
                    case 0: v = iop_##SUF(0, v, s); break;                                  \
                    case 1: v = iop_##SUF(1, v, s); break;                                  \
                    case 2: v = iop_##SUF(2, v, s); break;                                  \
                    case 3: v = v / s; break;                                               \
                    case 4: v = iop_##SUF(1, s, v); break;                                  \
                    case 5: v = s / v; break;                                               \
                    case 6: v = (T)(0 - (U)v); break;                                       \
                    case 7: v = v < 0 ? (T)(0 - (U)v) : v; break;                           \
                    case 8: v = iop_##SUF(2, v, v); break;                                  \
                    case 15: v = (T)(v > 0 ? 1 : (v == 0 ? 0 : -1)); break;                 \
                    case 16: v = v < s ? s : v; break;                                      \
                    case 17: v = v > s ? s : v; break;                                      \
                    default: break;                                                         \
                }                                                                           \
            }                                                                               \
            y[i] = v;                                                                       \
        }                                                                                   \
    }
DEFINE_INT_ELEM(int64_t, uint64_t, i64, 64)
DEFINE_INT_ELEM(int32_t, uint32_t, i32, 32)
DEFINE_INT_ELEM(uint64_t, uint64_t, u64, 64)
DEFINE_INT_ELEM(uint32_t, uint32_t, u32, 32)

/* expr/fused_eval.rs:605-658 eval_eager: the tree evaluated node by node, each node one rounded
 * operation (Binary -> the broadcast arithmetic, ScalarRhs / ScalarLhs -> the scalar ops, Unary ->
 * neg / abs / sqrt, Fma -> multiply then add), as a postfix program of (kind, op, scalar) nodes
 * (kinds 0 leaf, 1 step, 2 binary, 3 fma; ops the library's NRK_SOP_* / NRK_BOP_* codes). IEEE
 * operations only (the elementary functions are not restated). One element at a time. */
#define DEFINE_EXPR(T, SUF, SQRT, FABS, FMOD, COPYSIGN)                                      \
    void ref_expr_##SUF(const int32_t *kind, const int32_t *op, const double *sc, int len,   \
                        const T *const *leaves, T *y, size_t n) {                             \
        for (size_t i = 0; i < n; ++i) {                                                    \
            T st[16];                                                                       \
            int sp = 0;                                                                     \
            for (int k = 0; k < len; ++k) {                                                 \
                if (kind[k] == 0) { st[sp++] = leaves[op[k]][i]; continue; }                \
                if (kind[k] == 1) {                                                         \
                    T x = st[sp - 1], s = (T)sc[k];                                         \
                    switch (op[k]) {                                                        \
                        case 0: x = x + s; break;                                           \
                        case 1: x = x - s; break;                                           \
                        case 2: x = x * s; break;                                           \
                        case 3: x = x / s; break;                                           \
                        case 4: x = s - x; break;                                           \
                        case 5: x = s / x; break;                                           \
                        case 6: x = -x; break;                                              \
                        case 7: x = FABS(x); break;                                         \
                        case 9: x = SQRT(x); break;                                         \
                        default: break;                                                     \
                    }                                                                       \
                    st[sp - 1] = x;                                                         \
                    continue;                                                               \
                }                                                                           \
                if (kind[k] == 2) {                                                         \
                    T r = st[--sp], l = st[sp - 1], v;                                      \
                    switch (op[k]) {                                                        \
                        case 0: v = l + r; break;                                           \
                        case 1: v = l - r; break;                                           \
                        case 2: v = l * r; break;                                           \
                        case 3: v = l / r; break;                                           \
                        case 4: v = FMOD(l, r); break;                                      \
                        case 5: v = l > r ? l : r; break;                                   \
                        case 6: v = l < r ? l : r; break;                                   \
                        case 7: v = COPYSIGN(l, r); break;                                  \
                        default: v = l > 0 ? l : r * l; break; /* 11: prelu */              \
                    }                                                                       \
                    st[sp - 1] = v;                                                         \
                    continue;                                                               \
                }                                                                           \
                { T c = st[--sp], b = st[--sp], a = st[sp - 1]; T p = a * b; st[sp - 1] = p + c; } \
            }                                                                               \
            y[i] = st[0];                                                                   \
        }                                                                                   \
    }
DEFINE_EXPR(double, f64, sqrt, fabs, fmod, copysign)
DEFINE_EXPR(float, f32, sqrtf, fabsf, fmodf, copysignf)

/* ufuncs.rs:864-891 fma: a.mul_add(b, c), one rounding (C's fma is the IEEE fusedMultiplyAdd);
 * ufuncs.rs:893-912 lerp: a + t * (b - a), three roundings */
#define DEFINE_TRI(T, SUF, FMA)                                                             \
    void ref_fma_##SUF(const T *a, const T *b, const T *c, T *y, size_t n) {                \
        for (size_t i = 0; i < n; ++i) y[i] = FMA(a[i], b[i], c[i]);                        \
    }                                                                                       \
    void ref_lerp_##SUF(const T *a, const T *b, T t, T *y, size_t n) {                      \
        for (size_t i = 0; i < n; ++i) y[i] = a[i] + t * (b[i] - a[i]);                     \
    }
DEFINE_TRI(double, f64, fma)
DEFINE_TRI(float, f32, fmaf)

DEFINE_EW(double, f64, sqrt, fmod, fabs, copysign)
DEFINE_EW(float, f32, sqrtf, fmodf, fabsf, copysignf)

/* ------------------------------------------------------------------------------------------ */
/* GEMM restatement                                                                           */
/* ------------------------------------------------------------------------------------------ */
/* kernels/gemm.rs:408-445 gemm_generic: zero C, then 64-blocked i-k-j, accumulating in T.
 * For every C[i][j] the k order is 0..K-1 ascending, one rounding per multiply and add. */
#define DEFINE_GEMM_GENERIC(T, SUF)                                                          \
    void ref_gemm_generic_##SUF(size_t m, size_t k, size_t n, const T *a, const T *b, T *c) { \
        for (size_t i = 0; i < m * n; ++i) c[i] = 0;                                         \
        const size_t BS = 64;                                                                \
        for (size_t ib = 0; ib < m; ib += BS) {                                              \
            size_t ie = ib + BS < m ? ib + BS : m;                                           \
            for (size_t kb = 0; kb < k; kb += BS) {                                          \
                size_t ke = kb + BS < k ? kb + BS : k;                                       \
                for (size_t jb = 0; jb < n; jb += BS) {                                      \
                    size_t je = jb + BS < n ? jb + BS : n;                                   \
                    for (size_t i = ib; i < ie; ++i)                                         \
                        for (size_t kk = kb; kk < ke; ++kk) {                                \
                            T aik = a[i * k + kk];                                           \
                            for (size_t j = jb; j < je; ++j) c[i * n + j] = c[i * n + j] + aik * b[kk * n + j]; \
                        }                                                                    \
                }                                                                            \
            }                                                                                \
        }                                                                                    \
    }
DEFINE_GEMM_GENERIC(double, f64)
DEFINE_GEMM_GENERIC(float, f32)
/* Integers: Rust's release arithmetic wraps, so the restatement forms every product and sum in the
 * unsigned type of the same width (defined wrap-around in C), with the same loop order. */
#define DEFINE_GEMM_GENERIC_INT(T, U, SUF)                                                   \
    void ref_gemm_generic_##SUF(size_t m, size_t k, size_t n, const T *a, const T *b, T *c) { \
        U *cu = (U *)c;                                                                      \
        for (size_t i = 0; i < m * n; ++i) cu[i] = 0;                                        \
        const size_t BS = 64;                                                                \
        for (size_t ib = 0; ib < m; ib += BS) {                                              \
            size_t ie = ib + BS < m ? ib + BS : m;                                           \
            for (size_t kb = 0; kb < k; kb += BS) {                                          \
                size_t ke = kb + BS < k ? kb + BS : k;                                       \
                for (size_t jb = 0; jb < n; jb += BS) {                                      \
                    size_t je = jb + BS < n ? jb + BS : n;                                   \
                    for (size_t i = ib; i < ie; ++i)                                         \
                        for (size_t kk = kb; kk < ke; ++kk) {                                \
                            U aik = (U)a[i * k + kk];                                        \
                            for (size_t j = jb; j < je; ++j)                                 \
                                cu[i * n + j] = (U)(cu[i * n + j] + (U)(aik * (U)b[kk * n + j])); \
                        }                                                                    \
                }                                                                            \
            }                                                                                \
        }                                                                                    \
    }
DEFINE_GEMM_GENERIC_INT(int64_t, uint64_t, i64)
DEFINE_GEMM_GENERIC_INT(int32_t, uint32_t, i32)

/* CPU-baseline stand-in for the reference's f64/f32 tier (gemm.rs:290-381): rows split into
 * ceil(m / threads) chunks across the pool, K never split; each chunk runs a packed, cache-
 * blocked micro-kernel standing in for matrixmultiply's (third-party, not vendored). Used only
 * for timing: its rounding is not claimed to equal matrixmultiply's. */
#define MR 4
#define NR 8
#define KC 256
#define NC 512
typedef struct {
    size_t m, k, n, row0, rows;
    const double *a, *b;
    double *c;
} gemm_task;

static void gemm_block_f64(size_t rows, size_t k, size_t n, const double *a, const double *b, double *c) {
    double *bp = (double *)aligned_alloc(64, sizeof(double) * KC * NC);
    for (size_t i = 0; i < rows * n; ++i) c[i] = 0.0;
    for (size_t jc = 0; jc < n; jc += NC) {
        size_t nc = jc + NC < n ? NC : n - jc;
        for (size_t pc = 0; pc < k; pc += KC) {
            size_t kc = pc + KC < k ? KC : k - pc;
            for (size_t p = 0; p < kc; ++p) memcpy(bp + p * nc, b + (pc + p) * n + jc, nc * sizeof(double));
            for (size_t i = 0; i < rows; i += MR) {
                size_t mr = i + MR < rows ? MR : rows - i;
                for (size_t j = 0; j < nc; j += NR) {
                    size_t nr = j + NR < nc ? NR : nc - j;
                    double acc[MR][NR] = {{0}};
                    for (size_t p = 0; p < kc; ++p) {
                        const double *brow = bp + p * nc + j;
                        for (size_t ii = 0; ii < mr; ++ii) {
                            double av = a[(i + ii) * k + pc + p];
                            for (size_t jj = 0; jj < nr; ++jj) acc[ii][jj] += av * brow[jj];
                        }
                    }
                    for (size_t ii = 0; ii < mr; ++ii)
                        for (size_t jj = 0; jj < nr; ++jj) c[(i + ii) * n + jc + j + jj] += acc[ii][jj];
                }
            }
        }
    }
    free(bp);
}

static void gemm_chunk(const void *ctx, size_t chunk, void *out) {
    (void)out;
    const gemm_task *t = (const gemm_task *)ctx;
    size_t r0 = chunk * t->rows;
    if (r0 >= t->m) return;
    size_t rows = r0 + t->rows <= t->m ? t->rows : t->m - r0;
    gemm_block_f64(rows, t->k, t->n, t->a + r0 * t->k, t->b, t->c + r0 * t->n);
}

void ref_gemm_rowsplit_f64(size_t m, size_t k, size_t n, const double *a, const double *b, double *c) {
    if (m == 0 || n == 0) return;
    if (k == 0) {
        for (size_t i = 0; i < m * n; ++i) c[i] = 0;
        return;
    }
    size_t threads = (size_t)g_threads;
    size_t rows = (m + threads - 1) / threads;  /* gemm.rs:368 chunk_rows = ceil(m / threads) */
    if (rows == 0) rows = 1;
    size_t nchunks = (m + rows - 1) / rows;
    gemm_task t = {m, k, n, 0, rows, a, b, c};
    char dummy[1024];
    par_chunks(&t, nchunks, gemm_chunk, dummy, 0);
}

/* ------------------------------------------------------------------------------------------ */
/* high-precision references                                                                  */
/* ------------------------------------------------------------------------------------------ */
typedef struct { double hi, lo; } dd;

static inline dd two_sum(double a, double b) {
    double s = a + b, v = s - a;
    dd r = {s, (a - (s - v)) + (b - v)};
    return r;
}
static inline dd dd_add_d(dd x, double y) {
    dd s = two_sum(x.hi, y);
    s.lo += x.lo;
    return two_sum(s.hi, s.lo);
}
static inline dd dd_add(dd x, dd y) {
    dd s = two_sum(x.hi, y.hi);
    s.lo += x.lo + y.lo;
    return two_sum(s.hi, s.lo);
}
static inline dd two_prod(double a, double b) {
    double p = a * b;
    dd r = {p, fma(a, b, -p)};
    return r;
}
static inline dd dd_mul_d(dd x, double y) {
    dd p = two_prod(x.hi, y);
    p.lo += x.lo * y;
    return two_sum(p.hi, p.lo);
}
static inline dd dd_sub(dd x, dd y) {
    dd ny = {-y.hi, -y.lo};
    return dd_add(x, ny);
}
static inline dd dd_div_d(dd x, double y) {
    double q1 = x.hi / y;
    dd r = dd_sub(x, two_prod(q1, y));
    double q2 = (r.hi + r.lo) / y;
    return two_sum(q1, q2);
}
static inline dd dd_mul(dd x, dd y) {
    dd p = two_prod(x.hi, y.hi);
    p.lo += x.hi * y.lo + x.lo * y.hi;
    return two_sum(p.hi, p.lo);
}

/* Sums, mean and variance (two-pass, ddof divisor) of x[0..n) in double-double. The values
 * are read as double; f32 inputs are widened exactly by the caller. out: [sum_hi, sum_lo,
 * mean, var(ddof)] (mean and var rounded to double). */
void dd_stats_f64(const double *x, size_t n, size_t ddof, double *out) {
    dd s = {0, 0};
    for (size_t i = 0; i < n; ++i) s = dd_add_d(s, x[i]);
    out[0] = s.hi;
    out[1] = s.lo;
    if (n == 0) {
        out[2] = out[3] = 0;
        return;
    }
    dd mean = dd_div_d(s, (double)n);
    dd q = {0, 0};
    for (size_t i = 0; i < n; ++i) {
        dd d = dd_sub((dd){x[i], 0}, mean);
        q = dd_add(q, dd_mul(d, d));
    }
    out[2] = mean.hi + mean.lo;
    dd v = dd_div_d(q, (double)n - (double)ddof);
    out[3] = v.hi + v.lo;
}
void dd_stats_f32(const float *x, size_t n, size_t ddof, double *out) {
    double *w = (double *)malloc((n ? n : 1) * sizeof(double));
    for (size_t i = 0; i < n; ++i) w[i] = (double)x[i];
    dd_stats_f64(w, n, ddof, out);
    free(w);
}

/* Row statistics of a (rows, cols) matrix: out rows x 4 as in dd_stats_f64. */
void dd_row_stats_f64(const double *x, size_t rows, size_t cols, size_t ddof, double *out) {
    for (size_t r = 0; r < rows; ++r) dd_stats_f64(x + r * cols, cols, ddof, out + 4 * r);
}
void dd_row_stats_f32(const float *x, size_t rows, size_t cols, size_t ddof, double *out) {
    for (size_t r = 0; r < rows; ++r) dd_stats_f32(x + r * cols, cols, ddof, out + 4 * r);
}

/* Exact integer sum (no wrap) as a 128-bit value split into (hi, lo) 64-bit halves. */
void exact_sum_i64(const int64_t *a, size_t n, int64_t *hi, uint64_t *lo) {
    __int128 s = 0;
    for (size_t i = 0; i < n; ++i) s += a[i];
    *hi = (int64_t)(s >> 64);
    *lo = (uint64_t)s;
}

/* C = A * B in double-double, rounded to double, for the `count` entries (rows[i], cols[i])
 * only (full products are too slow on the CPU at 4096^3+). A is m x k, B is k x n, row-major. */
void dd_gemm_entries_f64(size_t k, size_t n, const double *a, const double *b, size_t count, const int64_t *rows,
                         const int64_t *cols, double *out) {
    for (size_t e = 0; e < count; ++e) {
        const double *ar = a + (size_t)rows[e] * k;
        dd s = {0, 0};
        for (size_t p = 0; p < k; ++p) s = dd_add(s, two_prod(ar[p], b[p * n + (size_t)cols[e]]));
        out[e] = s.hi + s.lo;
    }
}
void dd_gemm_entries_f32(size_t k, size_t n, const float *a, const float *b, size_t count, const int64_t *rows,
                         const int64_t *cols, double *out) {
    for (size_t e = 0; e < count; ++e) {
        const float *ar = a + (size_t)rows[e] * k;
        dd s = {0, 0};
        for (size_t p = 0; p < k; ++p) s = dd_add(s, two_prod((double)ar[p], (double)b[p * n + (size_t)cols[e]]));
        out[e] = s.hi + s.lo;
    }
}

/* Full double-double GEMM (small sizes): C rounded to double. */
void dd_gemm_f64(size_t m, size_t k, size_t n, const double *a, const double *b, double *c) {
    for (size_t i = 0; i < m; ++i)
        for (size_t j = 0; j < n; ++j) {
            dd s = {0, 0};
            for (size_t p = 0; p < k; ++p) s = dd_add(s, two_prod(a[i * k + p], b[p * n + j]));
            c[i * n + j] = s.hi + s.lo;
        }
}

/* sum_k |A[i][k]| |B[k][j]| for the given entries: the scale that GEMM error bounds use. */
void abs_gemm_entries_f64(size_t k, size_t n, const double *a, const double *b, size_t count, const int64_t *rows,
                          const int64_t *cols, double *out) {
    for (size_t e = 0; e < count; ++e) {
        const double *ar = a + (size_t)rows[e] * k;
        double s = 0;
        for (size_t p = 0; p < k; ++p) s += fabs(ar[p]) * fabs(b[p * n + (size_t)cols[e]]);
        out[e] = s;
    }
}
