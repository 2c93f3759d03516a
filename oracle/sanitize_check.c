/* TEST INFRASTRUCTURE ONLY: a driver that runs every function of the oracle (oracle.c) on
 * exactly-sized heap buffers, at the sizes around every tier boundary of the restatement and at
 * the integer edge values of Rust's release arithmetic, so that an AddressSanitizer /
 * UndefinedBehaviorSanitizer build of the two files (oracle/build.py --sanitize) reports any
 * out-of-bounds access, use of freed memory or undefined behaviour in the checker itself.
 * Nothing here checks values: the oracle's values are pinned by tests/test_oracle.py. */
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

void oracle_set_threads(int n);
#define DECL_REDUCE(T, SUF)                                                                        \
    T ref_sum_##SUF(const T *, size_t);                                                            \
    T ref_mean_##SUF(const T *, size_t);                                                           \
    T ref_sum_sq_dev_##SUF(const T *, size_t, T);                                                  \
    T ref_var_##SUF(const T *, size_t, size_t);                                                    \
    T ref_max_##SUF(const T *, size_t);                                                            \
    T ref_min_##SUF(const T *, size_t);                                                            \
    T ref_prod_##SUF(const T *, size_t);                                                           \
    uint64_t ref_argmax_##SUF(const T *, size_t);                                                  \
    uint64_t ref_argmin_##SUF(const T *, size_t);                                                  \
    void ref_axis_##SUF(int, const T *, size_t, size_t, size_t, size_t, void *);                   \
    void ref_map_##SUF(const int32_t *, const double *, int, const T *, T *, size_t);              \
    void ref_zip_##SUF(int, const T *, const T *, T *, size_t);                                    \
    void ref_expr_##SUF(const int32_t *, const int32_t *, const double *, int, const T *const *, T *, size_t); \
    void ref_fma_##SUF(const T *, const T *, const T *, T *, size_t);                              \
    void ref_lerp_##SUF(const T *, const T *, T, T *, size_t);                                     \
    void ref_gemm_generic_##SUF(size_t, size_t, size_t, const T *, const T *, T *);                \
    void dd_stats_##SUF(const T *, size_t, size_t, double *);                                      \
    void dd_row_stats_##SUF(const T *, size_t, size_t, size_t, double *);                          \
    void dd_gemm_entries_##SUF(size_t, size_t, const T *, const T *, size_t, const int64_t *, const int64_t *, double *);
DECL_REDUCE(double, f64)
DECL_REDUCE(float, f32)
int64_t ref_sum_i64(const int64_t *, size_t);
void exact_sum_i64(const int64_t *, size_t, int64_t *, uint64_t *);
void ref_gemm_rowsplit_f64(size_t, size_t, size_t, const double *, const double *, double *);
void dd_gemm_f64(size_t, size_t, size_t, const double *, const double *, double *);
void abs_gemm_entries_f64(size_t, size_t, const double *, const double *, size_t, const int64_t *, const int64_t *, double *);
#define DECL_INT(T, SUF)                                                                           \
    void ref_zip_int_##SUF(int, const T *, const T *, T *, size_t);                               \
    void ref_map_int_##SUF(const int32_t *, const int64_t *, int, const T *, T *, size_t);
DECL_INT(int64_t, i64)
DECL_INT(int32_t, i32)
DECL_INT(uint64_t, u64)
DECL_INT(uint32_t, u32)
void ref_gemm_generic_i64(size_t, size_t, size_t, const int64_t *, const int64_t *, int64_t *);
void ref_gemm_generic_i32(size_t, size_t, size_t, const int32_t *, const int32_t *, int32_t *);

static uint64_t g_state = 0x9E3779B97F4A7C15ull;
static uint64_t next_u64(void) {  /* splitmix64 */
    uint64_t z = (g_state += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}
static double next_unit(void) { return (double)(next_u64() >> 11) * 0x1p-53 * 2.0 - 1.0; }
static void *xmalloc(size_t bytes) {  /* exactly `bytes`: the sanitizer sees any access past them */
    void *p = malloc(bytes ? bytes : 1);
    if (!p) {
        fprintf(stderr, "out of memory\n");
        exit(2);
    }
    return p;
}

#define RUN_FLOAT(T, SUF)                                                                              \
    static void run_##SUF(void) {                                                                      \
        static const size_t sizes[] = {0, 1, 3, 4, 5, 7, 8191, 8192, 8193, 9999, 10000, 10001, 39999,  \
                                       40000, 40001, 79999, 80000, 80001, 100003};                   \
        for (size_t si = 0; si < sizeof(sizes) / sizeof(sizes[0]); ++si) {                             \
            const size_t n = sizes[si];                                                                \
            T *x = (T *)xmalloc(n * sizeof(T)), *y = (T *)xmalloc(n * sizeof(T));                      \
            T *z = (T *)xmalloc(n * sizeof(T)), *w = (T *)xmalloc(n * sizeof(T));                      \
            for (size_t i = 0; i < n; ++i) {                                                           \
                x[i] = (T)(next_unit() * 100.0);                                                       \
                y[i] = (T)(next_unit() + 1.5);                                                         \
                z[i] = (T)next_unit();                                                                 \
            }                                                                                          \
            if (n > 10) x[n / 2] = (T)NAN;                                                             \
            volatile T sink = ref_sum_##SUF(x, n) + ref_mean_##SUF(x, n) + ref_var_##SUF(x, n, 0) +   \
                              ref_var_##SUF(x, n, 1) + ref_sum_sq_dev_##SUF(x, n, (T)0.5) +             \
                              ref_max_##SUF(x, n) + ref_min_##SUF(x, n) + ref_prod_##SUF(y, n);        \
            (void)sink;                                                                                \
            if (n) {                                                                                   \
                volatile uint64_t a = ref_argmax_##SUF(x, n) + ref_argmin_##SUF(x, n);                 \
                (void)a;                                                                               \
            }                                                                                          \
            const int32_t ops[] = {0, 1, 2, 3, 4, 5, 6, 7, 8, 9, 42};                                  \
            const double sc[] = {0.5, -0.25, 1.5, 3.0, 2.0, 1.0, 0, 0, 0, 0, 0};                      \
            ref_map_##SUF(ops, sc, 11, y, w, n);                                                       \
            for (int op = 0; op < 12; ++op) ref_zip_##SUF(op, x, y, w, n);                             \
            ref_fma_##SUF(x, y, z, w, n);                                                              \
            ref_lerp_##SUF(x, y, (T)0.37, w, n);                                                       \
            const T *leaves[4] = {x, y, z, x};                                                         \
            /* a valid postfix program: fma(x, y, z) * 1.5 - x, then ((x + y) max z) */              \
            const int32_t kind[] = {0, 0, 0, 3, 1, 0, 2};                                              \
            const int32_t eop[] = {0, 1, 2, 0, 2, 3, 1};                                               \
            const double esc[] = {0, 0, 0, 0, 1.5, 0, 0};                                              \
            ref_expr_##SUF(kind, eop, esc, 7, leaves, w, n);                                           \
            const int32_t kind2[] = {0, 0, 2, 0, 2};                                                   \
            const int32_t eop2[] = {0, 1, 0, 2, 5};                                                    \
            const double esc2[] = {0, 0, 0, 0, 0};                                                     \
            ref_expr_##SUF(kind2, eop2, esc2, 5, leaves, w, n);                                        \
            double *dd = (double *)xmalloc(4 * sizeof(double));                                        \
            dd_stats_##SUF(x, n, 0, dd);                                                               \
            free(dd);                                                                                  \
            /* axis folds over (outer, len, inner) views of the same n elements */                    \
            const size_t shapes[][3] = {{1, n, 1}, {n, 1, 1}, {1, 1, n}};                              \
            for (int sh = 0; sh < 3 && n; ++sh) {                                                      \
                const size_t o = shapes[sh][0], l = shapes[sh][1], in = shapes[sh][2];                 \
                void *out = xmalloc(o * in * sizeof(uint64_t));                                        \
                for (int op = 0; op <= 6; ++op) ref_axis_##SUF(op, x, o, l, in, 0, out);              \
                free(out);                                                                             \
            }                                                                                          \
            free(x); free(y); free(z); free(w);                                                        \
        }                                                                                              \
        /* row statistics and GEMMs on ragged shapes, zero dimensions included */                     \
        static const size_t gs[][3] = {{0, 3, 4}, {3, 0, 4}, {3, 4, 0}, {1, 1, 1}, {65, 70, 129}, {7, 130, 3}}; \
        for (size_t gi = 0; gi < sizeof(gs) / sizeof(gs[0]); ++gi) {                                  \
            const size_t m = gs[gi][0], k = gs[gi][1], nn = gs[gi][2];                                 \
            T *a = (T *)xmalloc(m * k * sizeof(T)), *b = (T *)xmalloc(k * nn * sizeof(T));             \
            T *c = (T *)xmalloc(m * nn * sizeof(T));                                                   \
            for (size_t i = 0; i < m * k; ++i) a[i] = (T)next_unit();                                  \
            for (size_t i = 0; i < k * nn; ++i) b[i] = (T)next_unit();                                 \
            ref_gemm_generic_##SUF(m, k, nn, a, b, c);                                                 \
            double *rs = (double *)xmalloc(4 * m * sizeof(double));                                    \
            dd_row_stats_##SUF(a, m, k, 0, rs);                                                        \
            free(rs);                                                                                  \
            if (m && nn) {                                                                             \
                int64_t rows[3] = {0, (int64_t)(m - 1), (int64_t)(m / 2)};                             \
                int64_t cols[3] = {(int64_t)(nn - 1), 0, (int64_t)(nn / 2)};                           \
                double *e = (double *)xmalloc(3 * sizeof(double));                                     \
                dd_gemm_entries_##SUF(k, nn, a, b, 3, rows, cols, e);                                  \
                free(e);                                                                               \
            }                                                                                          \
            free(a); free(b); free(c);                                                                 \
        }                                                                                              \
    }
RUN_FLOAT(double, f64)
RUN_FLOAT(float, f32)

#define RUN_INT(T, SUF, MINV, MAXV)                                                                    \
    static void run_int_##SUF(void) {                                                                  \
        const T edge[] = {(T)0, (T)1, (T)-1, (T)2, (T)7, (T)(MAXV), (T)(MINV), (T)((MAXV) - 1), (T)((MINV) + 1)}; \
        const size_t ne = sizeof(edge) / sizeof(edge[0]);                                              \
        const size_t n = ne * ne;                                                                      \
        T *a = (T *)xmalloc(n * sizeof(T)), *b = (T *)xmalloc(n * sizeof(T)), *y = (T *)xmalloc(n * sizeof(T)); \
        for (size_t i = 0; i < ne; ++i)                                                                \
            for (size_t j = 0; j < ne; ++j) {                                                          \
                a[i * ne + j] = edge[i];                                                               \
                b[i * ne + j] = edge[j];                                                               \
            }                                                                                          \
        for (int op = 0; op < 18; ++op) {                                                              \
            if (op == 3 || op == 4) {  /* the callers never pass a zero divisor or MIN / -1 */        \
                T *bb = (T *)xmalloc(n * sizeof(T));                                                   \
                for (size_t i = 0; i < n; ++i) {                                                       \
                    bb[i] = b[i];                                                                      \
                    if (bb[i] == 0 || (a[i] == (T)(MINV) && bb[i] == (T)-1)) bb[i] = (T)3;             \
                }                                                                                      \
                ref_zip_int_##SUF(op, a, bb, y, n);                                                    \
                free(bb);                                                                              \
            } else {                                                                                   \
                ref_zip_int_##SUF(op, a, b, y, n);                                                     \
            }                                                                                          \
        }                                                                                              \
        const int32_t ops[] = {0, 1, 2, 4, 6, 7, 8, 15, 16, 17};                                       \
        const int64_t sc[] = {(int64_t)(MAXV), (int64_t)(MINV), 3, 5, 0, 0, 0, 0, -9, 9};             \
        ref_map_int_##SUF(ops, sc, 10, a, y, n);                                                       \
        free(a); free(b); free(y);                                                                     \
    }
RUN_INT(int64_t, i64, INT64_MIN, INT64_MAX)
RUN_INT(int32_t, i32, INT32_MIN, INT32_MAX)
RUN_INT(uint64_t, u64, 0, UINT64_MAX)
RUN_INT(uint32_t, u32, 0, UINT32_MAX)

static void run_misc(void) {
    const size_t n = 100003;
    int64_t *v = (int64_t *)xmalloc(n * sizeof(int64_t));
    for (size_t i = 0; i < n; ++i) v[i] = (int64_t)next_u64();
    int64_t hi;
    uint64_t lo;
    exact_sum_i64(v, n, &hi, &lo);
    volatile int64_t s = ref_sum_i64(v, n);
    (void)s;
    free(v);
    static const size_t gs[][3] = {{0, 2, 2}, {1, 1, 1}, {33, 65, 17}};
    for (size_t gi = 0; gi < 3; ++gi) {
        const size_t m = gs[gi][0], k = gs[gi][1], nn = gs[gi][2];
        int64_t *a = (int64_t *)xmalloc(m * k * 8), *b = (int64_t *)xmalloc(k * nn * 8), *c = (int64_t *)xmalloc(m * nn * 8);
        int32_t *a32 = (int32_t *)xmalloc(m * k * 4), *b32 = (int32_t *)xmalloc(k * nn * 4), *c32 = (int32_t *)xmalloc(m * nn * 4);
        double *ad = (double *)xmalloc(m * k * 8), *bd = (double *)xmalloc(k * nn * 8), *cd = (double *)xmalloc(m * nn * 8);
        for (size_t i = 0; i < m * k; ++i) {
            a[i] = (int64_t)next_u64();
            a32[i] = (int32_t)next_u64();
            ad[i] = next_unit();
        }
        for (size_t i = 0; i < k * nn; ++i) {
            b[i] = (int64_t)next_u64();
            b32[i] = (int32_t)next_u64();
            bd[i] = next_unit();
        }
        ref_gemm_generic_i64(m, k, nn, a, b, c);  /* products and sums that overflow: wrap */
        ref_gemm_generic_i32(m, k, nn, a32, b32, c32);
        ref_gemm_rowsplit_f64(m, k, nn, ad, bd, cd);
        dd_gemm_f64(m, k, nn, ad, bd, cd);
        if (m && nn) {
            int64_t rows[2] = {0, (int64_t)(m - 1)}, cols[2] = {(int64_t)(nn - 1), 0};
            double e[2];
            abs_gemm_entries_f64(k, nn, ad, bd, 2, rows, cols, e);
        }
        free(a); free(b); free(c); free(a32); free(b32); free(c32); free(ad); free(bd); free(cd);
    }
}

int main(void) {
    for (int threads = 1; threads <= 4; threads += 3) {  /* the serial path and the thread pool */
        oracle_set_threads(threads);
        run_f64();
        run_f32();
    }
    run_int_i64();
    run_int_i32();
    run_int_u64();
    run_int_u32();
    run_misc();
    printf("oracle sanitizer check: ok\n");
    return 0;
}
