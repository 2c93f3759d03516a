// Internal launchers (device pointers, one stream) behind the extern "C" entry points.
#pragma once

#include "common.hpp"

namespace nrk {

// Fused chain -> {sum, mean, var, min, max, argmin, argmax} along axis 1 of a (rows, cols)
// matrix and/or over all of it, in one pass (reduce.hip). ext: NULL for the stream's own
// scratch (allocated on demand), else the caller's.
template <typename T>
nrk_status seg_reduce(const Chain &chain, const T *x, size_t rows, size_t cols, double ddof,
                      const nrk_reduce_outputs *row_out, const nrk_reduce_outputs *full_out, hipStream_t stream,
                      const ExtScratch *ext = nullptr);

// The same outputs with every sum, mean and variance in the reference's own order of rounded
// additions (reduce_plain.hip): bit-identical to the reference's CPU kernels.
template <typename T>
nrk_status plain_reduce(const Chain &chain, const T *x, size_t rows, size_t cols, size_t ddof,
                        const nrk_reduce_outputs *row_out, const nrk_reduce_outputs *full_out, hipStream_t stream);

// The fixed device scratch of the calling thread's device (set up by nrk_init, runtime.hip).
nrk_status static_scratch(ExtScratch *out);

// A compiled expression (expr.hip): postfix instructions over up to NRK_MAX_LEAVES operands with a
// stack of at most EXPR_DEPTH values; FMA is (a * b) + c with two roundings.
constexpr int EXPR_DEPTH = 4, EXPR_MAX_INS = NRK_MAX_EXPR;
enum { EXPR_I_LEAF = 0, EXPR_I_STEP = 1, EXPR_I_BIN = 2, EXPR_I_FMA = 3 };
struct ExprProg {
    int len, n_leaves, math;  // math: an EXP or LOG step is present
    int kind[EXPR_MAX_INS];
    int op[EXPR_MAX_INS];
    double s[EXPR_MAX_INS];
};
nrk_status compile_expr(const nrk_expr_node *nodes, int32_t len, ExprProg *out);
template <typename T>
nrk_status expr_eval(const ExprProg &pg, const T *const *leaves, T *y, size_t n, hipStream_t stream);
// The fused reduction of an expression (reduce.hip): one pass when the program is one operand
// with scalar steps, or two operands combined by one binary op, each with scalar steps (any
// order of pushes), on the lean path; otherwise the expression is evaluated into the stream's
// workspace first.
template <typename T>
nrk_status expr_reduce(const ExprProg &pg, const T *const *leaves, size_t rows, size_t cols, double ddof,
                       const nrk_reduce_outputs *row_out, const nrk_reduce_outputs *full_out, hipStream_t stream);

// The tile configuration the selector picks for a dispatch with no forced tile (selector.hip):
// a recorded decision, the registered cost model's choice inside its trained shapes (timed when
// the model is unsure), or -1 when neither applies and the caller keeps its fixed rule.
// family: NRK_DT_*; GEMMs pass (m, n, k).
int select_config(int family, size_t d0, size_t d1, size_t d2, hipStream_t stream);

// Workgroups per CU of the segment reducer's persistent grid (result bits are independent of it).
int reduce_blocks_per_cu();
void set_reduce_blocks_per_cu(int b);

// Tile-configuration dispatch for the tuner and the learned selector (configs.hip).
template <typename T>
nrk_status gemm_config(int config, size_t m, size_t k, size_t n, const T *a, const T *b, T *c, hipStream_t stream);
nrk_status gemm_bf16_config(int config, const nrk_gemm_desc &d, hipStream_t stream);

// Reductions along a strided axis (inner > 1): one thread per output, left-to-right fold in the
// reference's order (reduce_axis.hip).
template <typename T>
nrk_status strided_reduce(int op, const T *x, size_t outer, size_t len, size_t inner, double ddof, void *out,
                          hipStream_t stream);
// sums / means along a strided axis with few outputs, reading the data once
template <typename T>
nrk_status axis_sum(const T *x, size_t outer, size_t len, size_t inner, bool mean, T *out, hipStream_t stream);
// max / min / argmax / argmin along a strided axis with few outputs, reading the data once
template <typename T>
nrk_status axis_extrema(int op, const T *x, size_t outer, size_t len, size_t inner, void *out, hipStream_t stream);
// sample variance (divisor len - ddof) along a strided axis with few outputs, reading the data once
template <typename T>
nrk_status axis_var(const T *x, size_t outer, size_t len, size_t inner, double ddof, T *out, hipStream_t stream);

nrk_status sum_i64(const int64_t *x, size_t n, int64_t *out, hipStream_t stream);
template <typename T>
nrk_status prod(const T *x, size_t n, T *out, hipStream_t stream);
template <typename T>
nrk_status dot(const T *a, const T *b, size_t n, T *out, hipStream_t stream);
template <typename T>
nrk_status sqrt_scalar(T *p, hipStream_t stream);
// inclusive scans along `rows` contiguous rows of `len` (op: NRK_BOP_ADD/SUB/MUL/DIV/MAXIMUM/MINIMUM)
template <typename T>
nrk_status accumulate(int op, const T *x, T *y, size_t rows, size_t len, hipStream_t stream);
// a[indices[k]][:] = a[indices[k]][:] op b[k][:] in k order (ufunc_at), a of n0 rows of row_len
template <typename T>
nrk_status ufunc_at(int op, T *a, size_t n0, size_t row_len, const uint64_t *indices, size_t m, const T *b,
                    hipStream_t stream);
// segmented reductions at indices along contiguous rows (ufunc_reduceat's segment rule)
template <typename T>
nrk_status reduceat(int op, const T *x, size_t rows, size_t len, const uint64_t *idx, size_t nidx, T *y, hipStream_t stream);
// ms[0] = max x (NaN ignored), ms[1] = sum e^(x - max), in one read of x
template <typename T>
nrk_status softmax_stats(const T *x, size_t n, T *ms, hipStream_t stream);

template <typename T>
nrk_status map(const Chain &chain, const T *x, T *y, size_t n, hipStream_t stream);
template <typename T>
nrk_status zip(int op, const T *a, const T *b, T *y, size_t n, hipStream_t stream);
// ufunc_outer / ufunc_where (op: NRK_BOP_ADD/SUB/MUL/DIV/MAXIMUM/MINIMUM, NaN-propagating max/min)
template <typename T>
nrk_status outer(int op, const T *a, size_t na, const T *b, size_t nb, T *y, hipStream_t stream);
template <typename T>
nrk_status where_op(int op, const T *a, const T *b, const uint8_t *mask, T *y, size_t n, hipStream_t stream);
// softmax passes: mode 0 y = exp(x - ms[0]); 1 y = y / ms[1]; 2 y = (x - ms[0]) - ln(ms[1])
template <typename T>
nrk_status softmax_pass(int mode, const T *x, T *y, const T *ms, size_t n, hipStream_t stream);
// op 0: y = fma(a, b, c); op 1: y = a + t * (b - a) (c unused)
template <typename T>
nrk_status tri(int op, const T *a, const T *b, const T *c, T t, T *y, size_t n, hipStream_t stream);

// Plain GEMM on MFMA: C(m,n) = A(m,k) * B(k,n), row-major, overwrite (gemm.hip).
template <typename T>
nrk_status gemm(size_t m, size_t k, size_t n, const T *a, const T *b, T *c, hipStream_t stream);

// gemm_generic's arithmetic (ascending k, acc + a * b in T; integers wrap) over `batch` panels with
// element strides sa, sb (0 broadcasts) and C panels of m*n (gemm_generic.hip); T = f32, f64, i32,
// i64. host_gemm_generic is the reference's host-slice contract: C zeroed over c_len, computed
// only when every slice is long enough.
template <typename T>
nrk_status gemm_generic(size_t batch, size_t m, size_t k, size_t n, const T *a, size_t sa, const T *b, size_t sb, T *c,
                        hipStream_t stream);
template <typename T>
nrk_status host_gemm_generic(size_t m, size_t k, size_t n, const T *a, size_t a_len, const T *b, size_t b_len, T *c,
                             size_t c_len);

// Split-precision GEMMs with an explicit tile configuration (-1: built-in choice):
// FP32 as BF16 split products (gemm_bf16.hip), FP64 by the Ozaki scheme II (gemm_ozaki.hip).
nrk_status gemm_f32_bf16x(size_t m, size_t k, size_t n, const float *a, const float *b, float *c, int products, int tile,
                          hipStream_t stream);
nrk_status gemm_f64_ozaki(size_t m, size_t k, size_t n, const double *a, const double *b, double *c, int moduli, int tile,
                          hipStream_t stream);

}  // namespace nrk
