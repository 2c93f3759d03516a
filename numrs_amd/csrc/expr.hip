// Expression fusion (expr/fused_eval.rs:421-697): an ExprNode tree of up to 4 operands evaluated in
// one pass over the data, each node rounded once, never contracted to an FMA.
//
// The tree arrives as a postfix program of nrk_expr_node (include/numrs_amd.h): LEAF j pushes
// operand j, STEP applies a scalar op or a unary function to the top (ScalarRhs / ScalarLhs /
// Unary), BINARY pops two and pushes l op r (Binary), FMA pops three and pushes (a * b) + c with
// two roundings (the reference's Fma node, fused_eval.rs:489-497). Every specialised shape of
// fused_specialized (Leaf op Leaf, (Leaf op Leaf) op Leaf, Leaf op (Leaf op Leaf), the axpy
// shapes, (Leaf op Leaf) op (Leaf op Leaf), Fma, the scalar and unary forms) and every other tree
// of at most 4 operands and stack depth 4 runs as one kernel, whose rounding sequence is the
// eager evaluation's (eval_eager, :605-658: one rounding per node) -- so fused and eager give the
// same bits, as the reference's own tests require.
//
// The kernel is a stack machine over register batches: the program is the same for every thread,
// so each instruction is one uniform branch followed by straight-line work on U 16-byte vectors
// per operand (the stack slot is also uniform: a switch picks the register array).
#include "ops.hpp"

namespace nrk {

// STEP ops an expression may use: the reference's ScalarRhs / ScalarLhs ops and UnaryOp.
static bool expr_step_ok(int op) {
    switch (op) {
        case NRK_SOP_ADD: case NRK_SOP_SUB: case NRK_SOP_MUL: case NRK_SOP_DIV: case NRK_SOP_RSUB: case NRK_SOP_RDIV:
        case NRK_SOP_NEG: case NRK_SOP_ABS: case NRK_SOP_SQRT: case NRK_SOP_EXP: case NRK_SOP_LOG: return true;
        default: return false;
    }
}
static bool expr_bop_ok(int op) { return (op >= NRK_BOP_ADD && op <= NRK_BOP_COPYSIGN) || op == NRK_BOP_PRELU; }

nrk_status compile_expr(const nrk_expr_node *nodes, int32_t len, ExprProg *out) {
    if (len < 1 || len > NRK_MAX_EXPR) return fail(NRK_ERR_INVALID_ARGUMENT, "expression length %d outside [1,%d]", len, NRK_MAX_EXPR);
    if (!nodes) return fail(NRK_ERR_INVALID_ARGUMENT, "expression is NULL");
    ExprProg p{};
    int sp = 0, depth = 0, k = 0;
    for (int i = 0; i < len; ++i) {
        const nrk_expr_node &nd = nodes[i];
        switch (nd.kind) {
            case NRK_EXPR_LEAF:
                if (nd.op < 0 || nd.op >= NRK_MAX_LEAVES) return fail(NRK_ERR_INVALID_ARGUMENT, "node %d: leaf %d outside [0,%d)", i, nd.op, NRK_MAX_LEAVES);
                p.n_leaves = std::max(p.n_leaves, nd.op + 1);
                ++sp;
                break;
            case NRK_EXPR_STEP:
                if (!expr_step_ok(nd.op)) return fail(NRK_ERR_INVALID_ARGUMENT, "node %d: step %d is not a scalar or unary expression op", i, nd.op);
                if (sp < 1) return fail(NRK_ERR_INVALID_ARGUMENT, "node %d: step on an empty stack", i);
                if (nd.op == NRK_SOP_EXP || nd.op == NRK_SOP_LOG) p.math = 1;
                break;
            case NRK_EXPR_BINARY:
                if (!expr_bop_ok(nd.op)) return fail(NRK_ERR_INVALID_ARGUMENT, "node %d: op %d is not a binary expression op", i, nd.op);
                if (sp < 2) return fail(NRK_ERR_INVALID_ARGUMENT, "node %d: binary op needs two operands", i);
                --sp;
                break;
            case NRK_EXPR_FMA:
                if (sp < 3) return fail(NRK_ERR_INVALID_ARGUMENT, "node %d: fma needs three operands", i);
                sp -= 2;
                break;
            default: return fail(NRK_ERR_INVALID_ARGUMENT, "node %d: kind %d is not an NRK_EXPR_* code", i, nd.kind);
        }
        depth = std::max(depth, sp);
        if (depth > EXPR_DEPTH) return fail(NRK_ERR_UNSUPPORTED, "expression needs a stack deeper than %d", EXPR_DEPTH);
        if (nd.kind == NRK_EXPR_FMA) {  // (a * b) + c over the top three, two roundings
            p.kind[k] = EXPR_I_FMA;
            p.op[k] = 0;
            p.s[k] = 0.0;
            ++k;
            continue;
        }
        p.kind[k] = nd.kind == NRK_EXPR_LEAF ? EXPR_I_LEAF : (nd.kind == NRK_EXPR_STEP ? EXPR_I_STEP : EXPR_I_BIN);
        p.op[k] = nd.op;
        p.s[k] = nd.scalar;
        ++k;
    }
    if (sp != 1) return fail(NRK_ERR_INVALID_ARGUMENT, "expression leaves %d values on the stack (1 expected)", sp);
    p.len = k;
    *out = p;
    return NRK_OK;
}

namespace {

constexpr int XB = 256;
constexpr int XU = 2;  // 16-byte vectors per operand per thread and pass

template <typename T, int N, int MATH>
__device__ __forceinline__ void expr_step(int op, double sd, T (&v)[N]) {
    const T s = (T)sd;
    switch (op) {
#define NRK_XSTEP(OP, EXPR)                             \
    case OP:                                            \
        _Pragma("unroll") for (int i = 0; i < N; ++i) { \
            const T x = v[i];                           \
            v[i] = (EXPR);                              \
        }                                               \
        break;
        NRK_XSTEP(NRK_SOP_ADD, rn_add(x, s))
        NRK_XSTEP(NRK_SOP_SUB, rn_sub(x, s))
        NRK_XSTEP(NRK_SOP_MUL, rn_mul(x, s))
        NRK_XSTEP(NRK_SOP_DIV, rn_div(x, s))
        NRK_XSTEP(NRK_SOP_RSUB, rn_sub(s, x))
        NRK_XSTEP(NRK_SOP_RDIV, rn_div(s, x))
        NRK_XSTEP(NRK_SOP_NEG, -x)
        NRK_XSTEP(NRK_SOP_ABS, dev_fabs(x))
        NRK_XSTEP(NRK_SOP_SQRT, rn_sqrt(x))
        default:
            if constexpr (MATH) {
                switch (op) {
                    NRK_XSTEP(NRK_SOP_EXP, fm::exp_t(x))
                    NRK_XSTEP(NRK_SOP_LOG, (T)fm::log((double)x))
                    default: break;
                }
            }
            break;
#undef NRK_XSTEP
    }
}

template <typename T, int N>
__device__ __forceinline__ void expr_bop(int op, T (&a)[N], const T (&b)[N]) {
    switch (op) {
#define NRK_XBOP(OP, EXPR)                                         \
    case OP:                                                       \
        _Pragma("unroll") for (int i = 0; i < N; ++i) a[i] = EXPR; \
        break;
        NRK_XBOP(NRK_BOP_ADD, rn_add(a[i], b[i]))
        NRK_XBOP(NRK_BOP_SUB, rn_sub(a[i], b[i]))
        NRK_XBOP(NRK_BOP_MUL, rn_mul(a[i], b[i]))
        NRK_XBOP(NRK_BOP_DIV, rn_div(a[i], b[i]))
        NRK_XBOP(NRK_BOP_REM, __builtin_fmod(a[i], b[i]))
        NRK_XBOP(NRK_BOP_MAXIMUM, a[i] > b[i] ? a[i] : b[i])
        NRK_XBOP(NRK_BOP_MINIMUM, a[i] < b[i] ? a[i] : b[i])
        NRK_XBOP(NRK_BOP_COPYSIGN, dev_copysign(a[i], b[i]))
        NRK_XBOP(NRK_BOP_PRELU, a[i] > (T)0 ? a[i] : rn_mul(b[i], a[i]))
#undef NRK_XBOP
        default: break;
    }
}

// (a * b) + c on the top three slots, two roundings
template <typename T, int N>
__device__ __forceinline__ void expr_fma(T (&a)[N], const T (&b)[N], const T (&c)[N]) {
#pragma unroll
    for (int i = 0; i < N; ++i) a[i] = rn_add(rn_mul(a[i], b[i]), c[i]);
}

struct Leaves {
    const void *p[NRK_MAX_LEAVES];
};

// Load N consecutive elements of operand j starting at element i0 (VEC: as 16-byte vectors).
template <typename T, int N, bool VEC>
__device__ __forceinline__ void expr_load(const Leaves &L, int j, size_t i0, T (&v)[N]) {
    const T *src = (const T *)L.p[j];
    if constexpr (VEC) {
        typedef typename Vec16<T>::type V;
        constexpr int VN = Vec16<T>::N;
#pragma unroll
        for (int u = 0; u < N / VN; ++u) {
            const V x = __builtin_nontemporal_load(reinterpret_cast<const V *>(src + i0) + u);
#pragma unroll
            for (int e = 0; e < VN; ++e) v[u * VN + e] = x[e];
        }
    } else {
#pragma unroll
        for (int i = 0; i < N; ++i) v[i] = src[i0 + i];
    }
}

// Run the program on N elements starting at i0; the result is left in s0.
template <typename T, int N, int MATH, bool VEC>
__device__ __forceinline__ void expr_run(const ExprProg &pg, const Leaves &L, size_t i0, T (&s0)[N]) {
    T s1[N], s2[N], s3[N];
    int sp = 0;
    for (int k = 0; k < pg.len; ++k) {
        const int kind = pg.kind[k], op = pg.op[k];
        if (kind == EXPR_I_LEAF) {
            switch (sp) {
                case 0: expr_load<T, N, VEC>(L, op, i0, s0); break;
                case 1: expr_load<T, N, VEC>(L, op, i0, s1); break;
                case 2: expr_load<T, N, VEC>(L, op, i0, s2); break;
                default: expr_load<T, N, VEC>(L, op, i0, s3); break;
            }
            ++sp;
        } else if (kind == EXPR_I_STEP) {
            switch (sp) {
                case 1: expr_step<T, N, MATH>(op, pg.s[k], s0); break;
                case 2: expr_step<T, N, MATH>(op, pg.s[k], s1); break;
                case 3: expr_step<T, N, MATH>(op, pg.s[k], s2); break;
                default: expr_step<T, N, MATH>(op, pg.s[k], s3); break;
            }
        } else if (kind == EXPR_I_BIN) {
            switch (sp) {
                case 2: expr_bop<T, N>(op, s0, s1); break;
                case 3: expr_bop<T, N>(op, s1, s2); break;
                default: expr_bop<T, N>(op, s2, s3); break;
            }
            --sp;
        } else {  // EXPR_I_FMA
            switch (sp) {
                case 3: expr_fma<T, N>(s0, s1, s2); break;
                default: expr_fma<T, N>(s1, s2, s3); break;
            }
            sp -= 2;
        }
    }
}

template <typename T, int MATH>
__global__ __launch_bounds__(XB) void expr_kernel(ExprProg pg, Leaves L, T *y, size_t n, int vec_ok) {
    constexpr int VN = Vec16<T>::N;
    constexpr int N = XU * VN;
    const size_t tid = blockIdx.x * (size_t)XB + threadIdx.x, nth = (size_t)gridDim.x * XB;
    size_t done = 0;
    if (vec_ok) {
        const size_t nb = n / N;  // batches of N elements per thread
        for (size_t b = tid; b < nb; b += nth) {
            T r[N];
            expr_run<T, N, MATH, true>(pg, L, b * N, r);
            typedef typename Vec16<T>::type V;
#pragma unroll
            for (int u = 0; u < XU; ++u) {
                V v;
#pragma unroll
                for (int e = 0; e < VN; ++e) v[e] = r[u * VN + e];
                __builtin_nontemporal_store(v, reinterpret_cast<V *>(y + b * N) + u);
            }
        }
        done = nb * N;
    }
    for (size_t i = done + tid; i < n; i += nth) {
        T r[1];
        expr_run<T, 1, MATH, false>(pg, L, i, r);
        y[i] = r[0];
    }
}

}  // namespace

template <typename T>
nrk_status expr_eval(const ExprProg &pg, const T *const *leaves, T *y, size_t n, hipStream_t stream) {
    NRK_TRY(ensure_device());
    if (n == 0) return NRK_OK;
    if (!y || !leaves) return fail(NRK_ERR_INVALID_ARGUMENT, "y or leaves is NULL");
    Leaves L{};
    int vec_ok = (uintptr_t)y % 16 == 0;
    for (int j = 0; j < pg.n_leaves; ++j) {
        if (!leaves[j]) return fail(NRK_ERR_INVALID_ARGUMENT, "leaf %d is NULL", j);
        L.p[j] = leaves[j];
        vec_ok &= (uintptr_t)leaves[j] % 16 == 0;
    }
    const size_t batches = n / (XU * Vec16<T>::N) + 1;
    const int grid = (int)std::min<size_t>(ceil_div(batches, (size_t)XB), (size_t)cu_count() * 32);
    if (pg.math)
        hipLaunchKernelGGL((expr_kernel<T, 1>), dim3(grid), dim3(XB), 0, stream, pg, L, y, n, vec_ok);
    else
        hipLaunchKernelGGL((expr_kernel<T, 0>), dim3(grid), dim3(XB), 0, stream, pg, L, y, n, vec_ok);
    NRK_HIP(hipGetLastError());
    return NRK_OK;
}
template nrk_status expr_eval<float>(const ExprProg &, const float *const *, float *, size_t, hipStream_t);
template nrk_status expr_eval<double>(const ExprProg &, const double *const *, double *, size_t, hipStream_t);

// Host form: operands staged through the device, result copied back.
template <typename T>
static nrk_status host_expr_eval(const nrk_expr_node *nodes, int32_t len, const T *const *leaves, int32_t n_leaves, T *y,
                                 size_t n) {
    NRK_TRY(ensure_device());
    ExprProg pg;
    NRK_TRY(compile_expr(nodes, len, &pg));
    if (n_leaves < pg.n_leaves) return fail(NRK_ERR_INVALID_ARGUMENT, "the expression reads %d operands, %d given", pg.n_leaves, n_leaves);
    if (n == 0) return NRK_OK;
    if (!leaves || !y) return fail(NRK_ERR_INVALID_ARGUMENT, "leaves or y is NULL");
    void *ws = nullptr;
    const size_t span = align_up(n * sizeof(T), 256);
    NRK_TRY(staging(3, span * (pg.n_leaves + 1), &ws));
    hipStream_t s = nullptr;
    const T *dl[NRK_MAX_LEAVES] = {nullptr, nullptr, nullptr, nullptr};
    for (int j = 0; j < pg.n_leaves; ++j) {
        if (!leaves[j]) return fail(NRK_ERR_INVALID_ARGUMENT, "leaf %d is NULL", j);
        T *d = (T *)((char *)ws + span * (j + 1));
        NRK_TRY(copy_h2d(d, leaves[j], n * sizeof(T), s));
        dl[j] = d;
    }
    NRK_TRY(expr_eval<T>(pg, dl, (T *)ws, n, s));
    NRK_TRY(copy_d2h(y, ws, n * sizeof(T), s));
    NRK_HIP(hipStreamSynchronize(s));
    return NRK_OK;
}

}  // namespace nrk

using namespace nrk;

extern "C" {

nrk_status nrk_expr_eval_f64(const nrk_expr_node *nodes, int32_t len, const double *const *leaves, int32_t n_leaves, double *y,
                             size_t n) {
    return host_expr_eval<double>(nodes, len, leaves, n_leaves, y, n);
}
nrk_status nrk_expr_eval_f32(const nrk_expr_node *nodes, int32_t len, const float *const *leaves, int32_t n_leaves, float *y,
                             size_t n) {
    return host_expr_eval<float>(nodes, len, leaves, n_leaves, y, n);
}
nrk_status nrk_dev_expr_eval_f64(const nrk_expr_node *nodes, int32_t len, const double *const *leaves, int32_t n_leaves,
                                 double *y, size_t n, nrk_stream stream) {
    ExprProg pg;
    NRK_TRY(compile_expr(nodes, len, &pg));
    if (n_leaves < pg.n_leaves) return fail(NRK_ERR_INVALID_ARGUMENT, "the expression reads %d operands, %d given", pg.n_leaves, n_leaves);
    return expr_eval<double>(pg, leaves, y, n, (hipStream_t)stream);
}
nrk_status nrk_dev_expr_eval_f32(const nrk_expr_node *nodes, int32_t len, const float *const *leaves, int32_t n_leaves,
                                 float *y, size_t n, nrk_stream stream) {
    ExprProg pg;
    NRK_TRY(compile_expr(nodes, len, &pg));
    if (n_leaves < pg.n_leaves) return fail(NRK_ERR_INVALID_ARGUMENT, "the expression reads %d operands, %d given", pg.n_leaves, n_leaves);
    return expr_eval<float>(pg, leaves, y, n, (hipStream_t)stream);
}

}  // extern "C"
