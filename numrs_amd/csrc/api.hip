// extern "C" entry points for reductions and elementwise ops: argument checking, the
// host-pointer staging forms, and dispatch onto the launchers in ops.hpp.
#include "ops.hpp"

using namespace nrk;

namespace {

inline hipStream_t S(nrk_stream s) { return (hipStream_t)s; }

// Runs `dev(d_in, d_out, stream)` on a device copy of a host slice and copies `out_count`
// results of type OutT back. Blocks until the result is on the host.
template <typename T, typename OutT, typename F>
nrk_status host_call(const T *a, size_t n, OutT *out, size_t out_count, F dev) {
    NRK_TRY(ensure_device());
    if (!out) return fail(NRK_ERR_INVALID_ARGUMENT, "out is NULL");
    if (n > 0 && !a) return fail(NRK_ERR_INVALID_ARGUMENT, "input is NULL");
    void *da = nullptr, *dout = nullptr;
    NRK_TRY(staging(0, n * sizeof(T) + 16, &da));
    NRK_TRY(staging(1, out_count * sizeof(OutT) + 16, &dout));
    hipStream_t s = nullptr;
    if (n) NRK_TRY(copy_h2d(da, a, n * sizeof(T), s));
    NRK_TRY(dev((const T *)da, (OutT *)dout, s));
    if (out_count) NRK_TRY(copy_d2h(out, dout, out_count * sizeof(OutT), s));
    NRK_HIP(hipStreamSynchronize(s));
    return NRK_OK;
}

template <typename T>
nrk_status full_reduce(const T *a, size_t n, double ddof, nrk_reduce_outputs o, hipStream_t s) {
    Chain c;
    NRK_TRY(make_chain(nullptr, 0, &c));
    return seg_reduce<T>(c, a, 1, n, ddof, nullptr, &o, s);
}

template <typename T>
nrk_status dev_sum(const T *a, size_t n, T *out, hipStream_t s) {
    nrk_reduce_outputs o{};
    o.sum = out;
    return full_reduce<T>(a, n, 0.0, o, s);
}
template <typename T>
nrk_status dev_mean(const T *a, size_t n, T *out, hipStream_t s) {
    nrk_reduce_outputs o{};
    o.mean = out;
    return full_reduce<T>(a, n, 0.0, o, s);
}
template <typename T>
nrk_status dev_var(const T *a, size_t n, size_t ddof, T *out, hipStream_t s) {
    nrk_reduce_outputs o{};
    o.var = out;
    return full_reduce<T>(a, n, (double)ddof, o, s);
}
template <typename T>
nrk_status dev_min(const T *a, size_t n, T *out, hipStream_t s) {
    nrk_reduce_outputs o{};
    o.min = out;
    return full_reduce<T>(a, n, 0.0, o, s);
}
template <typename T>
nrk_status dev_max(const T *a, size_t n, T *out, hipStream_t s) {
    nrk_reduce_outputs o{};
    o.max = out;
    return full_reduce<T>(a, n, 0.0, o, s);
}
template <typename T>
nrk_status dev_argmax(const T *a, size_t n, uint64_t *out, hipStream_t s) {
    if (n == 0) return fail(NRK_ERR_INVALID_ARGUMENT, "Cannot find argmax of empty array");
    nrk_reduce_outputs o{};
    o.argmax = out;
    return full_reduce<T>(a, n, 0.0, o, s);
}
template <typename T>
nrk_status dev_argmin(const T *a, size_t n, uint64_t *out, hipStream_t s) {
    if (n == 0) return fail(NRK_ERR_INVALID_ARGUMENT, "Cannot find argmin of empty array");
    nrk_reduce_outputs o{};
    o.argmin = out;
    return full_reduce<T>(a, n, 0.0, o, s);
}
template <typename T>
nrk_status dev_sum_sq_dev(const T *a, size_t n, T mean, T *out, hipStream_t s) {
    nrk_scalar_op ops[2] = {{NRK_SOP_SUB, 0, (double)mean}, {NRK_SOP_SQUARE, 0, 0.0}};
    Chain c;
    NRK_TRY(make_chain(ops, 2, &c));
    nrk_reduce_outputs o{};
    o.sum = out;
    return seg_reduce<T>(c, a, 1, n, 0.0, nullptr, &o, s);
}

// ufuncs.rs:966-979: norm_l2 = sqrt(sum x*x), norm_l1 = sum |x| (one fused pass each);
// ufuncs.rs:1060 std = var.sqrt().
template <typename T>
nrk_status dev_norm(int32_t sop, bool root, const T *a, size_t n, T *out, hipStream_t s) {
    nrk_scalar_op ops[1] = {{sop, 0, 0.0}};
    Chain c;
    NRK_TRY(make_chain(ops, 1, &c));
    nrk_reduce_outputs o{};
    o.sum = out;
    NRK_TRY(seg_reduce<T>(c, a, 1, n, 0.0, nullptr, &o, s));
    return root ? sqrt_scalar<T>(out, s) : NRK_OK;
}
template <typename T>
nrk_status dev_std(const T *a, size_t n, size_t ddof, T *out, hipStream_t s) {
    NRK_TRY(dev_var<T>(a, n, ddof, out, s));
    return sqrt_scalar<T>(out, s);
}
// ufuncs.rs:1411-1457 softmax / log_softmax over the whole array: (max, sum of e^(x - max)) in
// one read, then y = e^(x - max) / sum or (x - max) - ln(sum) in a second (read x, write y). The
// max and sum stay on the device (two scalars in the stream's workspace), so nothing waits on
// the host.
template <typename T>
nrk_status dev_softmax(bool log_form, const T *x, T *y, size_t n, hipStream_t s) {
    NRK_TRY(ensure_device());
    if (n == 0) return NRK_OK;
    if (!x || !y) return fail(NRK_ERR_INVALID_ARGUMENT, "x or y is NULL");
    void *ws = nullptr;
    NRK_TRY(get_workspace(s, 2 * sizeof(T), &ws));
    T *ms = (T *)ws;
    NRK_TRY(softmax_stats<T>(x, n, ms, s));
    return softmax_pass<T>(log_form ? 2 : 3, x, y, ms, n, s);
}

template <typename T>
nrk_status host_dot(const T *a, const T *b, size_t n, T *out) {
    NRK_TRY(ensure_device());
    if (n && !b) return fail(NRK_ERR_INVALID_ARGUMENT, "b is NULL");
    void *db = nullptr;
    NRK_TRY(staging(2, n * sizeof(T) + 16, &db));
    if (n) NRK_TRY(copy_h2d(db, b, n * sizeof(T), nullptr));
    return host_call<T, T>(a, n, out, 1, [&](const T *d, T *dout, hipStream_t s) { return dot<T>(d, (const T *)db, n, dout, s); });
}

constexpr size_t AXIS_DIRECT_OUTPUTS = (size_t)1 << 18;  // outputs from which one thread per output fills the chip

template <typename T>
nrk_status dev_axis(int32_t op, const T *x, size_t outer, size_t len, size_t inner, size_t ddof, void *out,
                    hipStream_t s) {
    NRK_TRY(ensure_device());
    if (op < NRK_AXIS_SUM || op > NRK_AXIS_ARGMAX) return fail(NRK_ERR_INVALID_ARGUMENT, "op %d is not an NRK_AXIS_* code", op);
    if (!out && outer * inner > 0) return fail(NRK_ERR_INVALID_ARGUMENT, "out is NULL");
    if (len == 0 && op != NRK_AXIS_SUM && op != NRK_AXIS_MEAN)
        return fail(NRK_ERR_INVALID_ARGUMENT, "reduction along an empty axis");
    if (op == NRK_AXIS_VAR && len <= ddof)
        return fail(NRK_ERR_INVALID_ARGUMENT, "Degrees of freedom %zu >= axis size %zu", ddof, len);
    if (outer * inner == 0) return NRK_OK;
    if (inner > 1) {
        // Enough outputs to fill the chip: one thread per output, left to right in the reference's
        // own order (bit-exact). Fewer: column tiles over ranges of the axis read the data once
        // (compensated sums and Chan variances; extrema and their indices are the same bits).
        if (outer * inner >= AXIS_DIRECT_OUTPUTS || len <= 1)
            return strided_reduce<T>(op, x, outer, len, inner, (double)ddof, out, s);
        if (op == NRK_AXIS_SUM || op == NRK_AXIS_MEAN)
            return axis_sum<T>(x, outer, len, inner, op == NRK_AXIS_MEAN, (T *)out, s);
        if (op == NRK_AXIS_VAR) return axis_var<T>(x, outer, len, inner, (double)ddof, (T *)out, s);
        return axis_extrema<T>(op, x, outer, len, inner, out, s);
    }
    nrk_reduce_outputs o{};
    switch (op) {
        case NRK_AXIS_SUM: o.sum = out; break;
        case NRK_AXIS_MEAN: o.mean = out; break;
        case NRK_AXIS_VAR: o.var = out; break;
        case NRK_AXIS_MIN: o.min = out; break;
        case NRK_AXIS_MAX: o.max = out; break;
        case NRK_AXIS_ARGMIN: o.argmin = (uint64_t *)out; break;
        default: o.argmax = (uint64_t *)out; break;
    }
    Chain c;
    NRK_TRY(make_chain(nullptr, 0, &c));
    return seg_reduce<T>(c, x, outer, len, (double)ddof, &o, nullptr, s);
}

template <typename T>
nrk_status host_axis(int32_t op, const T *x, size_t outer, size_t len, size_t inner, size_t ddof, void *out) {
    const size_t nout = outer * inner;
    const bool idx = (op == NRK_AXIS_ARGMAX || op == NRK_AXIS_ARGMIN);
    if (idx) {
        return host_call<T, uint64_t>(x, outer * len * inner, (uint64_t *)out, nout, [&](const T *d, uint64_t *dout, hipStream_t s) {
            return dev_axis<T>(op, d, outer, len, inner, ddof, dout, s);
        });
    }
    return host_call<T, T>(x, outer * len * inner, (T *)out, nout, [&](const T *d, T *dout, hipStream_t s) {
        return dev_axis<T>(op, d, outer, len, inner, ddof, dout, s);
    });
}

template <typename T>
nrk_status host_map(const nrk_scalar_op *chain, int32_t chain_len, const T *x, T *y, size_t n) {
    Chain c;
    NRK_TRY(make_chain(chain, chain_len, &c, true));
    if (n && !y) return fail(NRK_ERR_INVALID_ARGUMENT, "y is NULL");
    return host_call<T, T>(x, n, y, n, [&](const T *d, T *dout, hipStream_t s) { return map<T>(c, d, dout, n, s); });
}

template <typename T>
nrk_status host_zip(int32_t op, const T *a, const T *b, T *y, size_t n) {
    NRK_TRY(ensure_device());
    if (n && (!a || !b || !y)) return fail(NRK_ERR_INVALID_ARGUMENT, "a, b or y is NULL");
    void *db = nullptr;
    NRK_TRY(staging(2, n * sizeof(T) + 16, &db));
    if (n) NRK_TRY(copy_h2d(db, b, n * sizeof(T), nullptr));
    return host_call<T, T>(a, n, y, n, [&](const T *d, T *dout, hipStream_t s) { return zip<T>(op, d, (const T *)db, dout, n, s); });
}

// fma (op 0) and lerp (op 1) on host slices: a, b (and c) staged, y copied back.
template <typename T>
nrk_status host_tri(int op, const T *a, const T *b, const T *c, T t, T *y, size_t n) {
    NRK_TRY(ensure_device());
    if (n && (!a || !b || !y || (op == 0 && !c))) return fail(NRK_ERR_INVALID_ARGUMENT, "an input or y is NULL");
    void *db = nullptr, *dc = nullptr;
    NRK_TRY(staging(2, n * sizeof(T) + 16, &db));
    NRK_TRY(staging(3, n * sizeof(T) + 16, &dc));
    if (n) NRK_TRY(copy_h2d(db, b, n * sizeof(T), nullptr));
    if (n && op == 0) NRK_TRY(copy_h2d(dc, c, n * sizeof(T), nullptr));
    return host_call<T, T>(a, n, y, n, [&](const T *d, T *dout, hipStream_t s) {
        return tri<T>(op, d, (const T *)db, (const T *)dc, t, dout, n, s);
    });
}

template <typename T>
nrk_status dev_map_reduce(const nrk_scalar_op *chain, int32_t chain_len, const T *x, size_t rows, size_t cols, size_t ddof,
                          const nrk_reduce_outputs *row_out, const nrk_reduce_outputs *full_out, hipStream_t s,
                          const ExtScratch *ext = nullptr) {
    Chain c;
    NRK_TRY(make_chain(chain, chain_len, &c));
    if (row_out && row_out->var && cols <= ddof) return fail(NRK_ERR_INVALID_ARGUMENT, "Degrees of freedom %zu >= axis size %zu", ddof, cols);
    return seg_reduce<T>(c, x, rows, cols, (double)ddof, row_out, full_out, s, ext);
}

template <typename T>
nrk_status dev_map_reduce_plain(const nrk_scalar_op *chain, int32_t chain_len, const T *x, size_t rows, size_t cols, size_t ddof,
                                const nrk_reduce_outputs *row_out, const nrk_reduce_outputs *full_out, hipStream_t s) {
    Chain c;
    NRK_TRY(make_chain(chain, chain_len, &c));
    return plain_reduce<T>(c, x, rows, cols, ddof, row_out, full_out, s);
}

template <typename T>
nrk_status host_outer(int32_t op, const T *a, size_t na, const T *b, size_t nb, T *y) {
    NRK_TRY(ensure_device());
    if (nb && !b) return fail(NRK_ERR_INVALID_ARGUMENT, "b is NULL");
    void *db = nullptr;
    NRK_TRY(staging(2, nb * sizeof(T) + 16, &db));
    if (nb) NRK_TRY(copy_h2d(db, b, nb * sizeof(T), nullptr));
    return host_call<T, T>(a, na, y, na * nb, [&](const T *d, T *o, hipStream_t s) { return outer<T>(op, d, na, (const T *)db, nb, o, s); });
}
template <typename T>
nrk_status host_where(int32_t op, const T *a, const T *b, const uint8_t *mask, T *y, size_t n) {
    NRK_TRY(ensure_device());
    if (n && (!b || !mask)) return fail(NRK_ERR_INVALID_ARGUMENT, "b or mask is NULL");
    void *db = nullptr, *dm = nullptr;
    NRK_TRY(staging(2, n * sizeof(T) + 16, &db));
    NRK_TRY(staging(3, n + 16, &dm));
    if (n) {
        NRK_TRY(copy_h2d(db, b, n * sizeof(T), nullptr));
        NRK_TRY(copy_h2d(dm, mask, n, nullptr));
    }
    return host_call<T, T>(a, n, y, n, [&](const T *d, T *o, hipStream_t s) { return where_op<T>(op, d, (const T *)db, (const uint8_t *)dm, o, n, s); });
}
template <typename T>
nrk_status host_reduceat(int32_t op, const T *x, size_t rows, size_t len, const uint64_t *idx, size_t nidx, T *y) {
    NRK_TRY(ensure_device());
    if (nidx && !idx) return fail(NRK_ERR_INVALID_ARGUMENT, "indices is NULL");
    for (size_t k = 0; k < nidx; ++k)  // ufunc_ops.rs:627-633
        if (idx[k] >= len) return fail(NRK_ERR_INVALID_ARGUMENT, "index %llu out of bounds for reduceat axis of size %zu", (unsigned long long)idx[k], len);
    void *di = nullptr;
    NRK_TRY(staging(2, nidx * sizeof(uint64_t) + 16, &di));
    if (nidx) NRK_TRY(copy_h2d(di, idx, nidx * sizeof(uint64_t), nullptr));
    return host_call<T, T>(x, rows * len, y, rows * nidx, [&](const T *d, T *o, hipStream_t s) {
        return reduceat<T>(op, d, rows, len, (const uint64_t *)di, nidx, o, s);
    });
}

template <typename T>
nrk_status host_at(int32_t op, T *a, size_t n0, size_t row_len, const uint64_t *idx, size_t m, const T *b) {
    NRK_TRY(ensure_device());
    const size_t na = n0 * row_len, nb = m * row_len;
    if ((na && !a) || (m && !idx) || (nb && !b)) return fail(NRK_ERR_INVALID_ARGUMENT, "a, indices or b is NULL");
    for (size_t k = 0; k < m; ++k)  // ufunc_ops.rs:794-800
        if (idx[k] >= n0) return fail(NRK_ERR_INVALID_ARGUMENT, "index %llu out of bounds for axis 0 with size %zu", (unsigned long long)idx[k], n0);
    if (na == 0 || nb == 0) return NRK_OK;
    void *da = nullptr, *di = nullptr, *db = nullptr;
    NRK_TRY(staging(0, na * sizeof(T) + 16, &da));
    NRK_TRY(staging(2, m * sizeof(uint64_t) + 16, &di));
    NRK_TRY(staging(3, nb * sizeof(T) + 16, &db));
    hipStream_t s = nullptr;
    NRK_TRY(copy_h2d(da, a, na * sizeof(T), s));
    NRK_TRY(copy_h2d(di, idx, m * sizeof(uint64_t), s));
    NRK_TRY(copy_h2d(db, b, nb * sizeof(T), s));
    NRK_TRY(ufunc_at<T>(op, (T *)da, n0, row_len, (const uint64_t *)di, m, (const T *)db, s));
    NRK_TRY(copy_d2h(a, da, na * sizeof(T), s));
    NRK_HIP(hipStreamSynchronize(s));
    return NRK_OK;
}

}  // namespace

extern "C" {

// ---- elementwise ----
nrk_status nrk_map_f64(const nrk_scalar_op *chain, int32_t chain_len, const double *x, double *y, size_t n) { return host_map<double>(chain, chain_len, x, y, n); }
nrk_status nrk_map_f32(const nrk_scalar_op *chain, int32_t chain_len, const float *x, float *y, size_t n) { return host_map<float>(chain, chain_len, x, y, n); }
nrk_status nrk_dev_map_f64(const nrk_scalar_op *chain, int32_t chain_len, const double *x, double *y, size_t n, nrk_stream stream) {
    Chain c;
    NRK_TRY(make_chain(chain, chain_len, &c, true));
    return map<double>(c, x, y, n, S(stream));
}
nrk_status nrk_dev_map_f32(const nrk_scalar_op *chain, int32_t chain_len, const float *x, float *y, size_t n, nrk_stream stream) {
    Chain c;
    NRK_TRY(make_chain(chain, chain_len, &c, true));
    return map<float>(c, x, y, n, S(stream));
}
nrk_status nrk_zip_f64(int32_t op, const double *a, const double *b, double *y, size_t n) { return host_zip<double>(op, a, b, y, n); }
nrk_status nrk_zip_f32(int32_t op, const float *a, const float *b, float *y, size_t n) { return host_zip<float>(op, a, b, y, n); }
nrk_status nrk_dev_zip_f64(int32_t op, const double *a, const double *b, double *y, size_t n, nrk_stream stream) { return zip<double>(op, a, b, y, n, S(stream)); }
nrk_status nrk_dev_zip_f32(int32_t op, const float *a, const float *b, float *y, size_t n, nrk_stream stream) { return zip<float>(op, a, b, y, n, S(stream)); }

// ---- outer, where (ufunc_ops.rs:583, 841) ----
nrk_status nrk_outer_f64(int32_t op, const double *a, size_t na, const double *b, size_t nb, double *y) { return host_outer<double>(op, a, na, b, nb, y); }
nrk_status nrk_outer_f32(int32_t op, const float *a, size_t na, const float *b, size_t nb, float *y) { return host_outer<float>(op, a, na, b, nb, y); }
nrk_status nrk_dev_outer_f64(int32_t op, const double *a, size_t na, const double *b, size_t nb, double *y, nrk_stream stream) { return outer<double>(op, a, na, b, nb, y, S(stream)); }
nrk_status nrk_dev_outer_f32(int32_t op, const float *a, size_t na, const float *b, size_t nb, float *y, nrk_stream stream) { return outer<float>(op, a, na, b, nb, y, S(stream)); }
nrk_status nrk_where_f64(int32_t op, const double *a, const double *b, const uint8_t *mask, double *y, size_t n) { return host_where<double>(op, a, b, mask, y, n); }
nrk_status nrk_where_f32(int32_t op, const float *a, const float *b, const uint8_t *mask, float *y, size_t n) { return host_where<float>(op, a, b, mask, y, n); }
nrk_status nrk_dev_where_f64(int32_t op, const double *a, const double *b, const uint8_t *mask, double *y, size_t n, nrk_stream stream) { return where_op<double>(op, a, b, mask, y, n, S(stream)); }
nrk_status nrk_dev_where_f32(int32_t op, const float *a, const float *b, const uint8_t *mask, float *y, size_t n, nrk_stream stream) { return where_op<float>(op, a, b, mask, y, n, S(stream)); }

// ---- at (ufunc_ops.rs:772) ----
nrk_status nrk_at_f64(int32_t op, double *a, size_t n0, size_t row_len, const uint64_t *indices, size_t m, const double *b) { return host_at<double>(op, a, n0, row_len, indices, m, b); }
nrk_status nrk_at_f32(int32_t op, float *a, size_t n0, size_t row_len, const uint64_t *indices, size_t m, const float *b) { return host_at<float>(op, a, n0, row_len, indices, m, b); }
nrk_status nrk_dev_at_f64(int32_t op, double *a, size_t n0, size_t row_len, const uint64_t *indices, size_t m, const double *b, nrk_stream stream) { return ufunc_at<double>(op, a, n0, row_len, indices, m, b, S(stream)); }
nrk_status nrk_dev_at_f32(int32_t op, float *a, size_t n0, size_t row_len, const uint64_t *indices, size_t m, const float *b, nrk_stream stream) { return ufunc_at<float>(op, a, n0, row_len, indices, m, b, S(stream)); }

// ---- reduceat (ufunc_ops.rs:713) ----
nrk_status nrk_reduceat_f64(int32_t op, const double *x, size_t rows, size_t len, const uint64_t *indices, size_t nidx, double *y) { return host_reduceat<double>(op, x, rows, len, indices, nidx, y); }
nrk_status nrk_reduceat_f32(int32_t op, const float *x, size_t rows, size_t len, const uint64_t *indices, size_t nidx, float *y) { return host_reduceat<float>(op, x, rows, len, indices, nidx, y); }
nrk_status nrk_dev_reduceat_f64(int32_t op, const double *x, size_t rows, size_t len, const uint64_t *indices, size_t nidx, double *y, nrk_stream stream) { return reduceat<double>(op, x, rows, len, indices, nidx, y, S(stream)); }
nrk_status nrk_dev_reduceat_f32(int32_t op, const float *x, size_t rows, size_t len, const uint64_t *indices, size_t nidx, float *y, nrk_stream stream) { return reduceat<float>(op, x, rows, len, indices, nidx, y, S(stream)); }

// ---- accumulate (ufunc_ops.rs:535) ----
nrk_status nrk_accumulate_f64(int32_t op, const double *x, size_t rows, size_t len, double *y) {
    return host_call<double, double>(x, rows * len, y, rows * len, [&](const double *d, double *o, hipStream_t s) { return accumulate<double>(op, d, o, rows, len, s); });
}
nrk_status nrk_accumulate_f32(int32_t op, const float *x, size_t rows, size_t len, float *y) {
    return host_call<float, float>(x, rows * len, y, rows * len, [&](const float *d, float *o, hipStream_t s) { return accumulate<float>(op, d, o, rows, len, s); });
}
nrk_status nrk_dev_accumulate_f64(int32_t op, const double *x, size_t rows, size_t len, double *y, nrk_stream stream) { return accumulate<double>(op, x, y, rows, len, S(stream)); }
nrk_status nrk_dev_accumulate_f32(int32_t op, const float *x, size_t rows, size_t len, float *y, nrk_stream stream) { return accumulate<float>(op, x, y, rows, len, S(stream)); }

// ---- softmax, log_softmax ----
nrk_status nrk_softmax_f64(const double *x, double *y, size_t n) {
    return host_call<double, double>(x, n, y, n, [&](const double *d, double *o, hipStream_t s) { return dev_softmax<double>(false, d, o, n, s); });
}
nrk_status nrk_softmax_f32(const float *x, float *y, size_t n) {
    return host_call<float, float>(x, n, y, n, [&](const float *d, float *o, hipStream_t s) { return dev_softmax<float>(false, d, o, n, s); });
}
nrk_status nrk_log_softmax_f64(const double *x, double *y, size_t n) {
    return host_call<double, double>(x, n, y, n, [&](const double *d, double *o, hipStream_t s) { return dev_softmax<double>(true, d, o, n, s); });
}
nrk_status nrk_log_softmax_f32(const float *x, float *y, size_t n) {
    return host_call<float, float>(x, n, y, n, [&](const float *d, float *o, hipStream_t s) { return dev_softmax<float>(true, d, o, n, s); });
}
nrk_status nrk_dev_softmax_f64(const double *x, double *y, size_t n, nrk_stream stream) { return dev_softmax<double>(false, x, y, n, S(stream)); }
nrk_status nrk_dev_softmax_f32(const float *x, float *y, size_t n, nrk_stream stream) { return dev_softmax<float>(false, x, y, n, S(stream)); }
nrk_status nrk_dev_log_softmax_f64(const double *x, double *y, size_t n, nrk_stream stream) { return dev_softmax<double>(true, x, y, n, S(stream)); }
nrk_status nrk_dev_log_softmax_f32(const float *x, float *y, size_t n, nrk_stream stream) { return dev_softmax<float>(true, x, y, n, S(stream)); }

// ---- fma, lerp ----
nrk_status nrk_fma_f64(const double *a, const double *b, const double *c, double *y, size_t n) { return host_tri<double>(0, a, b, c, 0.0, y, n); }
nrk_status nrk_fma_f32(const float *a, const float *b, const float *c, float *y, size_t n) { return host_tri<float>(0, a, b, c, 0.0f, y, n); }
nrk_status nrk_dev_fma_f64(const double *a, const double *b, const double *c, double *y, size_t n, nrk_stream stream) { return tri<double>(0, a, b, c, 0.0, y, n, S(stream)); }
nrk_status nrk_dev_fma_f32(const float *a, const float *b, const float *c, float *y, size_t n, nrk_stream stream) { return tri<float>(0, a, b, c, 0.0f, y, n, S(stream)); }
nrk_status nrk_lerp_f64(const double *a, const double *b, double t, double *y, size_t n) { return host_tri<double>(1, a, b, nullptr, t, y, n); }
nrk_status nrk_lerp_f32(const float *a, const float *b, float t, float *y, size_t n) { return host_tri<float>(1, a, b, nullptr, t, y, n); }
nrk_status nrk_dev_lerp_f64(const double *a, const double *b, double t, double *y, size_t n, nrk_stream stream) { return tri<double>(1, a, b, nullptr, t, y, n, S(stream)); }
nrk_status nrk_dev_lerp_f32(const float *a, const float *b, float t, float *y, size_t n, nrk_stream stream) { return tri<float>(1, a, b, nullptr, t, y, n, S(stream)); }

// ---- flat reductions, host slices ----
#define NRK_HOST_REDUCE(NAME, T, OUT_T, CALL)                                                  \
    nrk_status NAME(const T *a, size_t n, OUT_T *out) {                                        \
        return host_call<T, OUT_T>(a, n, out, 1, [&](const T *d, OUT_T *dout, hipStream_t s) { \
            return CALL(d, n, dout, s);                                                        \
        });                                                                                    \
    }
NRK_HOST_REDUCE(nrk_sum_f64, double, double, dev_sum<double>)
NRK_HOST_REDUCE(nrk_sum_f32, float, float, dev_sum<float>)
NRK_HOST_REDUCE(nrk_mean_f64, double, double, dev_mean<double>)
NRK_HOST_REDUCE(nrk_mean_f32, float, float, dev_mean<float>)
NRK_HOST_REDUCE(nrk_min_f64, double, double, dev_min<double>)
NRK_HOST_REDUCE(nrk_min_f32, float, float, dev_min<float>)
NRK_HOST_REDUCE(nrk_max_f64, double, double, dev_max<double>)
NRK_HOST_REDUCE(nrk_max_f32, float, float, dev_max<float>)
NRK_HOST_REDUCE(nrk_prod_f64, double, double, prod<double>)
NRK_HOST_REDUCE(nrk_prod_f32, float, float, prod<float>)
NRK_HOST_REDUCE(nrk_sum_i64, int64_t, int64_t, sum_i64)
NRK_HOST_REDUCE(nrk_argmax_f64, double, uint64_t, dev_argmax<double>)
NRK_HOST_REDUCE(nrk_argmax_f32, float, uint64_t, dev_argmax<float>)
NRK_HOST_REDUCE(nrk_argmin_f64, double, uint64_t, dev_argmin<double>)
NRK_HOST_REDUCE(nrk_argmin_f32, float, uint64_t, dev_argmin<float>)
#undef NRK_HOST_REDUCE

nrk_status nrk_var_f64(const double *a, size_t n, size_t ddof, double *out) {
    return host_call<double, double>(a, n, out, 1, [&](const double *d, double *o, hipStream_t s) { return dev_var<double>(d, n, ddof, o, s); });
}
nrk_status nrk_var_f32(const float *a, size_t n, size_t ddof, float *out) {
    return host_call<float, float>(a, n, out, 1, [&](const float *d, float *o, hipStream_t s) { return dev_var<float>(d, n, ddof, o, s); });
}
nrk_status nrk_sum_sq_dev_f64(const double *a, size_t n, double mean, double *out) {
    return host_call<double, double>(a, n, out, 1, [&](const double *d, double *o, hipStream_t s) { return dev_sum_sq_dev<double>(d, n, mean, o, s); });
}
nrk_status nrk_sum_sq_dev_f32(const float *a, size_t n, float mean, float *out) {
    return host_call<float, float>(a, n, out, 1, [&](const float *d, float *o, hipStream_t s) { return dev_sum_sq_dev<float>(d, n, mean, o, s); });
}

// ---- ufuncs.rs reductions: dot, norms, std ----
nrk_status nrk_dot_f64(const double *a, const double *b, size_t n, double *out) { return host_dot<double>(a, b, n, out); }
nrk_status nrk_dot_f32(const float *a, const float *b, size_t n, float *out) { return host_dot<float>(a, b, n, out); }
nrk_status nrk_dev_dot_f64(const double *a, const double *b, size_t n, double *out, nrk_stream stream) { return dot<double>(a, b, n, out, S(stream)); }
nrk_status nrk_dev_dot_f32(const float *a, const float *b, size_t n, float *out, nrk_stream stream) { return dot<float>(a, b, n, out, S(stream)); }
#define NRK_NORM(NAME, T, SOP, ROOT)                                                                              \
    nrk_status nrk_##NAME(const T *a, size_t n, T *out) {                                                         \
        return host_call<T, T>(a, n, out, 1, [&](const T *d, T *o, hipStream_t s) { return dev_norm<T>(SOP, ROOT, d, n, o, s); }); \
    }                                                                                                             \
    nrk_status nrk_dev_##NAME(const T *a, size_t n, T *out, nrk_stream stream) { return dev_norm<T>(SOP, ROOT, a, n, out, S(stream)); }
NRK_NORM(norm_l1_f64, double, NRK_SOP_ABS, false)
NRK_NORM(norm_l1_f32, float, NRK_SOP_ABS, false)
NRK_NORM(norm_l2_f64, double, NRK_SOP_SQUARE, true)
NRK_NORM(norm_l2_f32, float, NRK_SOP_SQUARE, true)
#undef NRK_NORM
nrk_status nrk_std_f64(const double *a, size_t n, size_t ddof, double *out) {
    return host_call<double, double>(a, n, out, 1, [&](const double *d, double *o, hipStream_t s) { return dev_std<double>(d, n, ddof, o, s); });
}
nrk_status nrk_std_f32(const float *a, size_t n, size_t ddof, float *out) {
    return host_call<float, float>(a, n, out, 1, [&](const float *d, float *o, hipStream_t s) { return dev_std<float>(d, n, ddof, o, s); });
}
nrk_status nrk_dev_std_f64(const double *a, size_t n, size_t ddof, double *out, nrk_stream stream) { return dev_std<double>(a, n, ddof, out, S(stream)); }
nrk_status nrk_dev_std_f32(const float *a, size_t n, size_t ddof, float *out, nrk_stream stream) { return dev_std<float>(a, n, ddof, out, S(stream)); }

// ---- flat reductions, device pointers ----
nrk_status nrk_dev_sum_f64(const double *a, size_t n, double *out, nrk_stream stream) { return dev_sum<double>(a, n, out, S(stream)); }
nrk_status nrk_dev_sum_f32(const float *a, size_t n, float *out, nrk_stream stream) { return dev_sum<float>(a, n, out, S(stream)); }
nrk_status nrk_dev_mean_f64(const double *a, size_t n, double *out, nrk_stream stream) { return dev_mean<double>(a, n, out, S(stream)); }
nrk_status nrk_dev_mean_f32(const float *a, size_t n, float *out, nrk_stream stream) { return dev_mean<float>(a, n, out, S(stream)); }
nrk_status nrk_dev_var_f64(const double *a, size_t n, size_t ddof, double *out, nrk_stream stream) { return dev_var<double>(a, n, ddof, out, S(stream)); }
nrk_status nrk_dev_var_f32(const float *a, size_t n, size_t ddof, float *out, nrk_stream stream) { return dev_var<float>(a, n, ddof, out, S(stream)); }
nrk_status nrk_dev_min_f64(const double *a, size_t n, double *out, nrk_stream stream) { return dev_min<double>(a, n, out, S(stream)); }
nrk_status nrk_dev_min_f32(const float *a, size_t n, float *out, nrk_stream stream) { return dev_min<float>(a, n, out, S(stream)); }
nrk_status nrk_dev_max_f64(const double *a, size_t n, double *out, nrk_stream stream) { return dev_max<double>(a, n, out, S(stream)); }
nrk_status nrk_dev_max_f32(const float *a, size_t n, float *out, nrk_stream stream) { return dev_max<float>(a, n, out, S(stream)); }
nrk_status nrk_dev_prod_f64(const double *a, size_t n, double *out, nrk_stream stream) { return prod<double>(a, n, out, S(stream)); }
nrk_status nrk_dev_prod_f32(const float *a, size_t n, float *out, nrk_stream stream) { return prod<float>(a, n, out, S(stream)); }
nrk_status nrk_dev_sum_i64(const int64_t *a, size_t n, int64_t *out, nrk_stream stream) { return sum_i64(a, n, out, S(stream)); }
nrk_status nrk_dev_argmax_f64(const double *a, size_t n, uint64_t *out, nrk_stream stream) { return dev_argmax<double>(a, n, out, S(stream)); }
nrk_status nrk_dev_argmax_f32(const float *a, size_t n, uint64_t *out, nrk_stream stream) { return dev_argmax<float>(a, n, out, S(stream)); }
nrk_status nrk_dev_argmin_f64(const double *a, size_t n, uint64_t *out, nrk_stream stream) { return dev_argmin<double>(a, n, out, S(stream)); }
nrk_status nrk_dev_argmin_f32(const float *a, size_t n, uint64_t *out, nrk_stream stream) { return dev_argmin<float>(a, n, out, S(stream)); }

// ---- axis reductions ----
nrk_status nrk_reduce_axis_f64(int32_t op, const double *x, size_t outer, size_t len, size_t inner, size_t ddof, void *out) { return host_axis<double>(op, x, outer, len, inner, ddof, out); }
nrk_status nrk_reduce_axis_f32(int32_t op, const float *x, size_t outer, size_t len, size_t inner, size_t ddof, void *out) { return host_axis<float>(op, x, outer, len, inner, ddof, out); }
nrk_status nrk_dev_reduce_axis_f64(int32_t op, const double *x, size_t outer, size_t len, size_t inner, size_t ddof, void *out, nrk_stream stream) { return dev_axis<double>(op, x, outer, len, inner, ddof, out, S(stream)); }
nrk_status nrk_dev_reduce_axis_f32(int32_t op, const float *x, size_t outer, size_t len, size_t inner, size_t ddof, void *out, nrk_stream stream) { return dev_axis<float>(op, x, outer, len, inner, ddof, out, S(stream)); }

// ---- fused map -> reduce ----
nrk_status nrk_dev_map_reduce_f64(const nrk_scalar_op *chain, int32_t chain_len, const double *x, size_t rows, size_t cols, size_t ddof,
                                  const nrk_reduce_outputs *row_out, const nrk_reduce_outputs *full_out, nrk_stream stream) {
    return dev_map_reduce<double>(chain, chain_len, x, rows, cols, ddof, row_out, full_out, S(stream));
}
nrk_status nrk_dev_map_reduce_f32(const nrk_scalar_op *chain, int32_t chain_len, const float *x, size_t rows, size_t cols, size_t ddof,
                                  const nrk_reduce_outputs *row_out, const nrk_reduce_outputs *full_out, nrk_stream stream) {
    return dev_map_reduce<float>(chain, chain_len, x, rows, cols, ddof, row_out, full_out, S(stream));
}
nrk_status nrk_dev_map_reduce_plain_f64(const nrk_scalar_op *chain, int32_t chain_len, const double *x, size_t rows, size_t cols,
                                        size_t ddof, const nrk_reduce_outputs *row_out, const nrk_reduce_outputs *full_out,
                                        nrk_stream stream) {
    return dev_map_reduce_plain<double>(chain, chain_len, x, rows, cols, ddof, row_out, full_out, S(stream));
}
nrk_status nrk_dev_map_reduce_plain_f32(const nrk_scalar_op *chain, int32_t chain_len, const float *x, size_t rows, size_t cols,
                                        size_t ddof, const nrk_reduce_outputs *row_out, const nrk_reduce_outputs *full_out,
                                        nrk_stream stream) {
    return dev_map_reduce_plain<float>(chain, chain_len, x, rows, cols, ddof, row_out, full_out, S(stream));
}
nrk_status nrk_dev_expr_reduce_f64(const nrk_expr_node *nodes, int32_t len, const double *const *leaves, int32_t n_leaves,
                                   size_t rows, size_t cols, size_t ddof, const nrk_reduce_outputs *row_out,
                                   const nrk_reduce_outputs *full_out, nrk_stream stream) {
    ExprProg pg;
    NRK_TRY(compile_expr(nodes, len, &pg));
    if (n_leaves < pg.n_leaves) return fail(NRK_ERR_INVALID_ARGUMENT, "the expression reads %d operands, %d given", pg.n_leaves, n_leaves);
    return expr_reduce<double>(pg, leaves, rows, cols, (double)ddof, row_out, full_out, S(stream));
}
nrk_status nrk_dev_expr_reduce_f32(const nrk_expr_node *nodes, int32_t len, const float *const *leaves, int32_t n_leaves,
                                   size_t rows, size_t cols, size_t ddof, const nrk_reduce_outputs *row_out,
                                   const nrk_reduce_outputs *full_out, nrk_stream stream) {
    ExprProg pg;
    NRK_TRY(compile_expr(nodes, len, &pg));
    if (n_leaves < pg.n_leaves) return fail(NRK_ERR_INVALID_ARGUMENT, "the expression reads %d operands, %d given", pg.n_leaves, n_leaves);
    return expr_reduce<float>(pg, leaves, rows, cols, (double)ddof, row_out, full_out, S(stream));
}
// The same kernels on the device's fixed scratch: no hipMalloc / hipFree, ever.
nrk_status nrk_dev_map_reduce_f64_noalloc(const nrk_scalar_op *chain, int32_t chain_len, const double *x, size_t rows,
                                          size_t cols, size_t ddof, const nrk_reduce_outputs *row_out,
                                          const nrk_reduce_outputs *full_out, nrk_stream stream) {
    ExtScratch ext;
    NRK_TRY(static_scratch(&ext));
    return dev_map_reduce<double>(chain, chain_len, x, rows, cols, ddof, row_out, full_out, S(stream), &ext);
}

}  // extern "C"
