// Reductions in the reference's own summation order ("plain" mode): every sum is the sequence of
// rounded additions the reference's CPU kernels perform, so the results are bit-identical to
// them (and to each other under any launch shape), at the accuracy of plain recursive summation.
//
// The orders (kernels/reduce.rs:204-314, kernels/mod.rs:74-92, math/aggregation.rs:548-561):
//   - whole-array sum, n < 10000: simd_sum of everything: four interleaved partial sums (element i
//     to partial i % 4) over the largest multiple of 4, combined (p0 + p1) + (p2 + p3), then the
//     last 0..3 elements added left to right;
//   - whole-array sum, n >= 10000: simd_sum of each 8192-element chunk, the chunk sums folded
//     left to right from zero; mean = sum / n;
//   - whole-array var, n < 40000: mean = simd_sum(all) / n, then sum((x - mean)^2) left to right
//     over everything; n >= 40000: mean from the chunked sum, then per-chunk left-to-right sums of
//     (x - mean)^2 folded left to right; var = that / (n - ddof);
//   - along axis 1 (rows): a left-to-right sum of each row from zero; mean = sum / cols;
//     var = (left-to-right sum of (x - mean)^2) / (cols - ddof).
// Every product and difference is rounded on its own (no contraction).
//
// One kernel covers all of these: a "fold" over rows of `len` elements (the chunks, the matrix
// rows, or the whole array as one row), where each row is reduced by Q = 1 or 4 sequential
// chains, each chain one lane (a chain is a dependent sequence of additions and cannot be split).
// A workgroup owns rpb rows. Its 256 threads stage the rows through LDS in 32 KiB slices (16-byte
// coalesced loads, the element map applied on the way in, the next slice's loads in flight while
// the chain lanes fold the current one); the LDS layout puts each chain's elements contiguously,
// so a chain lane reads four of them per LDS access.
#include "ops.hpp"

namespace nrk {
namespace {

constexpr int PT = 256;                           // threads per workgroup
constexpr size_t PLAIN_MIN_LEN = 10000;           // kernels/mod.rs:74
constexpr size_t PLAIN_CHUNK = 8192;              // kernels/mod.rs:92
constexpr size_t PLAIN_VAR_MIN_LEN = 40000;       // kernels/reduce.rs:202
constexpr size_t SLICE_BYTES = 32768;             // one LDS slice: rpb rows x k elements
constexpr int PAD = 4;                            // elements of padding after each chain's run

enum { PL_SUM = 0, PL_SSD = 1 };

template <typename T>
struct PlainParams {
    const T *x;         // row r is x[r len, min((r + 1) len, n))
    size_t n, len, rows;
    int rpb, k;         // rows per workgroup; elements per row per slice (a multiple of 16)
    Chain chain;        // applied to every element before it is folded (len 0: none)
    const T *mean;      // PL_SSD: mean of row r at mean[r * mean_stride]
    size_t mean_stride;
    T *o_sum;           // per row: the fold itself (the raw sum of squares for PL_SSD)
    T *o_mean, *o_mean2;  // per row: fold / mean_div (two destinations, either may be NULL)
    T *o_var;           // per row: fold / var_div
    T mean_div, var_div;
};

template <typename T>
__device__ __forceinline__ T fold_one(T a, T y, T m, int op) {
    if (op == PL_SUM) return rn_add(a, y);
    const T d = rn_sub(y, m);
    return rn_add(a, rn_mul(d, d));
}

// LDS offset of position p (slice-relative) of row rr: chain (rr, p % Q) holds its elements at
// consecutive offsets, each chain's run padded so that runs start on different banks.
template <int Q>
__device__ __forceinline__ int lds_at(int rr, int p, int kq) {
    return (rr * Q + p % Q) * (kq + PAD) + p / Q;
}

template <typename T, int Q, int OP, bool VEC>
__global__ void __launch_bounds__(PT) plain_fold_kernel(PlainParams<T> p) {
    extern __shared__ __align__(16) unsigned char plain_lds_raw[];
    T *lds = reinterpret_cast<T *>(plain_lds_raw);
    using V = typename Vec16<T>::type;
    constexpr int VN = Vec16<T>::N;
    constexpr int SLICE = (int)(SLICE_BYTES / sizeof(T));
    constexpr int LV = SLICE / VN / PT;  // 16-byte vectors per thread per slice
    constexpr int LS = SLICE / PT;       // elements per thread per slice (scalar loads)
    const int t = threadIdx.x;
    const size_t r0 = (size_t)blockIdx.x * p.rpb;
    const int nr = (int)min((size_t)p.rpb, p.rows - r0);
    const int k = p.k, kq = k / Q;

    // this lane's chain: row rr of the workgroup, residue q
    const int rr = t / Q, q = t % Q;
    const bool chain_lane = t < nr * Q;
    const size_t row = r0 + rr;
    const size_t len_r = chain_lane ? min(p.len, p.n - row * p.len) : 0;
    const size_t nfold = Q == 4 ? (len_r & ~(size_t)3) : len_r;  // positions the chains take
    const T m = (OP == PL_SSD && chain_lane) ? p.mean[row * p.mean_stride] : (T)0;
    T acc = 0;

    const size_t nslice = (p.len + k - 1) / k;
    const int vpr = k / VN;  // vectors per row per slice
    V vr[VEC ? LV : 1];
    T sr[VEC ? 1 : LS];

    auto row_len = [&](int r) -> size_t { return min(p.len, p.n - (r0 + r) * p.len); };
    auto load = [&](size_t s) {
        if constexpr (VEC) {
#pragma unroll
            for (int j = 0; j < LV; ++j) {
                const int v = t + j * PT;
                const int r = v / vpr, c = (v % vpr) * VN;
                vr[j] = V{};
                if (r < nr) {
                    const size_t at = s * k + c, lr = row_len(r);
                    const T *src = p.x + (r0 + r) * p.len + at;
                    if (at + VN <= lr) {
                        vr[j] = __builtin_nontemporal_load(reinterpret_cast<const V *>(src));
                    } else {
#pragma unroll
                        for (int e = 0; e < VN; ++e)
                            if (at + e < lr) vr[j][e] = src[e];
                    }
                }
            }
        } else {
#pragma unroll
            for (int j = 0; j < LS; ++j) {
                const int v = t + j * PT;
                const int r = v / k, c = v % k;
                sr[j] = 0;
                if (r < nr) {
                    const size_t at = s * k + c;
                    if (at < row_len(r)) sr[j] = p.x[(r0 + r) * p.len + at];
                }
            }
        }
    };
    auto stage = [&]() {  // map, then write into the chains' runs
        if constexpr (VEC) {
            T e[LV * VN];  // one pass over the chain's steps for all of them
#pragma unroll
            for (int j = 0; j < LV; ++j)
#pragma unroll
                for (int i = 0; i < VN; ++i) e[j * VN + i] = vr[j][i];
            apply_chain_n<T, LV * VN, MATH_NONE>(p.chain, e);
#pragma unroll
            for (int j = 0; j < LV; ++j) {
                const int v = t + j * PT;
                const int r = v / vpr, c = (v % vpr) * VN;
                if (r < nr) {
                    if constexpr (Q == 1) {
                        V w;
#pragma unroll
                        for (int i = 0; i < VN; ++i) w[i] = e[j * VN + i];
                        *reinterpret_cast<V *>(lds + lds_at<1>(r, c, kq)) = w;
                    } else {
#pragma unroll
                        for (int i = 0; i < VN; ++i) lds[lds_at<Q>(r, c + i, kq)] = e[j * VN + i];
                    }
                }
            }
        } else {
            apply_chain_n<T, LS, MATH_NONE>(p.chain, sr);
#pragma unroll
            for (int j = 0; j < LS; ++j) {
                const int v = t + j * PT;
                const int r = v / k, c = v % k;
                if (r < nr) lds[lds_at<Q>(r, c, kq)] = sr[j];
            }
        }
    };

    load(0);
    for (size_t s = 0; s < nslice; ++s) {
        stage();
        __syncthreads();
        if (s + 1 < nslice) load(s + 1);  // in flight while the chains fold this slice
        if (chain_lane) {
            const size_t lo = s * k;
            // this chain's elements in the slice: positions lo + q + Q i < nfold, i < kq
            const size_t lim = nfold > lo ? nfold - lo : 0;
            const int cnt = (int)min((size_t)kq, Q == 4 ? lim / 4 : lim);
            const T *run = lds + (rr * Q + q) * (kq + PAD);
            int i = 0;
            for (; i + 4 <= cnt; i += 4) {
                const V a = *reinterpret_cast<const V *>(run + i);
                if constexpr (VN == 4) {
                    acc = fold_one(acc, a[0], m, OP);
                    acc = fold_one(acc, a[1], m, OP);
                    acc = fold_one(acc, a[2], m, OP);
                    acc = fold_one(acc, a[3], m, OP);
                } else {
                    const V b = *reinterpret_cast<const V *>(run + i + 2);
                    acc = fold_one(acc, a[0], m, OP);
                    acc = fold_one(acc, a[1], m, OP);
                    acc = fold_one(acc, b[0], m, OP);
                    acc = fold_one(acc, b[1], m, OP);
                }
            }
            for (; i < cnt; ++i) acc = fold_one(acc, run[i], m, OP);
        }
        __syncthreads();
    }

    T res = acc;
    if constexpr (Q == 4) {  // (p0 + p1) + (p2 + p3): the four chains of a row are adjacent lanes
        const T h = rn_add(acc, __shfl_xor(acc, 1));  // lanes 0 and 2: p0 + p1, p2 + p3
        res = rn_add(h, __shfl_xor(h, 2));
        if (chain_lane && q == 0) {  // the last 0..3 elements, left to right
            for (size_t i = nfold; i < len_r; ++i) {
                T e[1] = {p.x[row * p.len + i]};
                apply_chain_n<T, 1, MATH_NONE>(p.chain, e);
                res = rn_add(res, e[0]);
            }
        }
    }
    if (chain_lane && q == 0) {
        if (p.o_sum) p.o_sum[row] = res;
        if (p.o_mean) p.o_mean[row] = rn_div(res, p.mean_div);
        if (p.o_mean2) p.o_mean2[row] = rn_div(res, p.mean_div);
        if (p.o_var) p.o_var[row] = rn_div(res, p.var_div);
    }
}

// Launch one fold over `rows` rows of `len` elements (row r: x[r len, min((r + 1) len, n))).
template <typename T, int Q, int OP>
nrk_status launch_fold(PlainParams<T> p, hipStream_t stream) {
    if (p.rows == 0) return NRK_OK;
    constexpr size_t SLICE = SLICE_BYTES / sizeof(T);
    // rows per workgroup: about 512 workgroups when there are enough rows, at most one chain per
    // thread; then the slice width k (a multiple of 16, at most the row rounded up)
    p.rpb = (int)std::max<size_t>(1, std::min<size_t>(PT / Q, p.rows / 512));
    size_t k = (SLICE / p.rpb) / 16 * 16;
    k = std::min(k, align_up(std::max<size_t>(p.len, 1), 16));
    p.k = (int)k;
    const size_t lds = (size_t)p.rpb * Q * (k / Q + PAD) * sizeof(T);
    const bool vec = (reinterpret_cast<uintptr_t>(p.x) % 16 == 0) && (p.rows == 1 || (p.len * sizeof(T)) % 16 == 0);
    const dim3 grid((unsigned)ceil_div(p.rows, (size_t)p.rpb));
    if (vec)
        hipLaunchKernelGGL((plain_fold_kernel<T, Q, OP, true>), grid, dim3(PT), lds, stream, p);
    else
        hipLaunchKernelGGL((plain_fold_kernel<T, Q, OP, false>), grid, dim3(PT), lds, stream, p);
    NRK_HIP(hipGetLastError());
    return NRK_OK;
}

template <typename T>
PlainParams<T> fold_params(const T *x, size_t n, size_t len, const Chain *chain) {
    PlainParams<T> p{};
    p.x = x;
    p.n = n;
    p.len = len;
    p.rows = len ? ceil_div(n, len) : 0;
    if (chain) p.chain = *chain;
    p.mean_div = p.var_div = (T)1;
    return p;
}

}  // namespace

template <typename T>
nrk_status plain_reduce(const Chain &chain, const T *x, size_t rows, size_t cols, size_t ddof,
                        const nrk_reduce_outputs *row_out, const nrk_reduce_outputs *full_out, hipStream_t stream) {
    NRK_TRY(ensure_device());
    const size_t n = rows * cols;
    if (row_out && row_out->var && cols <= ddof) return fail(NRK_ERR_INVALID_ARGUMENT, "Degrees of freedom %zu >= axis size %zu", ddof, cols);
    // the extrema and their indices do not depend on an order of additions: the fused kernel's
    // are the reference's bits already
    nrk_reduce_outputs rx{}, fx{};
    if (row_out) rx = nrk_reduce_outputs{nullptr, nullptr, nullptr, row_out->min, row_out->max, row_out->argmin, row_out->argmax};
    if (full_out) fx = nrk_reduce_outputs{nullptr, nullptr, nullptr, full_out->min, full_out->max, full_out->argmin, full_out->argmax};
    const bool row_x = row_out && (rx.min || rx.max || rx.argmin || rx.argmax);
    const bool full_x = full_out && (fx.min || fx.max || fx.argmin || fx.argmax);
    if (n == 0 || row_x || full_x) {
        // empty: every output is 0 (kernels/reduce.rs:207-421), as the fused path writes them
        nrk_reduce_outputs re = n == 0 && row_out ? *row_out : rx, fe = n == 0 && full_out ? *full_out : fx;
        NRK_TRY(seg_reduce<T>(chain, x, rows, cols, (double)ddof, (row_out && (n == 0 || row_x)) ? &re : nullptr,
                              (full_out && (n == 0 || full_x)) ? &fe : nullptr, stream));
        if (n == 0) return NRK_OK;
    }
    if (!x) return fail(NRK_ERR_INVALID_ARGUMENT, "x is NULL");
    T *r_sum = row_out ? (T *)row_out->sum : nullptr, *r_mean = row_out ? (T *)row_out->mean : nullptr;
    T *r_var = row_out ? (T *)row_out->var : nullptr;
    T *f_sum = full_out ? (T *)full_out->sum : nullptr, *f_mean = full_out ? (T *)full_out->mean : nullptr;
    T *f_var = full_out ? (T *)full_out->var : nullptr;
    const bool want_full = f_sum || f_mean || f_var;
    const bool want_row = r_sum || r_mean || r_var;
    if (!want_full && !want_row) return NRK_OK;

    // workspace: chunk partials, the whole-array mean, the row means
    const size_t nc = ceil_div(n, PLAIN_CHUNK);
    const size_t parts_at = 0, mean_at = align_up(nc * sizeof(T), 256), rmean_at = mean_at + 256;
    const size_t ws_bytes = rmean_at + (r_var ? rows * sizeof(T) : 0);
    void *ws = nullptr;
    NRK_TRY(get_workspace(stream, ws_bytes, &ws));
    T *parts = (T *)((char *)ws + parts_at);
    T *wmean = (T *)((char *)ws + mean_at);
    T *rmean = (T *)((char *)ws + rmean_at);

    if (want_full) {
        const T n_t = (T)n;
        const bool var_chunked = n >= PLAIN_VAR_MIN_LEN;
        // the sum and mean; below 10000 elements, and from 40000 on, also the variance's mean
        bool have_mean = false;
        if (f_sum || f_mean || (f_var && var_chunked)) {
            if (n < PLAIN_MIN_LEN) {
                PlainParams<T> p = fold_params<T>(x, n, n, &chain);
                p.o_sum = f_sum;
                p.o_mean = f_mean;
                p.o_mean2 = f_var ? wmean : nullptr;
                have_mean = f_var != nullptr;
                p.mean_div = n_t;
                NRK_TRY((launch_fold<T, 4, PL_SUM>(p, stream)));
            } else {
                PlainParams<T> p = fold_params<T>(x, n, PLAIN_CHUNK, &chain);
                p.o_sum = parts;
                NRK_TRY((launch_fold<T, 4, PL_SUM>(p, stream)));
                PlainParams<T> f = fold_params<T>(parts, nc, nc, nullptr);
                f.o_sum = f_sum;
                f.o_mean = f_mean;
                f.o_mean2 = f_var && var_chunked ? wmean : nullptr;
                have_mean = f_var && var_chunked;
                f.mean_div = n_t;
                NRK_TRY((launch_fold<T, 1, PL_SUM>(f, stream)));
            }
        }
        if (f_var) {
            const T div = (T)((T)n - (T)ddof);
            if (!var_chunked) {
                if (!have_mean) {  // this tier's mean is simd_sum of everything / n
                    PlainParams<T> p = fold_params<T>(x, n, n, &chain);
                    p.o_mean2 = wmean;
                    p.mean_div = n_t;
                    NRK_TRY((launch_fold<T, 4, PL_SUM>(p, stream)));
                }
                PlainParams<T> p = fold_params<T>(x, n, n, &chain);
                p.mean = wmean;
                p.o_var = f_var;
                p.var_div = div;
                NRK_TRY((launch_fold<T, 1, PL_SSD>(p, stream)));
            } else {
                PlainParams<T> p = fold_params<T>(x, n, PLAIN_CHUNK, &chain);
                p.mean = wmean;
                p.o_sum = parts;
                NRK_TRY((launch_fold<T, 1, PL_SSD>(p, stream)));
                PlainParams<T> f = fold_params<T>(parts, nc, nc, nullptr);
                f.o_var = f_var;
                f.var_div = div;
                NRK_TRY((launch_fold<T, 1, PL_SUM>(f, stream)));
            }
        }
    }
    if (want_row) {
        PlainParams<T> p = fold_params<T>(x, n, cols, &chain);
        p.o_sum = r_sum;
        p.o_mean = r_mean;
        p.o_mean2 = r_var ? rmean : nullptr;
        p.mean_div = (T)cols;
        NRK_TRY((launch_fold<T, 1, PL_SUM>(p, stream)));
        if (r_var) {
            PlainParams<T> v = fold_params<T>(x, n, cols, &chain);
            v.mean = rmean;
            v.mean_stride = 1;
            v.o_var = r_var;
            v.var_div = (T)(cols - ddof);
            NRK_TRY((launch_fold<T, 1, PL_SSD>(v, stream)));
        }
    }
    return NRK_OK;
}

template nrk_status plain_reduce<float>(const Chain &, const float *, size_t, size_t, size_t, const nrk_reduce_outputs *,
                                        const nrk_reduce_outputs *, hipStream_t);
template nrk_status plain_reduce<double>(const Chain &, const double *, size_t, size_t, size_t, const nrk_reduce_outputs *,
                                         const nrk_reduce_outputs *, hipStream_t);

}  // namespace nrk
