// The short-row reduction kernel and its launcher, instantiated per element type in
// reduce_rows_f32.hip and reduce_rows_f64.hip.
#pragma once
#include "reduce_lean.hpp"

#include <type_traits>

namespace nrk {

// ---------------- short rows ----------------
// Rows of at most 64 K elements, K slots per lane (K = 16 or 32 binary32, 8 or 16 binary64: the
// larger where the smaller would take 32 or more lanes per row): G lanes per row, the smallest
// power of two with G K >= cols, so a wavefront reduces a block of 64 / G rows at a time and
// wavefront v of the W in the grid takes blocks v, v + W, ... Lane j of a row's G lanes holds slot
// s = position ((s / VN) G + j) VN + s % VN of the row (VN-element vectors, the G lanes' side by
// side), its first nv slots being inside the row. Binary32 rows whose length is not a multiple of
// the vector are phased: row r's positions start at the aligned element r cols - ph, ph = (r cols)
// mod VN, positions [ph, ph + cols) are the row's (so every load is a 16-byte vector), and G K
// covers cols + VN - 1 positions. The whole row is in registers: the sum is
// pairwise per lane (in T), then a butterfly over the row's lanes (double, double-word for
// binary64); the variance is two-pass, sum((y - m)^2) with m the mean rounded to T, less
// n (mean - m)^2. Extrema: strict winners per lane in slot (= element) order, then the butterfly
// with equal values going to the lower element. Whole-array results: every lane folds the records
// of its rows in turn (equal counts, no division), the wavefront merges its lanes' records once at
// the end, then the group finish.
constexpr size_t ROWS_GROUPS = 1024;  // round robin over at most this many workgroups

template <typename T>
struct RowsParams {
    const T *x;
    double s0, s1;  // CH_MULADD's scalars
    Chain chain;
    size_t rows, cols, nblk;
    unsigned G, lg;  // lanes per row and its log2
    int phased;      // binary32, cols not a multiple of the vector: row r's positions start at the
                     // aligned element r cols - ph, ph = (r cols) mod VN; [ph, ph + cols) are the row's
    double ddof;
    OutPtrs<T> row;
    int need_group, uniform;
    Parts grp;
    OutPtrs<T> full;
    unsigned int *counter;
};

template <typename T, int K, int FLAGS, bool VEC, int CH>
__global__ __launch_bounds__(RB) void rows_kernel(RowsParams<T> p) {
    typedef typename Vec16<T>::type V;
    constexpr int VN = Vec16<T>::N;
    constexpr int VPT = K / VN;
    __shared__ FinishLds fin;
    const unsigned lane = threadIdx.x & 63;
    const unsigned wave = lane0((unsigned)(threadIdx.x >> 6));
    const unsigned G = p.G, R = 64u >> p.lg;
    const unsigned j = lane & (G - 1), rr = lane >> p.lg;
    const unsigned cols = (unsigned)p.cols;
    const size_t W = (size_t)gridDim.x * WAVES;
    const size_t v = blockIdx.x * (size_t)WAVES + wave;
    const size_t nit = v < p.nblk ? (p.nblk - 1 - v) / W + 1 : 0;
    auto off = [&](unsigned s) -> unsigned { return ((s / VN) * G + j) * VN + s % VN; };
    unsigned nv = 0;  // slots inside the row (a prefix)
#pragma unroll
    for (int s = 0; s < K; ++s) nv += off(s) < cols ? 1u : 0u;
    const bool every = !p.phased && cols == G * (unsigned)K;  // every slot of every lane is inside the row (uniform)
    auto row_phase = [&](size_t r) -> unsigned { return (unsigned)((r * p.cols) & (size_t)(VN - 1)); };
    const T s0 = (T)p.s0, s1 = (T)p.s1;
    const double ncols = (double)cols, inv_cols = 1.0 / ncols;
    Part lp = part_empty();  // this lane's rows so far (whole-array results)
    unsigned wp_fn = 0;
    // block b's row for this lane (rows past the end read the last row: a harmless reload). PH:
    // phased rows (positions from the aligned element before the row).
    auto load = [&](T(&yy)[K], size_t b, auto ph_tag) {
        constexpr bool PH = decltype(ph_tag)::value;
        const size_t r = min(b * R + rr, p.rows - 1);
        if constexpr (PH) {
            const unsigned ph = row_phase(r);
            const T *base = p.x + r * p.cols - ph;  // position 0 (16-byte aligned when x is)
#pragma unroll
            for (int q = 0; q < VPT; ++q) {
                const unsigned e0 = (q * G + j) * VN;
                if constexpr (VEC) {
                    // a vector starting inside the row lies in the row's 16-byte blocks; else position 0's
                    const V t = __builtin_nontemporal_load(reinterpret_cast<const V *>(base + (e0 < cols + ph ? e0 : 0u)));
#pragma unroll
                    for (int e = 0; e < VN; ++e) yy[q * VN + e] = t[e];
                } else {
#pragma unroll
                    for (int e = 0; e < VN; ++e) yy[q * VN + e] = base[e0 + e - ph < cols ? e0 + e : ph];
                }
            }
        } else {
            const T *base = p.x + r * p.cols;
#pragma unroll
            for (int q = 0; q < VPT; ++q) {
                const unsigned e0 = (q * G + j) * VN;
                if constexpr (VEC) {
                    const V t = __builtin_nontemporal_load(reinterpret_cast<const V *>(base + min(e0, cols - VN)));
#pragma unroll
                    for (int e = 0; e < VN; ++e) yy[q * VN + e] = t[e];
                } else {
#pragma unroll
                    for (int e = 0; e < VN; ++e) yy[q * VN + e] = base[min(e0 + e, cols - 1)];
                }
            }
        }
    };
    // the row's totals from this lane's slots (the first n of them, or all K)
    struct RowTot {
        double hi, lo, m2;
        T vmax, vmin;
        unsigned imax, imin, nan;
    };
    // MASK 0: every slot is the row's; 1: the first nv slots; 2: positions [ph, ph + cols)
    auto row_totals = [&](const T(&y)[K], auto masked, unsigned ph) -> RowTot {
        constexpr int MASK = decltype(masked)::value;
        auto in = [&](int i) -> bool {
            if constexpr (MASK == 0) return true;
            else if constexpr (MASK == 1) return (unsigned)i < nv;
            else return off((unsigned)i) - ph < cols;  // (unsigned: positions before ph wrap high)
        };
        RowTot t;
        const T st = pairwise<K, 0, T>([&](int i) { return in(i) ? y[i] : (T)0; });
        t.nan = 0;
        if (__builtin_isnan(st)) {
#pragma unroll
            for (int i = 0; i < K; ++i) t.nan |= (in(i) && __builtin_isnan(y[i])) ? 1u : 0u;
        }
        t.vmax = -dev_inf<T>();
        t.vmin = dev_inf<T>();
        t.imax = NO_CODE;
        t.imin = NO_CODE;
#pragma unroll
        for (int i = 0; i < K; ++i) {
            if (FLAGS & F_MAX) {
                const bool w = in(i) && y[i] > t.vmax;
                t.vmax = w ? y[i] : t.vmax;
                t.imax = w ? (unsigned)i : t.imax;
            }
            if (FLAGS & F_MIN) {
                const bool w = in(i) && y[i] < t.vmin;
                t.vmin = w ? y[i] : t.vmin;
                t.imin = w ? (unsigned)i : t.imin;
            }
        }
        if (FLAGS & F_MAX) t.imax = t.imax == NO_CODE ? NO_CODE : off(t.imax) - ph;
        if (FLAGS & F_MIN) t.imin = t.imin == NO_CODE ? NO_CODE : off(t.imin) - ph;
        t.hi = (double)st;
        t.lo = 0.0;
        for (unsigned m = G >> 1; m >= 1; m >>= 1) {  // uniform: G is the same for every lane
            if constexpr (sizeof(T) == 8) {
                const double bh = shfl_xor(t.hi, (int)m), bl = shfl_xor(t.lo, (int)m);
                dw_add(t.hi, t.lo, bh, bl);
            } else {
                t.hi += shfl_xor(t.hi, (int)m);
            }
            if (FLAGS & F_MAX) lean_take_max(t.vmax, t.imax, shfl_xor(t.vmax, (int)m), shfl_xor_u(t.imax, m));
            if (FLAGS & F_MIN) lean_take_min(t.vmin, t.imin, shfl_xor(t.vmin, (int)m), shfl_xor_u(t.imin, m));
            t.nan |= shfl_xor_u(t.nan, m);
        }
        t.m2 = 0.0;
        if (FLAGS & F_M2) {
            const double mu = (t.hi + t.lo) / ncols;  // divided: a constant row has m = its value, variance 0
            const T mt = (T)mu;
            T q4[4] = {(T)0, (T)0, (T)0, (T)0};
#pragma unroll
            for (int i = 0; i < K; ++i) {
                const T d = in(i) ? rn_sub(y[i], mt) : (T)0;
                q4[i & 3] = fma(d, d, q4[i & 3]);
            }
            double q = (double)rn_add(rn_add(q4[0], q4[1]), rn_add(q4[2], q4[3]));
            for (unsigned m = G >> 1; m >= 1; m >>= 1) q += shfl_xor(q, (int)m);
            const double dm = mu - (double)mt;
            const double m2 = q - ncols * dm * dm;
            t.m2 = m2 < 0.0 ? 0.0 : m2;  // (NaN stays NaN)
        }
        return t;
    };
    auto step = [&](T(&y)[K], size_t it, auto ph_tag) {
        constexpr bool PH = decltype(ph_tag)::value;
        const size_t r = (v + it * W) * R + rr;
        if constexpr (CH == CH_MULADD) {
#pragma unroll
            for (int i = 0; i < K; ++i) y[i] = rn_add(rn_mul(y[i], s0), s1);
        } else if constexpr (CH == CH_ANY) {
            apply_chain_n(p.chain, y);
        }
        unsigned fn;  // the row's element 0 is NaN: slot ph (0 unphased) of its lane 0
        RowTot t;
        if constexpr (PH) {
            const unsigned ph = row_phase(min(r, p.rows - 1));
            T y0 = y[0];
#pragma unroll
            for (int i = 1; i < VN; ++i) y0 = ph == (unsigned)i ? y[i] : y0;
            fn = (unsigned)__shfl(__builtin_isnan(y0) ? 1 : 0, (int)(lane & ~(G - 1)), 64);
            t = row_totals(y, std::integral_constant<int, 2>{}, ph);
        } else {
            fn = (unsigned)__shfl(__builtin_isnan(y[0]) ? 1 : 0, (int)(lane & ~(G - 1)), 64);
            t = every ? row_totals(y, std::integral_constant<int, 0>{}, 0u) : row_totals(y, std::integral_constant<int, 1>{}, 0u);
        }
        if (r >= p.rows) return;  // past the last row (the last block only)
        Part a;
        a.n = ncols;
        a.hi = t.hi;
        a.lo = t.lo;
        a.m2 = t.m2;
        a.nan = t.nan;
        a.vmax = t.imax != NO_CODE ? (double)t.vmax : -__builtin_inf();
        a.imax = t.imax != NO_CODE ? (uint64_t)t.imax : NO_INDEX;
        a.vmin = t.imin != NO_CODE ? (double)t.vmin : __builtin_inf();
        a.imin = t.imin != NO_CODE ? (uint64_t)t.imin : NO_INDEX;
        if (j == 0) write_outputs<T>(p.row, r, a, p.ddof, fn, 0);
        if (!p.need_group) return;  // uniform over the grid
        const uint64_t rb = (uint64_t)r * p.cols;
        if (a.imax != NO_INDEX) a.imax += rb;
        if (a.imin != NO_INDEX) a.imin += rb;
        if (it == 0) wp_fn = fn;
        // this lane's it-th row (every earlier block was whole): the block index is uniform, so the
        // weight table is read with a scalar load
        lean_merge_next<FLAGS>(lp, a, it, inv_cols);
    };
    // two blocks in turn: the next one's loads are in flight while one is folded
    auto run = [&](auto ph_tag) {
        if (nit == 0) return;
        T y[2][K];
        load(y[0], v, ph_tag);
        const size_t npair = nit / 2;
        for (size_t it = 0; it < npair; ++it) {
            load(y[1], v + (2 * it + 1) * W, ph_tag);
            step(y[0], 2 * it, ph_tag);
            load(y[0], v + (2 * it + 2) * W, ph_tag);
            step(y[1], 2 * it + 1, ph_tag);
        }
        if (nit & 1) step(y[0], nit - 1, ph_tag);
    };
    if constexpr (sizeof(T) == 4) {  // uniform: one loop or the other
        if (p.phased) run(std::true_type{});
        else run(std::false_type{});
    } else {
        run(std::false_type{});  // binary64 rows are not phased (two 8-byte loads cost less than the masks)
    }
    if (!p.need_group) return;
    // the wavefront's rows: lanes G, 2G, ... apart hold different rows (equal counts when uniform)
    if (p.uniform) {
        double w = 0.5 / lp.n;
        for (unsigned m = G; m < 64; m <<= 1) {
            const Part o = part_shfl<FLAGS>(lp, (int)m);
            part_merge_eq<FLAGS>(lp, o, w);
            w *= 0.5;
        }
    } else {
        for (unsigned m = G; m < 64; m <<= 1) {
            const Part o = part_shfl<FLAGS>(lp, (int)m);
            part_merge<FLAGS>(lp, o);
        }
    }
    wp_fn = (unsigned)__shfl((int)wp_fn, 0, 64);
    lean_group_finish<T, FLAGS>(fin, p.grp, p.counter, p.full, p.ddof, p.uniform, lp, wp_fn, []() {});
}

template <typename T, int FLAGS, int CH>
static void launch_rows_kernel(const RowsParams<T> &p, size_t ngroups, bool vec, bool big, hipStream_t stream) {
    constexpr int K1 = rows_k<T>(), K2 = 2 * rows_k<T>();
    const dim3 g((unsigned)ngroups), b(RB);
    if (big) {
        if (vec) hipLaunchKernelGGL((rows_kernel<T, K2, FLAGS, true, CH>), g, b, 0, stream, p);
        else hipLaunchKernelGGL((rows_kernel<T, K2, FLAGS, false, CH>), g, b, 0, stream, p);
    } else {
        if (vec) hipLaunchKernelGGL((rows_kernel<T, K1, FLAGS, true, CH>), g, b, 0, stream, p);
        else hipLaunchKernelGGL((rows_kernel<T, K1, FLAGS, false, CH>), g, b, 0, stream, p);
    }
}

template <typename T, int FLAGS>
static nrk_status launch_rows(const RowsParams<T> &p, size_t ngroups, bool big, hipStream_t stream) {
    constexpr int VN = Vec16<T>::N;
    const bool vec = (uintptr_t)p.x % 16 == 0 && (p.cols % VN == 0 || p.phased);  // phased rows start on the grid
    const Chain &c = p.chain;
    if (c.len == 0)
        launch_rows_kernel<T, FLAGS, CH_NONE>(p, ngroups, vec, big, stream);
    else if (c.len == 2 && c.op[0] == NRK_SOP_MUL && (c.op[1] == NRK_SOP_ADD || c.op[1] == NRK_SOP_SUB))
        launch_rows_kernel<T, FLAGS, CH_MULADD>(p, ngroups, vec, big, stream);
    else
        launch_rows_kernel<T, FLAGS, CH_ANY>(p, ngroups, vec, big, stream);
    NRK_HIP(hipGetLastError());
    return NRK_OK;
}

template <typename T>
nrk_status run_rows(const Chain &chain, const T *x, size_t rows, size_t cols, double ddof, const OutPtrs<T> &row,
                           const OutPtrs<T> &full, bool want_full, hipStream_t stream, const ExtScratch *ext) {
    RowsParams<T> p{};
    p.x = x;
    p.chain = chain;
    p.rows = rows;
    p.cols = cols;
    p.ddof = ddof;
    // binary32 rows whose length is not a multiple of the vector are phased: a row's positions run
    // over cols + ph < cols + VN. K slots per lane, doubled where the smaller K would take 32 or
    // more lanes per row.
    p.phased = sizeof(T) == 4 && rows > 1 && cols % Vec16<T>::N != 0;
    const size_t span = cols + (p.phased ? Vec16<T>::N - 1 : 0);
    const bool big = span > 16 * (size_t)rows_k<T>();
    const size_t K = big ? 2 * rows_k<T>() : rows_k<T>();
    unsigned lg = 0;
    while ((K << lg) < span) ++lg;
    p.G = 1u << lg;
    p.lg = lg;
    const size_t R = (size_t)64 >> lg;
    p.nblk = ceil_div(rows, R);
    // rows only with the doubled K (wide rows, few per wavefront): twice the workgroups (the row
    // results do not depend on the grid; the whole-array finish keeps ROWS_GROUPS)
    const size_t gmax = (!want_full && big) ? 2 * ROWS_GROUPS : ROWS_GROUPS;
    const size_t ngroups = std::min<size_t>(gmax, ceil_div(p.nblk, (size_t)WAVES));
    // every wavefront reduces the same number of whole blocks: the finish merges equal counts
    p.uniform = rows % R == 0 && p.nblk % (ngroups * WAVES) == 0;
    if (chain.len == 2) {  // CH_MULADD: y = (x * s0) + s1, with x - b as x + (-b)
        p.s0 = chain.s[0];
        p.s1 = chain.op[1] == NRK_SOP_SUB ? -chain.s[1] : chain.s[1];
    }
    p.row = row;
    p.need_group = want_full;
    p.full = want_full ? full : OutPtrs<T>{};
    if (want_full) NRK_TRY(group_scratch(ngroups, rows, cols, ext, stream, &p.counter, &p.grp));
    const bool m2 = (row.var || full.var);
    const bool mx = (row.max || row.argmax || full.max || full.argmax);
    const bool mn = (row.min || row.argmin || full.min || full.argmin);
    const int flags = (m2 ? F_M2 : 0) | (mx ? F_MAX : 0) | (mn ? F_MIN : 0);
#define NRK_ROWS(F) \
    case F: return launch_rows<T, F>(p, ngroups, big, stream);
    switch (flags) {
        NRK_ROWS(0) NRK_ROWS(1) NRK_ROWS(2) NRK_ROWS(3) NRK_ROWS(4) NRK_ROWS(5) NRK_ROWS(6)
        default: NRK_ROWS(7)
    }
#undef NRK_ROWS
}

}  // namespace nrk
