// Device pieces shared by the lean kernel (reduce_lean.hip) and the short-row kernel
// (reduce_rows.hpp): the equal-count merges, the agent-scope record I/O and the one-launch
// whole-array finish.
#pragma once
#include "reduce_common.hpp"

namespace nrk {

// ---------------- the lean kernel: full segments, contiguous chunk ranges ----------------
// Used for row results when cols is a multiple of the segment and there are at least
// LEAN_MIN_CHUNKS rows (a chunk is then a row), and for every whole-array reduction (chunks of cl
// whole segments, and the elements after the last whole chunk folded by one more workgroup).
// Every chunk holds exactly cl full segments, and:
//   - chunk c is the contiguous elements [c cl WSEG, (c + 1) cl WSEG); each wavefront reduces an
//     arithmetic sequence of chunks (round robin or a contiguous block, see lean_layout) with one
//     pointer, keeping the next segment's loads in flight while it folds the current one;
//   - every lane of a chunk has the same count (16 per segment), so the Chan updates need no
//     division: the weight of the j-th segment's merge, 1 / (16 j (j + 1)), comes from a constant
//     table, and the butterfly's levels merge equal counts with weight 1 / (2 n);
//   - binary32 lanes add their segment sums into plain FP64 (exact to 2^-29 of a binary32 ulp);
//     binary64 lanes keep the double word;
//   - a row's results wait in LDS and go out after the segment loop; each wavefront merges its
//     chunks in order, the workgroup its 4 wavefronts into one group record, and the last
//     workgroup to finish (counted with agent-scope atomics, no fence) merges the group records
//     in a fixed order: no second launch. Where every record holds the same count, these merges
//     take no division.
// The chunking depends only on the shape, so results are bit-identical across launches, grids
// and devices. The input is an expression of one or two operands (see LeanProg).
constexpr int LEAN_WTAB = 1024;  // chunks of at most this many segments
struct LeanWTab {
    double w[LEAN_WTAB];
    constexpr LeanWTab() : w() {
        for (int j = 1; j < LEAN_WTAB; ++j) w[j] = 1.0 / (16.0 * j * (j + 1));
    }
};
static_assert(NRK_REDUCE_M == 16, "the Chan weight table assumes 16 elements per lane and segment");
static __constant__ LeanWTab kLeanW = LeanWTab();


// The one-operand programs the kernel is specialised on (the chain then costs no dispatch and no
// scalar loads inside the segment loop): CH_NONE no step, CH_MULADD y = (x * s0) + s1 (the
// Array ops `&x * a + b`, two roundings; x - b is x + (-b) exactly), CH_ANY any chain.
enum { CH_ANY = 0, CH_MULADD = 1, CH_NONE = 2 };

template <typename T>
struct LeanParams {
    const T *x0, *x1;
    double s0, s1;    // CH_MULADD's scalars
    size_t nchunk, cl, cols;
    size_t A, B;      // wavefront v reduces chunks v A + k B: A = 1, B = W (round robin over the W
                      // wavefronts of the grid) or A = chunks per wavefront, B = 1 (contiguous blocks)
    LeanProg prog;
    double ddof;
    int row_direct;   // a chunk is a row and row results are wanted
    OutPtrs<T> row;
    int need_group;   // whole-array results wanted: store one record per workgroup
    int uniform;      // every wavefront reduces the same number of chunks (equal record counts)
    // whole-array mode of any length: elements [tail0, n_total) after the last whole chunk are
    // folded by one more workgroup, tail_group (none: ~0u), whose record joins the finish
    size_t tail0, n_total;
    unsigned tail_group;
    Parts grp;
    OutPtrs<T> full;         // whole-array results (need_group)
    unsigned int *counter;   // zero at launch; the last workgroup to finish resets it
};

template <typename T, int N>
__device__ __forceinline__ void bop_n(int op, T (&a)[N], const T (&b)[N]) {
    switch (op) {
#define NRK_BOPN(OP, EXPR)                                         \
    case OP:                                                       \
        _Pragma("unroll") for (int i = 0; i < N; ++i) a[i] = EXPR; \
        break;
        NRK_BOPN(NRK_BOP_ADD, rn_add(a[i], b[i]))
        NRK_BOPN(NRK_BOP_SUB, rn_sub(a[i], b[i]))
        NRK_BOPN(NRK_BOP_MUL, rn_mul(a[i], b[i]))
        NRK_BOPN(NRK_BOP_DIV, rn_div(a[i], b[i]))
        NRK_BOPN(NRK_BOP_REM, __builtin_fmod(a[i], b[i]))
        NRK_BOPN(NRK_BOP_MAXIMUM, a[i] > b[i] ? a[i] : b[i])
        NRK_BOPN(NRK_BOP_MINIMUM, a[i] < b[i] ? a[i] : b[i])
        NRK_BOPN(NRK_BOP_COPYSIGN, dev_copysign(a[i], b[i]))
        NRK_BOPN(NRK_BOP_PRELU, a[i] > (T)0 ? a[i] : rn_mul(b[i], a[i]))
#undef NRK_BOPN
        default: break;
    }
}

// Issue the loads of one full segment: lane l takes elements (j * 64 + l) VN + e, as 16-byte
// vectors when the operand is aligned (VEC), else one by one -- the same elements either way, so
// alignment changes the speed only, never the bits.
template <typename T, int VPT, bool VEC>
__device__ __forceinline__ void lean_load(const T *base, T (&y)[VPT * Vec16<T>::N]) {
    typedef typename Vec16<T>::type V;
    constexpr int VN = Vec16<T>::N;
    const unsigned lane = threadIdx.x & 63;
    if constexpr (VEC) {
        const V *bv = reinterpret_cast<const V *>(base) + lane;
#pragma unroll
        for (int j = 0; j < VPT; ++j) {
            // binary32 plain loads, binary64 non-temporal ones (8192 x 8192 fused f32: median
            // 48.2 -> 46.6 us; binary64 unchanged either way)
            V v;
            if constexpr (sizeof(T) == 4)
                v = bv[j * 64];
            else
                v = __builtin_nontemporal_load(bv + j * 64);
#pragma unroll
            for (int e = 0; e < VN; ++e) y[j * VN + e] = v[e];
        }
    } else {
#pragma unroll
        for (int j = 0; j < VPT; ++j)
#pragma unroll
            for (int e = 0; e < VN; ++e) y[j * VN + e] = base[(j * 64 + lane) * VN + e];
    }
}

// Fold the j-th segment of the current chunk to be processed, segment sj of the chunk (loaded
// into y, and z for the second operand). Every segment holds 16 elements per lane, so the order
// of the Chan merges needs only j; the winners' codes and the chunk's first element use sj.
template <typename T, int VPT, int FLAGS, int LEAVES, int CH>
__device__ __forceinline__ void lean_fold(const LeanProg &pg, T s0, T s1, T (&y)[VPT * Vec16<T>::N], T (&z)[VPT * Vec16<T>::N],
                                          unsigned j, unsigned sj, LaneAcc<T> &a, unsigned &fnan) {
    constexpr int M = VPT * Vec16<T>::N;
    if constexpr (CH == CH_MULADD) {
#pragma unroll
        for (int i = 0; i < M; ++i) y[i] = rn_add(rn_mul(y[i], s0), s1);
    } else if constexpr (CH == CH_ANY) {
        apply_chain_n(pg.pre0, y);
    }
    if constexpr (LEAVES == 2) {
        apply_chain_n(pg.pre1, z);
        bop_n(pg.op, y, z);
        apply_chain_n(pg.post, y);
    }
    if (sj == 0) fnan = __builtin_isnan(y[0]) ? 1u : 0u;  // lane 0's value is the chunk's first element
    const T st = pairwise<M, 0, T>([&](int i) { return y[i]; });
    if (__builtin_isnan(st)) {  // rare: tell a NaN element apart from inf - inf
#pragma unroll
        for (int i = 0; i < M; ++i) a.nan |= __builtin_isnan(y[i]) ? 1u : 0u;
    }
    const double sb = (double)st;
    if (FLAGS & F_M2) {
        const T mt = rn_mul(st, (T)(1.0 / M));  // the segment mean (exact: a power of two)
        T q4[4] = {(T)0, (T)0, (T)0, (T)0};
#pragma unroll
        for (int i = 0; i < M; ++i) {
            const T d = rn_sub(y[i], mt);
            q4[i & 3] = fma(d, d, q4[i & 3]);
        }
        const double qt = (double)rn_add(rn_add(q4[0], q4[1]), rn_add(q4[2], q4[3]));
        if (j == 0) {
            a.m2 = qt;
        } else {  // Chan with n_a = 16 j, n_b = 16: (S_a - j S_b)^2 / (16 j (j + 1))
            const double d = (a.hi + a.lo) - (double)j * sb;
            a.m2 = a.m2 + qt + d * d * kLeanW.w[j];
        }
    }
    if constexpr (sizeof(T) == 8) {
        double s, e;
        two_sum(a.hi, sb, s, e);
        e = e + a.lo;
        dw_fix(s, e, a.hi, a.lo);
    } else {
        a.hi += sb;
    }
    // extrema: the segment's first strict winner, then the lane's running winner (equal values go
    // to the lower code: segments need not come in index order)
    if (FLAGS & F_MAX) {
        T sv = -dev_inf<T>();
        unsigned ci = NO_CODE;
#pragma unroll
        for (int i = 0; i < M; ++i) {
            const bool gt = y[i] > sv;
            sv = gt ? y[i] : sv;
            ci = gt ? (unsigned)i : ci;
        }
        const unsigned code = sj * M + ci;
        const bool take = ci != NO_CODE && (sv > a.vmax || (sv == a.vmax && code < a.cmax));
        a.vmax = take ? sv : a.vmax;
        a.cmax = take ? code : a.cmax;
    }
    if (FLAGS & F_MIN) {
        T sv = dev_inf<T>();
        unsigned ci = NO_CODE;
#pragma unroll
        for (int i = 0; i < M; ++i) {
            const bool lt = y[i] < sv;
            sv = lt ? y[i] : sv;
            ci = lt ? (unsigned)i : ci;
        }
        const unsigned code = sj * M + ci;
        const bool take = ci != NO_CODE && (sv < a.vmin || (sv == a.vmin && code < a.cmin));
        a.vmin = take ? sv : a.vmin;
        a.cmin = take ? code : a.cmin;
    }
}

// A lane's record of one chunk, in the butterfly: the count is the same on every lane (16 per
// segment), so it is not carried; winners are chunk-relative element offsets (< 2^32); binary32
// sums have no low word. Only the fields the op set uses cross lanes.
template <typename T>
struct LeanRec {
    double hi, lo, m2;
    T vmax, vmin;
    unsigned imax, imin;  // chunk-relative offsets, NO_CODE: no winner
    unsigned nan;
};

__device__ __forceinline__ unsigned shfl_xor_u(unsigned v, unsigned m) { return (unsigned)__shfl_xor((int)v, (int)m, 64); }

// (value, offset) winners: larger (smaller) value wins, equal values go to the lower offset; NO_CODE
// (no winner) loses to every winner.
template <typename T>
__device__ __forceinline__ void lean_take_max(T &v, unsigned &i, T bv, unsigned bi) {
    const bool take = bi != NO_CODE && (i == NO_CODE || bv > v || (bv == v && bi < i));
    v = take ? bv : v;
    i = take ? bi : i;
}
template <typename T>
__device__ __forceinline__ void lean_take_min(T &v, unsigned &i, T bv, unsigned bi) {
    const bool take = bi != NO_CODE && (i == NO_CODE || bv < v || (bv == v && bi < i));
    v = take ? bv : v;
    i = take ? bi : i;
}

// Butterfly of records of equal counts (n per lane): M2 += (S_a - S_b)^2 / (2 n) at the first
// level, the weight halving with every level as the counts double.
template <typename T, int FLAGS>
__device__ __forceinline__ void lean_wave_merge(LeanRec<T> &a, double n) {
    double w = 0.5 / n;
#pragma unroll
    for (unsigned m = 32; m >= 1; m >>= 1) {
        LeanRec<T> b;
        b.hi = shfl_xor(a.hi, (int)m);
        if constexpr (sizeof(T) == 8) b.lo = shfl_xor(a.lo, (int)m);
        if (FLAGS & F_M2) b.m2 = shfl_xor(a.m2, (int)m);
        if (FLAGS & F_MAX) { b.vmax = shfl_xor(a.vmax, (int)m); b.imax = shfl_xor_u(a.imax, m); }
        if (FLAGS & F_MIN) { b.vmin = shfl_xor(a.vmin, (int)m); b.imin = shfl_xor_u(a.imin, m); }
        b.nan = shfl_xor_u(a.nan, m);
        a.nan |= b.nan;
        if (FLAGS & F_MAX) lean_take_max(a.vmax, a.imax, b.vmax, b.imax);
        if (FLAGS & F_MIN) lean_take_min(a.vmin, a.imin, b.vmin, b.imin);
        if constexpr (sizeof(T) == 8) {
            if (FLAGS & F_M2) {
                const double d = (a.hi + a.lo) - (b.hi + b.lo);
                a.m2 = a.m2 + b.m2 + d * d * w;
            }
            dw_add(a.hi, a.lo, b.hi, b.lo);
        } else {
            if (FLAGS & F_M2) {
                const double d = a.hi - b.hi;
                a.m2 = a.m2 + b.m2 + d * d * w;
            }
            a.hi += b.hi;
        }
        w *= 0.5;
    }
}

// a += b for records of equal counts (w = 1 / (2 n), n either count): M2 += (S_a - S_b)^2 w.
template <int FLAGS>
__device__ __forceinline__ void part_merge_eq(Part &a, const Part &b, double w) {
    a.nan |= b.nan;
    if (FLAGS & F_MAX) take_max(a.vmax, a.imax, b.vmax, b.imax);
    if (FLAGS & F_MIN) take_min(a.vmin, a.imin, b.vmin, b.imin);
    if (FLAGS & F_M2) {
        const double d = (a.hi + a.lo) - (b.hi + b.lo);
        a.m2 = a.m2 + b.m2 + d * d * w;
    }
    dw_add(a.hi, a.lo, b.hi, b.lo);
    a.n += b.n;
}

// wp (k chunks of n elements) += one more chunk of n: Chan with n_a = k n, n_b = n, whose term
// (n_b S_a - n_a S_b)^2 / (n_a n_b (n_a + n_b)) is (S_a - k S_b)^2 / (k (k + 1) n) -- no division
// (1 / (k (k + 1)) = 16 w[k] from the weight table, k < LEAN_WTAB; else divided).
template <int FLAGS>
__device__ __forceinline__ void lean_merge_next(Part &wp, const Part &b, size_t k, double inv_n) {
    if (k == 0) {
        wp = b;
        return;
    }
    wp.nan |= b.nan;
    if (FLAGS & F_MAX) take_max(wp.vmax, wp.imax, b.vmax, b.imax);
    if (FLAGS & F_MIN) take_min(wp.vmin, wp.imin, b.vmin, b.imin);
    if (FLAGS & F_M2) {
        const double d = (wp.hi + wp.lo) - (double)k * (b.hi + b.lo);
        const double wk = k < (size_t)LEAN_WTAB ? 16.0 * kLeanW.w[k] : 1.0 / ((double)k * (double)(k + 1));
        wp.m2 = wp.m2 + b.m2 + d * d * (wk * inv_n);
    }
    dw_add(wp.hi, wp.lo, b.hi, b.lo);
    wp.n += b.n;
}

// The whole-array finish, at most LEAN_MAX_GROUPS group records; arrival counters (the top one,
// then one per set of LEAN_SET groups) LEAN_CTR_STRIDE words apart.
constexpr int LEAN_FIN_LOADS = 4;
constexpr size_t LEAN_MAX_GROUPS = (size_t)RB * LEAN_FIN_LOADS;
constexpr unsigned LEAN_SET = 32;
constexpr unsigned LEAN_CTR_STRIDE = 64;
constexpr size_t LEAN_CTR_WORDS = LEAN_CTR_STRIDE * (1 + LEAN_MAX_GROUPS / LEAN_SET);
static_assert(1 + LEAN_MAX_GROUPS / LEAN_SET <= (size_t)RB, "the last workgroup resets every counter, one per thread");

// Row results wait in LDS (no global store inside the segment loop, so the compiler's load
// counting stays exact and the next segment really is in flight) and go out ROWBUF rows at a time.
constexpr int LEAN_ROWBUF = 16;
struct RowRec {
    Part part;
    uint64_t chunk;
    unsigned first_nan;
};

// The whole-array finish of a kernel whose wavefronts each hold a record wp (wp_fn: the record's
// first element is NaN): the group's record goes out, its arrival is counted, after_count() runs
// on every thread, and the last group to arrive merges the group records in a fixed order and
// writes the whole-array results.
struct FinishLds {
    Part grp[WAVES], grp2[WAVES];
    unsigned fn[WAVES];
    unsigned last;
};
template <typename T, int FLAGS, typename F>
__device__ __forceinline__ void lean_group_finish(FinishLds &L, const Parts &grp, unsigned int *counter, const OutPtrs<T> &full,
                                                  double ddof, int uniform, const Part &wp, unsigned wp_fn, F after_count) {
    const unsigned lane = threadIdx.x & 63;
    const unsigned wave = lane0((unsigned)(threadIdx.x >> 6));
    if (lane == 0) {
        L.grp[wave] = wp;
        L.fn[wave] = wp_fn;
    }
    // LDS-only barriers from here: nothing waits for the row results' global stores, which go
    // out after this group's arrival is counted
    lds_barrier();
    if (threadIdx.x == 0) {
        Part g;
        if (uniform) {  // equal counts: pairwise, without division
            static_assert(WAVES == 4, "the pairwise group merge takes four wavefronts");
            Part g23 = L.grp[2];
            g = L.grp[0];
            const double w = 0.5 / g.n;
            part_merge_eq<FLAGS>(g, L.grp[1], w);
            part_merge_eq<FLAGS>(g23, L.grp[3], w);
            part_merge_eq<FLAGS>(g, g23, 0.5 * w);
        } else {
            g = L.grp[0];
            for (int w = 1; w < WAVES; ++w) part_merge<FLAGS>(g, L.grp[w]);
        }
        if (gridDim.x == 1) {  // the only group: its record is the result (no counting, no reload)
            write_outputs<T>(full, 0, g, ddof, L.fn[0], 0);
            L.last = 0;
        } else {
            part_store_agent<FLAGS>(grp, blockIdx.x, g, L.fn[0]);
            // the record has reached the point of coherence before it is counted. Arrivals are
            // counted in two levels (a counter per set of LEAN_SET groups, each on its own 256-byte
            // line, then one for the sets), so no counter takes more than LEAN_SET atomics.
            NRK_WAIT_VMEM0();
            const unsigned set = blockIdx.x / LEAN_SET, nset = (gridDim.x + LEAN_SET - 1) / LEAN_SET;
            const unsigned in_set = min(LEAN_SET, gridDim.x - set * LEAN_SET);
            bool last = false;
            if (__hip_atomic_fetch_add(counter + LEAN_CTR_STRIDE * (1 + set), 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) ==
                in_set - 1)
                last = __hip_atomic_fetch_add(counter, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) == nset - 1;
            L.last = last;
        }
    }
    lds_barrier();
    after_count();
    if (!L.last) return;  // uniform over the workgroup
    // The last workgroup to arrive merges the group records in a fixed order, whichever workgroup
    // it is: thread t folds records t, t + RB, ..., then the wavefront butterflies, then
    // wavefront order. (Every load is issued before the first is used: indices past the last
    // record reload it and are then replaced by empty records.) With equal counts and RB or 2 RB
    // records, every merge is between equal counts and needs no division; the order is then
    // pairwise: (t, t + RB), the butterfly, (w0 w1)(w2 w3).
    Part r[LEAN_FIN_LOADS];
    const bool eq = uniform && (gridDim.x == RB || gridDim.x == 2 * RB);
    if (eq) {
        r[0] = part_load_agent<FLAGS>(grp, threadIdx.x);
        r[1] = part_load_agent<FLAGS>(grp, gridDim.x == RB ? threadIdx.x : threadIdx.x + RB);
        double w = 0.5 / r[0].n;
        if (gridDim.x == 2 * RB) {
            part_merge_eq<FLAGS>(r[0], r[1], w);
            w *= 0.5;
        }
#pragma unroll
        for (int m = 32; m >= 1; m >>= 1) {
            const Part b = part_shfl<FLAGS>(r[0], m);
            part_merge_eq<FLAGS>(r[0], b, w);
            w *= 0.5;
        }
    } else {
#pragma unroll
        for (int q = 0; q < LEAN_FIN_LOADS; ++q)
            r[q] = part_load_agent<FLAGS>(grp, min(threadIdx.x + (size_t)q * RB, (size_t)gridDim.x - 1));
#pragma unroll
        for (int q = 1; q < LEAN_FIN_LOADS; ++q)
            if (threadIdx.x + (size_t)q * RB >= gridDim.x) r[q] = part_empty();
        if (threadIdx.x >= gridDim.x) r[0] = part_empty();
#pragma unroll
        for (int q = 1; q < LEAN_FIN_LOADS; ++q) part_merge<FLAGS>(r[0], r[q]);
        wave_merge<FLAGS>(r[0]);
    }
    if (lane == 0) L.grp2[wave] = r[0];
    lds_barrier();
    if (threadIdx.x == 0) {
        Part g = L.grp2[0];
        if (eq) {
            Part g23 = L.grp2[2];
            const double w = 0.5 / g.n;
            part_merge_eq<FLAGS>(g, L.grp2[1], w);
            part_merge_eq<FLAGS>(g23, L.grp2[3], w);
            part_merge_eq<FLAGS>(g, g23, 0.5 * w);
        } else {
            for (int w = 1; w < WAVES; ++w) part_merge<FLAGS>(g, L.grp2[w]);
        }
        const unsigned fn = (ld_agent(grp.flags) >> 1) & 1u;  // group 0 holds the first element
        write_outputs<T>(full, 0, g, ddof, fn, 0);
    }
    // the zero-at-launch invariant for the next launch on this stream
    if (threadIdx.x <= (gridDim.x + LEAN_SET - 1) / LEAN_SET) st_agent(counter + LEAN_CTR_STRIDE * threadIdx.x, 0u);
}

// The arrival counters (zero) and ngroups group records of the whole-array finish, from the
// caller's scratch or the stream's (reduce_lean.hip).
nrk_status group_scratch(size_t ngroups, size_t rows, size_t cols, const ExtScratch *ext, hipStream_t stream,
                         unsigned int **counter, Parts *grp);

// Slots per lane of the short-row kernel (doubled for longer rows).
template <typename T>
constexpr int rows_k() {
    return sizeof(T) == 4 ? 16 : 8;
}

// Row results (and whole-array ones) of rows of at most 128 rows_k elements (reduce_rows.hpp).
template <typename T>
nrk_status run_rows(const Chain &chain, const T *x, size_t rows, size_t cols, double ddof, const OutPtrs<T> &row,
                    const OutPtrs<T> &full, bool want_full, hipStream_t stream, const ExtScratch *ext);

}  // namespace nrk
