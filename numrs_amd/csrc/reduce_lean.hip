// The lean streaming reduction (see the comment below) and its launcher; the general kernels and
// the host entry points are in reduce.hip.
#include "reduce_common.hpp"

namespace nrk {

// ---------------- the lean kernel: full segments, contiguous chunk ranges ----------------
// Used when every segment is whole (cols, or rows * cols for a whole-array reduction, is a
// multiple of the segment), there are at least LEAN_MIN_CHUNKS chunks, and a chunk is a whole row
// when row results are wanted. Then every chunk holds exactly cl full segments, and:
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
__constant__ LeanWTab kLeanW = LeanWTab();


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
            const V v = __builtin_nontemporal_load(bv + j * 64);
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

template <int FLAGS>
__device__ __forceinline__ void part_store_agent(const Parts &P, size_t s, const Part &a, unsigned first_nan) {
    st_agent(P.n + s, a.n); st_agent(P.hi + s, a.hi); st_agent(P.lo + s, a.lo);
    if (FLAGS & F_M2) st_agent(P.m2 + s, a.m2);
    if (FLAGS & F_MAX) { st_agent(P.vmax + s, a.vmax); st_agent(P.imax + s, a.imax); }
    if (FLAGS & F_MIN) { st_agent(P.vmin + s, a.vmin); st_agent(P.imin + s, a.imin); }
    st_agent(P.flags + s, a.nan | (first_nan << 1));
}
template <int FLAGS>
__device__ __forceinline__ Part part_load_agent(const Parts &P, size_t s) {
    Part a = part_empty();
    a.n = ld_agent(P.n + s); a.hi = ld_agent(P.hi + s); a.lo = ld_agent(P.lo + s);
    if (FLAGS & F_M2) a.m2 = ld_agent(P.m2 + s);
    if (FLAGS & F_MAX) { a.vmax = ld_agent(P.vmax + s); a.imax = ld_agent(P.imax + s); }
    if (FLAGS & F_MIN) { a.vmin = ld_agent(P.vmin + s); a.imin = ld_agent(P.imin + s); }
    a.nan = ld_agent(P.flags + s) & 1u;
    return a;
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

// The tail of a whole-array reduction (what follows the last whole chunk: fewer than a chunk and
// a segment of elements): lane l of wavefront w folds elements tail0 + 64 w + l, + RB, ... in
// index order as one-element records, then the wavefront's butterfly; wp receives the result.
// Loads go out LEAN_TAIL_BATCH at a time (values past the end are loaded as zeros and skipped).
constexpr int LEAN_TAIL_BATCH = 16;
template <typename T, int FLAGS, int LEAVES, int CH>
__device__ void lean_tail(const LeanParams<T> &p, unsigned wave, unsigned lane, Part &wp) {
    constexpr int NB = LEAN_TAIL_BATCH;
    Part a = part_empty();
    const T s0 = (T)p.s0, s1 = (T)p.s1;
    for (size_t i0 = p.tail0 + wave * 64 + lane; i0 < p.n_total; i0 += (size_t)NB * RB) {
        T y[NB], z[LEAVES == 2 ? NB : 1];
#pragma unroll
        for (int k = 0; k < NB; ++k) {
            const size_t i = i0 + (size_t)k * RB;
            y[k] = i < p.n_total ? p.x0[i] : (T)0;
            if constexpr (LEAVES == 2) z[k] = i < p.n_total ? p.x1[i] : (T)0;
        }
        if constexpr (CH == CH_MULADD) {
#pragma unroll
            for (int k = 0; k < NB; ++k) y[k] = rn_add(rn_mul(y[k], s0), s1);
        } else if constexpr (CH == CH_ANY) {
            apply_chain_n(p.prog.pre0, y);
        }
        if constexpr (LEAVES == 2) {
            apply_chain_n(p.prog.pre1, z);
            bop_n(p.prog.op, y, z);
            apply_chain_n(p.prog.post, y);
        }
#pragma unroll
        for (int k = 0; k < NB; ++k) {
            const size_t i = i0 + (size_t)k * RB;
            if (i >= p.n_total) break;
            const T v = y[k];
            Part e;
            e.n = 1.0;
            e.hi = (double)v;
            e.lo = 0.0;
            e.m2 = 0.0;
            e.nan = v != v ? 1u : 0u;
            const bool up = v > -dev_inf<T>(), dn = v < dev_inf<T>();  // both false for NaN
            e.vmax = up ? (double)v : -__builtin_inf();
            e.imax = up ? (uint64_t)i : NO_INDEX;  // -inf and NaN never win
            e.vmin = dn ? (double)v : __builtin_inf();
            e.imin = dn ? (uint64_t)i : NO_INDEX;
            part_merge<FLAGS>(a, e);
        }
    }
    wave_merge<FLAGS>(a);
    wp = a;
}

template <typename T, int VPT, int FLAGS, int LEAVES, bool VEC, int CH>
__global__ __launch_bounds__(RB) void lean_kernel(LeanParams<T> p) {
    constexpr int VN = Vec16<T>::N;
    constexpr int M = VPT * VN;
    constexpr int WSEG = 64 * M;
    constexpr int DZ = LEAVES == 2 ? 2 : 1;
    __shared__ FinishLds fin;
    __shared__ RowRec rowbuf[WAVES][LEAN_ROWBUF];
    const unsigned lane = threadIdx.x & 63;
    const unsigned wave = lane0((unsigned)(threadIdx.x >> 6));
    const size_t v = blockIdx.x * (size_t)WAVES + wave;
    const size_t c0 = v * p.A;  // chunks c0, c0 + B, ...
    const bool is_tail = blockIdx.x == p.tail_group;  // uniform over the workgroup
    const size_t nck = (is_tail || c0 >= p.nchunk) ? 0 : (p.A == 1 ? (p.nchunk - c0 + p.B - 1) / p.B : min(p.A, p.nchunk - c0));
    const size_t nseg = nck * p.cl;
    const size_t cstride = p.B * p.cl * WSEG;  // elements from one of this wavefront's chunks to the next
    Part wp = part_empty();  // this wavefront's chunks, merged in chunk order
    unsigned wp_fn = 0;
    unsigned nbuf = 0;  // row results waiting in rowbuf[wave]
    // lanes r < nbuf write the waiting row results (one row per lane)
    auto flush_rows = [&]() {
        if (lane < nbuf) {
            const RowRec &r = rowbuf[wave][lane];
            write_outputs<T>(p.row, r.chunk, r.part, p.ddof, r.first_nan, (uint64_t)r.chunk * p.cols);
        }
        nbuf = 0;
    };
    if (nseg > 0) {
        LaneAcc<T> a;
        acc_reset(a);
        unsigned fnan = 0, j = 0;
        size_t k = 0;
        T y[2][M], z[DZ][M];
        // segment t of this wavefront: chunk c0 + (t / cl) B, segment t % cl of it
        const T *n0 = p.x0 + c0 * p.cl * WSEG;  // the next segment to load
        const T *n1 = LEAVES == 2 ? p.x1 + c0 * p.cl * WSEG : nullptr;
        unsigned nj = 0;
        size_t left = nseg;  // segments not yet loaded
        auto load = [&](T(&yy)[M], T(&zz)[M]) {
            lean_load<T, VPT, VEC>(n0, yy);
            if constexpr (LEAVES == 2) lean_load<T, VPT, VEC>(n1, zz);
            if (left > 1) {  // advance (past the last segment the pointer stays: a harmless reload)
                --left;
                if (++nj == p.cl) {
                    nj = 0;
                    n0 += cstride - (size_t)(p.cl - 1) * WSEG;
                    if constexpr (LEAVES == 2) n1 += cstride - (size_t)(p.cl - 1) * WSEG;
                } else {
                    n0 += WSEG;
                    if constexpr (LEAVES == 2) n1 += WSEG;
                }
            }
        };
        const T s0 = (T)p.s0, s1 = (T)p.s1;
        const double inv_n = 1.0 / ((double)M * (double)p.cl * 64.0);  // 1 / (elements per chunk)
        auto step = [&](T(&yy)[M], T(&zz)[M]) {
            lean_fold<T, VPT, FLAGS, LEAVES, CH>(p.prog, s0, s1, yy, zz, j, j, a, fnan);
            if (++j == p.cl) {  // chunk c = c0 + k B is complete
                const size_t c = c0 + k * p.B;
                const uint64_t fbase = (uint64_t)c * p.cl * WSEG;
                auto flat = [&](unsigned code) -> unsigned {  // chunk-relative offset of a lane's code
                    const unsigned sl = code / M, i = code % M;
                    return sl * WSEG + ((i / VN) * 64 + lane) * VN + (i % VN);
                };
                LeanRec<T> r;
                r.hi = a.hi; r.lo = a.lo; r.m2 = a.m2; r.nan = a.nan;
                r.vmax = a.vmax; r.imax = a.cmax == NO_CODE ? NO_CODE : flat(a.cmax);
                r.vmin = a.vmin; r.imin = a.cmin == NO_CODE ? NO_CODE : flat(a.cmin);
                lean_wave_merge<T, FLAGS>(r, (double)M * (double)p.cl);
                Part pa;
                pa.n = (double)M * (double)p.cl * 64.0;
                pa.hi = r.hi; pa.lo = sizeof(T) == 8 ? r.lo : 0.0; pa.m2 = r.m2; pa.nan = r.nan;
                pa.vmax = r.imax == NO_CODE ? -__builtin_inf() : (double)r.vmax;
                pa.imax = r.imax == NO_CODE ? NO_INDEX : fbase + r.imax;
                pa.vmin = r.imin == NO_CODE ? __builtin_inf() : (double)r.vmin;
                pa.imin = r.imin == NO_CODE ? NO_INDEX : fbase + r.imin;
                const unsigned fn0 = (unsigned)__shfl(fnan, 0, 64);
                if (p.row_direct) {
                    if (lane == 0) rowbuf[wave][nbuf] = RowRec{pa, (uint64_t)c, fn0};
                    if (++nbuf == LEAN_ROWBUF) flush_rows();
                }
                if (k == 0) wp_fn = fn0;
                lean_merge_next<FLAGS>(wp, pa, k, inv_n);
                acc_reset(a);
                fnan = 0;
                j = 0;
                ++k;
            }
        };
        // two segments in turn: the next one's loads are in flight while one is folded. The steady
        // state issues every load and fold unconditionally (exact load counting); an odd last
        // segment is already loaded when it ends.
        load(y[0], z[0]);
        const size_t npair = nseg / 2;
        for (size_t it = 0; it < npair; ++it) {
            load(y[1], z[LEAVES == 2 ? 1 : 0]);
            step(y[0], z[0]);
            load(y[0], z[0]);
            step(y[1], z[LEAVES == 2 ? 1 : 0]);
        }
        if (nseg & 1) step(y[0], z[0]);
    }
    if (is_tail) lean_tail<T, FLAGS, LEAVES, CH>(p, wave, lane, wp);
    if (!p.need_group) {  // uniform over the grid
        if (p.row_direct) flush_rows();
        return;
    }
    lean_group_finish<T, FLAGS>(fin, p.grp, p.counter, p.full, p.ddof, p.uniform, wp, wp_fn, [&]() {
        if (p.row_direct) flush_rows();
    });
}

// ---------------- short rows ----------------
// Rows of at most 64 K elements (K slots per lane: 16 binary32, 8 binary64): G lanes per row, the
// smallest power of two with G K >= cols, so a wavefront reduces a block of 64 / G rows at a time
// and wavefront v of the W in the grid takes blocks v, v + W, ... Lane j of a row's G lanes holds
// slot s = element ((s / VN) G + j) VN + s % VN of the row (VN-element vectors, the G lanes' side by
// side), a prefix of its slots being inside the row. The whole row is in registers: the sum is
// pairwise per lane (in T), then a butterfly over the row's lanes (double, double-word for
// binary64); the variance is two-pass, sum((y - m)^2) with m the mean rounded to T, less
// n (mean - m)^2. Extrema: strict winners per lane in slot (= element) order, then the butterfly
// with equal values going to the lower element.
template <typename T>
constexpr int rows_k() {
    return sizeof(T) == 4 ? 16 : 8;
}
constexpr size_t ROWS_GROUPS = 512;  // round robin over at most this many workgroups

template <typename T>
struct RowsParams {
    const T *x;
    double s0, s1;  // CH_MULADD's scalars
    Chain chain;
    size_t rows, cols, nblk;
    unsigned G, lg;  // lanes per row and its log2
    double ddof;
    OutPtrs<T> row;
    int need_group, uniform;
    Parts grp;
    OutPtrs<T> full;
    unsigned int *counter;
};

template <typename T, int FLAGS, bool VEC, int CH>
__global__ __launch_bounds__(RB) void rows_kernel(RowsParams<T> p) {
    typedef typename Vec16<T>::type V;
    constexpr int VN = Vec16<T>::N;
    constexpr int K = rows_k<T>();
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
    unsigned vm = 0;  // bit s: slot s is inside the row
#pragma unroll
    for (int s = 0; s < K; ++s) vm |= (off(s) < cols ? 1u : 0u) << s;
    const T s0 = (T)p.s0, s1 = (T)p.s1;
    const double ncols = (double)cols, inv_cols = 1.0 / ncols;
    const double inv_blk = 1.0 / ((double)R * ncols);
    Part wp = part_empty();
    unsigned wp_fn = 0;
    // block b's row for this lane (rows past the end read the last row: a harmless reload)
    auto load = [&](T(&yy)[K], size_t b) {
        const size_t r = min(b * R + rr, p.rows - 1);
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
    };
    auto step = [&](T(&y)[K], size_t it) {
        const size_t b = v + it * W;
        const size_t r = b * R + rr;
        if constexpr (CH == CH_MULADD) {
#pragma unroll
            for (int i = 0; i < K; ++i) y[i] = rn_add(rn_mul(y[i], s0), s1);
        } else if constexpr (CH == CH_ANY) {
            apply_chain_n(p.chain, y);
        }
        const unsigned fn = (unsigned)__shfl(__builtin_isnan(y[0]) ? 1 : 0, (int)(lane & ~(G - 1)), 64);  // element 0
        const T st = pairwise<K, 0, T>([&](int i) { return ((vm >> i) & 1u) ? y[i] : (T)0; });
        unsigned nan = 0;
        if (__builtin_isnan(st)) {
#pragma unroll
            for (int i = 0; i < K; ++i) nan |= (((vm >> i) & 1u) && __builtin_isnan(y[i])) ? 1u : 0u;
        }
        T bmax = -dev_inf<T>(), bmin = dev_inf<T>();
        unsigned imax = NO_CODE, imin = NO_CODE;
#pragma unroll
        for (int i = 0; i < K; ++i) {
            const bool in = (vm >> i) & 1u;
            if (FLAGS & F_MAX) {
                const bool t = in && y[i] > bmax;
                bmax = t ? y[i] : bmax;
                imax = t ? (unsigned)i : imax;
            }
            if (FLAGS & F_MIN) {
                const bool t = in && y[i] < bmin;
                bmin = t ? y[i] : bmin;
                imin = t ? (unsigned)i : imin;
            }
        }
        if (FLAGS & F_MAX) imax = imax == NO_CODE ? NO_CODE : off(imax);
        if (FLAGS & F_MIN) imin = imin == NO_CODE ? NO_CODE : off(imin);
        double hi = (double)st, lo = 0.0;
        for (unsigned m = G >> 1; m >= 1; m >>= 1) {  // uniform: G is the same for every lane
            if constexpr (sizeof(T) == 8) {
                const double bh = shfl_xor(hi, (int)m), bl = shfl_xor(lo, (int)m);
                dw_add(hi, lo, bh, bl);
            } else {
                hi += shfl_xor(hi, (int)m);
            }
            if (FLAGS & F_MAX) lean_take_max(bmax, imax, shfl_xor(bmax, (int)m), shfl_xor_u(imax, m));
            if (FLAGS & F_MIN) lean_take_min(bmin, imin, shfl_xor(bmin, (int)m), shfl_xor_u(imin, m));
            nan |= shfl_xor_u(nan, m);
        }
        double m2 = 0.0;
        if (FLAGS & F_M2) {
            const double mu = (hi + lo) / ncols;  // divided: a constant row has m = its value, variance 0
            const T mt = (T)mu;
            T q4[4] = {(T)0, (T)0, (T)0, (T)0};
#pragma unroll
            for (int i = 0; i < K; ++i) {
                const T d = ((vm >> i) & 1u) ? rn_sub(y[i], mt) : (T)0;
                q4[i & 3] = fma(d, d, q4[i & 3]);
            }
            double q = (double)rn_add(rn_add(q4[0], q4[1]), rn_add(q4[2], q4[3]));
            for (unsigned m = G >> 1; m >= 1; m >>= 1) q += shfl_xor(q, (int)m);
            const double dm = mu - (double)mt;
            m2 = q - ncols * dm * dm;
            m2 = m2 < 0.0 ? 0.0 : m2;  // (NaN stays NaN)
        }
        const bool rv = r < p.rows;
        Part a;
        a.n = rv ? ncols : 0.0;
        a.hi = rv ? hi : 0.0;
        a.lo = rv ? lo : 0.0;
        a.m2 = rv ? m2 : 0.0;
        a.nan = rv ? nan : 0u;
        a.vmax = (rv && imax != NO_CODE) ? (double)bmax : -__builtin_inf();
        a.imax = (rv && imax != NO_CODE) ? (uint64_t)imax : NO_INDEX;
        a.vmin = (rv && imin != NO_CODE) ? (double)bmin : __builtin_inf();
        a.imin = (rv && imin != NO_CODE) ? (uint64_t)imin : NO_INDEX;
        if (j == 0 && rv) write_outputs<T>(p.row, r, a, p.ddof, fn, 0);
        if (!p.need_group) return;  // uniform over the grid
        // the block's rows merged (lanes G, 2G, ... apart), then into this wavefront's record
        const uint64_t rb = (uint64_t)r * p.cols;
        if (a.imax != NO_INDEX) a.imax += rb;
        if (a.imin != NO_INDEX) a.imin += rb;
        const bool whole = (b + 1) * R <= p.rows;  // uniform over the wavefront
        if (whole) {
            double w = 0.5 * inv_cols;
            for (unsigned m = G; m < 64; m <<= 1) {
                const Part o = part_shfl<FLAGS>(a, (int)m);
                part_merge_eq<FLAGS>(a, o, w);
                w *= 0.5;
            }
        } else {
            for (unsigned m = G; m < 64; m <<= 1) {
                const Part o = part_shfl<FLAGS>(a, (int)m);
                part_merge<FLAGS>(a, o);
            }
        }
        if (it == 0) wp_fn = (unsigned)__shfl((int)fn, 0, 64);
        if (whole)
            lean_merge_next<FLAGS>(wp, a, it, inv_blk);
        else
            part_merge<FLAGS>(wp, a);
    };
    // two blocks in turn: the next one's loads are in flight while one is folded
    if (nit > 0) {
        T y[2][K];
        load(y[0], v);
        const size_t npair = nit / 2;
        for (size_t it = 0; it < npair; ++it) {
            load(y[1], v + (2 * it + 1) * W);
            step(y[0], 2 * it);
            load(y[0], v + (2 * it + 2) * W);
            step(y[1], 2 * it + 1);
        }
        if (nit & 1) step(y[0], nit - 1);
    }
    if (!p.need_group) return;
    lean_group_finish<T, FLAGS>(fin, p.grp, p.counter, p.full, p.ddof, p.uniform, wp, wp_fn, []() {});
}

// Chunk length for a whole-array lean reduction of spr full segments: the smallest divisor of spr
// that leaves at most CHUNK_TARGET chunks (0 if none up to the weight table's length).
static size_t lean_chunk_len(size_t spr) {
    const size_t lo = std::max<size_t>(1, ceil_div(spr, CHUNK_TARGET));
    for (size_t cl = lo; cl < (size_t)LEAN_WTAB && cl <= spr; ++cl)
        if (spr % cl == 0) return cl;
    return 0;
}

// The chunk-to-wavefront layout, from the dtype and the shape only (it fixes which chunks a group
// record holds, hence the whole-array result's bits): binary32 wavefront v takes chunks v, v + W,
// ... (W the grid's wavefronts), binary64 8 consecutive chunks per wavefront.
constexpr size_t LEAN_MIN_CHUNKS = 1024;
constexpr size_t LEAN_GROUPS_F32 = 512;  // round robin over at most this many workgroups
constexpr size_t LEAN_BLOCK_F64 = 8;     // consecutive chunks per binary64 wavefront (at least)
static_assert(LEAN_GROUPS_F32 <= LEAN_MAX_GROUPS, "the finish merges at most LEAN_MAX_GROUPS group records");
struct LeanLayout {
    size_t groups, A, B;
};
template <typename T>
static LeanLayout lean_layout(size_t nchunk, size_t max_groups) {
    LeanLayout L;
    if (sizeof(T) == 4) {
        L.groups = std::min<size_t>(LEAN_GROUPS_F32, ceil_div(nchunk, (size_t)4 * WAVES));
        L.A = 1;
        L.B = L.groups * WAVES;
    } else {
        L.A = std::max<size_t>(LEAN_BLOCK_F64, ceil_div(nchunk, max_groups * WAVES));
        L.groups = ceil_div(nchunk, L.A * WAVES);
        L.B = 1;
    }
    return L;
}

template <typename T, int VPT, int FLAGS, int LEAVES, int CH>
static void launch_lean_kernel(const LeanParams<T> &p, size_t ngroups, bool vec, hipStream_t stream) {
    if (vec)
        hipLaunchKernelGGL((lean_kernel<T, VPT, FLAGS, LEAVES, true, CH>), dim3((unsigned)ngroups), dim3(RB), 0, stream, p);
    else
        hipLaunchKernelGGL((lean_kernel<T, VPT, FLAGS, LEAVES, false, CH>), dim3((unsigned)ngroups), dim3(RB), 0, stream, p);
}

template <typename T, int VPT, int FLAGS, int LEAVES>
static nrk_status launch_lean(const LeanParams<T> &p, size_t ngroups, hipStream_t stream) {
    const bool vec = (uintptr_t)p.x0 % 16 == 0 && (LEAVES == 1 || (uintptr_t)p.x1 % 16 == 0);
    if constexpr (LEAVES == 2) {
        launch_lean_kernel<T, VPT, FLAGS, 2, CH_ANY>(p, ngroups, vec, stream);
    } else {
        const Chain &c = p.prog.pre0;
        if (c.len == 0)
            launch_lean_kernel<T, VPT, FLAGS, 1, CH_NONE>(p, ngroups, vec, stream);
        else if (c.len == 2 && c.op[0] == NRK_SOP_MUL && (c.op[1] == NRK_SOP_ADD || c.op[1] == NRK_SOP_SUB))
            launch_lean_kernel<T, VPT, FLAGS, 1, CH_MULADD>(p, ngroups, vec, stream);
        else
            launch_lean_kernel<T, VPT, FLAGS, 1, CH_ANY>(p, ngroups, vec, stream);
    }
    NRK_HIP(hipGetLastError());
    return NRK_OK;
}

template <typename T, int FLAGS, int CH>
static void launch_rows_kernel(const RowsParams<T> &p, size_t ngroups, bool vec, hipStream_t stream) {
    if (vec)
        hipLaunchKernelGGL((rows_kernel<T, FLAGS, true, CH>), dim3((unsigned)ngroups), dim3(RB), 0, stream, p);
    else
        hipLaunchKernelGGL((rows_kernel<T, FLAGS, false, CH>), dim3((unsigned)ngroups), dim3(RB), 0, stream, p);
}

template <typename T, int FLAGS>
static nrk_status launch_rows(const RowsParams<T> &p, size_t ngroups, hipStream_t stream) {
    constexpr int VN = Vec16<T>::N;
    const bool vec = (uintptr_t)p.x % 16 == 0 && p.cols % VN == 0;
    const Chain &c = p.chain;
    if (c.len == 0)
        launch_rows_kernel<T, FLAGS, CH_NONE>(p, ngroups, vec, stream);
    else if (c.len == 2 && c.op[0] == NRK_SOP_MUL && (c.op[1] == NRK_SOP_ADD || c.op[1] == NRK_SOP_SUB))
        launch_rows_kernel<T, FLAGS, CH_MULADD>(p, ngroups, vec, stream);
    else
        launch_rows_kernel<T, FLAGS, CH_ANY>(p, ngroups, vec, stream);
    NRK_HIP(hipGetLastError());
    return NRK_OK;
}

__global__ void lean_clear_counter_kernel(unsigned int *c) { c[threadIdx.x * LEAN_CTR_STRIDE] = 0u; }

// The arrival counters (zero) and ngroups group records of the whole-array finish: the caller's
// scratch (counters first, cleared here; records after) or the stream's.
static nrk_status group_scratch(size_t ngroups, size_t rows, size_t cols, const ExtScratch *ext, hipStream_t stream,
                                unsigned int **counter, Parts *grp) {
    const size_t pbytes = parts_bytes(ngroups) + 256;
    char *b = nullptr;
    constexpr size_t cbytes = LEAN_CTR_WORDS * sizeof(unsigned);
    if (ext) {
        if (pbytes + cbytes > ext->bytes)
            return fail(NRK_ERR_INVALID_ARGUMENT, "(%zu, %zu) needs %zu bytes of scratch; the fixed scratch holds %zu", rows, cols,
                        pbytes + cbytes, ext->bytes);
        *counter = (unsigned int *)ext->base;
        b = (char *)ext->base + cbytes;
        hipLaunchKernelGGL(lean_clear_counter_kernel, dim3(1), dim3(LEAN_CTR_WORDS / LEAN_CTR_STRIDE), 0, stream, *counter);
        NRK_HIP(hipGetLastError());
    } else {
        Scratch s;
        NRK_TRY(get_scratch(stream, LEAN_CTR_WORDS, pbytes, &s));
        *counter = s.counters;
        b = (char *)s.partials;
    }
    *grp = carve_parts(b, ngroups);
    return NRK_OK;
}

// Row results (and whole-array ones) of rows of at most 64 K elements, one operand.
template <typename T>
static nrk_status run_rows(const Chain &chain, const T *x, size_t rows, size_t cols, double ddof, const OutPtrs<T> &row,
                           const OutPtrs<T> &full, bool want_full, hipStream_t stream, const ExtScratch *ext) {
    constexpr int K = rows_k<T>();
    RowsParams<T> p{};
    p.x = x;
    p.chain = chain;
    p.rows = rows;
    p.cols = cols;
    p.ddof = ddof;
    unsigned lg = 0;
    while (((size_t)K << lg) < cols) ++lg;
    p.G = 1u << lg;
    p.lg = lg;
    const size_t R = (size_t)64 >> lg;
    p.nblk = ceil_div(rows, R);
    const size_t ngroups = std::min<size_t>(ROWS_GROUPS, ceil_div(p.nblk, (size_t)WAVES));
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
    case F: return launch_rows<T, F>(p, ngroups, stream);
    switch (flags) {
        NRK_ROWS(0) NRK_ROWS(1) NRK_ROWS(2) NRK_ROWS(3) NRK_ROWS(4) NRK_ROWS(5) NRK_ROWS(6)
        default: NRK_ROWS(7)
    }
#undef NRK_ROWS
}

// The lean path, if the shape and operands allow it (else returns false in *taken). x1 != NULL:
// two operands (prog.op and the three chains), else one (prog.pre0).
template <typename T>
nrk_status try_lean(const LeanProg &prog, const T *x0, const T *x1, size_t rows, size_t cols, double ddof,
                           const OutPtrs<T> &row, const OutPtrs<T> &full, bool want_row, bool want_full, hipStream_t stream,
                           const ExtScratch *ext, bool *taken) {
    constexpr int M = NRK_REDUCE_M;
    constexpr int VPT = M / Vec16<T>::N;
    constexpr size_t WSEG = 64 * M;
    *taken = false;
    // the choice depends on the shape only (never on alignment or the device), so a given shape
    // always gets the same summation tree
    size_t cl, nchunk, tail0 = rows * cols;
    const bool row_chunks = cols % WSEG == 0 && cols / WSEG < (size_t)LEAN_WTAB && rows >= LEAN_MIN_CHUNKS;
    if (want_row && !row_chunks && !x1 && cols <= 64 * (size_t)rows_k<T>()) {  // short rows
        *taken = true;
        return run_rows<T>(prog.pre0, x0, rows, cols, ddof, row, full, want_full, stream, ext);
    }
    if (want_row) {  // a chunk is a row
        if (!row_chunks) return NRK_OK;
        cl = cols / WSEG;
        nchunk = rows;
    } else {
        // whole segments in chunks of an exact divisor when there is one; else (any length)
        // chunks of ceil(segments / CHUNK_TARGET) segments and the rest as a tail
        const size_t n = rows * cols, spr = n / WSEG;
        cl = n % WSEG == 0 ? lean_chunk_len(spr) : 0;
        if (cl == 0) {
            cl = std::max<size_t>(1, ceil_div(spr, CHUNK_TARGET));
            if (cl >= (size_t)LEAN_WTAB) return NRK_OK;
            tail0 = spr / cl * cl * WSEG;
        }
        nchunk = spr / cl;
        if (nchunk < LEAN_MIN_CHUNKS) return NRK_OK;
    }
    LeanParams<T> p{};
    p.x0 = x0;
    p.x1 = x1;
    p.nchunk = nchunk;
    p.cl = cl;
    p.cols = cols;
    const bool tail = tail0 < rows * cols;
    const LeanLayout L = lean_layout<T>(nchunk, LEAN_MAX_GROUPS - (tail ? 1 : 0));
    const size_t ngroups = L.groups + (tail ? 1 : 0);
    p.A = L.A;
    p.B = L.B;
    p.tail0 = tail0;
    p.n_total = rows * cols;
    p.tail_group = tail ? (unsigned)L.groups : ~0u;
    // equal chunk counts on every wavefront of the grid: round robin over a multiple of the
    // wavefront count, or contiguous blocks that fill every wavefront
    p.uniform = !tail && (L.A == 1 ? nchunk % (L.groups * WAVES) == 0 : nchunk == L.A * L.groups * WAVES);
    p.prog = prog;
    if (!x1 && prog.pre0.len == 2) {  // CH_MULADD: y = (x * s0) + s1, with x - b as x + (-b)
        p.s0 = prog.pre0.s[0];
        p.s1 = prog.pre0.op[1] == NRK_SOP_SUB ? -prog.pre0.s[1] : prog.pre0.s[1];
    }
    p.ddof = ddof;
    p.row_direct = want_row;
    p.row = want_row ? row : OutPtrs<T>{};
    p.need_group = want_full;
    p.full = want_full ? full : OutPtrs<T>{};
    if (want_full) NRK_TRY(group_scratch(ngroups, rows, cols, ext, stream, &p.counter, &p.grp));
    const bool m2 = (row.var || full.var);
    const bool mx = (row.max || row.argmax || full.max || full.argmax);
    const bool mn = (row.min || row.argmin || full.min || full.argmin);
    const int flags = (m2 ? F_M2 : 0) | (mx ? F_MAX : 0) | (mn ? F_MIN : 0);
    *taken = true;
#define NRK_LEAN(F)                                                                             \
    case F:                                                                                     \
        return x1 ? launch_lean<T, VPT, F, 2>(p, ngroups, stream) : launch_lean<T, VPT, F, 1>(p, ngroups, stream);
    switch (flags) {
        NRK_LEAN(0) NRK_LEAN(1) NRK_LEAN(2) NRK_LEAN(3) NRK_LEAN(4) NRK_LEAN(5) NRK_LEAN(6)
        default: NRK_LEAN(7)
    }
#undef NRK_LEAN
}

// Host launcher for the fused map -> reduce over a (rows, cols) matrix.
template nrk_status try_lean<float>(const LeanProg &, const float *, const float *, size_t, size_t, double,
                                    const OutPtrs<float> &, const OutPtrs<float> &, bool, bool, hipStream_t,
                                    const ExtScratch *, bool *);
template nrk_status try_lean<double>(const LeanProg &, const double *, const double *, size_t, size_t, double,
                                     const OutPtrs<double> &, const OutPtrs<double> &, bool, bool, hipStream_t,
                                     const ExtScratch *, bool *);

}  // namespace nrk
