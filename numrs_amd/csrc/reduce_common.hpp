// Shared pieces of the streaming reductions (reduce.hip, reduce_lean.hip): constants, output
// pointer sets, partial records and their merges (Chan, Golub & LeVeque for the second moment,
// double-word sums, (value, index) winners), and the record I/O.
#pragma once

#include "ops.hpp"

#include <algorithm>

namespace nrk {

constexpr int RB = 256;  // threads per streaming workgroup
constexpr int WAVES = RB / 64;
constexpr uint64_t NO_INDEX = ~0ull;
constexpr unsigned NO_CODE = 0xffffffffu;
constexpr int FB = 1024;                // threads of the single-workgroup finish kernels
constexpr size_t CHUNK_TARGET = 8192;   // chunks per matrix when rows are few (about one per wavefront)

enum : int { F_M2 = 1, F_MAX = 2, F_MIN = 4 };

// Elements per lane in one wavefront segment of the f32/f64 streaming kernel (16: 4 16-byte
// vectors of f32 or 8 of f64 per lane; a segment is 4 KiB of f32 or 8 KiB of f64). It fixes the
// segment boundaries and hence the summation tree, so results are reproducible for a given build.
#ifndef NRK_REDUCE_M
#define NRK_REDUCE_M 16
#endif


template <typename T>
struct OutPtrs {
    T *sum, *mean, *var, *min, *max;
    uint64_t *argmin, *argmax;
};

// A set of partial records, structure of arrays. Indices are flat (row * cols + col).
struct Parts {
    double *n, *hi, *lo, *m2, *vmax, *vmin;
    uint64_t *imax, *imin;
    unsigned int *flags;  // bit0: a NaN was seen; bit1: the set's first element is NaN
};

template <typename T>
struct SegParams {
    const T *x;
    size_t rows, cols;
    size_t spr;     // segments per row
    size_t cl;      // segments per chunk (the last chunk of a row may hold fewer)
    size_t cpr;     // chunks per row
    size_t nchunk;  // rows * cpr
    int phased;     // cols is not a multiple of the vector: row r's segments start at the vector-aligned
                    // element r cols - ph, ph = (r cols) mod VN, its first ph positions masked
    Chain chain;
    double ddof;
    int row_direct;  // cpr == 1 and row results wanted: the stream writes them
    int need_l1;     // chunk records are read by a tree level
    Parts l1;        // chunk records, chunk c = row * cpr + q
    OutPtrs<T> row;
};

// One level of the record tree: `nsets` sets of `per` consecutive input records; a wavefront
// merges up to TREE_SPAN consecutive records of one set into one output record.
constexpr int TREE_LOADS = 4;               // records per lane per level, loaded together
constexpr int TREE_SPAN = 64 * TREE_LOADS;  // records per wavefront per level
template <typename T>
struct TreeParams {
    Parts in, out;
    size_t nsets, per, per_out;  // per_out = ceil(per / TREE_SPAN)
    int results;                 // per_out == 1: write the set's results to `o`
    int keep;                    // write output records (a later level reads them)
    int rowmode;                 // set r is row r: flat index base r * cols
    size_t cols;
    double ddof;
    OutPtrs<T> o;
    // fused last level (one set, per_out <= TREE_SPAN): the workgroup that finishes last merges
    // the level's output records and writes the results; `counter` is zero at launch and is
    // reset to zero by that workgroup
    int fuse_last;
    OutPtrs<T> last_o;
    unsigned int *counter;
};

// ---------------- small device utilities ----------------
// Record fields stored and loaded coherently across the chip's L2 caches: relaxed agent-scope
// atomic accesses (the stores write through to the point of coherence, the loads read there).
template <typename U>
__device__ __forceinline__ void st_agent(U *a, U v) {
    __hip_atomic_store(a, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
template <typename U>
__device__ __forceinline__ U ld_agent(U *a) {
    return __hip_atomic_load(a, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
// s_waitcnt vmcnt(0), expcnt and lgkmcnt unconstrained (gfx9 encoding: vmcnt in bits 3:0 and
// 15:14, expcnt 6:4, lgkmcnt 11:8): every vector memory access of this wavefront has completed.
#define NRK_WAIT_VMEM0() __builtin_amdgcn_s_waitcnt((0x7 << 4) | (0xf << 8))

// s_waitcnt lgkmcnt(0) then a bare s_barrier: this wavefront's LDS accesses are complete; its
// global loads and stores may still be in flight (the fence of __syncthreads would wait for them).
// The empty asm statements keep the compiler from moving memory accesses across the barrier.
__device__ __forceinline__ void lds_barrier() {
    __builtin_amdgcn_s_waitcnt(0xf | (0x7 << 4) | (0x3 << 14));
    asm volatile("" ::: "memory");
    __builtin_amdgcn_s_barrier();
    asm volatile("" ::: "memory");
}

template <typename T>
__device__ __forceinline__ void two_sum(T a, T b, T &s, T &e) {
    s = rn_add(a, b);
    T v = rn_sub(s, a);
    e = rn_add(rn_sub(a, rn_sub(s, v)), rn_sub(b, v));
}

// (hi, lo) = s + e renormalised. A sum that is not finite stays the plain sum: hi = s (inf, or NaN
// for inf - inf; TwoSum's error term is NaN there) or the overflowed s + e, and lo = 0.
template <typename T>
__device__ __forceinline__ void dw_fix(T s, T e, T &hi, T &lo) {
    const T h = rn_add(s, e);
    hi = __builtin_isfinite(s) ? h : s;
    lo = __builtin_isfinite(h) ? rn_sub(e, rn_sub(h, s)) : (T)0;
}

// double-word addition (hi + lo) += (bhi + blo)
template <typename T>
__device__ __forceinline__ void dw_add(T &hi, T &lo, T bhi, T blo) {
    T s, e;
    two_sum(hi, bhi, s, e);
    e = rn_add(e, rn_add(lo, blo));
    dw_fix(s, e, hi, lo);
}

__device__ __forceinline__ float shfl_xor(float v, int m) { return __shfl_xor(v, m, 64); }
__device__ __forceinline__ double shfl_xor(double v, int m) { return __shfl_xor(v, m, 64); }
__device__ __forceinline__ uint64_t shfl_xor(uint64_t v, int m) {
    unsigned lo = __shfl_xor((unsigned)(v & 0xffffffffu), m, 64);
    unsigned hi = __shfl_xor((unsigned)(v >> 32), m, 64);
    return ((uint64_t)hi << 32) | lo;
}

// Lane 0's value, for every lane (the wavefront is fully active where this is used).
__device__ __forceinline__ unsigned lane0(unsigned v) { return (unsigned)__builtin_amdgcn_readfirstlane((int)v); }

// (value, index) winners: larger (smaller) value wins, equal values go to the lower index.
// Values are never NaN here (NaN elements never become winners).
template <typename T>
__device__ __forceinline__ void take_max(T &v, uint64_t &i, T bv, uint64_t bi) {
    if (bv > v || (bv == v && bi < i)) { v = bv; i = bi; }
}
template <typename T>
__device__ __forceinline__ void take_min(T &v, uint64_t &i, T bv, uint64_t bi) {
    if (bv < v || (bv == v && bi < i)) { v = bv; i = bi; }
}

// Pairwise (balanced binary tree) sum of f(OFF) .. f(OFF+N-1), written depth-first.
template <int N, int OFF, typename T, typename F>
__device__ __forceinline__ T pairwise(const F &f) {
    if constexpr (N == 1) {
        return f(OFF);
    } else {
        return rn_add(pairwise<N / 2, OFF, T>(f), pairwise<N / 2, OFF + N / 2, T>(f));
    }
}

// A partial record in registers; merge() = Chan, Golub & LeVeque (1979) for the second moment:
// M2 = M2a + M2b + (na nb / n) (mean_a - mean_b)^2, in FP64; double-word sums.
struct Part {
    double n, hi, lo, m2, vmax, vmin;
    uint64_t imax, imin;
    unsigned nan;
};

__device__ __forceinline__ Part part_empty() {
    Part a;
    a.n = 0.0; a.hi = 0.0; a.lo = 0.0; a.m2 = 0.0;
    a.vmax = -__builtin_inf(); a.vmin = __builtin_inf();
    a.imax = NO_INDEX; a.imin = NO_INDEX; a.nan = 0;
    return a;
}

template <int FLAGS>
__device__ __forceinline__ void part_merge(Part &a, const Part &b) {
    a.nan |= b.nan;
    if (FLAGS & F_MAX) take_max(a.vmax, a.imax, b.vmax, b.imax);
    if (FLAGS & F_MIN) take_min(a.vmin, a.imin, b.vmin, b.imin);
    if (b.n == 0.0) return;
    if (a.n == 0.0) {
        a.n = b.n; a.hi = b.hi; a.lo = b.lo; a.m2 = b.m2;
        return;
    }
    if (FLAGS & F_M2) {
        // (na nb / n)(mean_a - mean_b)^2 = (nb Sa - na Sb)^2 / (na nb n): one division
        const double d = b.n * (a.hi + a.lo) - a.n * (b.hi + b.lo);
        a.m2 = a.m2 + b.m2 + d * d / (a.n * b.n * (a.n + b.n));
    }
    dw_add(a.hi, a.lo, b.hi, b.lo);
    a.n += b.n;
}

template <int FLAGS>
__device__ __forceinline__ Part part_shfl(const Part &a, int m) {
    Part b;
    b.n = shfl_xor(a.n, m);
    b.hi = shfl_xor(a.hi, m);
    b.lo = shfl_xor(a.lo, m);
    b.m2 = (FLAGS & F_M2) ? shfl_xor(a.m2, m) : 0.0;
    b.vmax = (FLAGS & F_MAX) ? shfl_xor(a.vmax, m) : -__builtin_inf();
    b.imax = (FLAGS & F_MAX) ? shfl_xor(a.imax, m) : NO_INDEX;
    b.vmin = (FLAGS & F_MIN) ? shfl_xor(a.vmin, m) : __builtin_inf();
    b.imin = (FLAGS & F_MIN) ? shfl_xor(a.imin, m) : NO_INDEX;
    b.nan = (unsigned)__shfl_xor((int)a.nan, m, 64);
    return b;
}

template <int FLAGS>
__device__ __forceinline__ void wave_merge(Part &a) {
#pragma unroll
    for (int m = 32; m >= 1; m >>= 1) {
        const Part b = part_shfl<FLAGS>(a, m);
        part_merge<FLAGS>(a, b);
    }
}

// Record I/O touches only the fields the op set uses.
template <int FLAGS>
__device__ __forceinline__ Part part_load(const Parts &P, size_t s) {
    Part a = part_empty();
    a.n = P.n[s]; a.hi = P.hi[s]; a.lo = P.lo[s];
    if (FLAGS & F_M2) a.m2 = P.m2[s];
    if (FLAGS & F_MAX) { a.vmax = P.vmax[s]; a.imax = P.imax[s]; }
    if (FLAGS & F_MIN) { a.vmin = P.vmin[s]; a.imin = P.imin[s]; }
    a.nan = P.flags[s] & 1u;
    return a;
}

template <int FLAGS>
__device__ __forceinline__ void part_store(const Parts &P, size_t s, const Part &a, unsigned first_nan) {
    P.n[s] = a.n; P.hi[s] = a.hi; P.lo[s] = a.lo;
    if (FLAGS & F_M2) P.m2[s] = a.m2;
    if (FLAGS & F_MAX) { P.vmax[s] = a.vmax; P.imax[s] = a.imax; }
    if (FLAGS & F_MIN) { P.vmin[s] = a.vmin; P.imin[s] = a.imin; }
    P.flags[s] = a.nan | (first_nan << 1);
}

// Record I/O at agent scope (coherent across the XCDs' L2s without a fence), for records that
// another workgroup of the same launch reads.
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

// Write one set of results. `index_base` is subtracted from the flat winner indices (a row's
// first flat index for row results, 0 for whole-matrix results). No winner means every element
// is -inf (+inf) or NaN; the first one is then the answer, and index 0 is its position.
template <typename T>
__device__ __forceinline__ void write_outputs(const OutPtrs<T> &o, size_t at, const Part &a, double ddof, unsigned first_nan,
                                              uint64_t index_base) {
    const double s = a.hi + a.lo;
    if (o.sum) o.sum[at] = (T)s;
    if (o.mean) o.mean[at] = (T)(s / a.n);
    if (o.var) o.var[at] = (T)(a.m2 / (a.n - ddof));
    if (o.max) o.max[at] = a.nan ? dev_nan<T>() : (T)a.vmax;
    if (o.min) o.min[at] = a.nan ? dev_nan<T>() : (T)a.vmin;
    if (o.argmax) o.argmax[at] = first_nan ? 0 : (a.imax == NO_INDEX ? 0 : a.imax - index_base);
    if (o.argmin) o.argmin[at] = first_nan ? 0 : (a.imin == NO_INDEX ? 0 : a.imin - index_base);
}

// Per-lane running state over the segments of one chunk.
template <typename T>
struct LaneAcc {
    double n, hi, lo, m2;
    T vmax, vmin;
    unsigned cmax, cmin;  // winner codes: segment-in-chunk * M + element slot; NO_CODE: none yet
    unsigned nan;
};

template <typename T>
__device__ __forceinline__ void acc_reset(LaneAcc<T> &a) {
    a.n = 0.0; a.hi = 0.0; a.lo = 0.0; a.m2 = 0.0;
    a.vmax = -dev_inf<T>(); a.vmin = dev_inf<T>();
    a.cmax = NO_CODE; a.cmin = NO_CODE; a.nan = 0;
}

template <typename T>
static OutPtrs<T> to_ptrs(const nrk_reduce_outputs *o) {
    OutPtrs<T> p{};
    if (!o) return p;
    p.sum = (T *)o->sum;
    p.mean = (T *)o->mean;
    p.var = (T *)o->var;
    p.min = (T *)o->min;
    p.max = (T *)o->max;
    p.argmin = o->argmin;
    p.argmax = o->argmax;
    return p;
}

template <typename T>
static bool any_out(const OutPtrs<T> &o) {
    return o.sum || o.mean || o.var || o.min || o.max || o.argmin || o.argmax;
}

static Parts carve_parts(char *&b, size_t count) {
    auto take = [&](size_t bytes) {
        void *r = b;
        b += align_up(bytes, 256);
        return r;
    };
    Parts P;
    P.n = (double *)take(count * 8);
    P.hi = (double *)take(count * 8);
    P.lo = (double *)take(count * 8);
    P.m2 = (double *)take(count * 8);
    P.vmax = (double *)take(count * 8);
    P.vmin = (double *)take(count * 8);
    P.imax = (uint64_t *)take(count * 8);
    P.imin = (uint64_t *)take(count * 8);
    P.flags = (unsigned *)take(count * 4);
    return P;
}
static size_t parts_bytes(size_t count) { return 8 * align_up(count * 8, 256) + align_up(count * 4, 256); }

// y = post(pre0(x0) op pre1(x1)) with two operands, post(pre0(x0)) with one (post empty then).
struct LeanProg {
    Chain pre0, pre1, post;
    int op;
};

// The lean path (reduce_lean.hip), if the shape allows it (else *taken = false, nothing launched).
template <typename T>
nrk_status try_lean(const LeanProg &prog, const T *x0, const T *x1, size_t rows, size_t cols, double ddof,
                    const OutPtrs<T> &row, const OutPtrs<T> &full, bool want_row, bool want_full, hipStream_t stream,
                    const ExtScratch *ext, bool *taken);

}  // namespace nrk
