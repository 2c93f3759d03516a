// Streaming reductions for gfx950: sum / mean / var / min / max / argmin / argmax over rows of
// a row-major matrix and over the whole matrix, optionally after an elementwise scalar chain
// (the fused map -> reduce path), reading the input from HBM exactly once.
//
// Reference semantics restated here (paths under the reference's src/):
//   kernels/reduce.rs:207-421   sum/mean/var(ddof)/min/max; empty -> 0; any NaN -> min/max NaN
//   math/statistics/mod.rs:79-183   argmax: first strict maximum, interior NaN skipped, NaN at
//                                   index 0 pins the result to 0 (argmin symmetric, :218-322)
//   math/aggregation.rs:244-261     axis max/min: NaN short-circuits to NaN
//   math/statistics/mod.rs:1403-1494 axis var = sum((x - mean)^2) / (len - ddof)
//
// Work decomposition. Row r of `cols` elements is cut into fixed segments of WSEG elements
// (64 lanes x VPT 16-byte vectors), each owned by ONE wavefront of the persistent streaming
// kernel. The wavefront reduces its segment from registers -- pairwise sums per lane, the
// segment mean, then squared deviations about it, combined across lanes with xor butterflies of
// FP64 additions -- and writes one partial record. Wavefronts never wait for each other (no
// barrier, no LDS). Two small dependent kernels on the same stream finish: one wavefront per
// row (or per group of 64 segments) combines segment partials and writes row results, then one
// 1024-thread workgroup combines those for the whole-matrix result, using Chan, Golub & LeVeque's
// pairwise update for the second moment. The kernel boundaries provide the cross-workgroup
// visibility; there is no fence or atomic in the stream (an agent-scope release/acquire writes
// back / invalidates each XCD's L2 on gfx950, which cost 20x the stream itself when it was done
// per segment).
//
// Determinism. Every combine is a fixed tree: pairwise within a lane, an xor butterfly within
// a wavefront (lane 0's value is used), fixed per-lane strided sequences and butterflies in the
// finish. None of it depends on the grid size or on scheduling, so results are bit-identical
// across launches and devices, and across the tuner's grid-size configurations.
//
// Precision. Per-lane partials in T (pairwise), everything after in FP64 (double-word for
// sums); the second moment is combined with Chan's formula. max/min/argmax/argmin are exact:
// the winner is the largest value, ties going to the lowest index, and the value written is the
// element itself, so even the sign of a zero maximum matches the reference's left-to-right fold.
#include "common.hpp"

namespace nrk {

constexpr int RB = 256;  // threads per streaming workgroup
constexpr int WAVES = RB / 64;
constexpr uint64_t NO_INDEX = ~0ull;
constexpr int GROUP = 64;    // segments per finish wavefront when rows need no combining
constexpr int FB = 1024;     // threads of the final workgroup

enum : int { F_M2 = 1, F_MAX = 2, F_MIN = 4 };

// Persistent-grid size of the streaming kernels (workgroups per CU). Result bits do not depend on
// it, only speed, so the tuner may choose it per shape.
static thread_local int g_blocks_per_cu = 8;
int reduce_blocks_per_cu() { return g_blocks_per_cu; }
void set_reduce_blocks_per_cu(int b) { g_blocks_per_cu = b < 1 ? 1 : (b > 16 ? 16 : b); }

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
    size_t rows, cols, spr, nseg;
    Chain chain;
    double ddof;
    int want_row, want_full, need_partials;
    size_t ngroups;  // first-level finish groups (rows when spr > 1 and rows are wanted)
    Parts l1, l2;    // segment partials, group partials
    OutPtrs<T> row, full;
};

// ---------------- small device utilities ----------------
template <typename T>
__device__ __forceinline__ void two_sum(T a, T b, T &s, T &e) {
    s = rn_add(a, b);
    T v = rn_sub(s, a);
    e = rn_add(rn_sub(a, rn_sub(s, v)), rn_sub(b, v));
}

// double-word addition (hi + lo) += (bhi + blo)
template <typename T>
__device__ __forceinline__ void dw_add(T &hi, T &lo, T bhi, T blo) {
    T s, e;
    two_sum(hi, bhi, s, e);
    e = rn_add(e, rn_add(lo, blo));
    hi = rn_add(s, e);
    lo = rn_sub(e, rn_sub(hi, s));
}

__device__ __forceinline__ float shfl_xor(float v, int m) { return __shfl_xor(v, m, 64); }
__device__ __forceinline__ double shfl_xor(double v, int m) { return __shfl_xor(v, m, 64); }
__device__ __forceinline__ uint64_t shfl_xor(uint64_t v, int m) {
    unsigned lo = __shfl_xor((unsigned)(v & 0xffffffffu), m, 64);
    unsigned hi = __shfl_xor((unsigned)(v >> 32), m, 64);
    return ((uint64_t)hi << 32) | lo;
}

// (value, index) winners: larger (smaller) value wins, equal values go to the lower index.
// Values are never NaN here (NaN elements are excluded before this point).
template <typename T>
__device__ __forceinline__ void take_max(T &v, uint64_t &i, T bv, uint64_t bi) {
    if (bv > v || (bv == v && bi < i)) { v = bv; i = bi; }
}
template <typename T>
__device__ __forceinline__ void take_min(T &v, uint64_t &i, T bv, uint64_t bi) {
    if (bv < v || (bv == v && bi < i)) { v = bv; i = bi; }
}

template <typename T, bool IS_MAX>
__device__ __forceinline__ void wave_arg(T &v, uint64_t &i) {
#pragma unroll
    for (int m = 32; m >= 1; m >>= 1) {
        T bv = shfl_xor(v, m);
        uint64_t bi = shfl_xor(i, m);
        if (IS_MAX) take_max(v, i, bv, bi); else take_min(v, i, bv, bi);
    }
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
        const double nt = a.n + b.n;
        const double d = (a.hi + a.lo) / a.n - (b.hi + b.lo) / b.n;
        a.m2 = a.m2 + b.m2 + d * d * (a.n * b.n / nt);
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

__device__ __forceinline__ Part part_load(const Parts &P, size_t s) {
    Part a;
    a.n = P.n[s]; a.hi = P.hi[s]; a.lo = P.lo[s]; a.m2 = P.m2[s];
    a.vmax = P.vmax[s]; a.vmin = P.vmin[s]; a.imax = P.imax[s]; a.imin = P.imin[s];
    a.nan = P.flags[s] & 1u;
    return a;
}

__device__ __forceinline__ void part_store(const Parts &P, size_t s, const Part &a, unsigned first_nan) {
    P.n[s] = a.n; P.hi[s] = a.hi; P.lo[s] = a.lo; P.m2[s] = a.m2;
    P.vmax[s] = a.vmax; P.vmin[s] = a.vmin; P.imax[s] = a.imax; P.imin[s] = a.imin;
    P.flags[s] = a.nan | (first_nan << 1);
}

// Write one set of results. `index_base` is subtracted from the flat winner indices (a row's
// first flat index for row results, 0 for whole-matrix results).
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

// ---------------- the streaming kernel ----------------
// A segment is WSEG = 64 lanes x VPT 16-byte vectors of one row, owned by ONE wavefront: lane l
// holds element i = j*VN + e at segment position (j*64 + l)*VN + e. The wavefront computes its
// count, double-word sum, M2 (two passes over registers: the wavefront mean first, then the
// squared deviations about it, both combined with xor butterflies of plain FP64 additions) and
// extremum winners, and lane 0 writes one partial record. Wavefronts never wait for each other:
// the kernel has no barrier and no LDS.
template <typename T, int VPT, int FLAGS, bool FULLSEG, bool VEC>
__device__ __forceinline__ void reduce_segment(const SegParams<T> &p, size_t seg) {
    typedef typename Vec16<T>::type V;
    constexpr int VN = Vec16<T>::N;
    constexpr int M = VPT * VN;  // elements per lane (<= 32: one validity bit each)
    constexpr int WSEG = 64 * M;
    static_assert(M <= 32, "validity mask holds 32 elements");
    const unsigned lane = threadIdx.x & 63;
    const size_t r = seg / p.spr, c = seg - r * p.spr;
    const size_t col0 = c * WSEG;
    const unsigned len = FULLSEG ? (unsigned)WSEG : (unsigned)min((size_t)WSEG, p.cols - col0);
    const size_t flat0 = r * p.cols + col0;
    const T *base = p.x + flat0;  // uniform across the wavefront

    // ---- load
    T y[M];
    unsigned vmask = ~0u;
    if (VEC) {
        // aligned rows: 16-byte loads; a tail segment predicates whole vectors (len % VN == 0)
        const V *bv = reinterpret_cast<const V *>(base);
        const unsigned nvec = len / VN;
        if (!FULLSEG) vmask = 0u;
#pragma unroll
        for (int j = 0; j < VPT; ++j) {
            const unsigned jv = j * 64 + lane;
            const bool ok = FULLSEG || jv < nvec;
            V v = V{};
            if (ok) v = __builtin_nontemporal_load(bv + jv);
#pragma unroll
            for (int e = 0; e < VN; ++e) y[j * VN + e] = v[e];
            if (!FULLSEG) vmask |= (ok ? ((1u << VN) - 1u) : 0u) << (j * VN);
        }
    } else {
        if (!FULLSEG) vmask = 0u;
#pragma unroll
        for (int j = 0; j < VPT; ++j) {
#pragma unroll
            for (int e = 0; e < VN; ++e) {
                const unsigned pos = (j * 64 + lane) * VN + e;
                const bool ok = FULLSEG || pos < len;
                y[j * VN + e] = ok ? base[pos] : (T)0;
                if (!FULLSEG) vmask |= (ok ? 1u : 0u) << (j * VN + e);
            }
        }
    }
    auto valid = [&](int i) -> bool { return FULLSEG || ((vmask >> i) & 1u); };
    auto pos_of = [&](int i) -> unsigned { return ((i / VN) * 64 + lane) * VN + (i % VN); };

    // ---- elementwise chain, one rounding per step (uniform branch per step)
    for (int k = 0; k < p.chain.len; ++k) {
        const int op = p.chain.op[k];
        const T s = (T)p.chain.s[k];
#pragma unroll
        for (int i = 0; i < M; ++i) y[i] = apply_sop<T>(op, s, y[i]);
    }

    // ---- sum: pairwise per lane in T, then a double-word FP64 butterfly
    const T st = pairwise<M, 0, T>([&](int i) { return valid(i) ? y[i] : (T)0; });
    unsigned nan = 0;
    if (__builtin_isnan(st)) {  // rare: tell a NaN element apart from inf - inf
#pragma unroll
        for (int i = 0; i < M; ++i) nan |= (valid(i) && __builtin_isnan(y[i])) ? 1u : 0u;
    }
    double hi = (double)st, lo = 0.0;
#pragma unroll
    for (int m = 32; m >= 1; m >>= 1) {
        const double bh = shfl_xor(hi, m), bl = shfl_xor(lo, m);
        dw_add(hi, lo, bh, bl);
    }
    const double n = (double)len;
    // ---- second moment about the wavefront mean (values still in registers)
    double m2 = 0.0;
    if (FLAGS & F_M2) {
        const double mean = FULLSEG ? (hi + lo) * (1.0 / WSEG) : (hi + lo) / n;
        const T mt = (T)mean;
        const T qt = pairwise<M, 0, T>([&](int i) {
            const T d = valid(i) ? rn_sub(y[i], mt) : (T)0;
            return rn_mul(d, d);
        });
        double q = (double)qt;
#pragma unroll
        for (int m = 32; m >= 1; m >>= 1) q += shfl_xor(q, m);
        // q is taken about mt, the mean rounded to T: Q = M2 + n (mt - mean)^2
        const double dm = (double)mt - mean;
        m2 = fmax(q - n * dm * dm, 0.0);
    }
    // ---- extremum winners
    T vmax = -dev_inf<T>(), vmin = dev_inf<T>();
    uint64_t imax = NO_INDEX, imin = NO_INDEX;
    if (FLAGS & F_MAX) {
        T m = -dev_inf<T>();
#pragma unroll
        for (int i = 0; i < M; ++i) m = valid(i) ? dev_fmax(m, y[i]) : m;  // fmax skips NaN
        int first = -1;
#pragma unroll
        for (int i = M - 1; i >= 0; --i) first = (valid(i) && y[i] == m) ? i : first;
        if (first >= 0) {
            vmax = m;
#pragma unroll
            for (int i = 0; i < M; ++i) vmax = (i == first) ? y[i] : vmax;  // keeps the sign of a zero
            imax = flat0 + pos_of(first);
        }
        wave_arg<T, true>(vmax, imax);
    }
    if (FLAGS & F_MIN) {
        T m = dev_inf<T>();
#pragma unroll
        for (int i = 0; i < M; ++i) m = valid(i) ? dev_fmin(m, y[i]) : m;
        int first = -1;
#pragma unroll
        for (int i = M - 1; i >= 0; --i) first = (valid(i) && y[i] == m) ? i : first;
        if (first >= 0) {
            vmin = m;
#pragma unroll
            for (int i = 0; i < M; ++i) vmin = (i == first) ? y[i] : vmin;
            imin = flat0 + pos_of(first);
        }
        wave_arg<T, false>(vmin, imin);
    }
    nan = __any((int)nan) ? 1u : 0u;
    const unsigned first_nan = (c == 0 && __builtin_isnan(y[0])) ? 1u : 0u;  // lane 0 holds element 0
    if (lane == 0) {
        const Parts &P = p.l1;
        P.n[seg] = n;
        P.hi[seg] = hi;
        P.lo[seg] = lo;
        P.m2[seg] = m2;
        P.vmax[seg] = (double)vmax;
        P.vmin[seg] = (double)vmin;
        P.imax[seg] = imax;
        P.imin[seg] = imin;
        P.flags[seg] = nan | (first_nan << 1);
    }
}

template <typename T, int VPT, int FLAGS, bool VEC>
__global__ __launch_bounds__(RB) void seg_reduce_kernel(SegParams<T> p) {
    constexpr int WSEG = 64 * VPT * Vec16<T>::N;
    const size_t nwaves = (size_t)gridDim.x * WAVES;
    for (size_t seg = blockIdx.x * (size_t)WAVES + (threadIdx.x >> 6); seg < p.nseg; seg += nwaves) {
        const size_t c = seg % p.spr;
        if (c * WSEG + WSEG <= p.cols) {
            reduce_segment<T, VPT, FLAGS, true, VEC>(p, seg);
        } else {
            reduce_segment<T, VPT, FLAGS, false, VEC>(p, seg);
        }
    }
}

// ---------------- finish, level 1: one wavefront per group of segment partials ----------------
// Group g is row g's segments when row results are wanted, else GROUP consecutive segments.
// Lane l folds partials lo+l, lo+l+64, ... in order, then the xor butterfly.
template <typename T, int FLAGS>
__global__ __launch_bounds__(256) void finish_groups_kernel(SegParams<T> p) {
    const size_t g = (blockIdx.x * (size_t)blockDim.x + threadIdx.x) >> 6;
    const int lane = threadIdx.x & 63;
    if (g >= p.ngroups) return;  // whole wavefronts exit together
    const bool rowmode = p.want_row;
    const size_t per = rowmode ? p.spr : (size_t)GROUP;
    const size_t lo = g * per, hi = min(p.nseg, lo + per);
    Part a = part_empty();
    for (size_t s = lo + lane; s < hi; s += 64) {
        const Part b = part_load(p.l1, s);
        part_merge<FLAGS>(a, b);
    }
    wave_merge<FLAGS>(a);
    if (lane == 0) {
        const unsigned first_nan = (p.l1.flags[lo] >> 1) & 1u;
        if (rowmode) write_outputs<T>(p.row, g, a, p.ddof, first_nan, g * p.cols);
        if (p.want_full) part_store(p.l2, g, a, first_nan);
    }
}

// ---------------- finish, level 2: one workgroup over the group partials ----------------
template <typename T, int FLAGS>
__global__ __launch_bounds__(FB) void finish_full_kernel(SegParams<T> p) {
    __shared__ Part wp[FB / 64];
    const int t = threadIdx.x, w = t >> 6, lane = t & 63;
    Part a = part_empty();
    for (size_t s = t; s < p.ngroups; s += FB) {
        const Part b = part_load(p.l2, s);
        part_merge<FLAGS>(a, b);
    }
    wave_merge<FLAGS>(a);
    if (lane == 0) wp[w] = a;
    __syncthreads();
    if (t == 0) {
        Part tot = wp[0];
        for (int k = 1; k < FB / 64; ++k) part_merge<FLAGS>(tot, wp[k]);
        const unsigned first_nan = (p.l1.flags[0] >> 1) & 1u;
        write_outputs<T>(p.full, 0, tot, p.ddof, first_nan, 0);
    }
}

// Writes T zero into up to 7 outputs; used for empty inputs so that no hipMemsetAsync is ever
// recorded into a caller's graph.
template <typename T>
__global__ void fill_outputs_kernel(OutPtrs<T> o, size_t n, T value) {
    size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
    if (i >= n) return;
    if (o.sum) o.sum[i] = value;
    if (o.mean) o.mean[i] = value;
    if (o.var) o.var[i] = value;
    if (o.min) o.min[i] = value;
    if (o.max) o.max[i] = value;
    if (o.argmin) o.argmin[i] = 0;
    if (o.argmax) o.argmax[i] = 0;
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

template <typename T, int VPT, int FLAGS>
static nrk_status launch_flags(SegParams<T> &p, hipStream_t stream) {
    const int grid = (int)std::min<size_t>(ceil_div(p.nseg, (size_t)WAVES), (size_t)cu_count() * reduce_blocks_per_cu());
    if (((reinterpret_cast<uintptr_t>(p.x) % 16) == 0) && (p.cols % Vec16<T>::N == 0))
        hipLaunchKernelGGL((seg_reduce_kernel<T, VPT, FLAGS, true>), dim3(grid), dim3(RB), 0, stream, p);
    else
        hipLaunchKernelGGL((seg_reduce_kernel<T, VPT, FLAGS, false>), dim3(grid), dim3(RB), 0, stream, p);
    NRK_HIP(hipGetLastError());
    if (!p.need_partials) return NRK_OK;
    hipLaunchKernelGGL((finish_groups_kernel<T, FLAGS>), dim3((unsigned)ceil_div(p.ngroups, 4)), dim3(256), 0, stream, p);
    NRK_HIP(hipGetLastError());
    if (p.want_full) {
        hipLaunchKernelGGL((finish_full_kernel<T, FLAGS>), dim3(1), dim3(FB), 0, stream, p);
        NRK_HIP(hipGetLastError());
    }
    return NRK_OK;
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

// Host launcher for the fused map -> reduce over a (rows, cols) matrix.
template <typename T>
nrk_status seg_reduce(const Chain &chain, const T *x, size_t rows, size_t cols, double ddof,
                      const nrk_reduce_outputs *row_out, const nrk_reduce_outputs *full_out, hipStream_t stream) {
    constexpr int VPT = 8;
    constexpr int WSEG = 64 * VPT * Vec16<T>::N;  // elements per wavefront segment
    NRK_TRY(ensure_device());
    SegParams<T> p{};
    p.row = to_ptrs<T>(row_out);
    p.full = to_ptrs<T>(full_out);
    p.want_row = any_out(p.row);
    p.want_full = any_out(p.full);
    if (!p.want_row && !p.want_full) return NRK_OK;
    if (p.want_row && (p.row.argmax || p.row.argmin || p.row.max || p.row.min) && cols == 0)
        return fail(NRK_ERR_INVALID_ARGUMENT, "max/min/argmax/argmin along an empty axis");
    if (p.want_full && (p.full.argmax || p.full.argmin) && rows * cols == 0)
        return fail(NRK_ERR_INVALID_ARGUMENT, "argmax/argmin of an empty array");
    if (!x && rows * cols > 0) return fail(NRK_ERR_INVALID_ARGUMENT, "x is NULL");
    if (rows * cols == 0) {
        // reference: empty -> 0 for sum/mean/min/max/var (kernels/reduce.rs:207-421)
        if (p.want_full) {
            hipLaunchKernelGGL(fill_outputs_kernel<T>, dim3(1), dim3(64), 0, stream, p.full, (size_t)1, (T)0);
            NRK_HIP(hipGetLastError());
        }
        if (p.want_row && rows > 0) {
            hipLaunchKernelGGL(fill_outputs_kernel<T>, dim3((unsigned)ceil_div(rows, 256)), dim3(256), 0, stream, p.row, rows,
                               (T)0);
            NRK_HIP(hipGetLastError());
        }
        return NRK_OK;
    }
    p.x = x;
    p.rows = rows;
    p.cols = cols;
    p.chain = chain;
    p.ddof = ddof;
    p.spr = ceil_div(cols, WSEG);
    p.nseg = rows * p.spr;
    if (p.nseg >= (size_t)UINT32_MAX) return fail(NRK_ERR_INVALID_ARGUMENT, "too many segments (%zu)", p.nseg);
    p.need_partials = 1;  // every result goes through the finish kernels
    p.ngroups = p.want_row ? rows : ceil_div(p.nseg, (size_t)GROUP);
    const size_t rec = 8 * 8 + 4;  // bytes of one partial record
    const size_t pbytes = (p.nseg + p.ngroups) * rec + 2 * 9 * 256 + 4096;
    Scratch s;
    NRK_TRY(get_scratch(stream, 0, pbytes, &s));
    char *b = (char *)s.partials;
    p.l1 = carve_parts(b, p.nseg);
    p.l2 = carve_parts(b, p.ngroups);
    const bool m2 = (p.row.var || p.full.var);
    const bool mx = (p.row.max || p.row.argmax || p.full.max || p.full.argmax);
    const bool mn = (p.row.min || p.row.argmin || p.full.min || p.full.argmin);
    const int flags = (m2 ? F_M2 : 0) | (mx ? F_MAX : 0) | (mn ? F_MIN : 0);
    switch (flags) {
        case 0: return launch_flags<T, VPT, 0>(p, stream);
        case 1: return launch_flags<T, VPT, 1>(p, stream);
        case 2: return launch_flags<T, VPT, 2>(p, stream);
        case 3: return launch_flags<T, VPT, 3>(p, stream);
        case 4: return launch_flags<T, VPT, 4>(p, stream);
        case 5: return launch_flags<T, VPT, 5>(p, stream);
        case 6: return launch_flags<T, VPT, 6>(p, stream);
        default: return launch_flags<T, VPT, 7>(p, stream);
    }
}

template nrk_status seg_reduce<float>(const Chain &, const float *, size_t, size_t, double, const nrk_reduce_outputs *,
                                      const nrk_reduce_outputs *, hipStream_t);
template nrk_status seg_reduce<double>(const Chain &, const double *, size_t, size_t, double, const nrk_reduce_outputs *,
                                       const nrk_reduce_outputs *, hipStream_t);

// ---------------- i64 sum: wrapping integer adds, exact and order-independent ----------------
constexpr int I64_VPT = 8;                 // 16-byte vectors per thread per segment
constexpr int I64_SEG = RB * I64_VPT * 2;  // 4096 elements = 32 KiB

// Each workgroup adds all of its segments (integer adds are exact, so the grouping cannot change
// the result) and writes one partial.
__global__ __launch_bounds__(RB) void sum_i64_kernel(const int64_t *x, size_t n, size_t nseg, int vec_ok,
                                                     unsigned long long *partials) {
    typedef long long V2 __attribute__((ext_vector_type(2)));
    __shared__ unsigned long long sred[WAVES];
    const int t = threadIdx.x;
    unsigned long long acc = 0;  // unsigned: wrapping is defined, bits equal two's complement i64
    for (size_t seg = blockIdx.x; seg < nseg; seg += gridDim.x) {
        const size_t off = seg * I64_SEG;
        if (vec_ok && off + I64_SEG <= n) {
            const V2 *bv = reinterpret_cast<const V2 *>(x + off);
#pragma unroll
            for (int j = 0; j < I64_VPT; ++j) {
                V2 v = __builtin_nontemporal_load(bv + j * RB + t);
                acc += (unsigned long long)v[0] + (unsigned long long)v[1];
            }
        } else {
            for (int j = 0; j < I64_VPT * 2; ++j) {
                size_t pos = off + (size_t)j * RB + t;
                if (pos < n) acc += (unsigned long long)x[pos];
            }
        }
    }
#pragma unroll
    for (int m = 32; m >= 1; m >>= 1) acc += shfl_xor((uint64_t)acc, m);
    if ((t & 63) == 0) sred[t >> 6] = acc;
    __syncthreads();
    if (t == 0) {
        unsigned long long s = 0;
        for (int k = 0; k < WAVES; ++k) s += sred[k];
        partials[blockIdx.x] = s;
    }
}

__global__ __launch_bounds__(FB) void sum_u64_finish_kernel(const unsigned long long *partials, size_t count, int64_t *out) {
    __shared__ unsigned long long sred[FB / 64];
    const int t = threadIdx.x;
    unsigned long long s = 0;
    for (size_t k = t; k < count; k += FB) s += partials[k];
#pragma unroll
    for (int m = 32; m >= 1; m >>= 1) s += shfl_xor((uint64_t)s, m);
    if ((t & 63) == 0) sred[t >> 6] = s;
    __syncthreads();
    if (t == 0) {
        unsigned long long tot = 0;
        for (int k = 0; k < FB / 64; ++k) tot += sred[k];
        *out = (int64_t)tot;
    }
}

__global__ void store_i64_kernel(int64_t *out, int64_t v) { *out = v; }

nrk_status sum_i64(const int64_t *x, size_t n, int64_t *out, hipStream_t stream) {
    NRK_TRY(ensure_device());
    if (n == 0) {
        hipLaunchKernelGGL(store_i64_kernel, dim3(1), dim3(1), 0, stream, out, (int64_t)0);
        NRK_HIP(hipGetLastError());
        return NRK_OK;
    }
    if (!x) return fail(NRK_ERR_INVALID_ARGUMENT, "x is NULL");
    const size_t nseg = ceil_div(n, I64_SEG);
    const int grid = (int)std::min<size_t>(nseg, (size_t)cu_count() * reduce_blocks_per_cu());
    Scratch s;
    NRK_TRY(get_scratch(stream, 0, (size_t)grid * 8, &s));
    const int vec_ok = (reinterpret_cast<uintptr_t>(x) % 16) == 0;
    hipLaunchKernelGGL(sum_i64_kernel, dim3(grid), dim3(RB), 0, stream, x, n, nseg, vec_ok, (unsigned long long *)s.partials);
    NRK_HIP(hipGetLastError());
    hipLaunchKernelGGL(sum_u64_finish_kernel, dim3(1), dim3(FB), 0, stream, (const unsigned long long *)s.partials, (size_t)grid,
                       out);
    NRK_HIP(hipGetLastError());
    return NRK_OK;
}

// ---------------- product: fixed-tree product in T ----------------
template <typename T>
__global__ __launch_bounds__(RB) void prod_kernel(const T *x, size_t n, size_t nseg, T *partials) {
    constexpr int PER = 16;
    __shared__ T sred[2][WAVES];
    const int t = threadIdx.x;
    int parity = 0;
    for (size_t seg = blockIdx.x; seg < nseg; seg += gridDim.x, parity ^= 1) {
        const size_t off = seg * (size_t)(RB * PER);
        T acc = (T)1;
        for (int j = 0; j < PER; ++j) {
            size_t pos = off + (size_t)j * RB + t;
            if (pos < n) acc = rn_mul(acc, x[pos]);
        }
#pragma unroll
        for (int m = 32; m >= 1; m >>= 1) acc = rn_mul(acc, shfl_xor(acc, m));
        if ((t & 63) == 0) sred[parity][t >> 6] = acc;
        __syncthreads();
        if (t == 0) {
            T s = sred[parity][0];
            for (int k = 1; k < WAVES; ++k) s = rn_mul(s, sred[parity][k]);
            partials[seg] = s;
        }
    }
}

// One workgroup multiplies the per-segment partials in a fixed tree.
template <typename T>
__global__ __launch_bounds__(FB) void prod_finish_kernel(const T *partials, size_t nseg, T *out) {
    __shared__ T sred[FB / 64];
    const int t = threadIdx.x;
    T s = (T)1;
    for (size_t k = t; k < nseg; k += FB) s = rn_mul(s, partials[k]);
#pragma unroll
    for (int m = 32; m >= 1; m >>= 1) s = rn_mul(s, shfl_xor(s, m));
    if ((t & 63) == 0) sred[t >> 6] = s;
    __syncthreads();
    if (t == 0) {
        T tot = sred[0];
        for (int k = 1; k < FB / 64; ++k) tot = rn_mul(tot, sred[k]);
        *out = tot;
    }
}

template <typename T>
__global__ void store_kernel(T *out, T v) { *out = v; }

template <typename T>
nrk_status prod(const T *x, size_t n, T *out, hipStream_t stream) {
    NRK_TRY(ensure_device());
    if (n == 0) {
        hipLaunchKernelGGL(store_kernel<T>, dim3(1), dim3(1), 0, stream, out, (T)1);
        NRK_HIP(hipGetLastError());
        return NRK_OK;
    }
    if (!x) return fail(NRK_ERR_INVALID_ARGUMENT, "x is NULL");
    const size_t nseg = ceil_div(n, (size_t)RB * 16);
    Scratch s;
    NRK_TRY(get_scratch(stream, 0, nseg * sizeof(T), &s));
    int grid = (int)std::min<size_t>(nseg, (size_t)cu_count() * 8);
    hipLaunchKernelGGL(prod_kernel<T>, dim3(grid), dim3(RB), 0, stream, x, n, nseg, (T *)s.partials);
    NRK_HIP(hipGetLastError());
    hipLaunchKernelGGL(prod_finish_kernel<T>, dim3(1), dim3(FB), 0, stream, (const T *)s.partials, nseg, out);
    NRK_HIP(hipGetLastError());
    return NRK_OK;
}
template nrk_status prod<float>(const float *, size_t, float *, hipStream_t);
template nrk_status prod<double>(const double *, size_t, double *, hipStream_t);

}  // namespace nrk
