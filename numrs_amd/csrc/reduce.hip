// Streaming reductions for gfx950: sum / mean / var / min / max / argmin / argmax over rows of
// a row-major matrix and over the whole matrix, optionally after an elementwise scalar chain
// (the fused map -> reduce path), in ONE pass over HBM.
//
// Reference semantics restated here (paths under the reference's src/):
//   kernels/reduce.rs:207-421   sum/mean/var(ddof)/min/max; empty -> 0; any NaN -> min/max NaN
//   math/statistics/mod.rs:79-183   argmax: first strict maximum, interior NaN skipped, NaN at
//                                   index 0 pins the result to 0 (argmin symmetric, :218-322)
//   math/aggregation.rs:244-261     axis max/min: NaN short-circuits to NaN
//   math/statistics/mod.rs:1403-1494 axis var = sum((x - mean)^2) / (len - ddof)
//
// Work decomposition. Row r of `cols` elements is cut into fixed segments of SEG elements
// (SEG = 256 threads * VPT 16-byte vectors). One workgroup-iteration reduces one segment in
// registers: the values stay in VGPRs between the sum pass and the second-moment pass, so
// variance costs no second read. Each segment yields a partial record; rows that fit in one
// segment are finalised in place, longer rows and the whole-matrix result are finished by the
// workgroup that retires the last segment (agent-scope acq_rel counter, self-resetting).
//
// Determinism. Every sum is a fixed tree: a pairwise tree over each thread's elements, an xor
// butterfly across the wave, a fixed-order combine of the 4 waves, and a fixed strided
// assignment of partials to threads in the finish. None of it depends on the grid size or on
// which workgroup finishes last, so results are bit-identical across launches and devices.
//
// Precision ("compensated" mode, the only mode): per-thread pairwise partials in T,
// double-word (TwoSum) combination across the workgroup, FP64 double-word cross-segment
// finish, and the second moment combined in FP64 with Chan's shift formula. max/min/argmax/
// argmin are exact (order-independent): the winner is the largest value, ties going to the
// lowest index, and the value written is the element itself, so even the sign of a zero
// maximum matches the reference's left-to-right fold.
#include <cstring>

#include "common.hpp"

namespace nrk {

constexpr int RB = 256;  // threads per reduction workgroup
constexpr int WAVES = RB / 64;
constexpr uint64_t NO_INDEX = ~0ull;

enum : int { F_M2 = 1, F_MAX = 2, F_MIN = 4 };

// Persistent-grid size of the segment reducer (workgroups per CU). The result bits do not depend
// on it (fixed segment tree); only speed does, so the tuner may choose it per shape.
static thread_local int g_blocks_per_cu = 8;
int reduce_blocks_per_cu() { return g_blocks_per_cu; }
void set_reduce_blocks_per_cu(int b) { g_blocks_per_cu = b < 1 ? 1 : (b > 16 ? 16 : b); }

template <typename T>
struct OutPtrs {
    T *sum, *mean, *var, *min, *max;
    uint64_t *argmin, *argmax;
};

template <typename T>
struct SegParams {
    const T *x;
    size_t rows, cols, spr, nseg;
    Chain chain;
    double ddof;
    int vec_ok;
    int want_row, want_full, need_partials;
    // partial records, structure of arrays, nseg entries each
    double *p_hi, *p_lo, *p_q, *p_center, *p_max, *p_min;
    uint64_t *p_imax, *p_imin;
    unsigned int *p_flags;  // bit0 saw NaN, bit1 first element of the row is NaN
    OutPtrs<T> row, full;
};

// ---------------- small device utilities ----------------
template <typename T>
__device__ __forceinline__ void two_sum(T a, T b, T &s, T &e) {
    s = rn_add(a, b);
    T v = rn_sub(s, a);
    e = rn_add(rn_sub(a, rn_sub(s, v)), rn_sub(b, v));
}

// double-word addition (a_hi + a_lo) + (b_hi + b_lo), symmetric in its arguments
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

// (value, index) winner for argmax: larger value wins, equal values go to the lower index.
// Values are never NaN here (NaN elements are excluded before this point).
template <typename T>
__device__ __forceinline__ void take_max(T &v, uint64_t &i, T bv, uint64_t bi) {
    if (bv > v || (bv == v && bi < i)) { v = bv; i = bi; }
}
template <typename T>
__device__ __forceinline__ void take_min(T &v, uint64_t &i, T bv, uint64_t bi) {
    if (bv < v || (bv == v && bi < i)) { v = bv; i = bi; }
}

// Per-wave results exchanged through LDS; double-buffered by segment parity so that the next
// segment's writes never race with a slower wave still reading this segment's values.
template <typename T>
struct WaveSlots {
    double hi[WAVES], lo[WAVES], m2[WAVES];
    int cnt[WAVES];
    T vmax[WAVES], vmin[WAVES];
    uint64_t imax[WAVES], imin[WAVES];
    int nan[WAVES];
};

// Running statistics of a set of elements: count, double-word sum, and M2 = sum of squared
// deviations from the set's own mean. merge() is Chan, Golub & LeVeque's pairwise update
// (1979): M2 = M2a + M2b + (na nb / n) (mean_a - mean_b)^2, evaluated in FP64.
struct Moments {
    int n;
    double hi, lo, m2;
    __device__ __forceinline__ void merge(int nb, double bhi, double blo, double bm2) {
        if (nb == 0) return;
        if (n == 0) {
            n = nb; hi = bhi; lo = blo; m2 = bm2;
            return;
        }
        const double na = (double)n, nbd = (double)nb, nt = na + nbd;
        const double d = (hi + lo) / na - (bhi + blo) / nbd;
        m2 = m2 + bm2 + d * d * (na * nbd / nt);
        dw_add(hi, lo, bhi, blo);
        n += nb;
    }
};
struct FinishSlots {
    double hi[WAVES], lo[WAVES], q[WAVES], vmax[WAVES], vmin[WAVES];
    uint64_t imax[WAVES], imin[WAVES];
    int nan[WAVES];
};
template <typename T>
struct BlockShared {
    WaveSlots<T> w[2];
    FinishSlots f;
    int last;
};

// ---- wave-level reductions (xor butterfly: every lane ends with the same value) ----
template <typename T>
__device__ __forceinline__ void wave_dw_sum(T &hi, T &lo) {
#pragma unroll
    for (int m = 32; m >= 1; m >>= 1) {
        T bh = shfl_xor(hi, m), bl = shfl_xor(lo, m);
        dw_add(hi, lo, bh, bl);
    }
}
__device__ __forceinline__ double wave_sum_f64(double v) {
#pragma unroll
    for (int m = 32; m >= 1; m >>= 1) v += shfl_xor(v, m);
    return v;
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

template <typename T>
__device__ __forceinline__ void write_outputs(const OutPtrs<T> &o, size_t at, double s_hi, double s_lo, double q_true,
                                              double count, double ddof, int nan, int first_nan, T vmax, uint64_t imax,
                                              T vmin, uint64_t imin) {
    if (o.sum) o.sum[at] = (T)(s_hi + s_lo);
    if (o.mean) o.mean[at] = (T)((s_hi + s_lo) / count);
    if (o.var) o.var[at] = (T)(q_true / (count - ddof));
    if (o.max) o.max[at] = nan ? dev_nan<T>() : vmax;
    if (o.min) o.min[at] = nan ? dev_nan<T>() : vmin;
    if (o.argmax) o.argmax[at] = first_nan ? 0 : imax;
    if (o.argmin) o.argmin[at] = first_nan ? 0 : imin;
}

// Combine the partial records of segments [lo, hi) (a whole row, or every segment) with the
// whole workgroup and write one set of outputs. Partial s goes to thread (s - lo) % RB, each
// thread folds its partials in increasing s, then a fixed butterfly + fixed wave order:
// deterministic whichever workgroup runs it.
template <typename T, int FLAGS, int SEG>
__device__ __noinline__ void finish_range(const SegParams<T> &p, size_t lo, size_t hi, bool full, const OutPtrs<T> &o,
                                          size_t at, FinishSlots *fs) {
    const int t = threadIdx.x, w = t >> 6, lane = t & 63;
    const size_t count_total = full ? p.rows * p.cols : p.cols;
    double shi = 0.0, slo = 0.0;
    int nan = 0;
    double vmax = -__builtin_inf(), vmin = __builtin_inf();
    uint64_t imax = NO_INDEX, imin = NO_INDEX;
    for (size_t s = lo + t; s < hi; s += RB) {
        dw_add(shi, slo, p.p_hi[s], p.p_lo[s]);
        nan |= p.p_flags[s] & 1;
        const size_t base = full ? (s / p.spr) * p.cols : 0;
        if (FLAGS & F_MAX) {
            const uint64_t ix = p.p_imax[s];
            take_max(vmax, imax, p.p_max[s], ix == NO_INDEX ? NO_INDEX : base + ix);
        }
        if (FLAGS & F_MIN) {
            const uint64_t ix = p.p_imin[s];
            take_min(vmin, imin, p.p_min[s], ix == NO_INDEX ? NO_INDEX : base + ix);
        }
    }
    wave_dw_sum(shi, slo);
    nan = __any(nan) ? 1 : 0;
    if (FLAGS & F_MAX) wave_arg<double, true>(vmax, imax);
    if (FLAGS & F_MIN) wave_arg<double, false>(vmin, imin);
    __syncthreads();
    if (lane == 0) {
        fs->hi[w] = shi; fs->lo[w] = slo; fs->nan[w] = nan;
        fs->vmax[w] = vmax; fs->imax[w] = imax; fs->vmin[w] = vmin; fs->imin[w] = imin;
    }
    __syncthreads();
    shi = fs->hi[0]; slo = fs->lo[0]; nan = fs->nan[0];
    vmax = fs->vmax[0]; imax = fs->imax[0]; vmin = fs->vmin[0]; imin = fs->imin[0];
#pragma unroll
    for (int k = 1; k < WAVES; ++k) {
        dw_add(shi, slo, fs->hi[k], fs->lo[k]);
        nan |= fs->nan[k];
        take_max(vmax, imax, fs->vmax[k], fs->imax[k]);
        take_min(vmin, imin, fs->vmin[k], fs->imin[k]);
    }
    const double mu = (shi + slo) / (double)count_total;
    double q = 0.0;
    if (FLAGS & F_M2) {
        for (size_t s = lo + t; s < hi; s += RB) {
            const size_t c = s % p.spr;
            const double n_s = (double)min((size_t)SEG, p.cols - c * SEG);
            const double mu_s = (p.p_hi[s] + p.p_lo[s]) / n_s;
            const double m_s = p.p_center[s];
            // Q_s is taken about m_s; shift it to the global mean mu (Chan et al.)
            q += p.p_q[s] + n_s * (m_s - mu) * (2.0 * mu_s - mu - m_s);
        }
        q = wave_sum_f64(q);
        if (lane == 0) fs->q[w] = q;
        __syncthreads();
        q = fs->q[0];
#pragma unroll
        for (int k = 1; k < WAVES; ++k) q += fs->q[k];
        if (q < 0.0) q = 0.0;
    }
    if (t == 0) {
        const int first_nan = (p.p_flags[lo] >> 1) & 1;
        write_outputs<T>(o, at, shi, slo, q, (double)count_total, p.ddof, nan, first_nan, (T)vmax, imax, (T)vmin, imin);
    }
    __syncthreads();
}

template <typename T, int VPT, int FLAGS, bool FULLSEG, bool VEC>
__device__ __forceinline__ void reduce_segment(const SegParams<T> &p, size_t seg, int parity, BlockShared<T> *sh) {
    typedef typename Vec16<T>::type V;
    constexpr int VN = Vec16<T>::N;
    constexpr int M = VPT * VN;  // elements per thread (<= 32: one validity bit each)
    constexpr int SEG = RB * M;
    static_assert(M <= 32, "validity mask holds 32 elements");
    const unsigned t = threadIdx.x, w = t >> 6, lane = t & 63;
    const size_t r = seg / p.spr, c = seg - r * p.spr;
    const size_t col0 = c * SEG;
    const unsigned len = FULLSEG ? (unsigned)SEG : (unsigned)min((size_t)SEG, p.cols - col0);
    const T *base = p.x + r * p.cols + col0;  // uniform across the workgroup

    // ---- load: element i = j*VN + e of this thread sits at segment position (j*RB + t)*VN + e
    T y[M];
    unsigned vmask = ~0u;
    if (VEC) {
        // aligned rows: 16-byte loads; a tail segment predicates whole vectors (len % VN == 0)
        const V *bv = reinterpret_cast<const V *>(base);
        const unsigned nvec = len / VN;
        if (!FULLSEG) vmask = 0u;
#pragma unroll
        for (int j = 0; j < VPT; ++j) {
            const unsigned jv = j * RB + t;
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
                const unsigned pos = (j * RB + t) * VN + e;
                const bool ok = FULLSEG || pos < len;
                y[j * VN + e] = ok ? base[pos] : (T)0;
                if (!FULLSEG) vmask |= (ok ? 1u : 0u) << (j * VN + e);
            }
        }
    }
    auto valid = [&](int i) -> bool { return FULLSEG || ((vmask >> i) & 1u); };
    auto pos_of = [&](int i) -> unsigned { return ((i / VN) * RB + t) * VN + (i % VN); };

    // ---- elementwise chain, one rounding per step (uniform branch per step)
    for (int k = 0; k < p.chain.len; ++k) {
        const int op = p.chain.op[k];
        const T s = (T)p.chain.s[k];
#pragma unroll
        for (int i = 0; i < M; ++i) y[i] = apply_sop<T>(op, s, y[i]);
    }

    // ---- per-thread statistics straight from registers; nothing here waits for other threads:
    // count, pairwise sum, M2 about the thread's own mean, extremum winners, NaN flag.
    const int nt = FULLSEG ? M : __builtin_popcount(vmask);
    const T st = pairwise<M, 0, T>([&](int i) { return valid(i) ? y[i] : (T)0; });
    int nan = 0;
    if (__builtin_isnan(st)) {  // rare: tell a NaN element apart from inf - inf
#pragma unroll
        for (int i = 0; i < M; ++i) nan |= (valid(i) && __builtin_isnan(y[i])) ? 1 : 0;
    }
    double m2t = 0.0;
    if ((FLAGS & F_M2) && nt > 0) {
        // mean of M elements: an exact power-of-two scaling for full segments
        const T mt = FULLSEG ? rn_mul(st, (T)(1.0 / M)) : rn_div(st, (T)nt);
        const T qt = pairwise<M, 0, T>([&](int i) {
            const T d = valid(i) ? rn_sub(y[i], mt) : (T)0;
            return rn_mul(d, d);
        });
        // qt is taken about the rounded mean mt; shift it to the mean st/nt: Q = M2 + n (mt - mean)^2
        const double dm = (double)mt - (double)st / (double)nt;
        m2t = fmax((double)qt - (double)nt * dm * dm, 0.0);
    }
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
            imax = col0 + pos_of(first);
        }
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
            imin = col0 + pos_of(first);
        }
    }
    const int first_nan = (c == 0 && t == 0 && __builtin_isnan(y[0])) ? 1 : 0;

    // ---- wavefront combine (fixed xor butterfly; lane 0's result is the one used)
    Moments mo{nt, (double)st, 0.0, m2t};
#pragma unroll
    for (int m = 32; m >= 1; m >>= 1) {
        const int bn = __shfl_xor(mo.n, m, 64);
        const double bhi = shfl_xor(mo.hi, m), blo = shfl_xor(mo.lo, m);
        const double bm2 = (FLAGS & F_M2) ? shfl_xor(mo.m2, m) : 0.0;
        if (FLAGS & F_M2) {
            mo.merge(bn, bhi, blo, bm2);
        } else {
            dw_add(mo.hi, mo.lo, bhi, blo);
            mo.n += bn;
        }
    }
    nan = __any(nan) ? 1 : 0;
    if (FLAGS & F_MAX) wave_arg<T, true>(vmax, imax);
    if (FLAGS & F_MIN) wave_arg<T, false>(vmin, imin);
    WaveSlots<T> &ws = sh->w[parity];
    if (lane == 0) {
        ws.hi[w] = mo.hi; ws.lo[w] = mo.lo; ws.m2[w] = mo.m2; ws.cnt[w] = mo.n; ws.nan[w] = nan;
        if (FLAGS & F_MAX) { ws.vmax[w] = vmax; ws.imax[w] = imax; }
        if (FLAGS & F_MIN) { ws.vmin[w] = vmin; ws.imin[w] = imin; }
    }
    __syncthreads();  // the only barrier of the segment
    if (t != 0) return;
    // ---- workgroup combine in wave order, by thread 0
    Moments tot{ws.cnt[0], ws.hi[0], ws.lo[0], ws.m2[0]};
    nan = ws.nan[0];
    if (FLAGS & F_MAX) { vmax = ws.vmax[0]; imax = ws.imax[0]; }
    if (FLAGS & F_MIN) { vmin = ws.vmin[0]; imin = ws.imin[0]; }
#pragma unroll
    for (int k = 1; k < WAVES; ++k) {
        if (FLAGS & F_M2) {
            tot.merge(ws.cnt[k], ws.hi[k], ws.lo[k], ws.m2[k]);
        } else {
            dw_add(tot.hi, tot.lo, ws.hi[k], ws.lo[k]);
            tot.n += ws.cnt[k];
        }
        nan |= ws.nan[k];
        if (FLAGS & F_MAX) take_max(vmax, imax, ws.vmax[k], ws.imax[k]);
        if (FLAGS & F_MIN) take_min(vmin, imin, ws.vmin[k], ws.imin[k]);
    }
    if (p.spr == 1 && p.want_row) {
        // the whole row is this segment: finalise in place
        write_outputs<T>(p.row, r, tot.hi, tot.lo, tot.m2, (double)len, p.ddof, nan, first_nan, vmax, imax, vmin, imin);
    }
    if (!p.need_partials) return;
    p.p_hi[seg] = tot.hi;
    p.p_lo[seg] = tot.lo;
    p.p_q[seg] = tot.m2;                              // about the segment's own mean ...
    p.p_center[seg] = (tot.hi + tot.lo) / (double)len;  // ... which is this
    p.p_max[seg] = (double)vmax;
    p.p_min[seg] = (double)vmin;
    p.p_imax[seg] = imax;
    p.p_imin[seg] = imin;
    p.p_flags[seg] = (unsigned)nan | ((unsigned)first_nan << 1);
}

template <typename T, int VPT, int FLAGS, bool VEC>
__global__ __launch_bounds__(RB) void seg_reduce_kernel(SegParams<T> p) {
    constexpr int SEG = RB * VPT * Vec16<T>::N;
    __shared__ BlockShared<T> sh;
    int parity = 0;
    for (size_t seg = blockIdx.x; seg < p.nseg; seg += gridDim.x, parity ^= 1) {
        const size_t c = seg % p.spr;
        if (c * SEG + SEG <= p.cols) {
            reduce_segment<T, VPT, FLAGS, true, VEC>(p, seg, parity, &sh);
        } else {
            reduce_segment<T, VPT, FLAGS, false, VEC>(p, seg, parity, &sh);
        }
    }
}

// Cross-segment finish, launched on the same stream right after the streaming kernel (the
// kernel boundary makes every partial visible; no fence or atomic is needed inside the stream).
// Workgroup b < row_blocks finishes row b (rows of more than one segment); the last workgroup
// finishes the whole-matrix result. On gfx950 a per-workgroup agent-scope release/acquire
// (L2 write-back + invalidate on each XCD) costs far more than this extra launch.
template <typename T, int VPT, int FLAGS>
__global__ __launch_bounds__(RB) void seg_finish_kernel(SegParams<T> p, int row_blocks) {
    constexpr int SEG = RB * VPT * Vec16<T>::N;
    __shared__ FinishSlots fs;
    const size_t b = blockIdx.x;
    if ((int)b < row_blocks) {
        finish_range<T, FLAGS, SEG>(p, b * p.spr, (b + 1) * p.spr, false, p.row, b, &fs);
    } else {
        finish_range<T, FLAGS, SEG>(p, 0, p.nseg, true, p.full, 0, &fs);
    }
}

// Writes T zero (or NaN) into up to 7 scalar outputs; used for empty inputs so that no
// hipMemsetAsync is ever recorded into a caller's graph.
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
    int grid = (int)std::min<size_t>(p.nseg, (size_t)cu_count() * reduce_blocks_per_cu());
    if (p.vec_ok)
        hipLaunchKernelGGL((seg_reduce_kernel<T, VPT, FLAGS, true>), dim3(grid), dim3(RB), 0, stream, p);
    else
        hipLaunchKernelGGL((seg_reduce_kernel<T, VPT, FLAGS, false>), dim3(grid), dim3(RB), 0, stream, p);
    NRK_HIP(hipGetLastError());
    const int row_blocks = (p.spr > 1 && p.want_row) ? (int)p.rows : 0;
    const int fin = row_blocks + (p.want_full ? 1 : 0);
    if (fin > 0) {
        hipLaunchKernelGGL((seg_finish_kernel<T, VPT, FLAGS>), dim3(fin), dim3(RB), 0, stream, p, row_blocks);
        NRK_HIP(hipGetLastError());
    }
    return NRK_OK;
}

// Host launcher for the fused map -> reduce over a (rows, cols) matrix.
template <typename T>
nrk_status seg_reduce(const Chain &chain, const T *x, size_t rows, size_t cols, double ddof,
                      const nrk_reduce_outputs *row_out, const nrk_reduce_outputs *full_out, hipStream_t stream) {
    constexpr int VPT = 8;
    constexpr int SEG = RB * VPT * Vec16<T>::N;
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
    p.spr = ceil_div(cols, SEG);
    p.nseg = rows * p.spr;
    if (p.nseg >= (size_t)UINT32_MAX) return fail(NRK_ERR_INVALID_ARGUMENT, "too many segments (%zu)", p.nseg);
    p.vec_ok = ((reinterpret_cast<uintptr_t>(x) % 16) == 0) && (cols % Vec16<T>::N == 0);
    p.need_partials = p.want_full || (p.spr > 1 && p.want_row);
    const size_t rec = 8 * 8 + 4;
    const size_t pbytes = p.need_partials ? align_up(p.nseg * rec + 64 * 9, 256) : 0;
    Scratch s;
    NRK_TRY(get_scratch(stream, 0, pbytes, &s));
    if (p.need_partials) {
        char *b = (char *)s.partials;
        size_t off = 0;
        auto take = [&](size_t bytes) {
            void *r = b + off;
            off = align_up(off + bytes, 256);
            return r;
        };
        p.p_hi = (double *)take(p.nseg * 8);
        p.p_lo = (double *)take(p.nseg * 8);
        p.p_q = (double *)take(p.nseg * 8);
        p.p_center = (double *)take(p.nseg * 8);
        p.p_max = (double *)take(p.nseg * 8);
        p.p_min = (double *)take(p.nseg * 8);
        p.p_imax = (uint64_t *)take(p.nseg * 8);
        p.p_imin = (uint64_t *)take(p.nseg * 8);
        p.p_flags = (unsigned *)take(p.nseg * 4);
    }
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
constexpr int I64_VPT = 8;                     // 16-byte vectors per thread per segment
constexpr int I64_SEG = RB * I64_VPT * 2;      // 4096 elements = 32 KiB

__global__ __launch_bounds__(RB) void sum_i64_kernel(const int64_t *x, size_t n, size_t nseg, int vec_ok,
                                                     unsigned long long *partials) {
    typedef long long V2 __attribute__((ext_vector_type(2)));
    __shared__ unsigned long long sred[2][WAVES];
    const int t = threadIdx.x;
    int parity = 0;
    for (size_t seg = blockIdx.x; seg < nseg; seg += gridDim.x) {
        const size_t off = seg * I64_SEG;
        unsigned long long acc = 0;  // unsigned: wrapping is defined, bits equal two's complement i64
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
#pragma unroll
        for (int m = 32; m >= 1; m >>= 1) acc += shfl_xor((uint64_t)acc, m);
        if ((t & 63) == 0) sred[parity][t >> 6] = acc;
        __syncthreads();
        if (t == 0) {
            unsigned long long s = 0;
            for (int k = 0; k < WAVES; ++k) s += sred[parity][k];
            partials[seg] = s;
        }
        parity ^= 1;  // the next segment writes the other slot set: no second barrier needed
    }
}

// One workgroup adds the per-segment partials (integer adds: exact in any order).
__global__ __launch_bounds__(RB) void sum_u64_finish_kernel(const unsigned long long *partials, size_t nseg,
                                                            int64_t *out) {
    __shared__ unsigned long long sred[WAVES];
    const int t = threadIdx.x;
    unsigned long long s = 0;
    for (size_t k = t; k < nseg; k += RB) s += partials[k];
#pragma unroll
    for (int m = 32; m >= 1; m >>= 1) s += shfl_xor((uint64_t)s, m);
    if ((t & 63) == 0) sred[t >> 6] = s;
    __syncthreads();
    if (t == 0) {
        unsigned long long tot = 0;
        for (int k = 0; k < WAVES; ++k) tot += sred[k];
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
    if (nseg >= (size_t)UINT32_MAX) return fail(NRK_ERR_INVALID_ARGUMENT, "input too large");
    Scratch s;
    NRK_TRY(get_scratch(stream, 0, nseg * 8, &s));
    const int vec_ok = (reinterpret_cast<uintptr_t>(x) % 16) == 0;
    int grid = (int)std::min<size_t>(nseg, (size_t)cu_count() * reduce_blocks_per_cu());
    hipLaunchKernelGGL(sum_i64_kernel, dim3(grid), dim3(RB), 0, stream, x, n, nseg, vec_ok, (unsigned long long *)s.partials);
    NRK_HIP(hipGetLastError());
    hipLaunchKernelGGL(sum_u64_finish_kernel, dim3(1), dim3(RB), 0, stream, (const unsigned long long *)s.partials, nseg, out);
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
__global__ __launch_bounds__(RB) void prod_finish_kernel(const T *partials, size_t nseg, T *out) {
    __shared__ T sred[WAVES];
    const int t = threadIdx.x;
    T s = (T)1;
    for (size_t k = t; k < nseg; k += RB) s = rn_mul(s, partials[k]);
#pragma unroll
    for (int m = 32; m >= 1; m >>= 1) s = rn_mul(s, shfl_xor(s, m));
    if ((t & 63) == 0) sred[t >> 6] = s;
    __syncthreads();
    if (t == 0) {
        T tot = sred[0];
        for (int k = 1; k < WAVES; ++k) tot = rn_mul(tot, sred[k]);
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
    hipLaunchKernelGGL(prod_finish_kernel<T>, dim3(1), dim3(RB), 0, stream, (const T *)s.partials, nseg, out);
    NRK_HIP(hipGetLastError());
    return NRK_OK;
}
template nrk_status prod<float>(const float *, size_t, float *, hipStream_t);
template nrk_status prod<double>(const double *, size_t, double *, hipStream_t);

}  // namespace nrk
