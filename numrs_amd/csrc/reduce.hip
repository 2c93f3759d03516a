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
// Work decomposition. Row r of `cols` elements is cut into segments of WSEG elements (64 lanes x
// VPT 16-byte vectors) and consecutive segments of a row are grouped into chunks of `cl`
// segments (a whole row when there are enough rows to fill the GPU). One wavefront owns a chunk:
// each lane keeps running accumulators over its elements of every segment of the chunk (count,
// double-word FP64 sum, second moment, max/min with the winner's position), so the per-element
// work is a handful of VALU operations and nothing crosses lanes until the chunk ends. Then one
// xor butterfly merges the 64 lanes, and lane `k` of the wavefront keeps the record of its k-th
// chunk in registers; the records (and the row results when a chunk is a row) are stored 64 at a
// time, so the streaming loop issues loads only. Wavefronts never wait for each other: no
// barrier, no LDS, no atomic, no fence. Chunk records are combined by a fixed tree of small
// dependent kernels on the same stream (up to 256 records per wavefront per level): rows split
// into several chunks first, then the whole-matrix result. The kernel boundaries provide the
// cross-workgroup visibility (an agent-scope release/acquire writes back / invalidates each XCD's
// L2 on gfx950, which cost 20x the stream itself when it was done per segment).
//
// Determinism. The chunking depends on the shape only, each chunk is reduced in a fixed order
// by one wavefront (per-lane sequences, then a butterfly whose lane-0 value is used) and the
// record tree is fixed, so results are bit-identical across launches, grid sizes and devices.
//
// Precision. Per lane and segment: pairwise sums in T; across segments and lanes: FP64
// double-word sums, and the second moment is combined with Chan, Golub & LeVeque's pairwise
// update (each segment's squared deviations are taken about its own lane mean). max/min/argmax/
// argmin are exact: a strict comparison scan per lane in index order (NaN never wins, ties keep
// the earlier element), then the larger value with ties to the lower index, and the value
// written is the element itself, so even the sign of a zero maximum matches the reference's
// left-to-right fold.
#include "reduce_common.hpp"

namespace nrk {

// Persistent-grid size of the streaming kernels (workgroups per CU). Result bits do not depend on
// it, only speed, so the tuner may choose it per shape.
static thread_local int g_blocks_per_cu = 8;
int reduce_blocks_per_cu() { return g_blocks_per_cu; }
void set_reduce_blocks_per_cu(int b) { g_blocks_per_cu = b < 1 ? 1 : (b > 16 ? 16 : b); }

// ---------------- the streaming kernel ----------------
// A segment is WSEG = 64 lanes x VPT 16-byte vectors of one row: lane l holds element
// i = j*VN + e at segment position (j*64 + l)*VN + e, so every load instruction of the
// wavefront reads 1 KiB of consecutive memory.

// Issue the loads of the segment at `base` into y; its positions [lo, len) belong to the row (lo > 0
// only in a phased row's first segment; len <= lo: none, and base is then a safe address). The
// others hold values of other elements; the fold reads them only through seg_mask's bits. VEC: base
// is 16-byte aligned, so whole vectors. Every lane issues every load,
// straight into y, so a segment is always exactly VPT (or VPT * VN) loads with no
// branch around them and no copy after them: the compiler can then wait for one segment's loads
// while the next segment's are in flight, where a data-dependent branch around a load (or a
// select that frees the load registers for reuse) makes it wait for all of them.
template <typename T, int VPT, bool VEC>
__device__ __forceinline__ void seg_load(const T *base, int64_t len, unsigned lo, T (&y)[VPT * Vec16<T>::N]) {
    typedef typename Vec16<T>::type V;
    constexpr int VN = Vec16<T>::N;
    const unsigned lane = threadIdx.x & 63;
#pragma unroll
    for (int j = 0; j < VPT; ++j) {
        const unsigned jv = j * 64 + lane;
        const int64_t p0 = (int64_t)jv * VN;
        if (VEC) {
            // a vector starting inside the row lies in the row's 16-byte blocks
            const V v = __builtin_nontemporal_load(reinterpret_cast<const V *>(base) + (p0 < len ? jv : 0u));
#pragma unroll
            for (int e = 0; e < VN; ++e) y[j * VN + e] = v[e];
        } else {
#pragma unroll
            for (int e = 0; e < VN; ++e) y[j * VN + e] = base[(p0 + e >= (int64_t)lo && p0 + e < len) ? p0 + e : (int64_t)lo];
        }
    }
}

// Which of a segment's positions belong to the row: [lo, len), one bit per element slot.
template <typename T, int VPT>
__device__ __forceinline__ unsigned seg_mask(int64_t len, unsigned lo) {
    constexpr int VN = Vec16<T>::N;
    static_assert(VPT * VN <= 32, "validity mask holds 32 elements");
    const unsigned lane = threadIdx.x & 63;
    unsigned vmask = 0u;
#pragma unroll
    for (int j = 0; j < VPT; ++j) {
        const int64_t p0 = (int64_t)(j * 64 + lane) * VN;
        const int64_t dh = len - p0, dl = (int64_t)lo - p0;
        const unsigned nhi = dh <= 0 ? 0u : (dh >= VN ? (unsigned)VN : (unsigned)dh);
        const unsigned nlo = dl <= 0 ? 0u : (dl >= VN ? (unsigned)VN : (unsigned)dl);
        vmask |= (((1u << nhi) - 1u) & ~((1u << nlo) - 1u)) << (j * VN);
    }
    return vmask;
}

// Fold one loaded segment (segment `sl` of its chunk) into the lane accumulators. FULLSEG: every
// element exists (vmask is all ones).
template <typename T, int VPT, int FLAGS, bool FULLSEG>
__device__ __forceinline__ void seg_body(const Chain &chain, T (&y)[VPT * Vec16<T>::N], unsigned vmask, unsigned sl,
                                         LaneAcc<T> &a, unsigned &fnan, bool row_start, unsigned ph) {
    constexpr int M = VPT * Vec16<T>::N;
    constexpr int VN = Vec16<T>::N;
    auto valid = [&](int i) -> bool { return FULLSEG || ((vmask >> i) & 1u); };

    // ---- elementwise chain, one rounding per step
    apply_chain_n(chain, y);
    if (row_start) {  // lane 0's y[ph] is the row's element 0
        T y0 = y[0];
#pragma unroll
        for (int i = 1; i < VN; ++i) y0 = ph == (unsigned)i ? y[i] : y0;
        fnan = __builtin_isnan(y0) ? 1u : 0u;
    }

    // ---- sum: pairwise in T over the lane's elements, then into the lane's FP64 double word
    const T st = pairwise<M, 0, T>([&](int i) { return valid(i) ? y[i] : (T)0; });
    if (__builtin_isnan(st)) {  // rare: tell a NaN element apart from inf - inf
#pragma unroll
        for (int i = 0; i < M; ++i) a.nan |= (valid(i) && __builtin_isnan(y[i])) ? 1u : 0u;
    }
    const double cnt = FULLSEG ? (double)M : (double)__builtin_popcount(vmask);
    const double sb = (double)st;
    // ---- second moment of this segment's lane elements about their own mean, merged (Chan)
    if ((FLAGS & F_M2) && (FULLSEG || cnt > 0.0)) {
        const double mb = FULLSEG ? sb * (1.0 / M) : sb / cnt;
        const T mt = (T)mb;
        // sum of squared deviations: four interleaved FMA chains, then a pairwise sum of them
        T q4[4] = {(T)0, (T)0, (T)0, (T)0};
#pragma unroll
        for (int i = 0; i < M; ++i) {
            const T d = valid(i) ? rn_sub(y[i], mt) : (T)0;
            q4[i & 3] = fma(d, d, q4[i & 3]);
        }
        const T qt = rn_add(rn_add(q4[0], q4[1]), rn_add(q4[2], q4[3]));
        // qt is taken about mt, the mean rounded to T: Q = M2 + cnt (mt - mean)^2
        const double dm = (double)mt - mb;
        double m2b = (double)qt - cnt * dm * dm;
        m2b = m2b < 0.0 ? 0.0 : m2b;  // rounding can leave a tiny negative; NaN passes through
        if (a.n == 0.0) {
            a.m2 = m2b;
        } else {
            const double sa = a.hi + a.lo;
            const double d = cnt * sa - a.n * sb;  // = na nb (mean_a - mean_b)
            a.m2 = a.m2 + m2b + d * d / (a.n * cnt * (a.n + cnt));
        }
    }
    {
        double s, e;
        two_sum(a.hi, sb, s, e);
        e = e + a.lo;
        dw_fix(s, e, a.hi, a.lo);
    }
    a.n += cnt;
    // ---- extremum winners: strict comparison in index order (NaN never compares greater). The
    // scan records only the winning slot i of this segment (a literal operand, no register copy
    // of sl * M + i per element); the code is formed once per segment.
    if (FLAGS & F_MAX) {
        unsigned ci = NO_CODE;
#pragma unroll
        for (int i = 0; i < M; ++i) {
            const bool gt = valid(i) && y[i] > a.vmax;
            a.vmax = gt ? y[i] : a.vmax;
            ci = gt ? (unsigned)i : ci;
        }
        a.cmax = ci != NO_CODE ? sl * M + ci : a.cmax;
    }
    if (FLAGS & F_MIN) {
        unsigned ci = NO_CODE;
#pragma unroll
        for (int i = 0; i < M; ++i) {
            const bool lt = valid(i) && y[i] < a.vmin;
            a.vmin = lt ? y[i] : a.vmin;
            ci = lt ? (unsigned)i : ci;
        }
        a.cmin = ci != NO_CODE ? sl * M + ci : a.cmin;
    }
}

// Row r's phase: its segments start ph = (r cols) mod VN elements before it (0 unless phased).
template <typename T>
__device__ __forceinline__ unsigned row_phase(const SegParams<T> &p, size_t r) {
    return p.phased ? (unsigned)((r * p.cols) & (size_t)(Vec16<T>::N - 1)) : 0u;
}

// Position of a wavefront in its work: chunk c = row r, chunk-of-row q, segments [s0, s1).
struct ChunkPos {
    size_t c, r, q, s0, s1;
};
template <typename T>
__device__ __forceinline__ ChunkPos chunk_pos(const SegParams<T> &p, size_t c) {
    ChunkPos k;
    k.c = c;
    k.r = c / p.cpr;
    k.q = c - k.r * p.cpr;
    k.s0 = k.q * p.cl;
    k.s1 = min(p.spr, k.s0 + p.cl);
    return k;
}


// Per-wavefront batch of up to 64 chunk records, kept in LDS (slot k: the k-th chunk of the
// batch) until they are stored together. Only the owning wavefront touches its slots.
struct LdsRecs {
    double n[64], hi[64], lo[64], m2[64], vmax[64], vmin[64];
    uint64_t imax[64], imin[64];
    unsigned flags[64];
};
struct Batch {
    LdsRecs *lds;
    unsigned nb;   // chunks in the batch
    size_t cbase;  // first chunk of the batch
};

// Lanes 0..nb-1 store the records of chunks cbase, cbase + W, ... (and the row results when chunks
// are rows).
template <typename T, int FLAGS>
__device__ __forceinline__ void flush_batch(const SegParams<T> &p, Batch &bt, size_t W) {
    const unsigned lane = threadIdx.x & 63;
    __builtin_amdgcn_wave_barrier();  // lane 0's slot writes are ordered before the other lanes' reads
    if (lane < bt.nb) {
        const LdsRecs &L = *bt.lds;
        Part r;
        r.n = L.n[lane]; r.hi = L.hi[lane]; r.lo = L.lo[lane];
        r.m2 = (FLAGS & F_M2) ? L.m2[lane] : 0.0;
        r.vmax = (FLAGS & F_MAX) ? L.vmax[lane] : -__builtin_inf();
        r.imax = (FLAGS & F_MAX) ? L.imax[lane] : NO_INDEX;
        r.vmin = (FLAGS & F_MIN) ? L.vmin[lane] : __builtin_inf();
        r.imin = (FLAGS & F_MIN) ? L.imin[lane] : NO_INDEX;
        r.nan = L.flags[lane] & 1u;
        const unsigned first = L.flags[lane] >> 1;
        const size_t c = bt.cbase + lane * W;
        if (p.row_direct) write_outputs<T>(p.row, c, r, p.ddof, first, (uint64_t)c * p.cols);
        if (p.need_l1) part_store<FLAGS>(p.l1, c, r, first);
    }
    __builtin_amdgcn_wave_barrier();  // ... and the reads before the next batch's writes
}

// The last segment of chunk `k` has been folded: merge the lanes and put the record in slot nb.
template <typename T, int VPT, int FLAGS>
__device__ __forceinline__ void chunk_end(const SegParams<T> &p, const ChunkPos &k, LaneAcc<T> &a, unsigned &fnan, Batch &bt,
                                          size_t W) {
    constexpr int VN = Vec16<T>::N;
    constexpr int M = VPT * VN;
    constexpr int WSEG = 64 * M;
    const unsigned lane = threadIdx.x & 63;
    const uint64_t fbase = (uint64_t)k.r * p.cols - row_phase<T>(p, k.r) + (uint64_t)k.s0 * WSEG;
    auto flat = [&](unsigned code) -> uint64_t {
        const unsigned sl = code / M, i = code % M;
        return fbase + (uint64_t)sl * WSEG + (uint64_t)(((i / VN) * 64 + lane) * VN + (i % VN));
    };
    Part pa;
    pa.n = a.n; pa.hi = a.hi; pa.lo = a.lo; pa.m2 = a.m2; pa.nan = a.nan;
    pa.vmax = a.cmax == NO_CODE ? -__builtin_inf() : (double)a.vmax;
    pa.imax = a.cmax == NO_CODE ? NO_INDEX : flat(a.cmax);
    pa.vmin = a.cmin == NO_CODE ? __builtin_inf() : (double)a.vmin;
    pa.imin = a.cmin == NO_CODE ? NO_INDEX : flat(a.cmin);
    wave_merge<FLAGS>(pa);
    if (lane == 0) {  // lane 0's value is the chunk's record; fnan of lane 0 is the row's element 0
        LdsRecs &L = *bt.lds;
        const unsigned s = bt.nb;
        L.n[s] = pa.n; L.hi[s] = pa.hi; L.lo[s] = pa.lo;
        if (FLAGS & F_M2) L.m2[s] = pa.m2;
        if (FLAGS & F_MAX) { L.vmax[s] = pa.vmax; L.imax[s] = pa.imax; }
        if (FLAGS & F_MIN) { L.vmin[s] = pa.vmin; L.imin[s] = pa.imin; }
        L.flags[s] = pa.nan | ((k.q == 0 ? fnan : 0u) << 1);
    }
    if (++bt.nb == 64) {
        flush_batch<T, FLAGS>(p, bt, W);
        bt.nb = 0;
        bt.cbase = k.c + W;
    }
    acc_reset(a);
    fnan = 0;
}

// Fold segment s of chunk k (already loaded into y) and close the chunk after its last segment.
template <typename T, int VPT, int FLAGS>
__device__ __forceinline__ void seg_step(const SegParams<T> &p, const ChunkPos &k, size_t s, T (&y)[VPT * Vec16<T>::N],
                                         LaneAcc<T> &a, unsigned &fnan, Batch &bt, size_t W) {
    constexpr int WSEG = 64 * VPT * Vec16<T>::N;
    const unsigned sl = (unsigned)(s - k.s0);
    const unsigned ph = row_phase<T>(p, k.r);
    if ((s > 0 || ph == 0) && s * WSEG + WSEG <= p.cols + ph) {
        seg_body<T, VPT, FLAGS, true>(p.chain, y, ~0u, sl, a, fnan, s == 0, ph);
    } else {
        const unsigned vmask = seg_mask<T, VPT>((int64_t)(p.cols + ph) - (int64_t)(s * WSEG), s == 0 ? ph : 0u);
        seg_body<T, VPT, FLAGS, false>(p.chain, y, vmask, sl, a, fnan, s == 0, ph);
    }
    if (s + 1 == k.s1) chunk_end<T, VPT, FLAGS>(p, k, a, fnan, bt, W);
}

// Wavefront w of the grid (W wavefronts) reduces chunks w, w + W, w + 2W, ..., segment by
// segment, with two register buffers: the loads of the next segment (possibly of its next
// chunk) are in flight while the current one is folded.
template <typename T, int VPT, int FLAGS, bool VEC>
__global__ __launch_bounds__(RB) void stream_kernel(SegParams<T> p) {
    constexpr int M = VPT * Vec16<T>::N;
    constexpr int WSEG = 64 * M;
    __shared__ LdsRecs recs[WAVES];
    const size_t W = (size_t)gridDim.x * WAVES;
    const unsigned wave = lane0((unsigned)(threadIdx.x >> 6));
    const size_t c0 = blockIdx.x * (size_t)WAVES + wave;
    if (c0 >= p.nchunk) return;  // whole wavefronts exit together
    Batch bt;
    bt.lds = &recs[wave];
    bt.nb = 0;
    bt.cbase = c0;
    LaneAcc<T> a;
    acc_reset(a);
    unsigned fnan = 0;
    // segment s of row r: positions from element r cols - ph + s WSEG; the row's own are [lo, len).
    // A segment past a phased row's end (len <= 0) reads from the row's first vector instead.
    auto len_of = [&](const ChunkPos &k, size_t s) {
        return (int64_t)(p.cols + row_phase<T>(p, k.r)) - (int64_t)(s * WSEG);
    };
    auto base_of = [&](const ChunkPos &k, size_t s) {
        const T *b = p.x + k.r * p.cols - row_phase<T>(p, k.r);
        return len_of(k, s) > 0 ? b + s * WSEG : b;
    };
    auto lo_of = [&](const ChunkPos &k, size_t s) { return s == 0 ? row_phase<T>(p, k.r) : 0u; };
    // advance (k, s) to the wavefront's next segment
    auto next = [&](ChunkPos &k, size_t &s) {
        if (s + 1 < k.s1) {
            ++s;
        } else {
            k = chunk_pos(p, k.c + W);
            s = k.s0;
        }
    };
    ChunkPos ka = chunk_pos(p, c0);
    size_t sa = ka.s0;
    T ya[M], yb[M];
    seg_load<T, VPT, VEC>(base_of(ka, sa), len_of(ka, sa), lo_of(ka, sa), ya);
    // The next segment's loads are issued unconditionally (past the wavefront's last segment they
    // reread the current one and are never used): no branch around them.
    for (;;) {
        ChunkPos kb = ka;
        size_t sb = sa;
        next(kb, sb);
        const bool hb = kb.c < p.nchunk;
        seg_load<T, VPT, VEC>(hb ? base_of(kb, sb) : base_of(ka, sa), hb ? len_of(kb, sb) : len_of(ka, sa),
                              hb ? lo_of(kb, sb) : lo_of(ka, sa), yb);
        seg_step<T, VPT, FLAGS>(p, ka, sa, ya, a, fnan, bt, W);
        if (!hb) break;
        ka = kb;
        sa = sb;
        next(ka, sa);
        const bool ha = ka.c < p.nchunk;
        seg_load<T, VPT, VEC>(ha ? base_of(ka, sa) : base_of(kb, sb), ha ? len_of(ka, sa) : len_of(kb, sb),
                              ha ? lo_of(ka, sa) : lo_of(kb, sb), ya);
        seg_step<T, VPT, FLAGS>(p, kb, sb, yb, a, fnan, bt, W);
        if (!ha) break;
    }
    flush_batch<T, FLAGS>(p, bt, W);
}

// ---------------- the record tree ----------------
// Wavefront g merges input records lo .. hi-1 of its set: lane l folds lo+l, lo+l+64, ... (all
// loads issued first) in order, then the xor butterfly; lane 0 writes.
template <typename T, int FLAGS>
__global__ __launch_bounds__(256) void tree_kernel(TreeParams<T> t) {
    const size_t g = (blockIdx.x * (size_t)blockDim.x + threadIdx.x) >> 6;
    const unsigned lane = threadIdx.x & 63;
    if (g < t.nsets * t.per_out) {  // whole wavefronts take the same branch
        const size_t set = g / t.per_out, b = g - set * t.per_out;
        const size_t first = set * t.per;
        const size_t lo = first + b * TREE_SPAN, hi = min(first + t.per, lo + TREE_SPAN);
        Part v[TREE_LOADS];
#pragma unroll
        for (int k = 0; k < TREE_LOADS; ++k) {
            const size_t s = lo + lane + 64 * k;
            v[k] = s < hi ? part_load<FLAGS>(t.in, s) : part_empty();
        }
        Part a = v[0];
#pragma unroll
        for (int k = 1; k < TREE_LOADS; ++k) part_merge<FLAGS>(a, v[k]);
        wave_merge<FLAGS>(a);
        if (lane == 0) {
            const unsigned fn = (t.in.flags[lo] >> 1) & 1u;
            if (t.results) write_outputs<T>(t.o, set, a, t.ddof, fn, t.rowmode ? (uint64_t)set * t.cols : 0);
            if (t.keep) {
                if (t.fuse_last)
                    part_store_agent<FLAGS>(t.out, g, a, fn);
                else
                    part_store<FLAGS>(t.out, g, a, fn);
            }
        }
    }
    if (!t.fuse_last) return;  // uniform over the grid
    // The last workgroup to finish merges the per_out records of this level (one set) in a fixed
    // order -- lane l folds records l, l+64, ..., then the butterfly -- whichever workgroup it is.
    // Records go out at agent scope and are acknowledged before the arrival is counted (relaxed
    // agent-scope atomic); the last workgroup reads them at agent scope: no fence.
    __shared__ unsigned is_last;
    NRK_WAIT_VMEM0();
    lds_barrier();
    if (threadIdx.x == 0)
        is_last = __hip_atomic_fetch_add(t.counter, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) == gridDim.x - 1;
    lds_barrier();
    if (!is_last) return;  // uniform over the workgroup
    if (threadIdx.x < 64) {
        Part v[TREE_LOADS];
#pragma unroll
        for (int k = 0; k < TREE_LOADS; ++k) {
            const size_t s = lane + 64 * k;
            v[k] = s < t.per_out ? part_load_agent<FLAGS>(t.out, s) : part_empty();
        }
        Part a = v[0];
#pragma unroll
        for (int k = 1; k < TREE_LOADS; ++k) part_merge<FLAGS>(a, v[k]);
        wave_merge<FLAGS>(a);
        if (lane == 0) {
            const unsigned fn = (ld_agent(t.out.flags) >> 1) & 1u;
            write_outputs<T>(t.last_o, 0, a, t.ddof, fn, 0);
            st_agent(t.counter, 0u);  // the zero-at-launch invariant for the next kernel on this stream
        }
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


// Workgroups of `kernel` that fit on one CU at once (register / LDS limited), cached per kernel.
template <typename K>
static int resident_blocks(K kernel, int *cache) {
    if (*cache == 0) {
        int b = 0;
        if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&b, kernel, RB, 0) != hipSuccess || b < 1) b = 1;
        *cache = b;
    }
    return *cache;
}

// The grid is persistent: at most the workgroups that are resident together (and at most the
// tuner's workgroups per CU), each wavefront then looping over its chunks.
template <typename T, int VPT, int FLAGS>
static nrk_status launch_stream(const SegParams<T> &p, hipStream_t stream) {
    const bool vec = (reinterpret_cast<uintptr_t>(p.x) % 16) == 0;  // phased rows start on the grid too
    static int occ[2] = {0, 0};
    const int res = vec ? resident_blocks(stream_kernel<T, VPT, FLAGS, true>, &occ[1])
                        : resident_blocks(stream_kernel<T, VPT, FLAGS, false>, &occ[0]);
    const size_t per_cu = (size_t)std::min(res, reduce_blocks_per_cu());
    const int grid = (int)std::min<size_t>(ceil_div(p.nchunk, (size_t)WAVES), (size_t)cu_count() * per_cu);
    if (vec)
        hipLaunchKernelGGL((stream_kernel<T, VPT, FLAGS, true>), dim3(grid), dim3(RB), 0, stream, p);
    else
        hipLaunchKernelGGL((stream_kernel<T, VPT, FLAGS, false>), dim3(grid), dim3(RB), 0, stream, p);
    NRK_HIP(hipGetLastError());
    return NRK_OK;
}

// Reduce `nsets` sets of `per` records (`in`) to one result per set, level by level. With `keep`
// the last level also stores the per-set records into `final_out` (index = set).
template <typename T, int FLAGS>
static nrk_status run_tree(Parts in, size_t nsets, size_t per, Parts tmp0, Parts tmp1, bool rowmode, size_t cols, double ddof,
                           const OutPtrs<T> &o, bool keep, Parts final_out, unsigned int *counter, hipStream_t stream) {
    int flip = 0;
    for (;;) {
        TreeParams<T> t{};
        t.in = in;
        t.nsets = nsets;
        t.per = per;
        t.per_out = ceil_div(per, (size_t)TREE_SPAN);
        t.results = t.per_out == 1;
        t.keep = !t.results || keep;
        t.out = t.results ? final_out : (flip ? tmp1 : tmp0);
        t.rowmode = rowmode;
        t.cols = cols;
        t.ddof = ddof;
        t.o = t.results ? o : OutPtrs<T>{};
        // one set whose next level would be the last: let this launch finish it (saves a launch)
        t.fuse_last = counter && nsets == 1 && !keep && !t.results && t.per_out <= (size_t)TREE_SPAN;
        t.last_o = t.fuse_last ? o : OutPtrs<T>{};
        t.counter = counter;
        const size_t waves = nsets * t.per_out;
        hipLaunchKernelGGL((tree_kernel<T, FLAGS>), dim3((unsigned)ceil_div(waves, (size_t)4)), dim3(256), 0, stream, t);
        NRK_HIP(hipGetLastError());
        if (t.results || t.fuse_last) return NRK_OK;
        in = t.out;
        per = t.per_out;
        flip ^= 1;
    }
}

template <typename T, int VPT, int FLAGS>
static nrk_status launch_all(SegParams<T> &p, bool want_row, bool want_full, const OutPtrs<T> &row, const OutPtrs<T> &full,
                             const Parts &rowrec, const Parts &tmp0, const Parts &tmp1, unsigned int *counter,
                             hipStream_t stream) {
    NRK_TRY((launch_stream<T, VPT, FLAGS>(p, stream)));
    const bool split_rows = want_row && p.cpr > 1;
    if (split_rows)
        NRK_TRY((run_tree<T, FLAGS>(p.l1, p.rows, p.cpr, tmp0, tmp1, true, p.cols, p.ddof, row, want_full, rowrec, nullptr,
                                    stream)));
    if (want_full) {
        const Parts in = split_rows ? rowrec : p.l1;
        const size_t per = split_rows ? p.rows : p.nchunk;
        NRK_TRY((run_tree<T, FLAGS>(in, 1, per, tmp0, tmp1, false, p.cols, p.ddof, full, false, Parts{}, counter, stream)));
    }
    return NRK_OK;
}

__global__ void clear_counter_kernel(unsigned int *c) { *c = 0u; }

// x1 != NULL (with prog2): the two-operand lean form; *no_lean is set (and nothing launched)
// when the lean path cannot take it.
template <typename T>
static nrk_status seg_reduce_impl(const Chain &chain, const T *x, size_t rows, size_t cols, double ddof,
                                  const nrk_reduce_outputs *row_out, const nrk_reduce_outputs *full_out, hipStream_t stream,
                                  const ExtScratch *ext, const LeanProg *prog2, const T *x1, bool *no_lean) {
    constexpr int M = NRK_REDUCE_M;
    constexpr int VPT = M / Vec16<T>::N;
    static_assert(VPT * Vec16<T>::N == M, "elements per lane must be whole 16-byte vectors");
    constexpr int WSEG = 64 * M;  // elements per wavefront segment
    NRK_TRY(ensure_device());
    const OutPtrs<T> row = to_ptrs<T>(row_out);
    const OutPtrs<T> full = to_ptrs<T>(full_out);
    const bool want_row = any_out(row);
    const bool want_full = any_out(full);
    if (!want_row && !want_full) return NRK_OK;
    if (want_row && (row.argmax || row.argmin || row.max || row.min) && cols == 0)
        return fail(NRK_ERR_INVALID_ARGUMENT, "max/min/argmax/argmin along an empty axis");
    if (want_full && (full.argmax || full.argmin) && rows * cols == 0)
        return fail(NRK_ERR_INVALID_ARGUMENT, "argmax/argmin of an empty array");
    if ((!x || (prog2 && !x1)) && rows * cols > 0) return fail(NRK_ERR_INVALID_ARGUMENT, "x is NULL");
    if (rows * cols == 0) {
        // reference: empty -> 0 for sum/mean/min/max/var (kernels/reduce.rs:207-421)
        if (want_full) {
            hipLaunchKernelGGL(fill_outputs_kernel<T>, dim3(1), dim3(64), 0, stream, full, (size_t)1, (T)0);
            NRK_HIP(hipGetLastError());
        }
        if (want_row && rows > 0) {
            hipLaunchKernelGGL(fill_outputs_kernel<T>, dim3((unsigned)ceil_div(rows, 256)), dim3(256), 0, stream, row, rows, (T)0);
            NRK_HIP(hipGetLastError());
        }
        return NRK_OK;
    }
    {
        LeanProg prog{};
        prog.pre0 = chain;
        bool taken = false;
        NRK_TRY(try_lean<T>(prog2 ? *prog2 : prog, x, prog2 ? x1 : nullptr, rows, cols, ddof, row, full, want_row, want_full,
                            stream, ext, &taken));
        if (taken) return NRK_OK;
        if (prog2) {
            *no_lean = true;
            return NRK_OK;
        }
    }
    if (!want_row) {  // a whole-array reduction: the rows are contiguous, so take them as one
        cols *= rows;
        rows = 1;
    }
    SegParams<T> p{};
    p.x = x;
    p.rows = rows;
    p.cols = cols;
    p.chain = chain;
    p.ddof = ddof;
    // rows whose length is not a multiple of the vector are phased: a row's positions run over
    // cols + ph < cols + VN, so up to one more segment per row
    p.phased = rows > 1 && cols % Vec16<T>::N != 0;
    p.spr = ceil_div(cols + (p.phased ? Vec16<T>::N - 1 : 0), (size_t)WSEG);
    // chunks: whole rows when there are enough of them (CHUNK_TARGET / 4 rows; the row results then
    // come straight from the stream, no record tree; measured against split rows from 1024 to 4096
    // rows), else rows split into ~CHUNK_TARGET chunks in total; depends on the shape only.
    size_t cpr = rows >= CHUNK_TARGET / 4 ? 1 : std::min(p.spr, std::max<size_t>(1, ceil_div(CHUNK_TARGET, rows)));
    p.cl = ceil_div(p.spr, cpr);
    p.cpr = ceil_div(p.spr, p.cl);
    p.nchunk = rows * p.cpr;
    if (p.cl >= ((size_t)1 << 26)) return fail(NRK_ERR_INVALID_ARGUMENT, "row of %zu elements is too long", cols);
    p.row_direct = want_row && p.cpr == 1;
    p.need_l1 = want_full || (want_row && p.cpr > 1);
    p.row = p.row_direct ? row : OutPtrs<T>{};
    // scratch: chunk records, row records, two ping-pong buffers for the tree levels
    const size_t tmp_count = std::max(rows * ceil_div(p.cpr, (size_t)TREE_SPAN), ceil_div(std::max(rows, p.nchunk), (size_t)TREE_SPAN)) + 1;
    const size_t pbytes = parts_bytes(p.nchunk) + parts_bytes(rows) + 2 * parts_bytes(tmp_count) + 256;
    Scratch s;
    if (ext) {  // the caller's scratch: counter in the first 256 bytes, cleared here; records after
        if (pbytes + 256 > ext->bytes)
            return fail(NRK_ERR_INVALID_ARGUMENT, "(%zu, %zu) needs %zu bytes of scratch; the fixed scratch holds %zu", rows, cols,
                        pbytes + 256, ext->bytes);
        s.counters = (unsigned int *)ext->base;
        s.n_counters = 1;
        s.partials = (char *)ext->base + 256;
        s.partial_bytes = ext->bytes - 256;
        hipLaunchKernelGGL(clear_counter_kernel, dim3(1), dim3(1), 0, stream, s.counters);
        NRK_HIP(hipGetLastError());
    } else {
        NRK_TRY(get_scratch(stream, 1, pbytes, &s));
    }
    char *b = (char *)s.partials;
    p.l1 = carve_parts(b, p.nchunk);
    const Parts rowrec = carve_parts(b, rows);
    const Parts tmp0 = carve_parts(b, tmp_count);
    const Parts tmp1 = carve_parts(b, tmp_count);
    const bool m2 = (row.var || full.var);
    const bool mx = (row.max || row.argmax || full.max || full.argmax);
    const bool mn = (row.min || row.argmin || full.min || full.argmin);
    const int flags = (m2 ? F_M2 : 0) | (mx ? F_MAX : 0) | (mn ? F_MIN : 0);
#define NRK_LAUNCH_FLAGS(F) \
    return launch_all<T, VPT, F>(p, want_row, want_full, row, full, rowrec, tmp0, tmp1, s.counters, stream)
    switch (flags) {
        case 0: NRK_LAUNCH_FLAGS(0);
        case 1: NRK_LAUNCH_FLAGS(1);
        case 2: NRK_LAUNCH_FLAGS(2);
        case 3: NRK_LAUNCH_FLAGS(3);
        case 4: NRK_LAUNCH_FLAGS(4);
        case 5: NRK_LAUNCH_FLAGS(5);
        case 6: NRK_LAUNCH_FLAGS(6);
        default: NRK_LAUNCH_FLAGS(7);
    }
#undef NRK_LAUNCH_FLAGS
}

template <typename T>
nrk_status seg_reduce(const Chain &chain, const T *x, size_t rows, size_t cols, double ddof,
                      const nrk_reduce_outputs *row_out, const nrk_reduce_outputs *full_out, hipStream_t stream,
                      const ExtScratch *ext) {
    return seg_reduce_impl<T>(chain, x, rows, cols, ddof, row_out, full_out, stream, ext, nullptr, nullptr, nullptr);
}

// LEAF a STEP* [LEAF b STEP* BIN STEP*] with IEEE steps only: the lean kernel's program form.
static bool lean_form(const ExprProg &pg, LeanProg *lp, int *first, int *second) {
    int k = 0;
    auto steps = [&](Chain &c) -> bool {
        c.len = 0;
        for (int i = 0; i < NRK_MAX_CHAIN; ++i) {
            c.op[i] = -1;
            c.s[i] = 0.0;
        }
        while (k < pg.len && pg.kind[k] == EXPR_I_STEP) {
            if (sop_math_class(pg.op[k]) != MATH_NONE || c.len == NRK_MAX_CHAIN) return false;
            c.op[c.len] = pg.op[k];
            c.s[c.len] = pg.s[k];
            ++c.len;
            ++k;
        }
        return true;
    };
    *lp = LeanProg{};
    if (pg.len < 1 || pg.kind[0] != EXPR_I_LEAF) return false;
    *first = pg.op[0];
    *second = -1;
    k = 1;
    if (!steps(lp->pre0)) return false;
    if (k == pg.len) {
        steps(lp->pre1);
        steps(lp->post);
        return true;
    }
    if (pg.kind[k] != EXPR_I_LEAF) return false;
    *second = pg.op[k++];
    if (!steps(lp->pre1)) return false;
    if (k >= pg.len || pg.kind[k] != EXPR_I_BIN) return false;
    lp->op = pg.op[k++];
    if (!steps(lp->post)) return false;
    return k == pg.len;
}

template <typename T>
nrk_status expr_reduce(const ExprProg &pg, const T *const *leaves, size_t rows, size_t cols, double ddof,
                       const nrk_reduce_outputs *row_out, const nrk_reduce_outputs *full_out, hipStream_t stream) {
    NRK_TRY(ensure_device());
    if (!leaves) return fail(NRK_ERR_INVALID_ARGUMENT, "leaves is NULL");
    for (int j = 0; j < pg.n_leaves; ++j)
        if (!leaves[j] && rows * cols > 0) return fail(NRK_ERR_INVALID_ARGUMENT, "leaf %d is NULL", j);
    if (row_out && row_out->var && cols <= (size_t)ddof)
        return fail(NRK_ERR_INVALID_ARGUMENT, "Degrees of freedom %zu >= axis size %zu", (size_t)ddof, cols);
    LeanProg lp;
    int a = 0, b = -1;
    if (lean_form(pg, &lp, &a, &b)) {
        if (b < 0) return seg_reduce<T>(lp.pre0, leaves[a], rows, cols, ddof, row_out, full_out, stream);
        bool no_lean = false;
        NRK_TRY(seg_reduce_impl<T>(lp.pre0, leaves[a], rows, cols, ddof, row_out, full_out, stream, nullptr, &lp, leaves[b],
                                   &no_lean));
        if (!no_lean) return NRK_OK;
    }
    // any other tree (or shape): evaluate it into the stream's workspace, then reduce that
    const size_t n = rows * cols;
    void *ws = nullptr;
    NRK_TRY(get_workspace(stream, std::max<size_t>(n, 1) * sizeof(T), &ws));
    NRK_TRY(expr_eval<T>(pg, leaves, (T *)ws, n, stream));
    Chain none;
    NRK_TRY(make_chain(nullptr, 0, &none));
    return seg_reduce<T>(none, (const T *)ws, rows, cols, ddof, row_out, full_out, stream);
}
template nrk_status expr_reduce<float>(const ExprProg &, const float *const *, size_t, size_t, double,
                                       const nrk_reduce_outputs *, const nrk_reduce_outputs *, hipStream_t);
template nrk_status expr_reduce<double>(const ExprProg &, const double *const *, size_t, size_t, double,
                                        const nrk_reduce_outputs *, const nrk_reduce_outputs *, hipStream_t);

template nrk_status seg_reduce<float>(const Chain &, const float *, size_t, size_t, double, const nrk_reduce_outputs *,
                                      const nrk_reduce_outputs *, hipStream_t, const ExtScratch *);
template nrk_status seg_reduce<double>(const Chain &, const double *, size_t, size_t, double, const nrk_reduce_outputs *,
                                       const nrk_reduce_outputs *, hipStream_t, const ExtScratch *);

// ---------------- i64 sum: wrapping integer adds, exact and order-independent ----------------
constexpr int I64_VPT = 8;                 // 16-byte vectors per thread per segment
constexpr int I64_SEG = RB * I64_VPT * 2;  // 4096 elements = 32 KiB
// Persistent grid of 2 workgroups per CU: on 2^28 elements it streams at 7.3 TB/s, against
// 7.0 TB/s with 3-4 and 6.8-6.9 TB/s with 5-16 (fewer concurrent streams, same bytes in flight
// per CU: 8 x 16 B per thread); 1 per CU falls to 5.0 TB/s. Plain (temporal) loads cost 10 %.
constexpr int I64_BLOCKS_PER_CU = 2;

// Each workgroup adds all of its segments (integer adds are exact, so the grouping cannot change
// the result) and writes one partial.
// The i64 sum: a persistent grid walks 32 KiB segments (non-temporal 16-byte loads, wrapping
// unsigned adds: exact and order-independent, so any grouping gives the same bits); one partial
// per workgroup, stored agent-scope and counted with one agent-scope atomic; the last workgroup to
// arrive adds the partials and writes the result (no second launch).
constexpr int I64_MAX_GROUPS = 4 * RB;
__global__ __launch_bounds__(RB) void sum_i64_kernel(const int64_t *x, size_t n, size_t nseg, int vec_ok,
                                                     unsigned long long *partials, unsigned *counter, int64_t *out) {
    typedef long long V2 __attribute__((ext_vector_type(2)));
    __shared__ unsigned long long sred[WAVES];
    __shared__ unsigned s_last;
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
    lds_barrier();
    if (t == 0) {
        unsigned long long g = 0;
        for (int k = 0; k < WAVES; ++k) g += sred[k];
        st_agent(partials + blockIdx.x, g);
        NRK_WAIT_VMEM0();  // the partial has reached the point of coherence before it is counted
        s_last = __hip_atomic_fetch_add(counter, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) == gridDim.x - 1;
    }
    lds_barrier();
    if (!s_last) return;  // uniform over the workgroup
    unsigned long long v[4];
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        const size_t i = t + (size_t)q * RB;
        v[q] = ld_agent(partials + min(i, (size_t)gridDim.x - 1));
        if (i >= gridDim.x) v[q] = 0;
    }
    unsigned long long tot = (v[0] + v[1]) + (v[2] + v[3]);
#pragma unroll
    for (int m = 32; m >= 1; m >>= 1) tot += shfl_xor((uint64_t)tot, m);
    if ((t & 63) == 0) sred[t >> 6] = tot;
    lds_barrier();
    if (t == 0) {
        unsigned long long r = 0;
        for (int k = 0; k < WAVES; ++k) r += sred[k];
        *out = (int64_t)r;
        st_agent(counter, 0u);  // the zero-at-launch invariant for the next launch on this stream
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
    const int grid = (int)std::min<size_t>({nseg, (size_t)cu_count() * I64_BLOCKS_PER_CU, (size_t)I64_MAX_GROUPS});
    Scratch s;
    NRK_TRY(get_scratch(stream, 1, (size_t)grid * 8, &s));
    const int vec_ok = (reinterpret_cast<uintptr_t>(x) % 16) == 0;
    hipLaunchKernelGGL(sum_i64_kernel, dim3(grid), dim3(RB), 0, stream, x, n, nseg, vec_ok, (unsigned long long *)s.partials,
                       s.counters, out);
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

// ---------------- dot: sum of rounded products a[i] * b[i] in a fixed tree ----------------
// ufuncs.rs:946-963 sums the rounded products in T. Each segment of RB * DOT_PER elements gives
// one partial (per-thread sums over DOT_PER products, then the wave and workgroup trees). Thread
// t of a segment sums the elements off + (j RB + t) VN + e, j < DOT_PER / VN, e < VN, in (j, e)
// order, whether they arrive as 16-byte vectors (both operands aligned) or one by one: the tree
// depends only on n, so the result bits are the same for every grid and every alignment.
constexpr int DOT_PER = 16;

template <typename T>
__global__ __launch_bounds__(RB) void dot_kernel(const T *a, const T *b, size_t n, size_t nseg, int vec_ok, T *partials) {
    __shared__ T sred[2][WAVES];
    const int t = threadIdx.x;
    int parity = 0;
    for (size_t seg = blockIdx.x; seg < nseg; seg += gridDim.x, parity ^= 1) {
        const size_t off = seg * (size_t)(RB * DOT_PER);
        T acc = (T)0;
        typedef typename Vec16<T>::type V;
        constexpr int VN = Vec16<T>::N, NV = DOT_PER / VN;
        if (vec_ok && off + (size_t)RB * DOT_PER <= n) {
            // 16-byte non-temporal loads: vector j of thread t holds elements off + (j RB + t) VN ..
            const V *av = reinterpret_cast<const V *>(a + off), *bv = reinterpret_cast<const V *>(b + off);
            V pa[NV], pb[NV];
#pragma unroll
            for (int j = 0; j < NV; ++j) {
                pa[j] = __builtin_nontemporal_load(av + j * RB + t);
                pb[j] = __builtin_nontemporal_load(bv + j * RB + t);
            }
#pragma unroll
            for (int j = 0; j < NV; ++j)
#pragma unroll
                for (int e = 0; e < VN; ++e) acc = rn_add(acc, rn_mul(pa[j][e], pb[j][e]));
        } else if (off + (size_t)RB * DOT_PER <= n) {
            // the same elements in the same order as the vector path, one by one
            T pa[DOT_PER], pb[DOT_PER];
#pragma unroll
            for (int j = 0; j < NV; ++j)
#pragma unroll
                for (int e = 0; e < VN; ++e) {
                    const size_t pos = off + ((size_t)j * RB + t) * VN + e;
                    pa[j * VN + e] = a[pos];
                    pb[j * VN + e] = b[pos];
                }
#pragma unroll
            for (int i = 0; i < DOT_PER; ++i) acc = rn_add(acc, rn_mul(pa[i], pb[i]));
        } else {
            for (int j = 0; j < NV; ++j)
                for (int e = 0; e < VN; ++e) {
                    const size_t pos = off + ((size_t)j * RB + t) * VN + e;
                    if (pos < n) acc = rn_add(acc, rn_mul(a[pos], b[pos]));
                }
        }
#pragma unroll
        for (int m = 32; m >= 1; m >>= 1) acc = rn_add(acc, shfl_xor(acc, m));
        if ((t & 63) == 0) sred[parity][t >> 6] = acc;
        __syncthreads();
        if (t == 0) {
            T s = sred[parity][0];
            for (int k = 1; k < WAVES; ++k) s = rn_add(s, sred[parity][k]);
            partials[seg] = s;
        }
    }
}

template <typename T>
__global__ __launch_bounds__(FB) void sum_finish_kernel(const T *partials, size_t nseg, T *out) {
    __shared__ T sred[FB / 64];
    const int t = threadIdx.x;
    T s = (T)0;
    for (size_t k = t; k < nseg; k += FB) s = rn_add(s, partials[k]);
#pragma unroll
    for (int m = 32; m >= 1; m >>= 1) s = rn_add(s, shfl_xor(s, m));
    if ((t & 63) == 0) sred[t >> 6] = s;
    __syncthreads();
    if (t == 0) {
        T tot = sred[0];
        for (int k = 1; k < FB / 64; ++k) tot = rn_add(tot, sred[k]);
        *out = tot;
    }
}

template <typename T>
nrk_status dot(const T *a, const T *b, size_t n, T *out, hipStream_t stream) {
    NRK_TRY(ensure_device());
    if (!out) return fail(NRK_ERR_INVALID_ARGUMENT, "out is NULL");
    if (n == 0) {  // the empty sum
        hipLaunchKernelGGL(store_kernel<T>, dim3(1), dim3(1), 0, stream, out, (T)0);
        NRK_HIP(hipGetLastError());
        return NRK_OK;
    }
    if (!a || !b) return fail(NRK_ERR_INVALID_ARGUMENT, "a or b is NULL");
    const size_t nseg = ceil_div(n, (size_t)RB * DOT_PER);
    Scratch s;
    NRK_TRY(get_scratch(stream, 0, nseg * sizeof(T), &s));
    const int grid = (int)std::min<size_t>(nseg, (size_t)cu_count() * 8);
    // vector loads where both operands are 16-byte aligned; the scalar path sums the same
    // elements in the same order, so alignment changes only the speed, never the bits
    const int vec_ok = (reinterpret_cast<uintptr_t>(a) % 16 == 0) && (reinterpret_cast<uintptr_t>(b) % 16 == 0);
    hipLaunchKernelGGL(dot_kernel<T>, dim3(grid), dim3(RB), 0, stream, a, b, n, nseg, vec_ok, (T *)s.partials);
    NRK_HIP(hipGetLastError());
    hipLaunchKernelGGL(sum_finish_kernel<T>, dim3(1), dim3(FB), 0, stream, (const T *)s.partials, nseg, out);
    NRK_HIP(hipGetLastError());
    return NRK_OK;
}
template nrk_status dot<float>(const float *, const float *, size_t, float *, hipStream_t);
template nrk_status dot<double>(const double *, const double *, size_t, double *, hipStream_t);

// ---------------- softmax statistics: (max, sum of e^(x - max)) in one read ----------------
// ufuncs.rs:1411-1457 forms max, then sum exp(x - max). Here every lane keeps a running pair
// (m, s): per segment its values' max m' (exact, NaN never wins), the running sum rescaled once by
// e^(m - m') when m' > m, then e^(v - m') added for each value -- one exp per element, no
// communication inside the loop. Pairs merge as (max(m1, m2), s1 e^(m1 - m) + s2 e^(m2 - m)):
// lanes by the xor butterfly, wavefronts in order, then the workgroups' pairs in a fixed order by
// the last workgroup to finish (no second launch). -inf values add 0 (as e^(-inf - M) does in the
// reference), a pair with s = 0 adds nothing, and NaN or +inf values make s NaN, as they make the
// reference's sum NaN.
// values per thread per segment: 16 whole 16-byte vectors of T (256 bytes in flight per lane)
template <typename T>
constexpr int sm_per() { return 16 * Vec16<T>::N; }

template <typename T>
__device__ __forceinline__ T sm_exp(T v) { return fm::exp_t(v); }
// the terms of the sum in binary32: e^v (v <= 0) to 2^-22 relative (degree 6 after the
// reduction), ample for a sum whose own bound is tens of ulps; binary64 keeps the full routine
__device__ __forceinline__ float sm_term(float v) {
    const double d = (double)v;
    return v != v ? v : (float)fm::f32::exp_reduced<6>(d < -110.0 ? -110.0 : d);  // NaN reaches the sum
}
__device__ __forceinline__ double sm_term(double v) { return fm::exp(v); }

template <typename T>
__device__ __forceinline__ void sm_merge(T &m, T &s, T m2, T s2) {
    const T mm = m > m2 ? m : (m2 == m2 ? m2 : m);  // max ignoring NaN
    const T a = s == (T)0 ? (T)0 : rn_mul(s, sm_exp(rn_sub(m, mm)));
    const T b = s2 == (T)0 ? (T)0 : rn_mul(s2, sm_exp(rn_sub(m2, mm)));
    m = mm;
    s = rn_add(a, b);
}

constexpr int SM_MAX_GROUPS = 4 * RB;  // the finish merges at most this many workgroup pairs

// VEC (x 16-byte aligned): a thread's SM_PER values are SM_PER / VN whole vectors of the segment.
template <typename T, bool VEC>
__global__ __launch_bounds__(RB) void softmax_stats_kernel(const T *x, size_t n, size_t nseg, T *partials, unsigned *counter,
                                                           T *ms) {
    typedef typename Vec16<T>::type V;
    constexpr int VN = Vec16<T>::N;
    __shared__ T sm[WAVES], ss[WAVES];
    __shared__ unsigned s_last;
    const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
    const T NEG_INF = -__builtin_huge_valf();
    constexpr int SM_PER = sm_per<T>();
    // the segment's values of this thread; the next segment's are loaded while this one is reduced
    auto load = [&](size_t seg, T (&v)[SM_PER]) {
        const size_t off = seg * (size_t)(RB * SM_PER);
        if constexpr (VEC) {
#pragma unroll
            for (int j = 0; j < SM_PER / VN; ++j) {
                const size_t pos = off + ((size_t)j * RB + t) * VN;
                if (pos + VN <= n) {  // a plain (cached) load: the output pass reads x again next
                    const V w = *reinterpret_cast<const V *>(x + pos);
#pragma unroll
                    for (int e = 0; e < VN; ++e) v[j * VN + e] = w[e];
                } else {
#pragma unroll
                    for (int e = 0; e < VN; ++e) v[j * VN + e] = pos + e < n ? x[pos + e] : NEG_INF;
                }
            }
        } else {
#pragma unroll
            for (int j = 0; j < SM_PER; ++j) {
                const size_t pos = off + (size_t)j * RB + t;
                v[j] = pos < n ? x[pos] : NEG_INF;
            }
        }
    };
    T lm = NEG_INF, ls = (T)0;  // this lane's running pair
    T v[SM_PER], vn[SM_PER];
    if (blockIdx.x < nseg) load(blockIdx.x, v);
    for (size_t seg = blockIdx.x; seg < nseg; seg += gridDim.x) {
        if (seg + gridDim.x < nseg) load(seg + gridDim.x, vn);
        T m = lm;
#pragma unroll
        for (int j = 0; j < SM_PER; ++j) m = v[j] > m ? v[j] : m;
        if (m > lm) {  // a larger max: rescale the running sum once
            ls = ls == (T)0 ? (T)0 : rn_mul(ls, sm_exp(rn_sub(lm, m)));
            lm = m;
        }
#pragma unroll
        for (int j = 0; j < SM_PER; ++j) ls = rn_add(ls, v[j] == NEG_INF ? (T)0 : sm_term(rn_sub(v[j], lm)));
#pragma unroll
        for (int j = 0; j < SM_PER; ++j) v[j] = vn[j];
    }
#pragma unroll
    for (int k = 32; k >= 1; k >>= 1) sm_merge(lm, ls, shfl_xor(lm, k), shfl_xor(ls, k));
    if (lane == 0) {
        sm[wave] = lm;
        ss[wave] = ls;
    }
    lds_barrier();
    if (t == 0) {  // the workgroup's pair, then its arrival
        T M = sm[0], S = ss[0];
        for (int k = 1; k < WAVES; ++k) sm_merge(M, S, sm[k], ss[k]);
        st_agent(partials + 2 * blockIdx.x, M);
        st_agent(partials + 2 * blockIdx.x + 1, S);
        NRK_WAIT_VMEM0();  // the pair has reached the point of coherence before it is counted
        s_last = __hip_atomic_fetch_add(counter, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) == gridDim.x - 1;
    }
    lds_barrier();
    if (!s_last) return;  // uniform over the workgroup
    // the last workgroup: thread t merges pairs t, t + RB, ... in order, then the butterfly,
    // then the wavefronts in order
    T pm[4], ps[4];
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        const size_t i = min(t + (size_t)q * RB, (size_t)gridDim.x - 1);
        pm[q] = ld_agent(partials + 2 * i);
        ps[q] = ld_agent(partials + 2 * i + 1);
    }
    T M = NEG_INF, S = (T)0;
#pragma unroll
    for (int q = 0; q < 4; ++q)
        if (t + (size_t)q * RB < gridDim.x) sm_merge(M, S, pm[q], ps[q]);
#pragma unroll
    for (int k = 32; k >= 1; k >>= 1) sm_merge(M, S, shfl_xor(M, k), shfl_xor(S, k));
    if (lane == 0) {
        sm[wave] = M;
        ss[wave] = S;
    }
    lds_barrier();
    if (t == 0) {
        T Mf = sm[0], Sf = ss[0];
        for (int k = 1; k < WAVES; ++k) sm_merge(Mf, Sf, sm[k], ss[k]);
        ms[0] = Mf;
        ms[1] = Sf;
        st_agent(counter, 0u);  // the zero-at-launch invariant for the next launch on this stream
    }
}

template <typename T>
nrk_status softmax_stats(const T *x, size_t n, T *ms, hipStream_t stream) {
    NRK_TRY(ensure_device());
    if (n == 0 || !x || !ms) return fail(NRK_ERR_INVALID_ARGUMENT, "softmax_stats: empty input or NULL pointer");
    const size_t nseg = ceil_div(n, (size_t)RB * sm_per<T>());
    // one pair per workgroup; as many workgroups as are resident at once (at most SM_MAX_GROUPS)
    const bool vec = reinterpret_cast<uintptr_t>(x) % 16 == 0;
    int resident = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&resident, vec ? (const void *)softmax_stats_kernel<T, true>
                                                                    : (const void *)softmax_stats_kernel<T, false>,
                                                     RB, 0) != hipSuccess || resident < 1)
        resident = 2;
    const int grid = (int)std::min<size_t>({nseg, (size_t)cu_count() * resident, (size_t)SM_MAX_GROUPS});
    Scratch s;
    NRK_TRY(get_scratch(stream, 1, 2 * (size_t)grid * sizeof(T), &s));
    if (vec)
        hipLaunchKernelGGL((softmax_stats_kernel<T, true>), dim3(grid), dim3(RB), 0, stream, x, n, nseg, (T *)s.partials, s.counters, ms);
    else
        hipLaunchKernelGGL((softmax_stats_kernel<T, false>), dim3(grid), dim3(RB), 0, stream, x, n, nseg, (T *)s.partials, s.counters, ms);
    NRK_HIP(hipGetLastError());
    return NRK_OK;
}
template nrk_status softmax_stats<float>(const float *, size_t, float *, hipStream_t);
template nrk_status softmax_stats<double>(const double *, size_t, double *, hipStream_t);

// ---------------- accumulate (ufunc_ops.rs:535-580): inclusive scans along contiguous rows -------
// op ADD (cumsum, statistics/mod.rs:584), MUL (cumprod, :795), MAXIMUM / MINIMUM (NaN-propagating
// UfuncOp::apply, ufunc_ops.rs:156), SUB and DIV (a0 - a1 - ... as ADD over a0, -a1, -a2, ...;
// a0 / a1 / ... as MUL over a0, 1/a1, ...). Rows are cut into tiles of SC_TILE elements: a tile's
// aggregate (K1), the rows of aggregates scanned by this same routine (recursively), then each
// tile rescanned from its carry (K3). Every combination order is fixed by the shape.
constexpr int SC_PER = 8;
constexpr int SC_TILE = RB * SC_PER;

template <typename T, int OP>
__device__ __forceinline__ T acc_apply(T a, T b) {
    if constexpr (OP == 0) return rn_add(a, b);
    else if constexpr (OP == 2) return rn_mul(a, b);
    else if constexpr (OP == 5) return (a != a || b != b) ? dev_nan<T>() : (a > b ? a : b);
    else return (a != a || b != b) ? dev_nan<T>() : (a < b ? a : b);
}
template <typename T, int OP>
__device__ __forceinline__ T acc_identity() {
    if constexpr (OP == 0) return (T)0;
    else if constexpr (OP == 2) return (T)1;
    else if constexpr (OP == 5) return -(T)__builtin_huge_valf();
    else return (T)__builtin_huge_valf();
}
// Workgroup-wide inclusive scan of one value per thread, in a fixed order: the wavefront's
// Hillis-Steele steps, then the wavefront totals from LDS. Its barriers are LDS-only: global loads
// of the next tile stay in flight across them.
template <typename T, int OP>
__device__ __forceinline__ T block_scan(T v, T *lds) {
    const int t = threadIdx.x, lane = t & 63, w = t >> 6;
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
        const T u = __shfl_up(v, d, 64);
        if (lane >= d) v = acc_apply<T, OP>(u, v);
    }
    if (lane == 63) lds[w] = v;
    lds_barrier();
    T pre = acc_identity<T, OP>();
    for (int k = 0; k < w; ++k) pre = acc_apply<T, OP>(pre, lds[k]);
    lds_barrier();
    return w ? acc_apply<T, OP>(pre, v) : v;
}

// A tile moves between HBM and the threads through LDS: the global loads and stores are
// coalesced, and thread t reads or writes its SC_PER consecutive elements in LDS. The padded
// index (one slot per SC_PER elements) keeps those per-thread accesses on distinct banks.
__device__ __forceinline__ int sc_pad(int i) { return i + i / SC_PER; }
constexpr int SC_LDS = SC_TILE + SC_TILE / SC_PER;

// The global side of a tile, held in registers between its loads and its LDS writes, so the next
// tile's loads are in flight while this one is scanned. VEC (operands 16-byte aligned, rows a
// whole number of vectors): thread t holds 16-byte vectors j RB + t of the tile; else elements
// j RB + t. Elements past the row's end hold the identity.
template <typename T, bool VEC>
struct ScRegs {
    static constexpr int VN = VEC ? Vec16<T>::N : 1;
    static constexpr int NL = SC_PER / VN;  // loads per thread and tile
    T v[SC_PER];
};

template <typename T, int OP, int XF, bool VEC>
__device__ __forceinline__ void sc_load(const T *x, size_t r, size_t len, size_t tr, ScRegs<T, VEC> &g) {
    constexpr int VN = ScRegs<T, VEC>::VN, NL = ScRegs<T, VEC>::NL;
    const size_t c0 = tr * SC_TILE;
#pragma unroll
    for (int j = 0; j < NL; ++j) {
        const size_t col = c0 + ((size_t)j * RB + threadIdx.x) * VN;  // first element of this load
        if constexpr (VEC) {
            typedef typename Vec16<T>::type V;
            if (col + VN <= len) {
                const V w = __builtin_nontemporal_load(reinterpret_cast<const V *>(x + r * len + col));
#pragma unroll
                for (int e = 0; e < VN; ++e) g.v[j * VN + e] = w[e];
            } else {
#pragma unroll
                for (int e = 0; e < VN; ++e) g.v[j * VN + e] = (T)0;  // past the end (rows are whole vectors)
            }
        } else {
            g.v[j] = col < len ? x[r * len + col] : (T)0;
        }
    }
}

// registers -> LDS, with the element as the scan sees it (SUB negates and DIV inverts every element
// after a row's first; identity past the row's end)
template <typename T, int OP, int XF, bool VEC>
__device__ __forceinline__ void sc_put(const ScRegs<T, VEC> &g, size_t len, size_t tr, T *buf) {
    constexpr int VN = ScRegs<T, VEC>::VN, NL = ScRegs<T, VEC>::NL;
    const size_t c0 = tr * SC_TILE;
#pragma unroll
    for (int j = 0; j < NL; ++j)
#pragma unroll
        for (int e = 0; e < VN; ++e) {
            const int i = (j * RB + (int)threadIdx.x) * VN + e;  // position in the tile
            const size_t col = c0 + i;
            T v = g.v[j * VN + e];
            if constexpr (XF == 1) v = col == 0 ? v : -v;
            else if constexpr (XF == 2) v = col == 0 ? v : rn_div((T)1, v);
            else if constexpr (OP == 0) v = rn_add(v, (T)0);  // -0 -> +0: cumsum's running sum starts at +0
            buf[sc_pad(i)] = col < len ? v : acc_identity<T, OP>();
        }
}

// K1: one aggregate per tile (thread folds its SC_PER consecutive elements, then the block scan's
// last value).
template <typename T, int OP, int XF, bool VEC>
__global__ __launch_bounds__(RB) void scan_tile_agg_kernel(const T *x, size_t rows, size_t len, size_t tpr, T *agg) {
    __shared__ T lds[WAVES];
    __shared__ T buf[SC_LDS];
    const size_t tiles = rows * tpr;
    ScRegs<T, VEC> g;
    if (blockIdx.x < tiles) sc_load<T, OP, XF, VEC>(x, blockIdx.x / tpr, len, blockIdx.x % tpr, g);
    for (size_t tile = blockIdx.x; tile < tiles; tile += gridDim.x) {
        const size_t r = tile / tpr, tr = tile - r * tpr;
        sc_put<T, OP, XF, VEC>(g, len, tr, buf);
        lds_barrier();
        const size_t nt = tile + gridDim.x;  // the next tile's loads, in flight while this one folds
        if (nt < tiles) sc_load<T, OP, XF, VEC>(x, nt / tpr, len, nt % tpr, g);
        T a = acc_identity<T, OP>();
#pragma unroll
        for (int j = 0; j < SC_PER; ++j) a = acc_apply<T, OP>(a, buf[sc_pad(threadIdx.x * SC_PER + j)]);
        a = block_scan<T, OP>(a, lds);  // its barriers also retire this tile's reads of buf
        if (threadIdx.x == RB - 1) agg[tile] = a;
    }
}

// K3: each tile rescanned: y = carry (x) thread prefix (x) in-thread prefix, where carry is the
// inclusive scan of the previous tiles' aggregates of the row (none for a row's first tile).
template <typename T, int OP, int XF, bool VEC>
__global__ __launch_bounds__(RB) void scan_tile_kernel(const T *x, T *y, size_t rows, size_t len, size_t tpr,
                                                       const T *agg_scan) {
    constexpr int VN = ScRegs<T, VEC>::VN, NL = ScRegs<T, VEC>::NL;
    __shared__ T lds[WAVES];
    __shared__ T buf[SC_LDS];
    const size_t tiles = rows * tpr;
    ScRegs<T, VEC> g;
    if (blockIdx.x < tiles) sc_load<T, OP, XF, VEC>(x, blockIdx.x / tpr, len, blockIdx.x % tpr, g);
    for (size_t tile = blockIdx.x; tile < tiles; tile += gridDim.x) {
        const size_t r = tile / tpr, tr = tile - r * tpr;
        sc_put<T, OP, XF, VEC>(g, len, tr, buf);
        lds_barrier();
        const size_t nt = tile + gridDim.x;  // the next tile's loads, in flight while this one scans
        if (nt < tiles) sc_load<T, OP, XF, VEC>(x, nt / tpr, len, nt % tpr, g);
        T v[SC_PER];
#pragma unroll
        for (int j = 0; j < SC_PER; ++j) v[j] = buf[sc_pad(threadIdx.x * SC_PER + j)];
#pragma unroll
        for (int j = 1; j < SC_PER; ++j) v[j] = acc_apply<T, OP>(v[j - 1], v[j]);
        const T incl = block_scan<T, OP>(v[SC_PER - 1], lds);
        // the exclusive prefix: the previous thread's inclusive value; for a wavefront's first
        // thread that is the fold of the earlier wavefronts' totals, in block_scan's own order
        T pre = __shfl_up(incl, 1, 64);
        if ((threadIdx.x & 63) == 0 && threadIdx.x) {
            T p = acc_identity<T, OP>();
            for (int k = 0; k < (int)(threadIdx.x >> 6); ++k) p = acc_apply<T, OP>(p, lds[k]);
            pre = p;
        }
        const bool has_carry = tr > 0;
        const T carry = has_carry ? agg_scan[tile - 1] : acc_identity<T, OP>();
        const bool first = threadIdx.x == 0;
        // block_scan's barriers have retired every read of buf: the results go back through it
#pragma unroll
        for (int j = 0; j < SC_PER; ++j) {
            T o = v[j];
            if (!first) o = acc_apply<T, OP>(pre, o);
            if (has_carry) o = acc_apply<T, OP>(carry, o);
            buf[sc_pad(threadIdx.x * SC_PER + j)] = o;
        }
        lds_barrier();
        const size_t c0 = tr * SC_TILE;
#pragma unroll
        for (int j = 0; j < NL; ++j) {
            const int i0 = (j * RB + (int)threadIdx.x) * VN;
            const size_t col = c0 + i0;
            if constexpr (VEC) {
                typedef typename Vec16<T>::type V;
                if (col + VN <= len) {
                    V w;
#pragma unroll
                    for (int e = 0; e < VN; ++e) w[e] = buf[sc_pad(i0 + e)];
                    __builtin_nontemporal_store(w, reinterpret_cast<V *>(y + r * len + col));
                }
            } else {
                if (col < len) y[r * len + col] = buf[sc_pad(i0)];
            }
        }
        lds_barrier();  // lds and buf are reused by the next tile
    }
}

template <typename T, int OP, int XF>
static nrk_status scan_rows(const T *x, T *y, size_t rows, size_t len, T *ws, hipStream_t stream) {
    const size_t tpr = ceil_div(len, (size_t)SC_TILE);
    const size_t tiles = rows * tpr;
    const int grid = (int)std::min<size_t>(tiles, (size_t)cu_count() * 16);
    // 16-byte loads and stores when every row starts on a 16-byte boundary (the summation tree is
    // the same either way)
    const bool vec = ((reinterpret_cast<uintptr_t>(x) | reinterpret_cast<uintptr_t>(y)) & 15) == 0 && len % Vec16<T>::N == 0;
    const T *agg = nullptr;
    if (tpr > 1) {
        T *a = ws;
        if (vec) hipLaunchKernelGGL((scan_tile_agg_kernel<T, OP, XF, true>), dim3(grid), dim3(RB), 0, stream, x, rows, len, tpr, a);
        else hipLaunchKernelGGL((scan_tile_agg_kernel<T, OP, XF, false>), dim3(grid), dim3(RB), 0, stream, x, rows, len, tpr, a);
        NRK_HIP(hipGetLastError());
        // the aggregates' own rows (rows x tpr), scanned in place; their transform is applied already
        NRK_TRY((scan_rows<T, OP, 0>(a, a, rows, tpr, ws + tiles, stream)));
        agg = a;
    }
    if (vec) hipLaunchKernelGGL((scan_tile_kernel<T, OP, XF, true>), dim3(grid), dim3(RB), 0, stream, x, y, rows, len, tpr, agg);
    else hipLaunchKernelGGL((scan_tile_kernel<T, OP, XF, false>), dim3(grid), dim3(RB), 0, stream, x, y, rows, len, tpr, agg);
    NRK_HIP(hipGetLastError());
    return NRK_OK;
}

template <typename T>
nrk_status accumulate(int op, const T *x, T *y, size_t rows, size_t len, hipStream_t stream) {
    NRK_TRY(ensure_device());
    if (op != NRK_BOP_ADD && op != NRK_BOP_SUB && op != NRK_BOP_MUL && op != NRK_BOP_DIV && op != NRK_BOP_MAXIMUM &&
        op != NRK_BOP_MINIMUM)
        return fail(NRK_ERR_INVALID_ARGUMENT, "accumulate: op %d is not ADD, SUB, MUL, DIV, MAXIMUM or MINIMUM", op);
    if (rows * len == 0) return NRK_OK;
    if (!x || !y) return fail(NRK_ERR_INVALID_ARGUMENT, "x or y is NULL");
    // workspace: every level's aggregates (each level 1/SC_TILE of the one before, at least one per row)
    size_t need = 0;
    for (size_t l = len; l > 1;) {
        const size_t tpr = ceil_div(l, (size_t)SC_TILE);
        if (tpr <= 1) break;
        need += rows * tpr;
        l = tpr;
    }
    void *wsv = nullptr;
    NRK_TRY(get_workspace(stream, (need + 1) * sizeof(T), &wsv));
    T *ws = (T *)wsv;
    switch (op) {
        case NRK_BOP_ADD: return scan_rows<T, 0, 0>(x, y, rows, len, ws, stream);
        case NRK_BOP_SUB: return scan_rows<T, 0, 1>(x, y, rows, len, ws, stream);
        case NRK_BOP_MUL: return scan_rows<T, 2, 0>(x, y, rows, len, ws, stream);
        case NRK_BOP_DIV: return scan_rows<T, 2, 2>(x, y, rows, len, ws, stream);
        case NRK_BOP_MAXIMUM: return scan_rows<T, 5, 0>(x, y, rows, len, ws, stream);
        default: return scan_rows<T, 6, 0>(x, y, rows, len, ws, stream);
    }
}
template nrk_status accumulate<float>(int, const float *, float *, size_t, size_t, hipStream_t);
template nrk_status accumulate<double>(int, const double *, double *, size_t, size_t, hipStream_t);

// ---------------- reduceat (ufunc_ops.rs:619-684) along contiguous rows ----------------
// Output k of row r reduces x[r][j .. end) with j = idx[k], end = idx[k + 1] (len for the last
// k), and just x[r][j] when end <= j. One workgroup per (row, segment): thread t folds elements
// j + t, j + t + RB, ... in order, then the wavefront butterfly and the wavefront totals in
// order. SUB / DIV fold x[j], -x[k] / 1/x[k] (k > j) with ADD / MUL. ADD starts from -0, the
// exact additive identity, so a segment of zeros keeps the reference's sign. An index >= len
// (which the host entry point rejects) yields NaN, never an out-of-bounds read.
template <typename T, int OP>
__device__ __forceinline__ T rat_identity() {
    if constexpr (OP == 0) return -(T)0;
    else return acc_identity<T, OP>();
}

template <typename T, int OP, int XF>
__global__ __launch_bounds__(RB) void reduceat_kernel(const T *x, size_t rows, size_t len, const uint64_t *idx, size_t nidx,
                                                      T *y) {
    __shared__ T lds[WAVES];
    const int t = threadIdx.x, lane = t & 63, w = t >> 6;
    for (size_t item = blockIdx.x; item < rows * nidx; item += gridDim.x) {
        const size_t r = item / nidx, k = item - r * nidx;
        const uint64_t j = idx[k];
        const uint64_t nxt = k + 1 < nidx ? idx[k + 1] : (uint64_t)len;
        const uint64_t e = nxt > j ? (nxt < len ? nxt : (uint64_t)len) : j + 1;
        T a = rat_identity<T, OP>();
        if (j < len) {
            // thread t folds the t-th contiguous run of the segment: every combine below keeps the
            // earlier elements on the left, so maximum / minimum return the element the reference's
            // sequential fold returns (a tie goes to the later operand, which fixes a zero's sign)
            const uint64_t per = (e - j + RB - 1) / RB;
            const uint64_t c0 = j + (uint64_t)t * per, c1 = c0 + per < e ? c0 + per : e;
            for (uint64_t c = c0; c < c1; ++c) {
                T v = x[r * len + c];
                if constexpr (XF == 1) v = c == j ? v : -v;
                else if constexpr (XF == 2) v = c == j ? v : rn_div((T)1, v);
                a = acc_apply<T, OP>(a, v);
            }
        }
#pragma unroll
        for (int m = 1; m < 64; m <<= 1) {
            const T o = shfl_xor(a, m);
            a = (lane & m) ? acc_apply<T, OP>(o, a) : acc_apply<T, OP>(a, o);  // lower lane's value first
        }
        if (lane == 0) lds[w] = a;
        __syncthreads();
        if (t == 0) {
            T s = lds[0];
            for (int q = 1; q < WAVES; ++q) s = acc_apply<T, OP>(s, lds[q]);
            y[item] = j < len ? s : (T)__builtin_nanf("");
        }
        __syncthreads();
    }
}

template <typename T>
nrk_status reduceat(int op, const T *x, size_t rows, size_t len, const uint64_t *idx, size_t nidx, T *y, hipStream_t stream) {
    NRK_TRY(ensure_device());
    if (op != NRK_BOP_ADD && op != NRK_BOP_SUB && op != NRK_BOP_MUL && op != NRK_BOP_DIV && op != NRK_BOP_MAXIMUM &&
        op != NRK_BOP_MINIMUM)
        return fail(NRK_ERR_INVALID_ARGUMENT, "reduceat: op %d is not ADD, SUB, MUL, DIV, MAXIMUM or MINIMUM", op);
    if (rows * nidx == 0) return NRK_OK;
    if (!x || !idx || !y) return fail(NRK_ERR_INVALID_ARGUMENT, "x, indices or y is NULL");
    const int grid = (int)std::min<size_t>(rows * nidx, (size_t)cu_count() * 16);
#define NRK_RAT(OPC, XFC) hipLaunchKernelGGL((reduceat_kernel<T, OPC, XFC>), dim3(grid), dim3(RB), 0, stream, x, rows, len, idx, nidx, y)
    switch (op) {
        case NRK_BOP_ADD: NRK_RAT(0, 0); break;
        case NRK_BOP_SUB: NRK_RAT(0, 1); break;
        case NRK_BOP_MUL: NRK_RAT(2, 0); break;
        case NRK_BOP_DIV: NRK_RAT(2, 2); break;
        case NRK_BOP_MAXIMUM: NRK_RAT(5, 0); break;
        default: NRK_RAT(6, 0); break;
    }
#undef NRK_RAT
    NRK_HIP(hipGetLastError());
    return NRK_OK;
}
template nrk_status reduceat<float>(int, const float *, size_t, size_t, const uint64_t *, size_t, float *, hipStream_t);
template nrk_status reduceat<double>(int, const double *, size_t, size_t, const uint64_t *, size_t, double *, hipStream_t);

// y = sqrt(y) for one device scalar (norm_l2 = sqrt of the sum of squares, std = sqrt of var).
template <typename T>
__global__ void sqrt_scalar_kernel(T *p) { *p = rn_sqrt(*p); }

template <typename T>
nrk_status sqrt_scalar(T *p, hipStream_t stream) {
    hipLaunchKernelGGL(sqrt_scalar_kernel<T>, dim3(1), dim3(1), 0, stream, p);
    NRK_HIP(hipGetLastError());
    return NRK_OK;
}
template nrk_status sqrt_scalar<float>(float *, hipStream_t);
template nrk_status sqrt_scalar<double>(double *, hipStream_t);

}  // namespace nrk
