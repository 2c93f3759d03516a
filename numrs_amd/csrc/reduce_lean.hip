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
//     pointer, loading the next segment while it folds the current one;
//   - every lane of a chunk has the same count (16 per segment), so the Chan updates need no
//     division: the weight of the j-th segment's merge, 1 / (16 j (j + 1)), comes from a constant
//     table, and the butterfly's levels merge equal counts with weight 1 / (2 n);
//   - binary32 lanes add their segment sums into plain FP64 (exact to 2^-29 of a binary32 ulp);
//     binary64 lanes keep the double word;
//   - lane 0 writes a row's results as soon as its chunk is merged; each wavefront merges its
//     chunks in order, the workgroup its 4 wavefronts into one group record, and one finish kernel
//     merges the group records in a fixed order.
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
    Parts grp;
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

// Fold segment j of the current chunk (loaded into y, and z for the second operand).
template <typename T, int VPT, int FLAGS, int LEAVES, int CH>
__device__ __forceinline__ void lean_fold(const LeanProg &pg, T s0, T s1, T (&y)[VPT * Vec16<T>::N], T (&z)[VPT * Vec16<T>::N],
                                          unsigned j, LaneAcc<T> &a, unsigned &fnan) {
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
    if (j == 0) fnan = __builtin_isnan(y[0]) ? 1u : 0u;  // lane 0's value is the chunk's first element
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
        a.hi = s + e;
        a.lo = e - (a.hi - s);
    } else {
        a.hi += sb;
    }
    if (FLAGS & F_MAX) {
        unsigned ci = NO_CODE;
#pragma unroll
        for (int i = 0; i < M; ++i) {
            const bool gt = y[i] > a.vmax;
            a.vmax = gt ? y[i] : a.vmax;
            ci = gt ? (unsigned)i : ci;
        }
        a.cmax = ci != NO_CODE ? j * M + ci : a.cmax;
    }
    if (FLAGS & F_MIN) {
        unsigned ci = NO_CODE;
#pragma unroll
        for (int i = 0; i < M; ++i) {
            const bool lt = y[i] < a.vmin;
            a.vmin = lt ? y[i] : a.vmin;
            ci = lt ? (unsigned)i : ci;
        }
        a.cmin = ci != NO_CODE ? j * M + ci : a.cmin;
    }
}

// Butterfly of records of equal counts n (every lane of a chunk): M2 += (S_a - S_b)^2 / (2 n).
template <typename T, int FLAGS>
__device__ __forceinline__ void lean_wave_merge(Part &a) {
    double w = 0.5 / a.n;
#pragma unroll
    for (int m = 32; m >= 1; m >>= 1) {
        const Part b = part_shfl<FLAGS>(a, m);
        a.nan |= b.nan;
        if (FLAGS & F_MAX) take_max(a.vmax, a.imax, b.vmax, b.imax);
        if (FLAGS & F_MIN) take_min(a.vmin, a.imin, b.vmin, b.imin);
        if (FLAGS & F_M2) {
            const double d = (a.hi + a.lo) - (b.hi + b.lo);
            a.m2 = a.m2 + b.m2 + d * d * w;
        }
        if constexpr (sizeof(T) == 8) dw_add(a.hi, a.lo, b.hi, b.lo);
        else a.hi += b.hi;
        a.n += b.n;
        w *= 0.5;
    }
}

template <typename T, int VPT, int FLAGS, int LEAVES, bool VEC, int CH>
__global__ __launch_bounds__(RB) void lean_kernel(LeanParams<T> p) {
    constexpr int VN = Vec16<T>::N;
    constexpr int M = VPT * VN;
    constexpr int WSEG = 64 * M;
    __shared__ Part grp_lds[WAVES];
    __shared__ unsigned grp_fn[WAVES];
    const unsigned lane = threadIdx.x & 63;
    const unsigned wave = lane0((unsigned)(threadIdx.x >> 6));
    const size_t v = blockIdx.x * (size_t)WAVES + wave;
    const size_t c0 = v * p.A;  // chunks c0, c0 + B, ...
    const size_t nck = c0 >= p.nchunk ? 0 : (p.A == 1 ? (p.nchunk - c0 + p.B - 1) / p.B : min(p.A, p.nchunk - c0));
    const size_t nseg = nck * p.cl;
    const size_t cstride = p.B * p.cl * WSEG;  // elements from one of this wavefront's chunks to the next
    Part wp = part_empty();  // this wavefront's chunks, merged in chunk order
    unsigned wp_fn = 0;
    if (nseg > 0) {
        LaneAcc<T> a;
        acc_reset(a);
        unsigned fnan = 0, j = 0;
        size_t k = 0;
        T ya[M], yb[M], za[M], zb[M];
        // segment t of this wavefront: chunk c0 + (t / cl) B, segment t % cl of it
        const T *n0 = p.x0 + c0 * p.cl * WSEG;  // the next segment to load
        const T *n1 = LEAVES == 2 ? p.x1 + c0 * p.cl * WSEG : nullptr;
        unsigned nj = 0;
        size_t left = nseg;  // segments not yet loaded
        auto load = [&](T(&y)[M], T(&z)[M]) {
            lean_load<T, VPT, VEC>(n0, y);
            if constexpr (LEAVES == 2) lean_load<T, VPT, VEC>(n1, z);
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
        auto step = [&](T(&y)[M], T(&z)[M]) {
            lean_fold<T, VPT, FLAGS, LEAVES, CH>(p.prog, s0, s1, y, z, j, a, fnan);
            if (++j == p.cl) {  // chunk c = c0 + k B is complete
                const size_t c = c0 + k * p.B;
                const uint64_t fbase = (uint64_t)c * p.cl * WSEG;
                auto flat = [&](unsigned code) -> uint64_t {
                    const unsigned sl = code / M, i = code % M;
                    return fbase + (uint64_t)sl * WSEG + (uint64_t)(((i / VN) * 64 + lane) * VN + (i % VN));
                };
                Part pa;
                pa.n = (double)M * (double)p.cl;
                pa.hi = a.hi; pa.lo = a.lo; pa.m2 = a.m2; pa.nan = a.nan;
                pa.vmax = a.cmax == NO_CODE ? -__builtin_inf() : (double)a.vmax;
                pa.imax = a.cmax == NO_CODE ? NO_INDEX : flat(a.cmax);
                pa.vmin = a.cmin == NO_CODE ? __builtin_inf() : (double)a.vmin;
                pa.imin = a.cmin == NO_CODE ? NO_INDEX : flat(a.cmin);
                lean_wave_merge<T, FLAGS>(pa);
                const unsigned fn0 = (unsigned)__shfl(fnan, 0, 64);
                if (p.row_direct && lane == 0) write_outputs<T>(p.row, c, pa, p.ddof, fn0, (uint64_t)c * p.cols);
                if (k == 0) wp_fn = fn0;
                part_merge<FLAGS>(wp, pa);
                acc_reset(a);
                fnan = 0;
                j = 0;
                ++k;
            }
        };
        load(ya, za);
        // the next segment's loads are issued before the current one is folded
        for (size_t s = 0;;) {
            load(yb, zb);
            step(ya, za);
            if (++s == nseg) break;
            load(ya, za);
            step(yb, zb);
            if (++s == nseg) break;
        }
    }
    if (!p.need_group) return;  // uniform over the grid
    if (lane == 0) {
        grp_lds[wave] = wp;
        grp_fn[wave] = wp_fn;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        Part g = grp_lds[0];
        for (int w = 1; w < WAVES; ++w) part_merge<FLAGS>(g, grp_lds[w]);
        part_store<FLAGS>(p.grp, blockIdx.x, g, grp_fn[0]);
    }
}

// The whole-array result from the group records, merged in a fixed order: thread t folds records
// t, t + LFB, ..., then the wavefront butterflies, then wavefront order.
constexpr int LFB = 512;
template <typename T, int FLAGS>
__global__ __launch_bounds__(LFB) void lean_finish_kernel(Parts in, size_t count, OutPtrs<T> o, double ddof) {
    __shared__ Part lds[LFB / 64];
    const unsigned t = threadIdx.x, lane = t & 63, w = t >> 6;
    Part a = part_empty();
    for (size_t i = t; i < count; i += LFB) part_merge<FLAGS>(a, part_load<FLAGS>(in, i));
    wave_merge<FLAGS>(a);
    if (lane == 0) lds[w] = a;
    __syncthreads();
    if (t == 0) {
        Part r = lds[0];
        for (int k = 1; k < LFB / 64; ++k) part_merge<FLAGS>(r, lds[k]);
        write_outputs<T>(o, 0, r, ddof, (in.flags[0] >> 1) & 1u, 0);
    }
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
// record holds, hence the whole-array result's bits). Measured on 8192 x 8192 (stream kernel
// alone): binary32 reads faster when wavefront v takes chunks v, v + W, ... (W the grid's
// wavefronts; the grid reads one compact window at a time): 41-45 us against 49-50 with
// contiguous blocks; binary64 the other way round (8 consecutive chunks per wavefront: 79-80 us
// against 84-86). At most 512 workgroups (the finish merges one record per thread).
constexpr size_t LEAN_MIN_CHUNKS = 1024;
struct LeanLayout {
    size_t groups, A, B;
};
template <typename T>
static LeanLayout lean_layout(size_t nchunk) {
    LeanLayout L;
    if (sizeof(T) == 4) {
        L.groups = std::min<size_t>(512, ceil_div(nchunk, (size_t)4 * WAVES));
        L.A = 1;
        L.B = L.groups * WAVES;
    } else {
        L.A = std::max<size_t>(8, ceil_div(nchunk, (size_t)512 * WAVES));
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
static nrk_status launch_lean(const LeanParams<T> &p, size_t ngroups, const OutPtrs<T> &full, hipStream_t stream) {
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
    if (p.need_group) {
        hipLaunchKernelGGL((lean_finish_kernel<T, FLAGS>), dim3(1), dim3(LFB), 0, stream, p.grp, ngroups, full, p.ddof);
        NRK_HIP(hipGetLastError());
    }
    return NRK_OK;
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
    size_t cl, nchunk;
    if (want_row) {  // a chunk is a row
        if (cols % WSEG != 0 || cols / WSEG >= (size_t)LEAN_WTAB || rows < LEAN_MIN_CHUNKS) return NRK_OK;
        cl = cols / WSEG;
        nchunk = rows;
    } else {
        const size_t n = rows * cols;
        if (n % WSEG != 0) return NRK_OK;
        cl = lean_chunk_len(n / WSEG);
        if (cl == 0) return NRK_OK;
        nchunk = n / WSEG / cl;
        if (nchunk < LEAN_MIN_CHUNKS) return NRK_OK;
    }
    LeanParams<T> p{};
    p.x0 = x0;
    p.x1 = x1;
    p.nchunk = nchunk;
    p.cl = cl;
    p.cols = cols;
    const LeanLayout L = lean_layout<T>(nchunk);
    const size_t ngroups = L.groups;
    p.A = L.A;
    p.B = L.B;
    p.prog = prog;
    if (!x1 && prog.pre0.len == 2) {  // CH_MULADD: y = (x * s0) + s1, with x - b as x + (-b)
        p.s0 = prog.pre0.s[0];
        p.s1 = prog.pre0.op[1] == NRK_SOP_SUB ? -prog.pre0.s[1] : prog.pre0.s[1];
    }
    p.ddof = ddof;
    p.row_direct = want_row;
    p.row = want_row ? row : OutPtrs<T>{};
    p.need_group = want_full;
    if (want_full) {
        const size_t pbytes = parts_bytes(ngroups) + 256;
        char *b = nullptr;
        if (ext) {
            if (pbytes + 256 > ext->bytes)
                return fail(NRK_ERR_INVALID_ARGUMENT, "(%zu, %zu) needs %zu bytes of scratch; the fixed scratch holds %zu", rows, cols,
                            pbytes + 256, ext->bytes);
            b = (char *)ext->base + 256;
        } else {
            Scratch s;
            NRK_TRY(get_scratch(stream, 1, pbytes, &s));
            b = (char *)s.partials;
        }
        p.grp = carve_parts(b, ngroups);
    }
    const bool m2 = (row.var || full.var);
    const bool mx = (row.max || row.argmax || full.max || full.argmax);
    const bool mn = (row.min || row.argmin || full.min || full.argmin);
    const int flags = (m2 ? F_M2 : 0) | (mx ? F_MAX : 0) | (mn ? F_MIN : 0);
    *taken = true;
#define NRK_LEAN(F)                                                                             \
    case F:                                                                                     \
        return x1 ? launch_lean<T, VPT, F, 2>(p, ngroups, full, stream) : launch_lean<T, VPT, F, 1>(p, ngroups, full, stream);
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
