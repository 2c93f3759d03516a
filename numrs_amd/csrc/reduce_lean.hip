// The lean streaming reduction kernel and its launcher (the design and the shared device pieces are
// in reduce_lean.hpp, the short-row kernel in reduce_rows.hpp; the general kernels and the host
// entry points are in reduce.hip).
#include "reduce_lean.hpp"

namespace nrk {

// The tail of a whole-array reduction (what follows the last whole chunk: fewer than a chunk and
// a segment of elements): lane l of wavefront w folds elements tail0 + 64 w + l, + RB, ... in
// index order as one-element records, then the wavefront's butterfly; wp receives the result.
// Loads go out LEAN_TAIL_BATCH at a time (values past the end are loaded as zeros and skipped).
constexpr int LEAN_TAIL_BATCH = 16;
template <typename T, int FLAGS, int LEAVES, int CH>
__device__ void lean_tail(const LeanParams<T> &p, unsigned wave, unsigned lane, Part &wp, unsigned &wp_fn) {
    constexpr int NB = LEAN_TAIL_BATCH;
    Part a = part_empty();
    const T s0 = (T)p.s0, s1 = (T)p.s1;
    unsigned fn = 0;  // the array's element 0 is NaN (lane 0 of wavefront 0 holds it when the tail starts at 0)
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
        if (i0 == 0) fn = y[0] != y[0] ? 1u : 0u;
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
    wp_fn = (unsigned)__shfl((int)fn, 0, 64);
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
    if (is_tail) lean_tail<T, FLAGS, LEAVES, CH>(p, wave, lane, wp, wp_fn);
    if (!p.need_group) {  // uniform over the grid
        if (p.row_direct) flush_rows();
        return;
    }
    lean_group_finish<T, FLAGS>(fin, p.grp, p.counter, p.full, p.ddof, p.uniform, wp, wp_fn, [&]() {
        if (p.row_direct) flush_rows();
    });
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
// ... (W the grid's wavefronts), binary64 8 consecutive chunks per wavefront (fewer below
// 8 x 256 x 4 chunks, for 256 workgroups); below LEAN_MIN_CHUNKS chunks, wavefront v takes chunk v.
constexpr size_t LEAN_MIN_CHUNKS = 1024;
constexpr size_t LEAN_GROUPS_F32 = 512;  // round robin over at most this many workgroups
constexpr size_t LEAN_BLOCK_F64 = 8;     // consecutive chunks per binary64 wavefront (at least) ...
constexpr size_t LEAN_GROUPS_F64 = 256;  // ... where that leaves at least this many workgroups
static_assert(LEAN_GROUPS_F32 <= LEAN_MAX_GROUPS, "the finish merges at most LEAN_MAX_GROUPS group records");
struct LeanLayout {
    size_t groups, A, B;
};
template <typename T>
static LeanLayout lean_layout(size_t nchunk, size_t max_groups) {
    LeanLayout L;
    if (nchunk < LEAN_MIN_CHUNKS) {  // a small array: one chunk per wavefront, the fewest dependent loads
        L.groups = ceil_div(nchunk, (size_t)WAVES);
        L.A = 1;
        L.B = L.groups * WAVES;
    } else if (sizeof(T) == 4) {
        L.groups = std::min<size_t>(LEAN_GROUPS_F32, ceil_div(nchunk, (size_t)4 * WAVES));
        L.A = 1;
        L.B = L.groups * WAVES;
    } else {
        // LEAN_BLOCK_F64 (or more) consecutive chunks per wavefront where that leaves a workgroup
        // per CU (LEAN_GROUPS_F64), fewer where it would not
        const size_t full = LEAN_BLOCK_F64 * LEAN_GROUPS_F64 * WAVES;
        L.A = nchunk >= full ? std::max<size_t>(LEAN_BLOCK_F64, ceil_div(nchunk, max_groups * WAVES))
                             : ceil_div(nchunk, LEAN_GROUPS_F64 * WAVES);
        L.groups = ceil_div(nchunk, L.A * WAVES);
        L.B = L.A == 1 ? L.groups * WAVES : 1;  // A = 1 is read as round robin (one chunk per wavefront)
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

__global__ void lean_clear_counter_kernel(unsigned int *c) { c[threadIdx.x * LEAN_CTR_STRIDE] = 0u; }

// The arrival counters (zero) and ngroups group records of the whole-array finish: the caller's
// scratch (counters first, cleared here; records after) or the stream's.
nrk_status group_scratch(size_t ngroups, size_t rows, size_t cols, const ExtScratch *ext, hipStream_t stream,
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
    // short rows: positions (cols, plus up to VN - 1 of phase for binary32) fit 64 lanes of
    // 2 rows_k slots
    const size_t span = cols + (sizeof(T) == 4 && rows > 1 && cols % Vec16<T>::N != 0 ? Vec16<T>::N - 1 : 0);
    if (want_row && !row_chunks && !x1 && span <= 128 * (size_t)rows_k<T>()) {
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
        nchunk = spr / cl;  // any count: fewer than a segment is all tail (one workgroup)
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
