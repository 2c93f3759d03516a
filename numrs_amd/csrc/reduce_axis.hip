// Reductions along a strided axis: the (outer, len, inner) view with inner > 1, e.g. axis 0 of
// a 2-D array. One thread owns one output and walks the axis left to right, so the fold order is
// exactly the reference's (math/aggregation.rs:548-561 sum; :244-261 max/min with the NaN
// short-circuit; math/statistics/mod.rs:1064-1071 mean = left-to-right sum / len;
// :1403-1494 var = mean(axis), then left-to-right sum of (x - mean)^2, / (len - ddof);
// :164-175 argmax, first strict maximum). Adjacent threads own adjacent outputs, so every step of
// the walk is one coalesced row read across the wavefront.
#include "ops.hpp"
#include "reduce_common.hpp"

namespace nrk {

// f(j, x[j]) for j = j0 .. len - 1 in order, the loads U at a time ahead of the fold (a strided
// column: one dependent load per element would leave the wavefront waiting on each).
template <int U, typename T, typename F>
__device__ __forceinline__ void strided_fold(const T *p, size_t inner, size_t j0, size_t len, F &&f) {
    size_t j = j0;
    for (; j + U <= len; j += U) {
        T v[U];
#pragma unroll
        for (int u = 0; u < U; ++u) v[u] = p[(j + u) * inner];
#pragma unroll
        for (int u = 0; u < U; ++u) f(j + u, v[u]);
    }
    for (; j < len; ++j) f(j, p[j * inner]);
}

template <typename T>
__global__ __launch_bounds__(256) void strided_reduce_kernel(int op, const T *x, size_t outer, size_t len, size_t inner,
                                                             T divisor, void *out) {
    // loads ahead of the fold: binary32 8 (64 x 2^20 sums 65 -> 51 us), binary64 2 (argmin 97 -> 89
    // us; 4 ahead: sums 88 -> 110 us, 8 ahead: 106 us)
    constexpr int U = sizeof(T) == 4 ? 8 : 2;
    const size_t id = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
    if (id >= outer * inner) return;
    const size_t o = id / inner, i = id % inner;
    const T *p = x + o * len * inner + i;
    switch (op) {
        case NRK_AXIS_SUM:
        case NRK_AXIS_MEAN:
        case NRK_AXIS_VAR: {
            T s = (T)0;
            strided_fold<U>(p, inner, 0, len, [&](size_t, T v) { s = rn_add(s, v); });
            if (op == NRK_AXIS_SUM) { ((T *)out)[id] = s; break; }
            const T mean = rn_div(s, (T)len);
            if (op == NRK_AXIS_MEAN) { ((T *)out)[id] = mean; break; }
            T q = (T)0;
            strided_fold<U>(p, inner, 0, len, [&](size_t, T v) {
                const T d = rn_sub(v, mean);
                q = rn_add(q, rn_mul(d, d));
            });
            ((T *)out)[id] = rn_div(q, divisor);
            break;
        }
        case NRK_AXIS_MAX:
        case NRK_AXIS_MIN: {
            // the first NaN is the result (the reference stops there); later elements change nothing
            T acc = p[0];
            bool nan = __builtin_isnan(acc);
            const bool mx = op == NRK_AXIS_MAX;
            strided_fold<U>(p, inner, 1, len, [&](size_t, T v) {
                if (nan) return;
                if (__builtin_isnan(v)) {
                    acc = v;
                    nan = true;
                } else if (mx ? (v > acc) : (v < acc)) {
                    acc = v;
                }
            });
            ((T *)out)[id] = acc;
            break;
        }
        case NRK_AXIS_ARGMAX:
        case NRK_AXIS_ARGMIN: {
            T best = p[0];
            uint64_t bi = 0;
            const bool mx = op == NRK_AXIS_ARGMAX;
            strided_fold<U>(p, inner, 1, len, [&](size_t j, T v) {
                if (mx ? (v > best) : (v < best)) {
                    best = v;
                    bi = j;
                }
            });
            ((uint64_t *)out)[id] = bi;
            break;
        }
        default: break;
    }
}

template <typename T>
nrk_status strided_reduce(int op, const T *x, size_t outer, size_t len, size_t inner, double ddof, void *out,
                          hipStream_t stream) {
    NRK_TRY(ensure_device());
    const size_t nout = outer * inner;
    if (nout == 0) return NRK_OK;
    if (!x || !out) return fail(NRK_ERR_INVALID_ARGUMENT, "x or out is NULL");
    const T divisor = (T)((double)len - ddof);
    hipLaunchKernelGGL(strided_reduce_kernel<T>, dim3((unsigned)ceil_div(nout, 256)), dim3(256), 0, stream, op, x, outer,
                       len, inner, divisor, out);
    NRK_HIP(hipGetLastError());
    return NRK_OK;
}

// Sums and means along a strided axis with few outputs, reading the data once. Workgroup
// (o, column tile, range) walks rows [range start, + rlen) of 64 adjacent outputs in tiles of 64
// rows; thread t sums rows 16 (t / 64) .. + 15 of every tile for output column t % 64 (pairwise in
// T, then binary64; a double word for binary64 elements), loading them straight from HBM (the
// wavefront's lanes read one row's 64 adjacent elements per load, no LDS tile, no barrier). The
// four row quarters are merged in LDS, and the range's partial goes to part[output][range]
// (contiguous per output, for the finish's lanes). The finish adds each output's partials: lane l
// takes ranges l, l + 64, ... in order, then the butterfly.
constexpr int AXS_C = 64, AXS_R = 64;

// Ranges of the axis for the column-tile kernels: about `groups` workgroups in all, at least one
// 64-row tile each, from the shape only (so the grouping of a sum is the same on every call).
// Fewer, longer ranges win while the chip stays full: per-op targets measured on 8192 x 8192,
// 2^20 x 64, 4096 x 1024 (x 16) and 100000 x 671 (sums 512; extrema 1024; variances 1024 binary32,
// 512 binary64; 2048 cost 5-20 %, 256 up to 2x).
struct AxisRanges {
    size_t nct, rlen, nr, nb;
};
static AxisRanges axis_ranges(size_t outer, size_t len, size_t inner, size_t groups) {
    AxisRanges a;
    a.nct = ceil_div(inner, (size_t)AXS_C);
    const size_t want = std::max<size_t>(1, groups / std::max<size_t>(1, outer * a.nct));
    const size_t nrange = std::max<size_t>(1, std::min(want, ceil_div(len, (size_t)AXS_R)));
    a.rlen = ceil_div(ceil_div(len, nrange), (size_t)AXS_R) * AXS_R;
    a.nr = ceil_div(len, a.rlen);  // ranges that hold rows
    a.nb = outer * a.nct * a.nr;
    return a;
}

// Rows q0 .. q0 + 15 of column c, zeros past j1 or past the row, straight from HBM: the lanes of a
// wavefront take 64 adjacent columns, so each load is one row's 64 contiguous elements, and the 16
// loads are in flight together.
template <typename T>
__device__ __forceinline__ void axis_rows16(const T *xb, size_t q0, size_t j1, size_t c, size_t inner, T (&v)[16]) {
    const bool in_row = c < inner;
#pragma unroll
    for (int i = 0; i < 16; ++i) {
        const size_t j = q0 + i;
        // binary32 plain loads (8192 x 8192 sums 51.7 -> 49.6 us), binary64 non-temporal ones
        // (plain: 82.9 -> 92.3 us)
        if constexpr (sizeof(T) == 4)
            v[i] = (in_row && j < j1) ? xb[j * inner + c] : (T)0;
        else
            v[i] = (in_row && j < j1) ? __builtin_nontemporal_load(xb + j * inner + c) : (T)0;
    }
}

template <typename T>
__global__ __launch_bounds__(256) void axis_sum_tile_kernel(const T *x, size_t len, size_t inner, size_t nct, size_t nrange,
                                                            size_t rlen, size_t nout, double *phi, double *plo) {
    __shared__ double qhi[4][AXS_C], qlo[4][AXS_C];
    size_t b = blockIdx.x;
    const size_t r = b % nrange;
    b /= nrange;
    const size_t ct = b % nct, o = b / nct;
    const size_t c0 = ct * AXS_C, j0 = r * rlen, j1 = min(len, j0 + rlen);
    const T *xb = x + o * len * inner;
    const unsigned t = threadIdx.x, col = t % AXS_C, quarter = t / AXS_C;
    double hi = 0.0, lo = 0.0;
    for (size_t jt = j0; jt < j1; jt += AXS_R) {
        T v[16];
        axis_rows16(xb, jt + quarter * 16, j1, c0 + col, inner, v);
        // 16 rows of this thread's quarter, pairwise (zeros past the range or the row are exact)
        const T st = pairwise<16, 0, T>([&](int i) { return v[i]; });
        if constexpr (sizeof(T) == 8) {
            double s2, e2;
            two_sum(hi, (double)st, s2, e2);
            dw_fix(s2, rn_add(e2, lo), hi, lo);
        } else {
            hi += (double)st;
        }
    }
    qhi[quarter][col] = hi;
    qlo[quarter][col] = lo;
    __syncthreads();
    if (t < AXS_C && c0 + t < inner) {
        double h = qhi[0][t], l = qlo[0][t];
        for (int q = 1; q < 4; ++q) dw_add(h, l, qhi[q][t], qlo[q][t]);
        const size_t out = o * inner + c0 + t;
        phi[out * nrange + r] = h;
        plo[out * nrange + r] = l;
    }
}

// One wavefront per output: lane l adds partials l, l + 64, ... in order, then the butterfly.
template <typename T>
__global__ __launch_bounds__(256) void axis_sum_finish_kernel(const double *phi, const double *plo, size_t nout, size_t nrange,
                                                              double len, int mean, T *out) {
    const size_t w = (blockIdx.x * (size_t)blockDim.x + threadIdx.x) >> 6;
    const unsigned lane = threadIdx.x & 63;
    if (w >= nout) return;  // whole wavefronts
    double h = 0.0, l = 0.0;
    for (size_t r0 = lane; r0 < nrange; r0 += 4 * 64) {  // four loads in flight, added in order
        double bh[4], bl[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            const size_t r = min(r0 + 64 * u, nrange - 1);
            bh[u] = phi[w * nrange + r];
            bl[u] = plo[w * nrange + r];
        }
#pragma unroll
        for (int u = 0; u < 4; ++u)
            if (r0 + 64 * u < nrange) dw_add(h, l, bh[u], bl[u]);
    }
#pragma unroll
    for (int m = 32; m >= 1; m >>= 1) {
        const double bh = __shfl_xor(h, m, 64), bl = __shfl_xor(l, m, 64);
        dw_add(h, l, bh, bl);
    }
    if (lane == 0) {
        const double sum = h + l;
        out[w] = (T)(mean ? sum / len : sum);
    }
}

template <typename T>
nrk_status axis_sum(const T *x, size_t outer, size_t len, size_t inner, bool mean, T *out, hipStream_t stream) {
    const size_t nout = outer * inner;
    if (nout == 0) return NRK_OK;
    const AxisRanges ar = axis_ranges(outer, len, inner, 512);
    const size_t nct = ar.nct, rlen = ar.rlen, nr = ar.nr, nb = ar.nb;
    if (nb > 0x7fffffffu) return fail(NRK_ERR_INVALID_ARGUMENT, "(%zu, %zu, %zu) has too many tiles", outer, len, inner);
    void *ws = nullptr;
    NRK_TRY(get_workspace(stream, 2 * nout * nr * sizeof(double), &ws));
    double *phi = (double *)ws, *plo = phi + nout * nr;
    hipLaunchKernelGGL(axis_sum_tile_kernel<T>, dim3((unsigned)nb), dim3(256), 0, stream, x, len, inner, nct, nr, rlen, nout, phi,
                       plo);
    NRK_HIP(hipGetLastError());
    hipLaunchKernelGGL(axis_sum_finish_kernel<T>, dim3((unsigned)ceil_div(nout * 64, (size_t)256)), dim3(256), 0, stream, phi, plo,
                       nout, nr, (double)len, mean ? 1 : 0, out);
    NRK_HIP(hipGetLastError());
    return NRK_OK;
}

// Extrema along a strided axis with few outputs, the same tiles and ranges: thread t keeps the
// strict winner (value, row) of its column over its quarter's rows in row order, NaN elements
// never winning but flagged (and whether row 0 is NaN); quarters and ranges merge with the lower
// row winning ties, so the result is the sequential fold's: max / min NaN where any element is,
// argmax / argmin 0 where the first element is NaN or no element wins (all -inf / +inf or NaN).
template <typename T, bool MX>
__global__ __launch_bounds__(256) void axis_ext_tile_kernel(const T *x, size_t len, size_t inner, size_t nct, size_t nrange,
                                                            size_t rlen, T *pv, uint64_t *pi, unsigned *pf) {
    __shared__ T qv[4][AXS_C];
    __shared__ uint64_t qi[4][AXS_C];
    __shared__ unsigned qf[4][AXS_C];
    size_t b = blockIdx.x;
    const size_t r = b % nrange;
    b /= nrange;
    const size_t ct = b % nct, o = b / nct;
    const size_t c0 = ct * AXS_C, j0 = r * rlen, j1 = min(len, j0 + rlen);
    const T *xb = x + o * len * inner;
    const unsigned t = threadIdx.x, col = t % AXS_C, quarter = t / AXS_C;
    T bv = MX ? -dev_inf<T>() : dev_inf<T>();
    uint64_t bj = NO_INDEX;
    unsigned flags = 0;  // bit 0: a NaN; bit 1: row 0 is NaN
    for (size_t jt = j0; jt < j1; jt += AXS_R) {
        T vv[16];
        axis_rows16(xb, jt + quarter * 16, j1, c0 + col, inner, vv);
#pragma unroll
        for (int i = 0; i < 16; ++i) {
            const size_t j = jt + quarter * 16 + i;
            if (j < j1 && c0 + col < inner) {
                const T v = vv[i];
                if (v != v) {
                    flags |= j == 0 ? 3u : 1u;
                } else if (MX ? v > bv : v < bv) {
                    bv = v;
                    bj = j;
                }
            }
        }
    }
    qv[quarter][col] = bv;
    qi[quarter][col] = bj;
    qf[quarter][col] = flags;
    __syncthreads();
    if (t < AXS_C && c0 + t < inner) {
        T v = qv[0][t];
        uint64_t ix = qi[0][t];
        unsigned f = qf[0][t];
        for (int q = 1; q < 4; ++q) {
            const T w = qv[q][t];
            const uint64_t wi = qi[q][t];
            f |= qf[q][t];
            if (wi != NO_INDEX && (ix == NO_INDEX || (MX ? w > v : w < v) || (w == v && wi < ix))) {
                v = w;
                ix = wi;
            }
        }
        const size_t out = o * inner + c0 + t;
        pv[out * nrange + r] = v;
        pi[out * nrange + r] = ix;
        pf[out * nrange + r] = f;
    }
}

template <typename T, bool MX>
__global__ __launch_bounds__(256) void axis_ext_finish_kernel(const T *pv, const uint64_t *pi, const unsigned *pf, size_t nout,
                                                              size_t nrange, int want_index, void *out) {
    const size_t w = (blockIdx.x * (size_t)blockDim.x + threadIdx.x) >> 6;
    const unsigned lane = threadIdx.x & 63;
    if (w >= nout) return;  // whole wavefronts
    T v = MX ? -dev_inf<T>() : dev_inf<T>();
    uint64_t ix = NO_INDEX;
    unsigned f = 0;
    auto take = [&](T bvv, uint64_t bi) {
        if (bi != NO_INDEX && (ix == NO_INDEX || (MX ? bvv > v : bvv < v) || (bvv == v && bi < ix))) {
            v = bvv;
            ix = bi;
        }
    };
    for (size_t r0 = lane; r0 < nrange; r0 += 4 * 64) {  // four loads in flight, merged in order
        T bv[4];
        uint64_t bi[4];
        unsigned bf[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            const size_t r = min(r0 + 64 * u, nrange - 1);
            bv[u] = pv[w * nrange + r];
            bi[u] = pi[w * nrange + r];
            bf[u] = pf[w * nrange + r];
        }
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            if (r0 + 64 * u < nrange) {
                take(bv[u], bi[u]);
                f |= bf[u];
            }
        }
    }
#pragma unroll
    for (int m = 32; m >= 1; m >>= 1) {
        const T bvv = __shfl_xor(v, m, 64);
        const unsigned lo = __shfl_xor((unsigned)(ix & 0xffffffffu), m, 64), hi = __shfl_xor((unsigned)(ix >> 32), m, 64);
        take(bvv, ((uint64_t)hi << 32) | lo);
        f |= (unsigned)__shfl_xor((int)f, m, 64);
    }
    if (lane == 0) {
        if (want_index)
            ((uint64_t *)out)[w] = (f & 2u) ? 0 : (ix == NO_INDEX ? 0 : ix);
        else
            ((T *)out)[w] = (f & 1u) ? dev_nan<T>() : v;
    }
}

template <typename T>
nrk_status axis_extrema(int op, const T *x, size_t outer, size_t len, size_t inner, void *out, hipStream_t stream) {
    const size_t nout = outer * inner;
    if (nout == 0) return NRK_OK;
    const bool mx = op == NRK_AXIS_MAX || op == NRK_AXIS_ARGMAX;
    const int want_index = (op == NRK_AXIS_ARGMAX || op == NRK_AXIS_ARGMIN) ? 1 : 0;
    const AxisRanges ar = axis_ranges(outer, len, inner, 1024);
    const size_t nct = ar.nct, rlen = ar.rlen, nr = ar.nr, nb = ar.nb;
    if (nb > 0x7fffffffu) return fail(NRK_ERR_INVALID_ARGUMENT, "(%zu, %zu, %zu) has too many tiles", outer, len, inner);
    void *ws = nullptr;
    const size_t np = nout * nr;
    NRK_TRY(get_workspace(stream, np * (sizeof(T) + sizeof(uint64_t) + sizeof(unsigned)) + 64, &ws));
    uint64_t *pi = (uint64_t *)ws;
    T *pv = (T *)(pi + np);
    unsigned *pf = (unsigned *)(pv + np);
    const dim3 g((unsigned)nb), bl(256), gf((unsigned)ceil_div(nout * 64, (size_t)256));
    if (mx) {
        hipLaunchKernelGGL((axis_ext_tile_kernel<T, true>), g, bl, 0, stream, x, len, inner, nct, nr, rlen, pv, pi, pf);
        hipLaunchKernelGGL((axis_ext_finish_kernel<T, true>), gf, bl, 0, stream, pv, pi, pf, nout, nr, want_index, out);
    } else {
        hipLaunchKernelGGL((axis_ext_tile_kernel<T, false>), g, bl, 0, stream, x, len, inner, nct, nr, rlen, pv, pi, pf);
        hipLaunchKernelGGL((axis_ext_finish_kernel<T, false>), gf, bl, 0, stream, pv, pi, pf, nout, nr, want_index, out);
    }
    NRK_HIP(hipGetLastError());
    return NRK_OK;
}

// Variances along a strided axis with few outputs, the same tiles and ranges: thread t turns its
// quarter's 16 rows of a tile into a record (count, pairwise sum, M2 about the block's mean, in
// binary64) and merges it into its running record (Chan, Golub & LeVeque, double-word sums); the
// quarters, the ranges (lane-strided per output) and the finish's lanes merge the same way.
template <typename T>
__global__ __launch_bounds__(256) void axis_var_tile_kernel(const T *x, size_t len, size_t inner, size_t nct, size_t nrange,
                                                            size_t rlen, double *pn, double *phi, double *plo, double *pm2) {
    __shared__ double qn[4][AXS_C], qhi[4][AXS_C], qlo[4][AXS_C], qm2[4][AXS_C];
    size_t b = blockIdx.x;
    const size_t r = b % nrange;
    b /= nrange;
    const size_t ct = b % nct, o = b / nct;
    const size_t c0 = ct * AXS_C, j0 = r * rlen, j1 = min(len, j0 + rlen);
    const T *xb = x + o * len * inner;
    const unsigned t = threadIdx.x, col = t % AXS_C, quarter = t / AXS_C;
    Part a = part_empty();
    for (size_t jt = j0; jt < j1; jt += AXS_R) {
        T v[16];
        axis_rows16(xb, jt + quarter * 16, j1, c0 + col, inner, v);
        const size_t q0 = jt + quarter * 16;
        const int m = q0 >= j1 ? 0 : (int)min((size_t)16, j1 - q0);
        if (m > 0) {
            Part bk = part_empty();
            const T st = pairwise<16, 0, T>([&](int i) { return v[i]; });
            const double c = (double)st / m;
            double q = 0.0;
#pragma unroll
            for (int i = 0; i < 16; ++i) {
                const double d = (double)v[i] - c;
                q = i < m ? fma(d, d, q) : q;
            }
            bk.n = m;
            bk.hi = st;
            bk.m2 = q;
            part_merge<F_M2>(a, bk);
        }
    }
    qn[quarter][col] = a.n;
    qhi[quarter][col] = a.hi;
    qlo[quarter][col] = a.lo;
    qm2[quarter][col] = a.m2;
    __syncthreads();
    if (t < AXS_C && c0 + t < inner) {
        Part v = part_empty();
        for (int q = 0; q < 4; ++q) {
            Part w = part_empty();
            w.n = qn[q][t];
            w.hi = qhi[q][t];
            w.lo = qlo[q][t];
            w.m2 = qm2[q][t];
            part_merge<F_M2>(v, w);
        }
        const size_t at = (o * inner + c0 + t) * nrange + r;
        pn[at] = v.n;
        phi[at] = v.hi;
        plo[at] = v.lo;
        pm2[at] = v.m2;
    }
}

template <typename T>
__global__ __launch_bounds__(256) void axis_var_finish_kernel(const double *pn, const double *phi, const double *plo,
                                                              const double *pm2, size_t nout, size_t nrange, double ddof, T *out) {
    const size_t w = (blockIdx.x * (size_t)blockDim.x + threadIdx.x) >> 6;
    const unsigned lane = threadIdx.x & 63;
    if (w >= nout) return;  // whole wavefronts
    Part a = part_empty();
    for (size_t r0 = lane; r0 < nrange; r0 += 4 * 64) {  // four records' loads in flight, merged in order
        Part b[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            const size_t r = min(r0 + 64 * u, nrange - 1);
            b[u] = part_empty();
            b[u].n = pn[w * nrange + r];
            b[u].hi = phi[w * nrange + r];
            b[u].lo = plo[w * nrange + r];
            b[u].m2 = pm2[w * nrange + r];
        }
#pragma unroll
        for (int u = 0; u < 4; ++u)
            if (r0 + 64 * u < nrange) part_merge<F_M2>(a, b[u]);
    }
    wave_merge<F_M2>(a);
    if (lane == 0) out[w] = (T)(a.m2 / (a.n - ddof));
}

template <typename T>
nrk_status axis_var(const T *x, size_t outer, size_t len, size_t inner, double ddof, T *out, hipStream_t stream) {
    const size_t nout = outer * inner;
    if (nout == 0) return NRK_OK;
    const AxisRanges ar = axis_ranges(outer, len, inner, sizeof(T) == 8 ? 512 : 1024);
    const size_t nct = ar.nct, rlen = ar.rlen, nr = ar.nr, nb = ar.nb;
    if (nb > 0x7fffffffu) return fail(NRK_ERR_INVALID_ARGUMENT, "(%zu, %zu, %zu) has too many tiles", outer, len, inner);
    void *ws = nullptr;
    const size_t np = nout * nr;
    NRK_TRY(get_workspace(stream, 4 * np * sizeof(double), &ws));
    double *pn = (double *)ws, *phi = pn + np, *plo = phi + np, *pm2 = plo + np;
    hipLaunchKernelGGL(axis_var_tile_kernel<T>, dim3((unsigned)nb), dim3(256), 0, stream, x, len, inner, nct, nr, rlen, pn, phi,
                       plo, pm2);
    NRK_HIP(hipGetLastError());
    hipLaunchKernelGGL(axis_var_finish_kernel<T>, dim3((unsigned)ceil_div(nout * 64, (size_t)256)), dim3(256), 0, stream, pn, phi,
                       plo, pm2, nout, nr, ddof, out);
    NRK_HIP(hipGetLastError());
    return NRK_OK;
}

template nrk_status axis_var<float>(const float *, size_t, size_t, size_t, double, float *, hipStream_t);
template nrk_status axis_var<double>(const double *, size_t, size_t, size_t, double, double *, hipStream_t);
template nrk_status axis_extrema<float>(int, const float *, size_t, size_t, size_t, void *, hipStream_t);
template nrk_status axis_extrema<double>(int, const double *, size_t, size_t, size_t, void *, hipStream_t);
template nrk_status axis_sum<float>(const float *, size_t, size_t, size_t, bool, float *, hipStream_t);
template nrk_status axis_sum<double>(const double *, size_t, size_t, size_t, bool, double *, hipStream_t);
template nrk_status strided_reduce<float>(int, const float *, size_t, size_t, size_t, double, void *, hipStream_t);
template nrk_status strided_reduce<double>(int, const double *, size_t, size_t, size_t, double, void *, hipStream_t);

}  // namespace nrk
