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

template <typename T>
__global__ __launch_bounds__(256) void strided_reduce_kernel(int op, const T *x, size_t outer, size_t len, size_t inner,
                                                             T divisor, void *out) {
    const size_t id = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
    if (id >= outer * inner) return;
    const size_t o = id / inner, i = id % inner;
    const T *p = x + o * len * inner + i;
    switch (op) {
        case NRK_AXIS_SUM:
        case NRK_AXIS_MEAN:
        case NRK_AXIS_VAR: {
            T s = (T)0;
            for (size_t j = 0; j < len; ++j) s = rn_add(s, p[j * inner]);
            if (op == NRK_AXIS_SUM) { ((T *)out)[id] = s; break; }
            const T mean = rn_div(s, (T)len);
            if (op == NRK_AXIS_MEAN) { ((T *)out)[id] = mean; break; }
            T q = (T)0;
            for (size_t j = 0; j < len; ++j) {
                const T d = rn_sub(p[j * inner], mean);
                q = rn_add(q, rn_mul(d, d));
            }
            ((T *)out)[id] = rn_div(q, divisor);
            break;
        }
        case NRK_AXIS_MAX:
        case NRK_AXIS_MIN: {
            T acc = p[0];
            bool nan = __builtin_isnan(acc);
            for (size_t j = 1; j < len && !nan; ++j) {
                const T v = p[j * inner];
                if (__builtin_isnan(v)) { acc = v; nan = true; break; }
                if (op == NRK_AXIS_MAX ? (v > acc) : (v < acc)) acc = v;
            }
            ((T *)out)[id] = acc;
            break;
        }
        case NRK_AXIS_ARGMAX:
        case NRK_AXIS_ARGMIN: {
            T best = p[0];
            uint64_t bi = 0;
            for (size_t j = 1; j < len; ++j) {
                const T v = p[j * inner];
                if (op == NRK_AXIS_ARGMAX ? (v > best) : (v < best)) { best = v; bi = j; }
            }
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

// (outer, len, inner) -> (outer, inner, len): 64 x 64 tiles through LDS (a padded row: no bank
// conflicts), rows of the tile read and written 64 consecutive elements at a time. One workgroup
// per tile, the tiles numbered (o, i-tile, j-tile) over a 1-D grid.
constexpr int AXIS_TILE = 64;
template <typename T>
__global__ __launch_bounds__(256) void axis_transpose_kernel(const T *x, T *y, size_t len, size_t inner, size_t ntile_i,
                                                             size_t ntile_j) {
    __shared__ T tile[AXIS_TILE][AXIS_TILE + 1];
    size_t b = blockIdx.x;
    const size_t jt = b % ntile_j;
    b /= ntile_j;
    const size_t it = b % ntile_i, o = b / ntile_i;
    const size_t i0 = it * AXIS_TILE, j0 = jt * AXIS_TILE;
    const T *xb = x + o * len * inner;
    T *yb = y + o * inner * len;
    const unsigned tx = threadIdx.x & 63, ty = threadIdx.x >> 6;  // 64 x 4 threads
    for (unsigned r = ty; r < AXIS_TILE; r += 4) {
        const size_t j = j0 + r, i = i0 + tx;
        if (j < len && i < inner) tile[r][tx] = xb[j * inner + i];
    }
    __syncthreads();
    for (unsigned r = ty; r < AXIS_TILE; r += 4) {
        const size_t i = i0 + r, j = j0 + tx;
        if (i < inner && j < len) yb[i * len + j] = tile[tx][r];
    }
}

// The same with 16-byte vectors in and out, where rows of x (stride inner) and of y (stride len)
// start on the 16-byte grid: a tile row is 64 / VN vectors, 256 / (64 / VN) rows per pass.
template <typename T>
__global__ __launch_bounds__(256) void axis_transpose_vec_kernel(const T *x, T *y, size_t len, size_t inner, size_t ntile_i,
                                                                 size_t ntile_j) {
    typedef typename Vec16<T>::type V;
    constexpr int VN = Vec16<T>::N;
    constexpr unsigned VPR = AXIS_TILE / VN, RPP = 256 / VPR;
    __shared__ T tile[AXIS_TILE][AXIS_TILE + 1];
    size_t b = blockIdx.x;
    const size_t jt = b % ntile_j;
    b /= ntile_j;
    const size_t it = b % ntile_i, o = b / ntile_i;
    const size_t i0 = it * AXIS_TILE, j0 = jt * AXIS_TILE;
    const T *xb = x + o * len * inner;
    T *yb = y + o * inner * len;
    const unsigned cv = threadIdx.x % VPR, r0 = threadIdx.x / VPR;
    for (unsigned r = r0; r < AXIS_TILE; r += RPP) {
        const size_t j = j0 + r, i = i0 + cv * VN;
        if (j < len && i < inner) {  // inner % VN == 0: the whole vector is inside the row
            const V v = __builtin_nontemporal_load(reinterpret_cast<const V *>(xb + j * inner + i));
#pragma unroll
            for (int e = 0; e < VN; ++e) tile[r][cv * VN + e] = v[e];
        }
    }
    __syncthreads();
    for (unsigned r = r0; r < AXIS_TILE; r += RPP) {
        const size_t i = i0 + r, j = j0 + cv * VN;
        if (i < inner && j < len) {  // len % VN == 0
            V v;
#pragma unroll
            for (int e = 0; e < VN; ++e) v[e] = tile[cv * VN + e][r];
            *reinterpret_cast<V *>(yb + i * len + j) = v;
        }
    }
}

template <typename T>
nrk_status axis_transpose(const T *x, T *y, size_t outer, size_t len, size_t inner, hipStream_t stream) {
    constexpr size_t VN = Vec16<T>::N;
    const size_t ni = ceil_div(inner, (size_t)AXIS_TILE), nj = ceil_div(len, (size_t)AXIS_TILE);
    const size_t nb = outer * ni * nj;
    if (nb == 0) return NRK_OK;
    if (nb > 0x7fffffffu) return fail(NRK_ERR_INVALID_ARGUMENT, "(%zu, %zu, %zu) has too many tiles", outer, len, inner);
    const bool vec = (uintptr_t)x % 16 == 0 && (uintptr_t)y % 16 == 0 && inner % VN == 0 && len % VN == 0;
    if (vec)
        hipLaunchKernelGGL(axis_transpose_vec_kernel<T>, dim3((unsigned)nb), dim3(256), 0, stream, x, y, len, inner, ni, nj);
    else
        hipLaunchKernelGGL(axis_transpose_kernel<T>, dim3((unsigned)nb), dim3(256), 0, stream, x, y, len, inner, ni, nj);
    NRK_HIP(hipGetLastError());
    return NRK_OK;
}

// Sums and means along a strided axis with few outputs, reading the data once. Workgroup
// (o, column tile, range) walks rows [range start, + rlen) of 64 adjacent outputs in tiles of 64
// rows staged through LDS (16-byte loads where inner and x allow); thread t sums rows
// 16 (t / 64) .. + 15 of every tile for output column t % 64 (pairwise in T, then binary64; a
// double word for binary64 elements), the four row quarters are merged in LDS, and the range's
// partial goes to part[output][range] (contiguous per output, for the finish's lanes). The finish
// adds each output's partials: lane l takes ranges l, l + 64, ... in order, then the butterfly.
constexpr int AXS_C = 64, AXS_R = 64;
template <typename T, bool VEC>
__global__ __launch_bounds__(256) void axis_sum_tile_kernel(const T *x, size_t len, size_t inner, size_t nct, size_t nrange,
                                                            size_t rlen, size_t nout, double *phi, double *plo) {
    typedef typename Vec16<T>::type V;
    constexpr int VN = Vec16<T>::N;
    __shared__ T tile[AXS_R][AXS_C + 1];
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
        if constexpr (VEC) {  // thread t: vector t % 16 of rows t / 16 + 16 k
            constexpr unsigned VPR = AXS_C / VN, RPP = 256 / VPR;
            const unsigned cv = t % VPR;
            for (unsigned rr = t / VPR; rr < AXS_R; rr += RPP) {
                const size_t j = jt + rr, c = c0 + cv * VN;
                if (j < j1 && c < inner) {  // inner % VN == 0: the whole vector is in the row
                    const V v = __builtin_nontemporal_load(reinterpret_cast<const V *>(xb + j * inner + c));
#pragma unroll
                    for (int e = 0; e < VN; ++e) tile[rr][cv * VN + e] = v[e];
                } else {
#pragma unroll
                    for (int e = 0; e < VN; ++e) tile[rr][cv * VN + e] = (T)0;
                }
            }
        } else {  // thread t: column t % 64 of rows t / 64 + 4 k
            for (unsigned rr = t / AXS_C; rr < AXS_R; rr += 256 / AXS_C) {
                const size_t j = jt + rr, c = c0 + col;
                tile[rr][col] = (j < j1 && c < inner) ? xb[j * inner + c] : (T)0;
            }
        }
        __syncthreads();
        // 16 rows of this thread's quarter, pairwise (zeros past the range or the row are exact)
        const T st = pairwise<16, 0, T>([&](int i) { return tile[quarter * 16 + i][col]; });
        if constexpr (sizeof(T) == 8) {
            double s2, e2;
            two_sum(hi, (double)st, s2, e2);
            dw_fix(s2, rn_add(e2, lo), hi, lo);
        } else {
            hi += (double)st;
        }
        __syncthreads();
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
    for (size_t r = lane; r < nrange; r += 64) dw_add(h, l, phi[w * nrange + r], plo[w * nrange + r]);
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
    constexpr size_t VN = Vec16<T>::N;
    const size_t nout = outer * inner;
    if (nout == 0) return NRK_OK;
    const bool vec = (uintptr_t)x % 16 == 0 && inner % VN == 0;
    const size_t nct = ceil_div(inner, (size_t)AXS_C);
    // ranges: about 2048 workgroups in all, at least one tile of rows each; from the shape only
    const size_t want = std::max<size_t>(1, 2048 / std::max<size_t>(1, outer * nct));
    const size_t nrange = std::max<size_t>(1, std::min(want, ceil_div(len, (size_t)AXS_R)));
    const size_t rlen = ceil_div(ceil_div(len, nrange), (size_t)AXS_R) * AXS_R;
    const size_t nr = ceil_div(len, rlen);  // ranges that hold rows
    const size_t nb = outer * nct * nr;
    if (nb > 0x7fffffffu) return fail(NRK_ERR_INVALID_ARGUMENT, "(%zu, %zu, %zu) has too many tiles", outer, len, inner);
    void *ws = nullptr;
    NRK_TRY(get_workspace(stream, 2 * nout * nr * sizeof(double), &ws));
    double *phi = (double *)ws, *plo = phi + nout * nr;
    if (vec)
        hipLaunchKernelGGL((axis_sum_tile_kernel<T, true>), dim3((unsigned)nb), dim3(256), 0, stream, x, len, inner, nct, nr, rlen,
                           nout, phi, plo);
    else
        hipLaunchKernelGGL((axis_sum_tile_kernel<T, false>), dim3((unsigned)nb), dim3(256), 0, stream, x, len, inner, nct, nr, rlen,
                           nout, phi, plo);
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
template <typename T, bool VEC, bool MX>
__global__ __launch_bounds__(256) void axis_ext_tile_kernel(const T *x, size_t len, size_t inner, size_t nct, size_t nrange,
                                                            size_t rlen, T *pv, uint64_t *pi, unsigned *pf) {
    typedef typename Vec16<T>::type V;
    constexpr int VN = Vec16<T>::N;
    __shared__ T tile[AXS_R][AXS_C + 1];
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
        if constexpr (VEC) {
            constexpr unsigned VPR = AXS_C / VN, RPP = 256 / VPR;
            const unsigned cv = t % VPR;
            for (unsigned rr = t / VPR; rr < AXS_R; rr += RPP) {
                const size_t j = jt + rr, c = c0 + cv * VN;
                if (j < j1 && c < inner) {
                    const V v = __builtin_nontemporal_load(reinterpret_cast<const V *>(xb + j * inner + c));
#pragma unroll
                    for (int e = 0; e < VN; ++e) tile[rr][cv * VN + e] = v[e];
                }
            }
        } else {
            for (unsigned rr = t / AXS_C; rr < AXS_R; rr += 256 / AXS_C) {
                const size_t j = jt + rr, c = c0 + col;
                if (j < j1 && c < inner) tile[rr][col] = xb[j * inner + c];
            }
        }
        __syncthreads();
#pragma unroll
        for (int i = 0; i < 16; ++i) {
            const size_t j = jt + quarter * 16 + i;
            if (j < j1 && c0 + col < inner) {
                const T v = tile[quarter * 16 + i][col];
                if (v != v) {
                    flags |= j == 0 ? 3u : 1u;
                } else if (MX ? v > bv : v < bv) {
                    bv = v;
                    bj = j;
                }
            }
        }
        __syncthreads();
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
    for (size_t r = lane; r < nrange; r += 64) {
        take(pv[w * nrange + r], pi[w * nrange + r]);
        f |= pf[w * nrange + r];
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
    constexpr size_t VN = Vec16<T>::N;
    const size_t nout = outer * inner;
    if (nout == 0) return NRK_OK;
    const bool vec = (uintptr_t)x % 16 == 0 && inner % VN == 0;
    const bool mx = op == NRK_AXIS_MAX || op == NRK_AXIS_ARGMAX;
    const int want_index = (op == NRK_AXIS_ARGMAX || op == NRK_AXIS_ARGMIN) ? 1 : 0;
    const size_t nct = ceil_div(inner, (size_t)AXS_C);
    const size_t want = std::max<size_t>(1, 2048 / std::max<size_t>(1, outer * nct));
    const size_t nrange = std::max<size_t>(1, std::min(want, ceil_div(len, (size_t)AXS_R)));
    const size_t rlen = ceil_div(ceil_div(len, nrange), (size_t)AXS_R) * AXS_R;
    const size_t nr = ceil_div(len, rlen);
    const size_t nb = outer * nct * nr;
    if (nb > 0x7fffffffu) return fail(NRK_ERR_INVALID_ARGUMENT, "(%zu, %zu, %zu) has too many tiles", outer, len, inner);
    void *ws = nullptr;
    const size_t np = nout * nr;
    NRK_TRY(get_workspace(stream, np * (sizeof(T) + sizeof(uint64_t) + sizeof(unsigned)) + 64, &ws));
    uint64_t *pi = (uint64_t *)ws;
    T *pv = (T *)(pi + np);
    unsigned *pf = (unsigned *)(pv + np);
    const dim3 g((unsigned)nb), bl(256), gf((unsigned)ceil_div(nout * 64, (size_t)256));
#define NRK_EXT(V, M)                                                                                                   \
    do {                                                                                                                \
        hipLaunchKernelGGL((axis_ext_tile_kernel<T, V, M>), g, bl, 0, stream, x, len, inner, nct, nr, rlen, pv, pi, pf); \
        hipLaunchKernelGGL((axis_ext_finish_kernel<T, M>), gf, bl, 0, stream, pv, pi, pf, nout, nr, want_index, out);    \
    } while (0)
    if (vec && mx) NRK_EXT(true, true);
    else if (vec) NRK_EXT(true, false);
    else if (mx) NRK_EXT(false, true);
    else NRK_EXT(false, false);
#undef NRK_EXT
    NRK_HIP(hipGetLastError());
    return NRK_OK;
}

template nrk_status axis_extrema<float>(int, const float *, size_t, size_t, size_t, void *, hipStream_t);
template nrk_status axis_extrema<double>(int, const double *, size_t, size_t, size_t, void *, hipStream_t);
template nrk_status axis_sum<float>(const float *, size_t, size_t, size_t, bool, float *, hipStream_t);
template nrk_status axis_sum<double>(const double *, size_t, size_t, size_t, bool, double *, hipStream_t);
template nrk_status axis_transpose<float>(const float *, float *, size_t, size_t, size_t, hipStream_t);
template nrk_status axis_transpose<double>(const double *, double *, size_t, size_t, size_t, hipStream_t);
template nrk_status strided_reduce<float>(int, const float *, size_t, size_t, size_t, double, void *, hipStream_t);
template nrk_status strided_reduce<double>(int, const double *, size_t, size_t, size_t, double, void *, hipStream_t);

}  // namespace nrk
