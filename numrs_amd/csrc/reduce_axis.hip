// Reductions along a strided axis: the (outer, len, inner) view with inner > 1, e.g. axis 0 of
// a 2-D array. One thread owns one output and walks the axis left to right, so the fold order is
// exactly the reference's (math/aggregation.rs:548-561 sum; :244-261 max/min with the NaN
// short-circuit; math/statistics/mod.rs:1064-1071 mean = left-to-right sum / len;
// :1403-1494 var = mean(axis), then left-to-right sum of (x - mean)^2, / (len - ddof);
// :164-175 argmax, first strict maximum). Adjacent threads own adjacent outputs, so every step of
// the walk is one coalesced row read across the wavefront.
#include "ops.hpp"

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

template nrk_status axis_transpose<float>(const float *, float *, size_t, size_t, size_t, hipStream_t);
template nrk_status axis_transpose<double>(const double *, double *, size_t, size_t, size_t, hipStream_t);
template nrk_status strided_reduce<float>(int, const float *, size_t, size_t, size_t, double, void *, hipStream_t);
template nrk_status strided_reduce<double>(int, const double *, size_t, size_t, size_t, double, void *, hipStream_t);

}  // namespace nrk
