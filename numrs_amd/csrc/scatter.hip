// ufunc_at (ufunc_ops.rs:772-838): a[indices[k], :] = a[indices[k], :] op b[k, :] for k = 0, 1, ...
// in order. Updates of different rows are independent, and those of one row follow k. A stable
// radix sort of the (index, k) pairs by index groups them; one thread per (group, column) then
// folds b's rows into a in ascending k, which is the reference's sequence of roundings exactly.
#include "ops.hpp"

#include <rocprim/device/device_radix_sort.hpp>

namespace nrk {
namespace {

constexpr int AT_B = 256;

// k_in[i] = i, and the sort keys: indices clamped to n0, so that an index past a's rows becomes
// the key n0 (sorted after every valid key, and skipped by the fold) instead of a value whose
// bits above end_bit would be ignored by the sort.
__global__ __launch_bounds__(AT_B) void at_prep_kernel(const uint64_t *indices, uint64_t *keys_in, uint32_t *k_in, size_t m,
                                                       uint64_t n0) {
    for (size_t i = blockIdx.x * (size_t)AT_B + threadIdx.x; i < m; i += (size_t)gridDim.x * AT_B) {
        const uint64_t v = indices[i];
        keys_in[i] = v < n0 ? v : n0;
        k_in[i] = (uint32_t)i;
    }
}

template <typename T, int OP>
__device__ __forceinline__ T at_apply(T a, T b) {
    if constexpr (OP == NRK_BOP_ADD) return rn_add(a, b);
    else if constexpr (OP == NRK_BOP_SUB) return rn_sub(a, b);
    else if constexpr (OP == NRK_BOP_MUL) return rn_mul(a, b);
    else if constexpr (OP == NRK_BOP_DIV) return rn_div(a, b);
    else if constexpr (OP == NRK_BOP_MAXIMUM) return (a != a || b != b) ? dev_nan<T>() : (a > b ? a : b);
    else return (a != a || b != b) ? dev_nan<T>() : (a < b ? a : b);
}

// Thread (p, col): if sorted position p starts a group (its key differs from p - 1), fold the
// group's b rows into a[key][col] in the sorted (ascending k) order.
template <typename T, int OP>
__global__ __launch_bounds__(AT_B) void at_fold_kernel(T *a, const T *b, size_t n0, size_t row_len, const uint64_t *keys,
                                                       const uint32_t *ks, size_t m) {
    const size_t total = m * row_len;
    for (size_t t = blockIdx.x * (size_t)AT_B + threadIdx.x; t < total; t += (size_t)gridDim.x * AT_B) {
        const size_t p = t / row_len, col = t - p * row_len;
        const uint64_t key = keys[p];
        if (key >= n0 || (p > 0 && keys[p - 1] == key)) continue;  // key n0: an index past a's rows, skipped
        T v = a[key * row_len + col];
        for (size_t q = p; q < m && keys[q] == key; ++q) v = at_apply<T, OP>(v, b[(size_t)ks[q] * row_len + col]);
        a[key * row_len + col] = v;
    }
}

}  // namespace

template <typename T>
nrk_status ufunc_at(int op, T *a, size_t n0, size_t row_len, const uint64_t *indices, size_t m, const T *b,
                    hipStream_t stream) {
    NRK_TRY(ensure_device());
    if (op != NRK_BOP_ADD && op != NRK_BOP_SUB && op != NRK_BOP_MUL && op != NRK_BOP_DIV && op != NRK_BOP_MAXIMUM &&
        op != NRK_BOP_MINIMUM)
        return fail(NRK_ERR_INVALID_ARGUMENT, "at: op %d is not ADD, SUB, MUL, DIV, MAXIMUM or MINIMUM", op);
    if (m == 0 || row_len == 0 || n0 == 0) return NRK_OK;
    if (!a || !b || !indices) return fail(NRK_ERR_INVALID_ARGUMENT, "a, indices or b is NULL");
    if (m >= ((size_t)1 << 32)) return fail(NRK_ERR_INVALID_ARGUMENT, "at: more than 2^32 - 1 indices");
    // the keys are clamped to n0, so the sort needs the bits of n0 itself: n0 < 2^end_bit
    int end_bit = 1;
    while (end_bit < 64 && ((uint64_t)1 << end_bit) <= (uint64_t)n0) ++end_bit;
    // workspace: clamped and sorted keys, k in and out, then rocPRIM's temporary storage
    const size_t kb = align_up(m * sizeof(uint64_t), 256), vb = align_up(m * sizeof(uint32_t), 256);
    size_t temp_bytes = 0;
    if (rocprim::radix_sort_pairs(nullptr, temp_bytes, (const uint64_t *)nullptr, (uint64_t *)nullptr, (const uint32_t *)nullptr,
                                  (uint32_t *)nullptr, m, 0, end_bit, stream) != hipSuccess)
        return fail(NRK_ERR_HIP, "at: radix sort size query failed");
    void *ws = nullptr;
    NRK_TRY(get_workspace(stream, 2 * kb + 2 * vb + temp_bytes + 256, &ws));
    char *w = (char *)ws;
    uint64_t *keys_in = (uint64_t *)w;
    uint64_t *keys = (uint64_t *)(w + kb);
    uint32_t *k_in = (uint32_t *)(w + 2 * kb);
    uint32_t *k_out = (uint32_t *)(w + 2 * kb + vb);
    void *temp = w + 2 * kb + 2 * vb;
    const int g1 = (int)std::min<size_t>(ceil_div(m, (size_t)AT_B), (size_t)cu_count() * 8);
    hipLaunchKernelGGL(at_prep_kernel, dim3(g1), dim3(AT_B), 0, stream, indices, keys_in, k_in, m, (uint64_t)n0);
    NRK_HIP(hipGetLastError());
    NRK_HIP(rocprim::radix_sort_pairs(temp, temp_bytes, (const uint64_t *)keys_in, keys, (const uint32_t *)k_in, k_out, m, 0,
                                      end_bit, stream));
    const int g2 = (int)std::min<size_t>(ceil_div(m * row_len, (size_t)AT_B), (size_t)cu_count() * 16);
    switch (op) {
#define NRK_AT(OPC) \
    case OPC: hipLaunchKernelGGL((at_fold_kernel<T, OPC>), dim3(g2), dim3(AT_B), 0, stream, a, b, n0, row_len, (const uint64_t *)keys, (const uint32_t *)k_out, m); break;
        NRK_AT(NRK_BOP_ADD) NRK_AT(NRK_BOP_SUB) NRK_AT(NRK_BOP_MUL) NRK_AT(NRK_BOP_DIV) NRK_AT(NRK_BOP_MAXIMUM) NRK_AT(NRK_BOP_MINIMUM)
#undef NRK_AT
    }
    NRK_HIP(hipGetLastError());
    return NRK_OK;
}
template nrk_status ufunc_at<float>(int, float *, size_t, size_t, const uint64_t *, size_t, const float *, hipStream_t);
template nrk_status ufunc_at<double>(int, double *, size_t, size_t, const uint64_t *, size_t, const double *, hipStream_t);

}  // namespace nrk
