// Ragged batches in ONE persistent launch: many independent bf16 GEMMs (fp32 accumulate and
// output) plus many ragged f32 row sums, scheduled as a queue of work items.
//
// Reference: array/linalg.rs:56-136 (batched_gemm / par_batch_f64: uniform panels, one task per
// panel). This generalises it to ragged shapes and adds row reductions to the same launch.
//
// Plan (host, once per batch shape): every GEMM is cut into BM x BN output tiles; every row sum
// becomes either a member of a "short-row group" (rows of <= SHORT_MAX elements packed until the
// group holds ~GROUP_ELEMS elements; each wavefront of the workgroup sums whole rows) or, for long
// rows, one workgroup per row (its 4 wavefronts take fixed quarters, combined in wave order, so
// nothing crosses workgroups and results are deterministic). Items are sorted by
// estimated cost, largest first (LPT), using the cost model the caller provides (or the default
// flop/byte estimate), and uploaded with the GEMM descriptors.
//
// Run: a persistent grid (CUs x blocks/CU) deals the ordered items in a fixed snake order
// (no shared queue state), so a plan can be replayed -- e.g. from a captured hipGraph -- with no
// host work and no memset nodes.
//
// GEMM tile: 64 x 64 outputs, K-step 32, 4 wavefronts of 32 x 32 (2 x 2 v_mfma_f32_16x16x32_bf16).
// Measured alternatives on that batch, all slower: 32 x 32 tiles where the tile selector picks
// them for a standalone launch (0.38 -> 0.75 ms for the GEMMs: 3.3x the items), 64-deep K steps
// (-> 1.2 ms), and the next K step's loads issued before this step's products (-> 0.83 ms).
// A (m x k) row-major is staged K-contiguous; B (k x n) row-major is transposed into LDS as
// Bt[n][k] so that each lane's MFMA operand is 8 contiguous bf16 along K.
#include <algorithm>
#include <cmath>
#include <numeric>
#include <vector>

#include "gemm_lowp.hpp"
#include "ops.hpp"

namespace nrk {

constexpr int GB = 256;          // threads per workgroup
constexpr int GT_M = 64, GT_N = 64, GT_K = 32;
constexpr int G_LDK = GT_K + 8;  // padded LDS row (bf16)
constexpr int SHORT_MAX = 2048;  // rows up to this length are summed whole by one wavefront
constexpr int GROUP_ELEMS = 8192;

enum : int { ITEM_GEMM = 0, ITEM_ROWS = 1, ITEM_LONGROW = 2 };

struct Item {
    int kind;
    int a;  // GEMM: problem index   ROWS: first row   LONGROW: row
    int b;  // GEMM: tile index      ROWS: end row
    int c;
};

struct GroupedArgs {
    const nrk_gemm_desc *descs;
    const Item *items;
    int n_items;
    const float *rows_x;
    const int64_t *rows_off;
    float *rows_out;
};

// ---------------- GEMM tile ----------------
__device__ __forceinline__ unsigned short ld_bf16(const unsigned short *p) { return *p; }

// 8 consecutive bf16 at src (2-byte aligned, all 8 inside the operand) as packed pairs, for
// operands whose rows do not start on 16-byte boundaries (any K or N not a multiple of 8, or a
// base that is not 16-byte aligned): the aligned 16-byte block holding src and, when src is not
// 16-byte aligned, the next one (it holds the chunk's last byte), shifted into place. Both blocks
// contain operand bytes, so they lie in the operand's own allocation (hipMalloc granularity is far
// above 16 bytes). Two 16-byte loads instead of eight 2-byte ones: measured on the 4096-GEMM
// batch (m, n, k log-uniform in [16, 512], operands packed back to back) 0.56 -> 0.38 ms.
__device__ __forceinline__ u32x4_t chunk8(const unsigned short *src) {
    const uintptr_t mis = reinterpret_cast<uintptr_t>(src) & 15u;
    const u32x4_t *q = reinterpret_cast<const u32x4_t *>(reinterpret_cast<const char *>(src) - mis);
    const u32x4_t lo = q[0];
    const u32x4_t hi = mis ? q[1] : u32x4_t{0u, 0u, 0u, 0u};
    const unsigned w[8] = {lo[0], lo[1], lo[2], lo[3], hi[0], hi[1], hi[2], hi[3]};
    const int d = (int)(mis >> 2), sh = (int)(mis & 3u) * 8;
    u32x4_t r;
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        unsigned a = w[i], b = w[i + 1];
#pragma unroll
        for (int t = 1; t < 4; ++t)
            if (d == t) { a = w[i + t]; b = w[i + t + 1 < 8 ? i + t + 1 : 7]; }
        r[i] = sh ? (a >> sh) | (b << (32 - sh)) : a;
    }
    return r;
}

// One BM x BN output tile of an NN bf16 problem, 4 wavefronts in a 2 x 2 arrangement, K-step 32.
template <int BM, int BN>
__device__ void gemm_tile_t(const nrk_gemm_desc &d, int tile, unsigned short *As, unsigned short *Bs) {
    constexpr int WM = BM / 2, WN = BN / 2, TM = WM / 16, TN = WN / 16;
    constexpr int ACH = BM * (GT_K / 8);   // 16-byte chunks of the A tile
    constexpr int BCH = GT_K * (BN / 8);   // 16-byte chunks of the B tile
    constexpr int LA = (ACH + GB - 1) / GB, LB = (BCH + GB - 1) / GB;
    const int M = d.m, N = d.n, K = d.k;
    const unsigned short *A = (const unsigned short *)d.a;
    const unsigned short *B = (const unsigned short *)d.b;
    const int nbn = (N + BN - 1) / BN;
    const int bm0 = (tile / nbn) * BM, bn0 = (tile % nbn) * BN;
    const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
    const int wm = wave >> 1, wn = wave & 1;

    f32x4_t acc[TM][TN];
#pragma unroll
    for (int i = 0; i < TM; ++i)
#pragma unroll
        for (int j = 0; j < TN; ++j) acc[i][j] = f32x4_t{0.f, 0.f, 0.f, 0.f};
    const bool a_vec = (K % 8 == 0) && ((reinterpret_cast<uintptr_t>(A) & 15) == 0);
    const bool b_vec = (N % 8 == 0) && ((reinterpret_cast<uintptr_t>(B) & 15) == 0);

    for (int k0 = 0; k0 < K; k0 += GT_K) {
        // ---- global -> registers (zero outside the problem)
        unsigned short av[LA][8], bv[LB][8];
#pragma unroll
        for (int s = 0; s < LA; ++s) {
            const int q = tid + s * GB;
            const int r = bm0 + q / (GT_K / 8), kk = k0 + (q % (GT_K / 8)) * 8;
            if (q < ACH && r < M && kk + 8 <= K) {
                const u32x4_t v = a_vec ? *reinterpret_cast<const u32x4_t *>(A + (size_t)r * K + kk) : chunk8(A + (size_t)r * K + kk);
#pragma unroll
                for (int e = 0; e < 4; ++e) {
                    av[s][2 * e] = (unsigned short)(v[e] & 0xffff);
                    av[s][2 * e + 1] = (unsigned short)(v[e] >> 16);
                }
            } else {
#pragma unroll
                for (int e = 0; e < 8; ++e) av[s][e] = (q < ACH && r < M && kk + e < K) ? ld_bf16(A + (size_t)r * K + kk + e) : 0;
            }
        }
#pragma unroll
        for (int s = 0; s < LB; ++s) {
            const int q = tid + s * GB;
            const int kr = k0 + q / (BN / 8), c = bn0 + (q % (BN / 8)) * 8;
            if (q < BCH && kr < K && c + 8 <= N) {
                const u32x4_t v = b_vec ? *reinterpret_cast<const u32x4_t *>(B + (size_t)kr * N + c) : chunk8(B + (size_t)kr * N + c);
#pragma unroll
                for (int e = 0; e < 4; ++e) {
                    bv[s][2 * e] = (unsigned short)(v[e] & 0xffff);
                    bv[s][2 * e + 1] = (unsigned short)(v[e] >> 16);
                }
            } else {
#pragma unroll
                for (int e = 0; e < 8; ++e) bv[s][e] = (q < BCH && kr < K && c + e < N) ? ld_bf16(B + (size_t)kr * N + c + e) : 0;
            }
        }
        __syncthreads();  // previous step's fragments are consumed
#pragma unroll
        for (int s = 0; s < LA; ++s) {
            const int q = tid + s * GB;
            if (q < ACH) {
                const int row = q / (GT_K / 8), ch = q % (GT_K / 8);
#pragma unroll
                for (int e = 0; e < 8; ++e) As[row * G_LDK + ch * 8 + e] = av[s][e];
            }
        }
#pragma unroll
        for (int s = 0; s < LB; ++s) {
            const int q = tid + s * GB;
            if (q < BCH) {
                const int krow = q / (BN / 8), ch = q % (BN / 8);
#pragma unroll
                for (int e = 0; e < 8; ++e) Bs[(ch * 8 + e) * G_LDK + krow] = bv[s][e];  // transpose to Bt[n][k]
            }
        }
        __syncthreads();
        const int koff = 8 * (lane >> 4);
        bf16x8_t fa[TM], fb[TN];
#pragma unroll
        for (int i = 0; i < TM; ++i)
            fa[i] = *reinterpret_cast<const bf16x8_t *>(As + (wm * WM + i * 16 + (lane & 15)) * G_LDK + koff);
#pragma unroll
        for (int j = 0; j < TN; ++j)
            fb[j] = *reinterpret_cast<const bf16x8_t *>(Bs + (wn * WN + j * 16 + (lane & 15)) * G_LDK + koff);
#pragma unroll
        for (int i = 0; i < TM; ++i)
#pragma unroll
            for (int j = 0; j < TN; ++j) acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(fa[i], fb[j], acc[i][j], 0, 0, 0);
    }
    float *C = d.c;
#pragma unroll
    for (int i = 0; i < TM; ++i)
#pragma unroll
        for (int j = 0; j < TN; ++j) {
            const int col = bn0 + wn * WN + j * 16 + (lane & 15);
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int row = bm0 + wm * WM + i * 16 + 4 * (lane >> 4) + r;
                if (row < M && col < N) C[(size_t)row * N + col] = acc[i][j][r];
            }
        }
}

__device__ __forceinline__ void gemm_tile(const nrk_gemm_desc &d, int tile, unsigned short *As, unsigned short *Bs) {
    gemm_tile_t<GT_M, GT_N>(d, tile, As, Bs);
}

// ---------------- row sums ----------------
__device__ __forceinline__ void two_sum_f(float a, float b, float &s, float &e) {
    s = __fadd_rn(a, b);
    const float v = __fsub_rn(s, a);
    e = __fadd_rn(__fsub_rn(a, __fsub_rn(s, v)), __fsub_rn(b, v));
}

// Double-word wave sum; every lane returns (hi, lo).
__device__ __forceinline__ void wave_dw(float &hi, float &lo) {
#pragma unroll
    for (int m = 32; m >= 1; m >>= 1) {
        const float bh = __shfl_xor(hi, m, 64), bl = __shfl_xor(lo, m, 64);
        float s, e;
        two_sum_f(hi, bh, s, e);
        e = __fadd_rn(e, __fadd_rn(lo, bl));
        const float h = __fadd_rn(s, e);  // a sum that is not finite stays the plain sum
        lo = __builtin_isfinite(h) ? __fsub_rn(e, __fsub_rn(h, s)) : 0.f;
        hi = __builtin_isfinite(s) ? h : s;
    }
}

// Sum of x[0..len) by one wavefront. The (at most 3) elements before the first 16-byte boundary
// and the (at most 3) after the last whole vector are added by single lanes; the body streams
// 16-byte non-temporal vectors, lane l taking vectors l, l+64, ..., 8 per lane per step (8 KiB
// in flight per wavefront), summed pairwise per lane and accumulated with TwoSum.
__device__ __forceinline__ void wave_row_sum(const float *x, int64_t len, float &hi, float &lo) {
    typedef float f4 __attribute__((ext_vector_type(4)));
    constexpr int NV = 8;
    const int lane = threadIdx.x & 63;
    float acc = 0.f, comp = 0.f;  // per-lane TwoSum accumulation
    auto add = [&](float p) {
        float s, e;
        two_sum_f(acc, p, s, e);
        acc = s;
        comp = __fadd_rn(comp, __builtin_isfinite(s) ? e : 0.f);  // inf stays inf (its error term is NaN)
    };
    const int64_t head = min(len, (int64_t)(((16u - ((unsigned)(uintptr_t)x & 15u)) & 15u) >> 2));
    if (lane < head) add(x[lane]);
    const f4 *xv = reinterpret_cast<const f4 *>(x + head);
    const int64_t nv = (len - head) >> 2;
    for (int64_t base = 0; base < nv; base += 64 * NV) {
        f4 v[NV];
#pragma unroll
        for (int e = 0; e < NV; ++e) {
            const int64_t i = base + e * 64 + lane;
            v[e] = f4{0.f, 0.f, 0.f, 0.f};
            if (i < nv) v[e] = __builtin_nontemporal_load(xv + i);
        }
        float q[NV];
#pragma unroll
        for (int e = 0; e < NV; ++e) q[e] = __fadd_rn(__fadd_rn(v[e][0], v[e][1]), __fadd_rn(v[e][2], v[e][3]));
        add(__fadd_rn(__fadd_rn(__fadd_rn(q[0], q[1]), __fadd_rn(q[2], q[3])),
                      __fadd_rn(__fadd_rn(q[4], q[5]), __fadd_rn(q[6], q[7]))));
    }
    const int64_t t0 = head + (nv << 2);
    if (t0 + lane < len) add(x[t0 + lane]);
    hi = acc;
    lo = comp;
    wave_dw(hi, lo);
}

__device__ void rows_item(const GroupedArgs &g, int r0, int r1) {
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    for (int r = r0 + wave; r < r1; r += GB / 64) {
        const int64_t off = g.rows_off[r], len = g.rows_off[r + 1] - off;
        float hi, lo;
        wave_row_sum(g.rows_x + off, len, hi, lo);
        if (lane == 0) g.rows_out[r] = (float)((double)hi + (double)lo);
    }
}

// A long row is summed whole by one workgroup: the 4 wavefronts take fixed contiguous quarters
// and their double-word results are combined in wave order. No partial ever leaves the
// workgroup, so no cross-workgroup synchronisation (and no agent-scope fence) is needed.
__device__ void long_row_item(const GroupedArgs &g, int row, double *sh) {
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int64_t off = g.rows_off[row], len = g.rows_off[row + 1] - off;
    const int64_t q = ((len + 3) / 4 + 63) / 64 * 64;  // quarter, rounded to whole wave strides
    const int64_t s0 = min(len, wave * q), s1 = min(len, s0 + q);
    float hi, lo;
    wave_row_sum(g.rows_x + off + s0, s1 - s0, hi, lo);
    __syncthreads();
    if (lane == 0) {
        sh[2 * wave] = (double)hi;
        sh[2 * wave + 1] = (double)lo;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        double t = 0.0;
        for (int w = 0; w < GB / 64; ++w) t += sh[2 * w] + sh[2 * w + 1];
        g.rows_out[row] = (float)t;
    }
}

// 6 waves per SIMD (80 VGPRs, no spills; 4 without the hint at 100 registers): more resident
// workgroups let row streaming overlap the latency-bound GEMM tiles (measured on the 4096-GEMM /
// 65536-row batch: 1.03 -> 0.96 ms; 8 waves would spill).
__global__ __launch_bounds__(GB) __attribute__((amdgpu_waves_per_eu(6, 8))) void grouped_kernel(GroupedArgs g) {
    __shared__ __attribute__((aligned(16))) unsigned short As[GT_M * G_LDK];
    __shared__ __attribute__((aligned(16))) unsigned short Bs[GT_N * G_LDK];
    __shared__ double sh[2 * GB / 64];
    // Static "snake" deal of the cost-sorted queue: in round r workgroup b takes item
    // r*G + b (r even) or r*G + G-1-b (r odd). Every quantity below derives from blockIdx/gridDim
    // and kernel arguments, so all control flow around the barriers inside the item bodies is
    // uniform, and there is no shared queue state to reset between replays.
    const int G = (int)gridDim.x, b = (int)blockIdx.x;
    for (int round = 0;; ++round) {
        const int it = round * G + ((round & 1) ? (G - 1 - b) : b);
        if (it >= g.n_items) break;
        const Item w = g.items[it];
        if (w.kind == ITEM_GEMM) {
            gemm_tile(g.descs[w.a], w.b, As, Bs);
        } else if (w.kind == ITEM_ROWS) {
            rows_item(g, w.a, w.b);
        } else {
            long_row_item(g, w.a, sh);
        }
        __syncthreads();  // the next item may reuse the LDS this one used
    }
}

// One plain bf16 GEMM per launch (the unbatched per-call path and the tuner's candidates):
// one workgroup per BM x BN tile.
template <int BM, int BN>
__global__ __launch_bounds__(GB) void single_gemm_kernel(nrk_gemm_desc d) {
    __shared__ __attribute__((aligned(16))) unsigned short As[BM * G_LDK];
    __shared__ __attribute__((aligned(16))) unsigned short Bs[BN * G_LDK];
    gemm_tile_t<BM, BN>(d, blockIdx.x, As, Bs);
}

template <int BM, int BN>
nrk_status launch_single_bf16(const nrk_gemm_desc &d, hipStream_t stream) {
    const int tiles = ((d.m + BM - 1) / BM) * ((d.n + BN - 1) / BN);
    hipLaunchKernelGGL((single_gemm_kernel<BM, BN>), dim3(tiles), dim3(GB), 0, stream, d);
    NRK_HIP(hipGetLastError());
    return NRK_OK;
}

// Tile configurations of the bf16 NN GEMM, indexed by the tuner / selector.
nrk_status gemm_bf16_config(int config, const nrk_gemm_desc &d, hipStream_t stream) {
    if (d.m <= 0 || d.n <= 0) return NRK_OK;
    switch (config) {
        case 0: return launch_single_bf16<32, 32>(d, stream);
        case 1: return launch_single_bf16<64, 64>(d, stream);
        case 2: return launch_single_bf16<128, 64>(d, stream);
        case 3: return launch_single_bf16<64, 128>(d, stream);
        case 4: return launch_single_bf16<128, 128>(d, stream);
        default: return fail(NRK_ERR_INVALID_ARGUMENT, "bf16 GEMM config %d outside [0,4]", config);
    }
}

}  // namespace nrk

// ---------------- plan object ----------------
struct nrk_grouped_plan {
    int device;
    std::vector<nrk_gemm_desc> h_descs;
    nrk_gemm_desc *d_descs = nullptr;
    nrk::Item *d_items = nullptr;
    int n_items = 0;
    int n_gemm_items = 0, n_row_items = 0;
    const float *rows_x = nullptr;
    int64_t *d_rows_off = nullptr;
    float *rows_out = nullptr;
    int n_rows = 0;
    int grid = 0;
};

using namespace nrk;

namespace {
void free_plan(nrk_grouped_plan *p) {
    if (!p) return;
    (void)hipFree(p->d_descs);
    (void)hipFree(p->d_items);
    (void)hipFree(p->d_rows_off);
    delete p;
}
}  // namespace

extern "C" {

nrk_status nrk_grouped_plan_create(int32_t gemm_count, const nrk_gemm_desc *descs, int32_t row_count, const float *rows_x,
                                   const int64_t *row_offsets, float *rows_out, const float *item_cost_ns,
                                   nrk_grouped_plan **out_plan) {
    NRK_TRY(ensure_device());
    if (!out_plan) return fail(NRK_ERR_INVALID_ARGUMENT, "out_plan is NULL");
    *out_plan = nullptr;
    if (gemm_count < 0 || row_count < 0) return fail(NRK_ERR_INVALID_ARGUMENT, "negative count");
    if (gemm_count > 0 && !descs) return fail(NRK_ERR_INVALID_ARGUMENT, "descs is NULL");
    if (row_count > 0 && (!rows_x || !row_offsets || !rows_out)) return fail(NRK_ERR_INVALID_ARGUMENT, "row arrays are NULL");
    auto *p = new nrk_grouped_plan();
    int dev = 0;
    (void)hipGetDevice(&dev);
    p->device = dev;
    std::vector<Item> items;
    std::vector<double> cost;
    // GEMM tiles; default cost: MFMA work of the tile plus its operand bytes
    for (int g = 0; g < gemm_count; ++g) {
        const nrk_gemm_desc &d = descs[g];
        if (d.m < 0 || d.n < 0 || d.k < 0 || (d.m * (int64_t)d.n > 0 && !d.c) || (d.k > 0 && d.m * (int64_t)d.n > 0 && (!d.a || !d.b))) {
            free_plan(p);
            return fail(NRK_ERR_INVALID_ARGUMENT, "descriptor %d is invalid", g);
        }
        const int tiles = ((d.m + GT_M - 1) / GT_M) * ((d.n + GT_N - 1) / GT_N);
        for (int t = 0; t < tiles; ++t) {
            items.push_back(Item{ITEM_GEMM, g, t, 0});
            cost.push_back(item_cost_ns ? (double)item_cost_ns[g] : 200.0 + (double)std::max(d.k, 1) * 2.0);
        }
    }
    p->n_gemm_items = (int)items.size();
    // rows: short rows grouped, long rows one workgroup each
    int r = 0;
    while (r < row_count) {
        const int64_t len = row_offsets[r + 1] - row_offsets[r];
        if (len < 0) {
            free_plan(p);
            return fail(NRK_ERR_INVALID_ARGUMENT, "row_offsets decreases at row %d", r);
        }
        if (len > SHORT_MAX) {
            items.push_back(Item{ITEM_LONGROW, r, 0, 0});
            cost.push_back(300.0 + (double)len * 0.15);
            ++r;
        } else {
            int r1 = r;
            int64_t acc = 0;
            while (r1 < row_count) {
                const int64_t l = row_offsets[r1 + 1] - row_offsets[r1];
                if (l > SHORT_MAX || (acc + l > GROUP_ELEMS && r1 > r)) break;
                acc += l;
                ++r1;
            }
            items.push_back(Item{ITEM_ROWS, r, r1, 0});
            cost.push_back(300.0 + acc * 0.6);
            r = r1;
        }
    }
    p->n_row_items = (int)items.size() - p->n_gemm_items;
    // Queue order: each class (GEMM tiles: MFMA / latency bound; rows: HBM bound) longest first,
    // the two merged so that both advance at the same fraction of their total cost. Every round
    // of the snake deal then mixes the classes, and the workgroups resident on a CU overlap
    // memory streaming with matrix work instead of all streaming rows first and all
    // multiplying afterwards.
    std::vector<int> order;
    order.reserve(items.size());
    {
        std::vector<int> og(p->n_gemm_items), orow(p->n_row_items);
        std::iota(og.begin(), og.end(), 0);
        std::iota(orow.begin(), orow.end(), p->n_gemm_items);
        auto by_cost = [&](int x, int y) { return cost[x] > cost[y]; };
        std::stable_sort(og.begin(), og.end(), by_cost);
        std::stable_sort(orow.begin(), orow.end(), by_cost);
        double tg = 0.0, tr = 0.0, cg = 0.0, cr = 0.0;
        for (int i : og) tg += cost[i];
        for (int i : orow) tr += cost[i];
        size_t i = 0, j = 0;
        while (i < og.size() || j < orow.size()) {
            const bool take_g = j >= orow.size() || (i < og.size() && cg * tr <= cr * tg);
            if (take_g) {
                order.push_back(og[i]);
                cg += cost[og[i++]];
            } else {
                order.push_back(orow[j]);
                cr += cost[orow[j++]];
            }
        }
    }
    std::vector<Item> sorted(items.size());
    for (size_t i = 0; i < order.size(); ++i) sorted[i] = items[order[i]];
    p->n_items = (int)sorted.size();
    p->h_descs.assign(descs, descs + gemm_count);
    p->rows_x = rows_x;
    p->rows_out = rows_out;
    p->n_rows = row_count;
    auto up = [&](void **dst, const void *src, size_t bytes) -> nrk_status {
        NRK_HIP(hipMalloc(dst, bytes ? bytes : 16));
        if (bytes) NRK_HIP(hipMemcpy(*dst, src, bytes, hipMemcpyHostToDevice));
        return NRK_OK;
    };
    nrk_status st = NRK_OK;
    if ((st = up((void **)&p->d_descs, descs, sizeof(nrk_gemm_desc) * gemm_count)) != NRK_OK ||
        (st = up((void **)&p->d_items, sorted.data(), sizeof(Item) * sorted.size())) != NRK_OK ||
        (st = up((void **)&p->d_rows_off, row_offsets, row_count ? sizeof(int64_t) * (row_count + 1) : 0)) != NRK_OK) {
        free_plan(p);
        return st;
    }
    // as many workgroups as are resident together: small GEMM tiles are latency bound, so every
    // resident workgroup is another tile in flight
    int resident = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&resident, grouped_kernel, GB, 0) != hipSuccess || resident < 1)
        resident = 4;
    p->grid = std::max(1, std::min(p->n_items, cu_count() * resident));
    *out_plan = p;
    return NRK_OK;
}

nrk_status nrk_grouped_plan_run(const nrk_grouped_plan *p, nrk_stream stream) {
    if (!p) return fail(NRK_ERR_INVALID_ARGUMENT, "plan is NULL");
    if (p->n_items == 0) return NRK_OK;
    GroupedArgs g{};
    g.descs = p->d_descs;
    g.items = p->d_items;
    g.n_items = p->n_items;
    g.rows_x = p->rows_x;
    g.rows_off = p->d_rows_off;
    g.rows_out = p->rows_out;
    hipLaunchKernelGGL(grouped_kernel, dim3(p->grid), dim3(GB), 0, (hipStream_t)stream, g);
    NRK_HIP(hipGetLastError());
    return NRK_OK;
}

nrk_status nrk_grouped_plan_info(const nrk_grouped_plan *p, int32_t *n_items, int32_t *n_gemm_tiles, int32_t *grid) {
    if (!p) return fail(NRK_ERR_INVALID_ARGUMENT, "plan is NULL");
    if (n_items) *n_items = p->n_items;
    if (n_gemm_tiles) *n_gemm_tiles = p->n_gemm_items;
    if (grid) *grid = p->grid;
    return NRK_OK;
}

nrk_status nrk_grouped_plan_destroy(nrk_grouped_plan *p) {
    free_plan(p);
    return NRK_OK;
}

nrk_status nrk_dev_gemm_bf16(const nrk_gemm_desc *desc, nrk_stream stream) {
    NRK_TRY(ensure_device());
    if (!desc) return fail(NRK_ERR_INVALID_ARGUMENT, "desc is NULL");
    const int cfg = select_config(NRK_DT_BF16, (size_t)std::max(desc->m, 0), (size_t)std::max(desc->n, 0),
                                  (size_t)std::max(desc->k, 0), (hipStream_t)stream);
    return gemm_bf16_config(cfg >= 0 ? cfg : 1, *desc, (hipStream_t)stream);  // built-in: 64 x 64
}

nrk_status nrk_dev_ragged_per_call(int32_t gemm_count, const nrk_gemm_desc *descs, int32_t row_count, const float *rows_x,
                                   const int64_t *row_offsets, float *rows_out, nrk_stream stream) {
    NRK_TRY(ensure_device());
    if (gemm_count < 0 || row_count < 0) return fail(NRK_ERR_INVALID_ARGUMENT, "negative count");
    if (gemm_count > 0 && !descs) return fail(NRK_ERR_INVALID_ARGUMENT, "descs is NULL");
    if (row_count > 0 && (!rows_x || !row_offsets || !rows_out)) return fail(NRK_ERR_INVALID_ARGUMENT, "row arrays are NULL");
    for (int g = 0; g < gemm_count; ++g) NRK_TRY(nrk_dev_gemm_bf16(&descs[g], stream));
    for (int r = 0; r < row_count; ++r) {
        const int64_t len = row_offsets[r + 1] - row_offsets[r];
        if (len < 0) return fail(NRK_ERR_INVALID_ARGUMENT, "row_offsets decreases at row %d", r);
        NRK_TRY(nrk_dev_sum_f32(rows_x + row_offsets[r], (size_t)len, rows_out + r, stream));
    }
    return NRK_OK;
}

// ---- hipGraph capture helpers ----
nrk_status nrk_graph_begin_capture(nrk_stream stream) {
    NRK_TRY(ensure_device());
    if (!stream) return fail(NRK_ERR_INVALID_ARGUMENT, "capture needs a created stream (not NULL)");
    NRK_HIP(hipStreamBeginCapture((hipStream_t)stream, hipStreamCaptureModeThreadLocal));
    return NRK_OK;
}

nrk_status nrk_graph_end_capture(nrk_stream stream, void **graph_exec) {
    if (!graph_exec) return fail(NRK_ERR_INVALID_ARGUMENT, "graph_exec is NULL");
    hipGraph_t g = nullptr;
    NRK_HIP(hipStreamEndCapture((hipStream_t)stream, &g));
    hipGraphExec_t e = nullptr;
    hipError_t err = hipGraphInstantiate(&e, g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    if (err != hipSuccess) return hip_fail(err, "hipGraphInstantiate", __FILE__, __LINE__);
    *graph_exec = (void *)e;
    return NRK_OK;
}

nrk_status nrk_graph_launch(void *graph_exec, nrk_stream stream) {
    NRK_HIP(hipGraphLaunch((hipGraphExec_t)graph_exec, (hipStream_t)stream));
    return NRK_OK;
}

nrk_status nrk_graph_destroy(void *graph_exec) {
    if (graph_exec) NRK_HIP(hipGraphExecDestroy((hipGraphExec_t)graph_exec));
    return NRK_OK;
}

}  // extern "C"
