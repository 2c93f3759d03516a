// FP32 GEMM as BF16 split products on gfx950 matrix cores (v_mfma_f32_16x16x32_bf16), and the
// plain bf16 -> fp32 GEMM tile that the ragged batch reuses.
//
// Split (Henry, Tang & Heinecke 2019, "Leveraging the bfloat16 Artificial Intelligence Datatype
// for Higher-Precision Computations"): x = x0 + x1 + x2 with x0 = bf16(x), x1 = bf16(x - x0),
// x2 = bf16(x - x0 - x1); each difference is exact in f32. The product keeps the pairs
//   BF16x3: A1*B0 + A0*B1 + A0*B0              (3 MFMA products, drops terms of order 2^-16)
//   BF16x6: A2*B0 + A1*B1 + A0*B2 + A1*B0 + A0*B1 + A0*B0   (drops terms of order 2^-24)
// accumulated smallest-first (BF16x6: A0*B0 in an accumulator of its own, see ACC_SPLIT).
//
// Data path: a split pass writes the terms as bf16 matrices zero-padded to the tile grid, A as
// (Mp, Kp) and B TRANSPOSED as (Np, Kp) (K contiguous for both, so every MFMA operand is one
// 16-byte LDS read), then one GEMM kernel streams all terms: global -> LDS by the load unit
// (three buffers), XCD-aware grouped tile order. Padding removes every bounds check from the main
// loop; only the C store is predicated.
#include "gemm_lowp.hpp"
#include "ops.hpp"

namespace nrk {

constexpr int SPLIT_TILE = 32;

// out[t][r][c] (or out[t][c][r] when TRANSPOSE) = term t of X[r][c]; zero in the padding.
template <int T, bool TRANSPOSE>
__global__ __launch_bounds__(256) void split_bf16_kernel(const float *__restrict__ X, int R, int Cc,
                                                        unsigned short *__restrict__ out, int out_rows, int out_cols,
                                                        size_t term_stride) {
    __shared__ unsigned short tile[T][SPLIT_TILE][SPLIT_TILE + 2];
    const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;  // 32 x 8
    const int r0 = blockIdx.y * SPLIT_TILE, c0 = blockIdx.x * SPLIT_TILE;
    for (int dy = ty; dy < SPLIT_TILE; dy += 8) {
        const int r = r0 + dy, c = c0 + tx;
        const float x = (r < R && c < Cc) ? X[(size_t)r * Cc + c] : 0.0f;
        unsigned short h[T];
        float rem = x;
#pragma unroll
        for (int t = 0; t < T; ++t) {
            h[t] = f32_to_bf16_rn(rem);
            rem = rem - bf16_to_f32(h[t]);  // exact
        }
        if (TRANSPOSE) {
#pragma unroll
            for (int t = 0; t < T; ++t) tile[t][tx][dy] = h[t];
        } else {
#pragma unroll
            for (int t = 0; t < T; ++t) tile[t][dy][tx] = h[t];
        }
    }
    __syncthreads();
    // TRANSPOSE: output row = input column
    const int o_r0 = TRANSPOSE ? c0 : r0, o_c0 = TRANSPOSE ? r0 : c0;
    for (int dy = ty; dy < SPLIT_TILE; dy += 8) {
        const int orow = o_r0 + dy, ocol = o_c0 + tx;
        if (orow < out_rows && ocol < out_cols) {
#pragma unroll
            for (int t = 0; t < T; ++t) out[t * term_stride + (size_t)orow * out_cols + ocol] = tile[t][dy][tx];
        }
    }
}

// Pair p of the product list (smallest terms first): A term pair_a(p) times B term pair_b(p).
template <int NPAIR>
__device__ __forceinline__ constexpr int pair_a(int p) {
    return NPAIR == 1 ? 0 : NPAIR == 3 ? (p == 0 ? 1 : 0) : (p == 0 ? 2 : (p == 1 || p == 3) ? 1 : 0);
}
template <int NPAIR>
__device__ __forceinline__ constexpr int pair_b(int p) {
    return NPAIR == 1 ? 0 : NPAIR == 3 ? (p == 1 ? 1 : 0) : (p == 2 ? 2 : (p == 1 || p == 4) ? 1 : 0);
}

// C(M,N) f32 = sum over pairs p of A[a_p] * Bt[b_p]^T. A terms: (Mp, Kp) each, Bt terms: (Np, Kp).
//
// K tiles of 32 bf16 (64 bytes per row) move global -> LDS by the load unit itself
// (global_load_lds), all terms of both operands in one buffer (NTA*BM A rows, then NTB*BN Bt
// rows, swizzled by lds_swz), NBUF buffers deep: tiles t+1 .. t+NBUF-2 are in flight while tile
// t is multiplied, one barrier per tile. Per output and K step the products are accumulated
// pairs smallest first.
constexpr int BF_BK = 32;  // bf16 values per K tile (one MFMA K step)

// Accumulation of the pairs (ACC): ACC_ONE, every pair into one FP32 accumulator per output;
// ACC_SPLIT, A0*B0 into its own accumulator and the smaller pairs into a second, added once at
// the end, so the small terms never round against the large running sum. Measured at K = 8192
// (N(0,1) operands, 2048 sampled outputs, error / sum_k |a_k b_k|): BF16x6 goes from 1.4e-7 max
// (2.5e-8 rms) with one accumulator to 4.9e-8 (9.8e-9 rms) split, against 1.8e-7 (2.8e-8) for
// native FP32. BF16x3's error (3.4e-7 max, 7.7e-8 rms) is the pairs it drops (order 2^-16 of
// each product) and is the same under either accumulation, so it keeps the one accumulator.
enum { ACC_ONE = 0, ACC_SPLIT = 1 };

template <int BM, int BN, int WM, int WN, int NTA, int NTB, int NPAIR, int NBUF, int ACC>
__global__ __launch_bounds__((BM / WM) * (BN / WN) * 64) void gemm_bf16_nt_kernel(
    int M, int N, int Kp, const unsigned short *__restrict__ A, size_t a_term_stride,
    const unsigned short *__restrict__ Bt, size_t b_term_stride, float *__restrict__ C, int ldc, int group_m) {
    constexpr int NT = (BM / WM) * (BN / WN) * 64;
    constexpr int WGN = BN / WN;
    constexpr int TM = WM / 16, TN = WN / 16;
    constexpr int ROWS = NTA * BM + NTB * BN;  // 64-byte LDS rows per buffer
    constexpr int BUF = ROWS * 64;             // bytes per buffer
    constexpr int R = ROWS * 4 / NT;           // 16-byte loads per thread per tile
    static_assert((ROWS * 4) % NT == 0, "tile/threads mismatch");

    extern __shared__ __attribute__((aligned(16))) unsigned char smem_bf[];
    const int tid = threadIdx.x, lane = tid & 63;
    const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
    const int wm = wave / WGN, wn = wave % WGN;
    int tm, tn;
    lowp_tile_of(blockIdx.x, (M + BM - 1) / BM, (N + BN - 1) / BN, group_m, tm, tn);
    const int bm0 = tm * BM, bn0 = tn * BN;

    // Load s of this thread covers chunk position p = s*NT + tid of the buffer: LDS row p/4 (A
    // term rows first, then Bt term rows), stored chunk p%4 holding logical chunk (p%4) ^ lds_swz.
    const unsigned short *src[R];
#pragma unroll
    for (int s = 0; s < R; ++s) {
        const int p = s * NT + tid, row = p >> 2, c = (p & 3) ^ lds_swz(row);
        if (row < NTA * BM)
            src[s] = A + (row / BM) * a_term_stride + (size_t)(bm0 + row % BM) * Kp + c * 8;
        else
            src[s] = Bt + ((row - NTA * BM) / BN) * b_term_stride + (size_t)(bn0 + (row - NTA * BM) % BN) * Kp + c * 8;
    }
    auto issue = [&](int t, int slot) {  // K tile t into LDS buffer slot
        unsigned char *base = smem_bf + slot * BUF + wave * 64 * 16;  // this wave's 1 KiB slot
        const size_t k0 = (size_t)t * BF_BK;
#pragma unroll
        for (int s = 0; s < R; ++s) lds_dma16(src[s] + k0, base + s * NT * 16);
    };

    f32x4_t acc[TM][TN];
    f32x4_t lo[ACC == ACC_SPLIT ? TM : 1][ACC == ACC_SPLIT ? TN : 1];  // ACC_SPLIT: the smaller pairs
#pragma unroll
    for (int i = 0; i < TM; ++i)
#pragma unroll
        for (int j = 0; j < TN; ++j) acc[i][j] = f32x4_t{0.f, 0.f, 0.f, 0.f};
    if constexpr (ACC == ACC_SPLIT) {
#pragma unroll
        for (int i = 0; i < TM; ++i)
#pragma unroll
            for (int j = 0; j < TN; ++j) lo[i][j] = f32x4_t{0.f, 0.f, 0.f, 0.f};
    }

    const int r16 = lane & 15;
    const int pc = ((lane >> 4) ^ lds_swz(r16)) * 16;  // stored position of this lane's chunk (row = r16 mod 16)
    const int nk = Kp / BF_BK;
    // Operand fragments in registers: the A row blocks (all terms) of the tile being multiplied,
    // each reloaded with the next tile's as soon as its products are issued, and the B column
    // blocks of two tiles (the next tile's are read while this one's products run), so the matrix
    // cores never wait on LDS latency. The scheduling fences keep the compiler from sinking the
    // reads to the next barrier, where their latency would be exposed.
    bf16x8_t fa[NTA][TM], fb0[NTB][TN], fb1[NTB][TN];
    auto load_a = [&](int i, int t) {
        const unsigned char *As = smem_bf + (t % NBUF) * BUF;
#pragma unroll
        for (int u = 0; u < NTA; ++u)
            fa[u][i] = *reinterpret_cast<const bf16x8_t *>(As + (u * BM + wm * WM + i * 16 + r16) * 64 + pc);
    };
    auto load_b = [&](bf16x8_t(&fb)[NTB][TN], int t) {
        const unsigned char *Bs = smem_bf + (t % NBUF) * BUF + NTA * BM * 64;
#pragma unroll
        for (int u = 0; u < NTB; ++u)
#pragma unroll
            for (int j = 0; j < TN; ++j)
                fb[u][j] = *reinterpret_cast<const bf16x8_t *>(Bs + (u * BN + wn * WN + j * 16 + r16) * 64 + pc);
    };
    // products of the tile held in fa / fb (per output: the pairs in order; the last pair of the
    // list is A0*B0), A row block i then reloaded from tile tn
    auto mma_reload = [&](const bf16x8_t(&fb)[NTB][TN], int tn) {
#pragma unroll
        for (int i = 0; i < TM; ++i) {
            if constexpr (ACC == ACC_ONE) {
#pragma unroll
                for (int p = 0; p < NPAIR; ++p)
#pragma unroll
                    for (int j = 0; j < TN; ++j)
                        acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(fa[pair_a<NPAIR>(p)][i], fb[pair_b<NPAIR>(p)][j],
                                                                            acc[i][j], 0, 0, 0);
            } else {
#pragma unroll
                for (int p = 0; p < NPAIR - 1; ++p)
#pragma unroll
                    for (int j = 0; j < TN; ++j)
                        lo[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(fa[pair_a<NPAIR>(p)][i], fb[pair_b<NPAIR>(p)][j],
                                                                           lo[i][j], 0, 0, 0);
#pragma unroll
                for (int j = 0; j < TN; ++j)
                    acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(fa[0][i], fb[0][j], acc[i][j], 0, 0, 0);
            }
            load_a(i, tn);
            __builtin_amdgcn_sched_barrier(0);
        }
    };
    // Every pass issues one group of loads (past the last tile, a harmless reload of the last
    // tile into the free buffer), so exactly NBUF-2 groups are in flight behind tile t and the
    // wait count is a constant: no branch around the barrier.
    auto sync_tile = [&](int t) {
        lowp_tile_barrier<(NBUF - 2) * R>();
        issue(min(t + NBUF - 1, nk - 1), (t + NBUF - 1) % NBUF);  // into the buffer tile t-1 used
    };
    // Kp is a multiple of 64, so nk is even and tiles pair up (t, t+1); the last pair also reads
    // the fragments of "tile nk" from a stale buffer and never uses them (no branch around the
    // MFMAs, which would give the accumulators a second register assignment).
#pragma unroll
    for (int t = 0; t < NBUF - 1; ++t) issue(min(t, nk - 1), t);
    sync_tile(0);
#pragma unroll
    for (int i = 0; i < TM; ++i) load_a(i, 0);
    load_b(fb0, 0);
    for (int t = 0; t < nk; t += 2) {
        sync_tile(t + 1);
        load_b(fb1, t + 1);
        __builtin_amdgcn_sched_barrier(0);
        mma_reload(fb0, t + 1);
        sync_tile(t + 2);
        load_b(fb0, t + 2);
        __builtin_amdgcn_sched_barrier(0);
        mma_reload(fb1, t + 2);
    }
    LOWP_WAIT_VM_LGKM0(0);  // the trailing reloads have landed before the workgroup ends
#pragma unroll
    for (int i = 0; i < TM; ++i)
#pragma unroll
        for (int j = 0; j < TN; ++j) {
            const int col = bn0 + wn * WN + j * 16 + r16;
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int row = bm0 + wm * WM + i * 16 + 4 * (lane >> 4) + r;
                if (row < M && col < N) {
                    if constexpr (ACC == ACC_SPLIT) C[(size_t)row * ldc + col] = acc[i][j][r] + lo[i][j][r];
                    else C[(size_t)row * ldc + col] = acc[i][j][r];
                }
            }
        }
}

template <int BM, int BN, int WM, int WN, int NTA, int NTB, int NPAIR>
static nrk_status launch_bf16_nt(int M, int N, int Kp, const unsigned short *A, size_t a_stride, const unsigned short *Bt,
                                 size_t b_stride, float *C, int ldc, hipStream_t stream) {
    constexpr int NBUF = 3;
    constexpr int threads = (BM / WM) * (BN / WN) * 64;
    constexpr size_t lds = (size_t)NBUF * (NTA * BM + NTB * BN) * 64;
    static_assert(lds <= 160 * 1024, "tile does not fit in LDS");
    // the six-product form runs four-wave tiles only (the eight-wave ones exceed the LDS), where
    // the second accumulator set fits; one ACC per product count keeps every tile's bits equal
    constexpr int ACC = NPAIR == 6 ? ACC_SPLIT : ACC_ONE;
    static_assert(ACC == ACC_ONE || threads <= 256, "split accumulators need a four-wave tile");
    auto kern = gemm_bf16_nt_kernel<BM, BN, WM, WN, NTA, NTB, NPAIR, NBUF, ACC>;
    static bool attr_set = false;
    if (!attr_set) {
        NRK_HIP(hipFuncSetAttribute((const void *)kern, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
        attr_set = true;
    }
    const int grid = ((M + BM - 1) / BM) * ((N + BN - 1) / BN);
    hipLaunchKernelGGL(kern, dim3(grid), dim3(threads), lds, stream, M, N, Kp, A, a_stride, Bt, b_stride, C, ldc, 8);
    NRK_HIP(hipGetLastError());
    return NRK_OK;
}

// Padding granularity of the split matrices: every tile config below divides it.
constexpr int PAD_MN = 256;
constexpr int PAD_K = 64;

template <int T>
static nrk_status split(const float *X, int R, int Cc, bool transpose, unsigned short *out, int out_rows, int out_cols,
                        hipStream_t stream) {
    const size_t stride = (size_t)out_rows * out_cols;
    // grid covers the OUTPUT (padded) extent so that the padding is written with zeros
    const int in_rows_cover = transpose ? out_cols : out_rows;
    const int in_cols_cover = transpose ? out_rows : out_cols;
    dim3 grid((in_cols_cover + SPLIT_TILE - 1) / SPLIT_TILE, (in_rows_cover + SPLIT_TILE - 1) / SPLIT_TILE);
    if (transpose)
        hipLaunchKernelGGL((split_bf16_kernel<T, true>), grid, dim3(256), 0, stream, X, R, Cc, out, out_rows, out_cols, stride);
    else
        hipLaunchKernelGGL((split_bf16_kernel<T, false>), grid, dim3(256), 0, stream, X, R, Cc, out, out_rows, out_cols, stride);
    NRK_HIP(hipGetLastError());
    return NRK_OK;
}

// Tile configurations of the split-product GEMM, BM x BN (wave tile): 0: 64x64 (32x32),
// 1: 128x128 (64x64), 2: 128x64 (64x32), 3: 256x128 (64x64, 8 waves), 4: 128x256 (64x64, 8 waves).
// Every term of both operands sits in LDS, three K tiles deep; for the six-product form
// configurations 3 and 4 exceed the LDS and run as configuration 1.
constexpr int SPLIT_TILE_CONFIGS = 5;
template <int NTA, int NTB, int NPAIR>
static nrk_status launch_split_tile(int tile, size_t m, size_t n, size_t Kp, const unsigned short *As, size_t sa,
                                    const unsigned short *Bts, size_t sb, float *c, hipStream_t stream) {
    constexpr bool wide_fits = (size_t)3 * (NTA * 256 + NTB * 128) * 64 <= 160 * 1024;
    switch (tile) {
        case 0: return launch_bf16_nt<64, 64, 32, 32, NTA, NTB, NPAIR>((int)m, (int)n, (int)Kp, As, sa, Bts, sb, c, (int)n, stream);
        case 2: return launch_bf16_nt<128, 64, 64, 32, NTA, NTB, NPAIR>((int)m, (int)n, (int)Kp, As, sa, Bts, sb, c, (int)n, stream);
        case 3:
            if constexpr (wide_fits)
                return launch_bf16_nt<256, 128, 64, 64, NTA, NTB, NPAIR>((int)m, (int)n, (int)Kp, As, sa, Bts, sb, c, (int)n, stream);
            break;
        case 4:
            if constexpr (wide_fits)
                return launch_bf16_nt<128, 256, 64, 64, NTA, NTB, NPAIR>((int)m, (int)n, (int)Kp, As, sa, Bts, sb, c, (int)n, stream);
            break;
        default: break;
    }
    return launch_bf16_nt<128, 128, 64, 64, NTA, NTB, NPAIR>((int)m, (int)n, (int)Kp, As, sa, Bts, sb, c, (int)n, stream);
}

nrk_status gemm_f32_bf16x(size_t m, size_t k, size_t n, const float *a, const float *b, float *c, int products, int tile,
                          hipStream_t stream) {
    NRK_TRY(ensure_device());
    if (tile < -1 || tile >= SPLIT_TILE_CONFIGS)
        return fail(NRK_ERR_INVALID_ARGUMENT, "split GEMM tile config %d outside [0,%d)", tile, SPLIT_TILE_CONFIGS);
    if (products != 3 && products != 6 && products != 1)
        return fail(NRK_ERR_INVALID_ARGUMENT, "products must be 1, 3 or 6 (got %d)", products);
    if (m == 0 || n == 0) return NRK_OK;
    if (k == 0) return gemm<float>(m, 0, n, a, b, c, stream);  // zero fill
    if (!a || !b || !c) return fail(NRK_ERR_INVALID_ARGUMENT, "A, B or C is NULL");
    if (m > (1u << 30) || n > (1u << 30) || k > (1u << 30)) return fail(NRK_ERR_INVALID_ARGUMENT, "dimension too large");
    const int T = products == 6 ? 3 : (products == 3 ? 2 : 1);
    const size_t Mp = align_up(m, PAD_MN), Np = align_up(n, PAD_MN), Kp = align_up(k, PAD_K);
    const size_t bytes = (size_t)T * (Mp + Np) * Kp * sizeof(unsigned short);
    void *ws = nullptr;
    NRK_TRY(get_workspace(stream, bytes, &ws));
    unsigned short *As = (unsigned short *)ws;
    unsigned short *Bts = As + (size_t)T * Mp * Kp;
    if (T == 3) {
        NRK_TRY(split<3>(a, (int)m, (int)k, false, As, (int)Mp, (int)Kp, stream));
        NRK_TRY(split<3>(b, (int)k, (int)n, true, Bts, (int)Np, (int)Kp, stream));
    } else if (T == 2) {
        NRK_TRY(split<2>(a, (int)m, (int)k, false, As, (int)Mp, (int)Kp, stream));
        NRK_TRY(split<2>(b, (int)k, (int)n, true, Bts, (int)Np, (int)Kp, stream));
    } else {
        NRK_TRY(split<1>(a, (int)m, (int)k, false, As, (int)Mp, (int)Kp, stream));
        NRK_TRY(split<1>(b, (int)k, (int)n, true, Bts, (int)Np, (int)Kp, stream));
    }
    const size_t sa = Mp * Kp, sb = Np * Kp;
    if (tile < 0 && products != 1) tile = select_config(products == 6 ? NRK_DT_F32_BF16X6 : NRK_DT_F32_BF16X3, m, n, k, stream);
    if (tile < 0) {  // the largest tiles that still give every CU a workgroup
        if (ceil_div(m, 128) * ceil_div(n, 256) >= (size_t)cu_count()) tile = 4;
        else if (ceil_div(m, 128) * ceil_div(n, 128) >= (size_t)cu_count()) tile = 1;
        else tile = 0;
    }
    if (products == 6) return launch_split_tile<3, 3, 6>(tile, m, n, Kp, As, sa, Bts, sb, c, stream);
    if (products == 3) return launch_split_tile<2, 2, 3>(tile, m, n, Kp, As, sa, Bts, sb, c, stream);
    return launch_split_tile<1, 1, 1>(tile, m, n, Kp, As, sa, Bts, sb, c, stream);
}

}  // namespace nrk

using namespace nrk;

extern "C" {

nrk_status nrk_dev_gemm_f32_bf16x(size_t m, size_t k, size_t n, const float *a, const float *b, float *c, int32_t products,
                                  nrk_stream stream) {
    return gemm_f32_bf16x(m, k, n, a, b, c, products, -1, (hipStream_t)stream);
}

nrk_status nrk_gemm_f32_bf16x(size_t m, size_t k, size_t n, const float *a, const float *b, float *c, int32_t products) {
    NRK_TRY(ensure_device());
    if (m == 0 || n == 0) return NRK_OK;
    void *da = nullptr, *db = nullptr, *dc = nullptr;
    NRK_TRY(staging(0, m * k * 4 + 16, &da));
    NRK_TRY(staging(2, k * n * 4 + 16, &db));
    NRK_TRY(staging(1, m * n * 4 + 16, &dc));
    hipStream_t s = nullptr;
    if (m * k) NRK_TRY(copy_h2d(da, a, m * k * 4, s));
    if (k * n) NRK_TRY(copy_h2d(db, b, k * n * 4, s));
    NRK_TRY(gemm_f32_bf16x(m, k, n, (const float *)da, (const float *)db, (float *)dc, products, -1, s));
    NRK_TRY(copy_d2h(c, dc, m * n * 4, s));
    NRK_HIP(hipStreamSynchronize(s));
    return NRK_OK;
}

}  // extern "C"
