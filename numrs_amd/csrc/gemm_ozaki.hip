// FP64 GEMM emulated on gfx950 INT8 matrix cores: the Ozaki scheme II (Ozaki, Uchino & Imamura,
// "Ozaki Scheme II: A GEMM-oriented emulation of floating-point matrix multiplication using an
// integer modular technique", 2024/25), built on the Ozaki-Ogita-Oishi-Rump error-free splitting
// idea and the INT8-unit DGEMM of Ootomo, Ozaki & Yokota (2024).
//
//   1. Scale rows of A and columns of B by powers of two, 2^s_i and 2^t_j, so that every scaled
//      row and column has a 2-norm of at most X, X^2 < M / 2 (Cauchy-Schwarz then bounds every
//      |C'_ij| by X^2), and round to integers: A' = round(2^s A), B' = round(B 2^t). X <= 2^62,
//      and an entry of 2^52 or more is already an integer, so A' and B' are exact doubles.
//   2. For each of N pairwise-coprime odd moduli m_j <= 255, take symmetric residues
//      a_j = A' mod m_j, b_j = B' mod m_j in [-127, 127] (int8, exact).
//   3. C_j = a_j * b_j^T on v_mfma_i32_16x16x64_i8: int32 accumulation is exact
//      (k * 127^2 < 2^31 for k <= 133 000), then c_j = C_j mod m_j (int8).
//   4. CRT: C' = sum_j c_j w_j mod M (w_j = (M/m_j) * ((M/m_j)^-1 mod m_j), M = prod m_j) is the
//      EXACT integer product A'B'. It is formed with B-bit limbs (B = 53 - ceil(log2(127 N + 1)),
//      42 for up to 16 moduli) whose partial sums are exact in FP64, reduced mod M, normalised,
//      and evaluated in double-double; C = C' 2^(-s_i - t_j).
// The only error is the rounding of A and B to integers in step 1 (at most 1/2 in units of
// |row|_2 / X); all integer work is exact.
//
// Inputs must be finite: a row of A or column of B holding Inf/NaN yields NaN in the
// corresponding outputs.
#include <algorithm>
#include <cmath>
#include <vector>

#include "gemm_lowp.hpp"
#include "ops.hpp"

namespace nrk {

constexpr int OZ_MAX_MODULI = 20;
constexpr int OZ_LIMBS = 4;       // 4 x 41-bit limbs = 164 bits > log2(M) for 20 moduli
constexpr int OZ_DEFAULT_MODULI = 16;
constexpr int OZ_TILE_CONFIGS = 5;  // INT8 GEMM tiles (workgroup tile, wave tile): 64x64 (32x32),
                                     // 128x128 (64x64), 128x256 (64x64), 256x256 (128x64),
                                     // 256x256 (128x128)
constexpr int OZ_BAD_SCALE = 100000;  // marks a row/column holding Inf/NaN

// Pairwise-coprime odd moduli <= 255, largest first (greedy).
static const int kModuli[OZ_MAX_MODULI] = {255, 253, 251, 247, 241, 239, 233, 229, 227, 223,
                                           217, 211, 199, 197, 193, 191, 181, 179, 173, 167};

struct OzConsts {
    int n;                                  // number of moduli
    int m[OZ_MAX_MODULI];
    double inv_m[OZ_MAX_MODULI];
    double c26[OZ_MAX_MODULI];              // 2^26 mod m_j
    double w[OZ_MAX_MODULI][OZ_LIMBS];      // CRT weights in B-bit limbs, least significant first
    double Mlimb[OZ_LIMBS];                 // M in B-bit limbs
    int limbs;                              // B-bit limbs of M (and of every weight)
    double base, inv_base;                  // 2^B, 2^-B
    double M_approx;                        // M as a double (rounded)
    double inv_M;                           // 1 / M_approx
};

// ---------------- host-side multiprecision for the CRT constants ----------------
struct Big {  // little-endian base 2^24 limbs, non-negative
    std::vector<uint64_t> d;
    explicit Big(uint64_t v = 0) {
        while (v) { d.push_back(v & 0xffffff); v >>= 24; }
    }
    void mul_small(uint64_t k) {
        uint64_t carry = 0;
        for (auto &x : d) {
            uint64_t t = x * k + carry;
            x = t & 0xffffff;
            carry = t >> 24;
        }
        while (carry) { d.push_back(carry & 0xffffff); carry >>= 24; }
    }
    double to_double() const {
        double r = 0.0;
        for (size_t i = d.size(); i-- > 0;) r = r * 16777216.0 + (double)d[i];
        return r;
    }
};

// Bits [lo, lo + len) of x (len <= 64).
static uint64_t big_bits(const Big &x, int lo, int len) {
    unsigned __int128 acc = 0;
    const int d0 = lo / 24, sh = lo % 24;
    for (int i = 3; i >= 0; --i) {
        const size_t di = (size_t)(d0 + i);
        acc = (acc << 24) | (unsigned __int128)(di < x.d.size() ? x.d[di] : 0u);
    }
    acc >>= sh;  // 96 - 23 >= 64 bits remain
    return len >= 64 ? (uint64_t)acc : (uint64_t)(acc & ((((unsigned __int128)1) << len) - 1));
}
static int big_bitlen(const Big &x) {
    if (x.d.empty()) return 0;
    int b = 0;
    for (uint64_t top = x.d.back(); top; top >>= 1) ++b;
    return (int)(x.d.size() - 1) * 24 + b;
}

static int modinv(int a, int m) {  // a^-1 mod m, gcd(a, m) = 1
    int t = 0, nt = 1, r = m, nr = ((a % m) + m) % m;
    while (nr) {
        int q = r / nr, tmp = t - q * nt;
        t = nt; nt = tmp;
        tmp = r - q * nr; r = nr; nr = tmp;
    }
    return ((t % m) + m) % m;
}

static nrk_status make_consts(int n, OzConsts *c) {
    if (n < 2 || n > OZ_MAX_MODULI) return fail(NRK_ERR_INVALID_ARGUMENT, "moduli count %d outside [2,%d]", n, OZ_MAX_MODULI);
    *c = OzConsts{};
    c->n = n;
    Big M(1);
    for (int j = 0; j < n; ++j) M.mul_small((uint64_t)kModuli[j]);
    // Limb width: a sum of n products c_j w_j[t] (|c_j| <= 127, w_j[t] < 2^B) plus a carry stays
    // below 2^53, and so does a normalised limb minus q Mlimb[t] (|q| <= 127 n).
    int B = 53;
    while ((double)(127 * n + 1) * std::ldexp(1.0, B) > 0x1p53) --B;
    const int bits = big_bitlen(M);
    int NL = std::max(1, (bits + B - 1) / B);
    // the top limb carries everything above B (NL - 1) bits: 127 n (2 M_top + 1) + 1 < 2^53
    while (NL < OZ_LIMBS && 127.0 * n * (2.0 * (double)big_bits(M, B * (NL - 1), 64) + 1.0) + 1.0 >= 0x1p53) ++NL;
    if (NL > OZ_LIMBS || B * NL < bits) return fail(NRK_ERR_UNSUPPORTED, "CRT limbs do not cover M");
    c->limbs = NL;
    c->base = std::ldexp(1.0, B);
    c->inv_base = std::ldexp(1.0, -B);
    auto limb = [&](const Big &x, int t) -> double {
        if (t >= NL) return 0.0;
        return (double)big_bits(x, B * t, t == NL - 1 ? 64 : B);
    };
    for (int t = 0; t < OZ_LIMBS; ++t) c->Mlimb[t] = limb(M, t);
    c->M_approx = M.to_double();
    c->inv_M = 1.0 / c->M_approx;
    for (int j = 0; j < n; ++j) {
        const int mj = kModuli[j];
        c->m[j] = mj;
        c->inv_m[j] = 1.0 / (double)mj;
        c->c26[j] = (double)((1u << 26) % (unsigned)mj);
        Big Mj(1);
        int Mj_mod = 1;
        for (int i = 0; i < n; ++i) {
            if (i == j) continue;
            Mj.mul_small((uint64_t)kModuli[i]);
            Mj_mod = (Mj_mod * (kModuli[i] % mj)) % mj;
        }
        const int y = modinv(Mj_mod, mj);
        Mj.mul_small((uint64_t)y);  // w_j = M_j * y_j  (< M)
        for (int t = 0; t < OZ_LIMBS; ++t) c->w[j][t] = limb(Mj, t);
    }
    return NRK_OK;
}

// ---------------- 1. per-row / per-column power-of-two scales ----------------
// The scales follow the Cauchy-Schwarz rule of Ozaki-II's fast mode: |C'_ij| <= |A'_i|_2 |B'_j|_2,
// so scaling every row of A and column of B to a 2-norm of at most X, X^2 < M/2, keeps the
// integer product inside the CRT range whatever k is. For flat rows (all |a_ik| equal) this is
// the same precision as bounding every entry by X / sqrt(k); for every other row it is more.
//
// The 2-norm is accumulated as ssq 2^(2e) with the squares taken of |x| 2^-e, e the exponent of
// the largest |x| seen so far, so no finite input overflows or underflows it.
struct SsqAcc {
    double s;  // sum of (|x| 2^-e)^2
    int e;
};
constexpr int OZ_NO_EXP = -4000;  // below every double exponent: an empty accumulator
constexpr int OZ_BATCH = 8;        // loads a thread issues before folding them in
__device__ __forceinline__ void ssq_add(SsqAcc &a, double x) {
    const double ax = fabs(x);
    if (!(ax > 0.0) || ax == __builtin_inf()) return;  // zero or non-finite (recorded separately)
    int ex;
    (void)frexp(ax, &ex);  // ax = f 2^ex, f in [0.5, 1)
    if (ex > a.e) {
        a.s = ldexp(a.s, 2 * (a.e - ex));
        a.e = ex;
    }
    const double y = ldexp(ax, -a.e);
    a.s = fma(y, y, a.s);
}
__device__ __forceinline__ SsqAcc ssq_merge(SsqAcc a, SsqAcc b) {
    if (b.e > a.e) {
        const SsqAcc t = a;
        a = b;
        b = t;
    }
    a.s += ldexp(b.s, 2 * (b.e - a.e));
    return a;
}
// The largest power of two 2^s with 2^s |x|_2 <= xeff (the 2^-30 covers the rounding of the sum
// of squares, sqrt and the division); 0 for a zero vector.
__device__ __forceinline__ int cs_scale(SsqAcc t, double xeff) {
    if (!(t.s > 0.0)) return 0;
    const double r = xeff / (sqrt(t.s) * (1.0 + 0x1p-30));
    int e;
    (void)frexp(r, &e);  // r in [2^(e-1), 2^e)
    return e - 1 - t.e;
}

// scale[r] for row r of X (R x Cc, row-major): one workgroup per row.
__global__ __launch_bounds__(256) void row_scale_kernel(const double *__restrict__ X, int R, int Cc, double xeff,
                                                        int *__restrict__ scale) {
    const int r = blockIdx.x;
    SsqAcc acc{0.0, OZ_NO_EXP};
    int bad = 0;
    for (int c0 = 0; c0 < Cc; c0 += 256 * OZ_BATCH) {  // OZ_BATCH loads in flight, then the updates
        double v[OZ_BATCH];
#pragma unroll
        for (int u = 0; u < OZ_BATCH; ++u) {
            const int c = c0 + u * 256 + (int)threadIdx.x;
            v[u] = c < Cc ? X[(size_t)r * Cc + c] : 0.0;
        }
#pragma unroll
        for (int u = 0; u < OZ_BATCH; ++u) {
            bad |= !__builtin_isfinite(v[u]);
            ssq_add(acc, v[u]);
        }
    }
    __shared__ double ss[256];
    __shared__ int se[256], sbad[256];
    ss[threadIdx.x] = acc.s;
    se[threadIdx.x] = acc.e;
    sbad[threadIdx.x] = bad;
    __syncthreads();
    for (int h = 128; h > 0; h >>= 1) {
        if (threadIdx.x < h) {
            const SsqAcc m = ssq_merge(SsqAcc{ss[threadIdx.x], se[threadIdx.x]}, SsqAcc{ss[threadIdx.x + h], se[threadIdx.x + h]});
            ss[threadIdx.x] = m.s;
            se[threadIdx.x] = m.e;
            sbad[threadIdx.x] |= sbad[threadIdx.x + h];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) scale[r] = sbad[0] ? OZ_BAD_SCALE : cs_scale(SsqAcc{ss[0], se[0]}, xeff);
}

// Column 2-norms of X (R x Cc, row-major) in two passes: chunks of COL_ROWS rows per thread
// (threads of a workgroup read consecutive columns, so every row read is coalesced), then one
// thread per column over the chunk accumulators. part holds the chunks' sums of squares (+inf
// for a chunk with a non-finite entry) followed by their exponents.
constexpr int COL_ROWS = 128;
__global__ __launch_bounds__(256) void col_ssq_partial_kernel(const double *__restrict__ X, int R, int Cc,
                                                              double *__restrict__ part) {
    const int c = blockIdx.x * 256 + threadIdx.x;
    if (c >= Cc) return;
    const int r0 = blockIdx.y * COL_ROWS, r1 = min(R, r0 + COL_ROWS);
    SsqAcc acc{0.0, OZ_NO_EXP};
    int bad = 0;
    for (int b = r0; b < r1; b += OZ_BATCH) {
        double v[OZ_BATCH];
#pragma unroll
        for (int u = 0; u < OZ_BATCH; ++u) v[u] = b + u < r1 ? X[(size_t)(b + u) * Cc + c] : 0.0;
#pragma unroll
        for (int u = 0; u < OZ_BATCH; ++u) {
            bad |= !__builtin_isfinite(v[u]);
            ssq_add(acc, v[u]);
        }
    }
    part[(size_t)blockIdx.y * Cc + c] = bad ? __builtin_inf() : acc.s;
    part[(size_t)(gridDim.y + blockIdx.y) * Cc + c] = (double)acc.e;
}

__global__ __launch_bounds__(256) void col_scale_kernel(const double *__restrict__ part, int nchunk, int Cc, double xeff,
                                                        int *__restrict__ scale) {
    const int c = blockIdx.x * 256 + threadIdx.x;
    if (c >= Cc) return;
    SsqAcc acc{0.0, OZ_NO_EXP};
    bool bad = false;
    for (int h0 = 0; h0 < nchunk; h0 += OZ_BATCH) {
        double sv[OZ_BATCH], ev[OZ_BATCH];
#pragma unroll
        for (int u = 0; u < OZ_BATCH; ++u) {
            const int h = h0 + u;
            sv[u] = h < nchunk ? part[(size_t)h * Cc + c] : 0.0;
            ev[u] = h < nchunk ? part[(size_t)(nchunk + h) * Cc + c] : (double)OZ_NO_EXP;
        }
#pragma unroll
        for (int u = 0; u < OZ_BATCH; ++u) {
            bad |= __builtin_isinf(sv[u]);
            acc = ssq_merge(acc, SsqAcc{bad ? 0.0 : sv[u], (int)ev[u]});
        }
    }
    scale[c] = bad ? OZ_BAD_SCALE : cs_scale(acc, xeff);
}

// ---------------- 2. residues ----------------
// Symmetric residue of the integer v = vh 2^26 + vl (|v| <= 2^62, so |vh| <= 2^36, and
// 0 <= vl < 2^26) modulo m: t = vh (2^26 mod m) + vl is congruent to v and below 2^45 in
// magnitude (an exact double), and the FP64 t / m is within 2^-13 of the true quotient while t / m
// is at least 1/(2m) >= 2^-9 from every half-integer (m odd), so rint picks the nearest integer
// quotient: r = t - m rint(t / m) lies in [-(m-1)/2, (m-1)/2], exactly.
__device__ __forceinline__ unsigned sym_mod_byte(double vh, double vl, double m, double inv_m, double c26) {
    const double t = fma(vh, c26, vl);  // exact
    const double r = fma(-rint(t * inv_m), m, t);
    return (unsigned)(int)r & 0xffu;
}

// Each thread produces 16 consecutive K positions of one output row for every modulus and
// stores them as one 16-byte vector per modulus: out[j] is (out_rows, out_cols) with out_cols
// (the padded K) a multiple of 16. Without TRANSPOSE X is (R, Cc) = (rows, K) and consecutive
// threads take consecutive chunks of a row. With TRANSPOSE X is (R, Cc) = (K, rows), output row o
// is input column o, and a workgroup covers 16 output rows x 16 chunks (lanes 0-15 read 16
// consecutive doubles of one row of X; 16 consecutive chunks of one output row are stored by
// 16 consecutive threads' groups). Zero in the padding.
template <bool TRANSPOSE>
__global__ __launch_bounds__(256) void residue_kernel(const double *__restrict__ X, int R, int Cc,
                                                     const int *__restrict__ scale, OzConsts k,
                                                     signed char *__restrict__ out, int out_rows, int out_cols) {
    typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));
    const int nkc = out_cols / 16;
    int orow, kc;
    if (TRANSPOSE) {
        const int nob = (out_rows + 15) / 16;
        const int ob = blockIdx.x % nob, kb = blockIdx.x / nob;
        orow = ob * 16 + (threadIdx.x & 15);
        kc = kb * 16 + (threadIdx.x >> 4);
    } else {
        const size_t t = blockIdx.x * (size_t)256 + threadIdx.x;
        orow = (int)(t / nkc);
        kc = (int)(t % nkc);
    }
    if (orow >= out_rows || kc >= nkc) return;
    const int k0 = kc * 16;
    const int nrows = TRANSPOSE ? Cc : R;  // logical output rows that exist
    const int klen = TRANSPOSE ? R : Cc;   // logical K
    const int s = orow < nrows ? scale[orow] : OZ_BAD_SCALE;
    double vh[16], vl[16];  // the scaled integer split as vh 2^26 + vl, 0 <= vl < 2^26 (exact)
#pragma unroll
    for (int e = 0; e < 16; ++e) {
        const int kk = k0 + e;
        double x = 0.0;
        if (s != OZ_BAD_SCALE && kk < klen) x = TRANSPOSE ? X[(size_t)kk * Cc + orow] : X[(size_t)orow * Cc + kk];
        const double v = rint(ldexp(x, s == OZ_BAD_SCALE ? 0 : s));
        vh[e] = floor(ldexp(v, -26));
        vl[e] = fma(-vh[e], 67108864.0, v);
    }
    const size_t stride = (size_t)out_rows * out_cols;
    signed char *o = out + (size_t)orow * out_cols + k0;
    for (int j = 0; j < k.n; ++j) {
        const double m = (double)k.m[j], im = k.inv_m[j], c = k.c26[j];
        unsigned w[4];
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            unsigned x = 0;
#pragma unroll
            for (int b = 0; b < 4; ++b) x |= sym_mod_byte(vh[4 * q + b], vl[4 * q + b], m, im, c) << (8 * b);
            w[q] = x;
        }
        *reinterpret_cast<u32x4 *>(o + j * stride) = u32x4{w[0], w[1], w[2], w[3]};
    }
}

// ---------------- 3. batched int8 GEMM with a mod-m epilogue ----------------
typedef signed char i8x16_t __attribute__((ext_vector_type(16)));

// Cres[j] = (Ares[j] (Mp x Kp) * Btres[j] (Np x Kp)^T) mod m_j for modulus j = blockIdx.y, stored
// in 16 x 16 blocks in the MFMA accumulator order (see crt_kernel).
//
// Operands move global -> LDS by the load unit itself (global_load_lds, 16 bytes per lane): no
// staging registers and no LDS write instructions, so a wave's registers go to accumulators and
// fragments. NBUF LDS buffers of one 64-deep K tile each: tiles t+1 .. t+NBUF-2 are in flight
// while tile t is multiplied, one barrier per tile.
//
// LDS layout of a buffer: A rows then Bt rows, 64 bytes per row, swizzled by lds_swz.
constexpr int OZ_DMA_BK = 64;

template <int BM, int BN, int WM, int WN, int NBUF>
__global__ __launch_bounds__((BM / WM) * (BN / WN) * 64) void gemm_i8_dma_kernel(
    int M, int N, int Kp, const signed char *__restrict__ A, size_t a_stride, const signed char *__restrict__ Bt,
    size_t b_stride, signed char *__restrict__ Cres, size_t c_stride, OzConsts k, int group_m) {
    constexpr int NT = (BM / WM) * (BN / WN) * 64;
    constexpr int WGN = BN / WN;
    constexpr int TM = WM / 16, TN = WN / 16;
    constexpr int BK = OZ_DMA_BK;
    constexpr int RA = BM * 4 / NT, RB = BN * 4 / NT;  // 16-byte loads per thread per tile
    constexpr int BUF = (BM + BN) * BK;
    static_assert((BM * 4) % NT == 0 && (BN * 4) % NT == 0, "tile/threads mismatch");

    extern __shared__ __attribute__((aligned(16))) signed char smem8[];
    const int j = blockIdx.y;
    A += j * a_stride;
    Bt += j * b_stride;
    const int tid = threadIdx.x, lane = tid & 63;
    const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
    const int wm = wave / WGN, wn = wave % WGN;
    int tm, tn;
    lowp_tile_of(blockIdx.x, (M + BM - 1) / BM, (N + BN - 1) / BN, group_m, tm, tn);
    const int bm0 = tm * BM, bn0 = tn * BN;

    // Load s of this thread covers LDS chunk position p = s*NT + tid of its operand: row p/4,
    // stored chunk p%4, which holds logical chunk (p%4) ^ lds_swz(row).
    const signed char *srcA[RA], *srcB[RB];
#pragma unroll
    for (int s = 0; s < RA; ++s) {
        const int p = s * NT + tid, row = p >> 2, c = (p & 3) ^ lds_swz(row);
        srcA[s] = A + (size_t)(bm0 + row) * Kp + c * 16;
    }
#pragma unroll
    for (int s = 0; s < RB; ++s) {
        const int p = s * NT + tid, row = p >> 2, c = (p & 3) ^ lds_swz(row);
        srcB[s] = Bt + (size_t)(bn0 + row) * Kp + c * 16;
    }
    auto issue = [&](int t, int slot) {  // K tile t into LDS buffer slot
        signed char *base = smem8 + slot * BUF + wave * 64 * 16;  // this wave's 1 KiB slot
        const size_t k0 = (size_t)t * BK;
#pragma unroll
        for (int s = 0; s < RA; ++s) lds_dma16(srcA[s] + k0, base + s * NT * 16);
#pragma unroll
        for (int s = 0; s < RB; ++s) lds_dma16(srcB[s] + k0, base + BM * BK + s * NT * 16);
    };

    i32x4_t acc[TM][TN];
#pragma unroll
    for (int i = 0; i < TM; ++i)
#pragma unroll
        for (int jj = 0; jj < TN; ++jj) acc[i][jj] = i32x4_t{0, 0, 0, 0};

    const int r16 = lane & 15;
    const int pc = ((lane >> 4) ^ lds_swz(r16)) * 16;  // stored position of this lane's chunk (row = r16 mod 16)
    const int nk = Kp / BK;
    // Operand fragments in registers: the A row blocks of the tile being multiplied (each reloaded
    // with the next tile's as soon as its products are issued) and the B column blocks of two
    // tiles (the next tile's are read while this one's products run), so the matrix cores never
    // wait on LDS latency. The scheduling fences keep the compiler from sinking the reads to the
    // next barrier, where their latency would be exposed.
    i8x16_t fa[TM], fb0[TN], fb1[TN];
    auto load_a = [&](int i, int t) {
        fa[i] = *reinterpret_cast<const i8x16_t *>(smem8 + (t % NBUF) * BUF + (wm * WM + i * 16 + r16) * BK + pc);
    };
    auto load_b = [&](i8x16_t(&fb)[TN], int t) {
        const signed char *Bs = smem8 + (t % NBUF) * BUF + BM * BK;
#pragma unroll
        for (int jj = 0; jj < TN; ++jj) fb[jj] = *reinterpret_cast<const i8x16_t *>(Bs + (wn * WN + jj * 16 + r16) * BK + pc);
    };
    // products of the tile held in fa / fb; A row block i then reloaded from tile tn
    auto mma_reload = [&](const i8x16_t(&fb)[TN], int tn) {
#pragma unroll
        for (int i = 0; i < TM; ++i) {
#pragma unroll
            for (int jj = 0; jj < TN; ++jj)
                acc[i][jj] = __builtin_amdgcn_mfma_i32_16x16x64_i8(__builtin_bit_cast(i32x4_t, fa[i]),
                                                                   __builtin_bit_cast(i32x4_t, fb[jj]), acc[i][jj], 0, 0, 0);
            load_a(i, tn);
            __builtin_amdgcn_sched_barrier(0);
        }
    };
    // Every pass issues one group of loads (past the last tile, a harmless reload of the last
    // tile into the free buffer), so exactly NBUF-2 groups are in flight behind tile t and the
    // wait count is a constant: no branch around the barrier.
    auto sync_tile = [&](int t) {
        lowp_tile_barrier<(NBUF - 2) * (RA + RB)>();
        issue(min(t + NBUF - 1, nk - 1), (t + NBUF - 1) % NBUF);  // into the buffer tile t-1 used
    };
    // Kp is a multiple of 128, so the tile count nk is even and tiles pair up (t, t+1). The loop
    // body has no branch around fragment reads or MFMAs (a branch would make the compiler copy
    // fragments and accumulators between register assignments): the last pair also reads the
    // fragments of "tile nk" from a buffer holding stale data, and never uses them.
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
    // Epilogue: symmetric residues mod m_j, written in 16 x 16 blocks (block (bi, bj) at byte
    // (bi * Nb + bj) * 256, lane l's word at + 4 l holding rows 4 (l/16) + 0..3 of column l % 16,
    // its accumulator order), so each block is one 256-byte store of the wave. Blocks outside
    // ceil(M/16) x ceil(N/16) are not stored.
    const int mj = k.m[j];
    const double inv_mj = k.inv_m[j];
    const int Mb = (M + 15) >> 4, Nb = (N + 15) >> 4;
    // v - m round(v / m): v / m is never a half-integer (m odd), and |v| < 2^31 keeps v / m far
    // more accurate in FP64 than its distance to one, so this is the residue in [-(m-1)/2, (m-1)/2].
    auto smod = [&](int v) -> unsigned { return (unsigned)(v - (int)rint((double)v * inv_mj) * mj) & 0xffu; };
    unsigned *Cj = reinterpret_cast<unsigned *>(Cres + j * c_stride);
#pragma unroll
    for (int i = 0; i < TM; ++i)
#pragma unroll
        for (int jj = 0; jj < TN; ++jj) {
            const int bi = (bm0 + wm * WM + i * 16) >> 4, bj = (bn0 + wn * WN + jj * 16) >> 4;
            if (bi < Mb && bj < Nb) {
                const i32x4_t v = acc[i][jj];
                Cj[((size_t)bi * Nb + bj) * 64 + lane] = smod(v[0]) | smod(v[1]) << 8 | smod(v[2]) << 16 | smod(v[3]) << 24;
            }
        }
}

// ---------------- 4. CRT reconstruction ----------------
struct dd_t {
    double hi, lo;
};
__device__ __forceinline__ dd_t dd_two_sum(double a, double b) {
    const double s = a + b, v = s - a;
    return dd_t{s, (a - (s - v)) + (b - v)};
}

// Thread t reads word t of every modulus' residue array: the residues of block t / 64 (16 x 16
// blocks, Nb = ceil(N/16) per block row) at lane l = t % 64, i.e. rows 4 (l/16) + 0..3 of column
// l % 16, and reconstructs those four outputs. A wavefront reads 256 contiguous bytes per modulus.
// NL >= k.limbs limbs of B bits: the weights, M and every intermediate fit (limbs above k.limbs are zero).
template <int NL>
__global__ __launch_bounds__(256) void crt_kernel(const signed char *__restrict__ Cres, size_t c_stride, int M, int N,
                                                  const int *__restrict__ srow, const int *__restrict__ scol, OzConsts k,
                                                  double *__restrict__ C) {
    const size_t w = blockIdx.x * (size_t)256 + threadIdx.x;
    const int Nb = (N + 15) >> 4;
    const size_t blk = w >> 6;
    const int l = (int)(w & 63);
    const int col = (int)(blk % Nb) * 16 + (l & 15);
    const int row0 = (int)(blk / Nb) * 16 + 4 * (l >> 4);
    if (row0 >= M || col >= N) return;
    // Lx[u][t]: limb t of X_raw = sum_j c_j w_j for output row row0 + u. Every modulus' word is
    // requested up front (fully unrolled over the maximum count, uniform branches past k.n), then
    // the moduli are accumulated in turn, each one's weights read once for the four outputs.
    const unsigned *src = reinterpret_cast<const unsigned *>(Cres) + w;
    const size_t wstride = c_stride / 4;
    unsigned words[OZ_MAX_MODULI];
#pragma unroll
    for (int j = 0; j < OZ_MAX_MODULI; ++j) words[j] = j < k.n ? src[j * wstride] : 0u;
    double Lx[4][NL];
#pragma unroll
    for (int u = 0; u < 4; ++u)
#pragma unroll
        for (int t = 0; t < NL; ++t) Lx[u][t] = 0.0;
#pragma unroll
    for (int j = 0; j < OZ_MAX_MODULI; ++j) {
        if (j >= k.n) break;
        double wj[NL];
#pragma unroll
        for (int t = 0; t < NL; ++t) wj[t] = k.w[j][t];
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            const double c = (double)(signed char)(words[j] >> (8 * u));
#pragma unroll
            for (int t = 0; t < NL; ++t) Lx[u][t] = fma(c, wj[t], Lx[u][t]);  // exact: |sum| < 2^36
        }
    }
    const int sc = scol[col];
#pragma unroll
    for (int u = 0; u < 4; ++u) {
        const int row = row0 + u;
        if (row >= M) break;
        const size_t idx = (size_t)row * N + col;
        const int sr = srow[row];
        if (sr == OZ_BAD_SCALE || sc == OZ_BAD_SCALE) {
            C[idx] = __builtin_nan("");
            continue;
        }
        double L[NL];
#pragma unroll
        for (int t = 0; t < NL; ++t) L[t] = Lx[u][t];
        // balanced carry normalisation: limbs below the top in [-2^(B-1), 2^(B-1)] (exact)
        auto normalise = [&]() {
#pragma unroll
            for (int t = 0; t < NL - 1; ++t) {
                const double carry = rint(L[t] * k.inv_base);
                L[t] = fma(-carry, k.base, L[t]);
                L[t + 1] += carry;
            }
        };
        normalise();
        // X_raw = sum_t L[t] 2^(B t); subtract q M with q = round(X_raw / M)
        double approx = 0.0;
#pragma unroll
        for (int t = NL - 1; t >= 0; --t) approx = fma(approx, k.base, L[t]);
        // |C'| <= (M/2)(1 - 2^-20) by the choice of the scales, and |X_raw| < 2^12 M, so the
        // FP64 quotient approx / M is within 2^-36 of X_raw / M and rounds to the right q
        const double q = rint(approx * k.inv_M);
#pragma unroll
        for (int t = 0; t < NL; ++t) L[t] = fma(-q, k.Mlimb[t], L[t]);  // exact: below 2^53
        normalise();
        // double-double Horner from the top limb, base 2^B (every limb is an exact double)
        dd_t x{L[NL - 1], 0.0};
#pragma unroll
        for (int t = NL - 2; t >= 0; --t) {
            x.hi *= k.base;  // exact
            x.lo *= k.base;
            dd_t s = dd_two_sum(x.hi, L[t]);
            s.lo += x.lo;
            x = dd_two_sum(s.hi, s.lo);
        }
        C[idx] = ldexp(x.hi + x.lo, -(sr + sc));
    }
}

// ---------------- host driver ----------------
template <int BM, int BN, int WM, int WN, int NBUF = 4>
static nrk_status launch_i8(int M, int N, int Kp, const signed char *A, size_t sa, const signed char *Bt, size_t sb,
                            signed char *Cres, size_t sc, const OzConsts &k, hipStream_t stream) {
    constexpr int threads = (BM / WM) * (BN / WN) * 64;
    constexpr size_t lds = (size_t)NBUF * (BM + BN) * OZ_DMA_BK;
    static_assert(lds <= 160 * 1024, "tile does not fit in LDS");
    auto kern = gemm_i8_dma_kernel<BM, BN, WM, WN, NBUF>;
    static bool attr_set = false;
    if (!attr_set) {
        NRK_HIP(hipFuncSetAttribute((const void *)kern, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
        attr_set = true;
    }
    dim3 grid(((M + BM - 1) / BM) * ((N + BN - 1) / BN), k.n);
    hipLaunchKernelGGL(kern, grid, dim3(threads), lds, stream, M, N, Kp, A, sa, Bt, sb, Cres, sc, k, 8);
    NRK_HIP(hipGetLastError());
    return NRK_OK;
}

nrk_status gemm_f64_ozaki(size_t m, size_t k, size_t n, const double *a, const double *b, double *c, int moduli, int tile,
                          hipStream_t stream) {
    NRK_TRY(ensure_device());
    if (tile < -1 || tile >= OZ_TILE_CONFIGS) return fail(NRK_ERR_INVALID_ARGUMENT, "INT8 tile config %d outside [0,%d)", tile, OZ_TILE_CONFIGS);
    if (moduli == 0) moduli = OZ_DEFAULT_MODULI;
    OzConsts kc;
    NRK_TRY(make_consts(moduli, &kc));
    if (m == 0 || n == 0) return NRK_OK;
    if (k == 0) return gemm<double>(m, 0, n, a, b, c, stream);
    if (!a || !b || !c) return fail(NRK_ERR_INVALID_ARGUMENT, "A, B or C is NULL");
    if (k > 130000) return fail(NRK_ERR_INVALID_ARGUMENT, "k = %zu exceeds the exact int32 accumulation bound", k);
    if (m > (1u << 28) || n > (1u << 28)) return fail(NRK_ERR_INVALID_ARGUMENT, "dimension too large");
    // 2-norm bound of the scaled rows / columns: X^2 <= (M/2)(1 - 2^-20), capped at 2^62 (the
    // residue split), less the rounding of k entries (|round(v) - v|_2 <= sqrt(k) / 2) and a unit
    const double X = std::min(std::sqrt(kc.M_approx * 0.5 * (1.0 - 0x1p-20)), 0x1p62);
    const double xeff = X - 0.5 * std::sqrt((double)k) - 1.0;
    // worst case (all entries of a row equal in magnitude) each entry keeps log2(xeff / sqrt(k)) bits
    if (std::log2(xeff) - 0.5 * std::log2((double)k) < 8.0)
        return fail(NRK_ERR_INVALID_ARGUMENT, "too few moduli (%d) for k = %zu", moduli, k);
    const size_t Mp = align_up(m, 256), Np = align_up(n, 256), Kp = align_up(k, 128);
    const size_t sa = Mp * Kp, sb = Np * Kp, sc = ceil_div(m, 16) * ceil_div(n, 16) * 256;  // residues in 16 x 16 blocks
    const size_t nchunk = ceil_div(k, (size_t)COL_ROWS);
    const size_t res_bytes = align_up((size_t)moduli * (sa + sb + sc), 256);
    const size_t bytes = res_bytes + align_up(2 * nchunk * n * sizeof(double), 256) + (m + n) * sizeof(int);
    void *ws = nullptr;
    NRK_TRY(get_workspace(stream, bytes, &ws));
    signed char *Ares = (signed char *)ws;
    signed char *Bres = Ares + (size_t)moduli * sa;
    signed char *Cres = Bres + (size_t)moduli * sb;
    double *part = (double *)((char *)ws + res_bytes);
    int *srow = (int *)((char *)part + align_up(2 * nchunk * n * sizeof(double), 256));
    int *scol = srow + m;

    hipLaunchKernelGGL(row_scale_kernel, dim3((unsigned)m), dim3(256), 0, stream, a, (int)m, (int)k, xeff, srow);
    NRK_HIP(hipGetLastError());
    hipLaunchKernelGGL(col_ssq_partial_kernel, dim3((unsigned)ceil_div(n, 256), (unsigned)nchunk), dim3(256), 0, stream, b,
                       (int)k, (int)n, part);
    NRK_HIP(hipGetLastError());
    hipLaunchKernelGGL(col_scale_kernel, dim3((unsigned)ceil_div(n, 256)), dim3(256), 0, stream, part, (int)nchunk, (int)n,
                       xeff, scol);
    NRK_HIP(hipGetLastError());
    const size_t nkc = Kp / 16;
    hipLaunchKernelGGL((residue_kernel<false>), dim3((unsigned)ceil_div(Mp * nkc, 256)), dim3(256), 0, stream, a, (int)m,
                       (int)k, srow, kc, Ares, (int)Mp, (int)Kp);
    NRK_HIP(hipGetLastError());
    hipLaunchKernelGGL((residue_kernel<true>), dim3((unsigned)(ceil_div(Np, 16) * ceil_div(nkc, 16))), dim3(256), 0, stream,
                       b, (int)k, (int)n, scol, kc, Bres, (int)Np, (int)Kp);
    NRK_HIP(hipGetLastError());
    if (tile < 0) tile = select_config(NRK_DT_F64_OZAKI, m, n, k, stream);
    if (tile < 0) {  // default: the largest tiles that still give every CU a workgroup
        if (ceil_div(m, 256) * ceil_div(n, 256) * moduli >= (size_t)cu_count()) tile = 3;
        else if (ceil_div(m, 128) * ceil_div(n, 128) * moduli >= (size_t)cu_count() * 2) tile = 1;
        else tile = 0;
    }
    switch (tile) {
        case 0: NRK_TRY((launch_i8<64, 64, 32, 32>((int)m, (int)n, (int)Kp, Ares, sa, Bres, sb, Cres, sc, kc, stream))); break;
        case 1: NRK_TRY((launch_i8<128, 128, 64, 64>((int)m, (int)n, (int)Kp, Ares, sa, Bres, sb, Cres, sc, kc, stream))); break;
        case 2: NRK_TRY((launch_i8<128, 256, 64, 64>((int)m, (int)n, (int)Kp, Ares, sa, Bres, sb, Cres, sc, kc, stream))); break;
        case 3: NRK_TRY((launch_i8<256, 256, 128, 64>((int)m, (int)n, (int)Kp, Ares, sa, Bres, sb, Cres, sc, kc, stream))); break;
        default: NRK_TRY((launch_i8<256, 256, 128, 128>((int)m, (int)n, (int)Kp, Ares, sa, Bres, sb, Cres, sc, kc, stream)));
    }
    {  // the limbs that hold M: 3 for 10 to 16 moduli
        const dim3 g((unsigned)ceil_div(sc / 4, 256));
        switch (kc.limbs) {
            case 1: hipLaunchKernelGGL(crt_kernel<1>, g, dim3(256), 0, stream, Cres, sc, (int)m, (int)n, srow, scol, kc, c); break;
            case 2: hipLaunchKernelGGL(crt_kernel<2>, g, dim3(256), 0, stream, Cres, sc, (int)m, (int)n, srow, scol, kc, c); break;
            case 3: hipLaunchKernelGGL(crt_kernel<3>, g, dim3(256), 0, stream, Cres, sc, (int)m, (int)n, srow, scol, kc, c); break;
            default: hipLaunchKernelGGL(crt_kernel<OZ_LIMBS>, g, dim3(256), 0, stream, Cres, sc, (int)m, (int)n, srow, scol, kc, c);
        }
    }
    NRK_HIP(hipGetLastError());
    return NRK_OK;
}

}  // namespace nrk

using namespace nrk;

extern "C" {

nrk_status nrk_dev_gemm_f64_ozaki(size_t m, size_t k, size_t n, const double *a, const double *b, double *c, int32_t moduli,
                                  nrk_stream stream) {
    return gemm_f64_ozaki(m, k, n, a, b, c, moduli, -1, (hipStream_t)stream);
}

nrk_status nrk_gemm_f64_ozaki(size_t m, size_t k, size_t n, const double *a, const double *b, double *c, int32_t moduli) {
    NRK_TRY(ensure_device());
    if (m == 0 || n == 0) return NRK_OK;
    void *da = nullptr, *db = nullptr, *dc = nullptr;
    NRK_TRY(staging(0, m * k * 8 + 16, &da));
    NRK_TRY(staging(2, k * n * 8 + 16, &db));
    NRK_TRY(staging(1, m * n * 8 + 16, &dc));
    hipStream_t s = nullptr;
    if (m * k) NRK_TRY(copy_h2d(da, a, m * k * 8, s));
    if (k * n) NRK_TRY(copy_h2d(db, b, k * n * 8, s));
    NRK_TRY(gemm_f64_ozaki(m, k, n, (const double *)da, (const double *)db, (double *)dc, moduli, -1, s));
    NRK_TRY(copy_d2h(c, dc, m * n * 8, s));
    NRK_HIP(hipStreamSynchronize(s));
    return NRK_OK;
}

}  // extern "C"
