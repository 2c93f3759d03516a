// Low-precision MFMA GEMM tile kernels shared by the split-precision GEMMs and the ragged batch.
//
// Both kernels take "NT" operands: A is (M, K) row-major and B is given TRANSPOSED, Bt (N, K)
// row-major, so that every MFMA operand a lane supplies is 16 contiguous bytes along K:
//   v_mfma_f32_16x16x32_bf16: lane l supplies A[l%16][8*(l/16) .. +7] and Bt[l%16][8*(l/16) .. +7]
//   v_mfma_i32_16x16x64_i8:   lane l supplies A[l%16][16*(l/16) .. +15] and Bt[l%16][16*(l/16) .. +15]
// and accumulator element r of lane l is D[4*(l/16) + r][l%16].
#pragma once

#include "common.hpp"

namespace nrk {

typedef __bf16 bf16x8_t __attribute__((ext_vector_type(8)));
typedef float f32x4_t __attribute__((ext_vector_type(4)));
typedef int i32x4_t __attribute__((ext_vector_type(4)));
typedef unsigned int u32x4_t __attribute__((ext_vector_type(4)));

// Round-to-nearest-even f32 -> bf16 bits (finite inputs).
__device__ __forceinline__ unsigned short f32_to_bf16_rn(float x) {
    unsigned int u = __float_as_uint(x);
    u += 0x7fffu + ((u >> 16) & 1u);
    return (unsigned short)(u >> 16);
}
__device__ __forceinline__ float bf16_to_f32(unsigned short h) { return __uint_as_float((unsigned int)h << 16); }

// Grouped, XCD-aware tile order (hardware puts workgroup g on XCD g % 8).
__device__ __forceinline__ void lowp_tile_of(int g, int nbm, int nbn, int group_m, int &tm, int &tn) {
    const int total = nbm * nbn;
    if (total % 8 == 0) {
        const int per = total / 8;
        g = (g % 8) * per + g / 8;
    }
    const int in_group = group_m * nbn;
    const int grp = g / in_group;
    const int first_m = grp * group_m;
    const int gm = min(nbm - first_m, group_m);
    const int r = g % in_group;
    tm = first_m + r % gm;
    tn = r / gm;
}

// ---- LDS tiles filled by the load unit (global_load_lds), shared by the INT8 and BF16 kernels ----
// A K tile is 64 bytes per row (four 16-byte chunks: 64 int8 or 32 bf16 values), rows of all
// operand terms stacked. Chunk c of row r is stored at chunk position c ^ lds_swz(r). A fragment
// read puts lane l on row l % 16, chunk l / 16. Measured on gfx950 (SQ_LDS_BANK_CONFLICT, INT8
// 8192^3 kernel with the loads switched off): xor with (r >> 1) & 3 makes these reads
// conflict-free and halves the LDS busy cycles; no swizzle, and xors with r & 3, (r >> 2) & 3 or
// ((r >> 1) & 3) ^ ((r >> 3) & 1), all leave them 2-way conflicted.
__device__ __forceinline__ int lds_swz(int r) { return (r >> 1) & 3; }

// One 16-byte global -> LDS load per lane: lane i's bytes land at l + 16 i (l must be the same
// for the whole wavefront; it goes to M0).
__device__ __forceinline__ void lds_dma16(const void *g, void *l) {
    __builtin_amdgcn_global_load_lds((const void __attribute__((address_space(1))) *)g,
                                     (void __attribute__((address_space(3))) *)l, 16, 0, 0);
}

// s_waitcnt vmcnt(n) lgkmcnt(0), expcnt unconstrained (gfx9 encoding: vmcnt in bits 3:0 and
// 15:14, expcnt 6:4, lgkmcnt 11:8).
#define LOWP_WAIT_VM_LGKM0(n) __builtin_amdgcn_s_waitcnt(((n) & 15) | ((((n) >> 4) & 3) << 14) | (0x7 << 4))

// Start of a K tile's pass in a pipeline that always has NPEND of this thread's loads issued
// after the tile's own. Before the barrier: this wave's loads of the tile have landed and its LDS
// reads of the previous tile are complete. After it, the tile is visible to every wave and nobody
// reads the previous tile's buffer any more, so the caller may stream a new tile into it. A bare
// s_barrier: the workgroup fence of __syncthreads would also wait for the later tiles' loads; the
// empty asm statements keep the compiler from moving LDS accesses across it.
template <int NPEND>
__device__ __forceinline__ void lowp_tile_barrier() {
    static_assert(NPEND >= 0 && NPEND <= 63, "vmcnt range");
    LOWP_WAIT_VM_LGKM0(NPEND);
    asm volatile("" ::: "memory");
    __builtin_amdgcn_s_barrier();
    asm volatile("" ::: "memory");
}

}  // namespace nrk
