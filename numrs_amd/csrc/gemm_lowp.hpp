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

}  // namespace nrk
