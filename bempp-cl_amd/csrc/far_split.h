// Device helpers shared by the frequency-batched and adjoint kernels: wavefront reductions
// and the FP32 reference-distance split of a well-separated element pair.
//
// For test/trial centroids c_e, c_f the offset d = c_f - c_e and r0 = |d| (rounded to FP32)
// come from FP64; a point pair then only needs t = |d + b_q - a_p|^2 - r0^2 in FP32
// (a, b: offsets of the quadrature points from the centroids) and delta = r - r0 =
// t / (r + r0), so phases k delta stay small and the FP32 rounding error is relative to the
// element size, not to the distance. The phase e^{i k r0} is reduced in FP64.
#pragma once

#include "bem_device.h"

namespace bem {

template <typename T>
__device__ __forceinline__ T wave_max64(T v) {
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) v = max(v, __shfl_xor(v, off, 64));
    return v;
}

template <typename T>
__device__ __forceinline__ T wave_sum64(T v) {
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
}

// Sum over the 64 lanes of a converged wave, on the VALU (DPP) instead of LDS permutes: row
// shifts by 1, 2, 4, 8 leave each 16-lane row's sum in its last lane, two row broadcasts carry
// rows 0 + 1 and rows 2 + 3 into lane 63, which is read back as a wave-uniform value.
template <int CTRL, int ROWS>
__device__ __forceinline__ float dpp_moved(float v) {
    return __builtin_bit_cast(float, __builtin_amdgcn_update_dpp(0, __builtin_bit_cast(int, v), CTRL, ROWS, 0xf, true));
}

__device__ __forceinline__ float wave_total_dpp(float v) {
    v += dpp_moved<0x111, 0xf>(v);  // row_shr:1
    v += dpp_moved<0x112, 0xf>(v);  // row_shr:2
    v += dpp_moved<0x114, 0xf>(v);  // row_shr:4
    v += dpp_moved<0x118, 0xf>(v);  // row_shr:8
    v += dpp_moved<0x142, 0xa>(v);  // row_bcast:15 into rows 1 and 3
    v += dpp_moved<0x143, 0xc>(v);  // row_bcast:31 into rows 2 and 3
    return __builtin_bit_cast(float, __builtin_amdgcn_readlane(__builtin_bit_cast(int, v), 63));
}

// Two wave sums at once: permlane32_swap leaves a's upper half beside b's lower half, so one
// add gives a summed over lanes i and i + 32 in lanes 0-31 and b likewise in lanes 32-63; row
// shifts and one row broadcast then finish a in lane 31 and b in lane 63.
__device__ __forceinline__ void wave_total2_dpp(float a, float b, float &sa, float &sb) {
    const auto sw = __builtin_amdgcn_permlane32_swap(__builtin_bit_cast(unsigned, a), __builtin_bit_cast(unsigned, b),
                                                     false, false);
    // copy the two halves out before the bit casts: a bit cast applied to sw[1] directly reads
    // element 0 with this compiler
    const unsigned lo = sw[0], hi = sw[1];
    float v = __builtin_bit_cast(float, lo) + __builtin_bit_cast(float, hi);
    v += dpp_moved<0x111, 0xf>(v);  // row_shr:1
    v += dpp_moved<0x112, 0xf>(v);  // row_shr:2
    v += dpp_moved<0x114, 0xf>(v);  // row_shr:4
    v += dpp_moved<0x118, 0xf>(v);  // row_shr:8
    v += dpp_moved<0x142, 0xa>(v);  // row_bcast:15 into rows 1 and 3
    sa = __builtin_bit_cast(float, __builtin_amdgcn_readlane(__builtin_bit_cast(int, v), 31));
    sb = __builtin_bit_cast(float, __builtin_amdgcn_readlane(__builtin_bit_cast(int, v), 63));
}

struct FarPair {
    float fd0, fd1, fd2;  // centroid offset d = c_f - c_e (FP32)
    float r0, r0sq, resid;
};

__device__ __forceinline__ FarPair far_pair_setup(const double *ce, const double *cf) {
    FarPair P;
    const double d0 = cf[0] - ce[0], d1 = cf[1] - ce[1], d2 = cf[2] - ce[2];
    const double dist2 = d0 * d0 + d1 * d1 + d2 * d2;
    P.r0 = sqrtf((float)dist2);
    P.r0sq = P.r0 * P.r0;
    P.resid = (float)(dist2 - (double)P.r0 * (double)P.r0);
    P.fd0 = (float)d0;
    P.fd1 = (float)d1;
    P.fd2 = (float)d2;
    return P;
}

// e^{i k r0} with the phase k r0 / (2 pi) reduced to one revolution in FP64; r0rev = r0 / (2 pi)
__device__ __forceinline__ void phase_rev_f32(double k, double r0rev, float &c, float &s) {
    double rev = k * r0rev;
    rev -= rint(rev);
    c = __builtin_amdgcn_cosf((float)rev);
    s = __builtin_amdgcn_sinf((float)rev);
}

// Integration-element gradient of a flat triangle: J = |e1 x e2| with e1 = v1 - v0,
// e2 = v2 - v0 gives dJ/dv1 = e2 x n, dJ/dv2 = n x e1, dJ/dv0 = -(dJ/dv1 + dJ/dv2).
__device__ __forceinline__ void jacobian_gradient(const double *g, double (&dJ)[3][3]) {
    const double e1[3] = {g[3], g[4], g[5]}, e2[3] = {g[6], g[7], g[8]}, n[3] = {g[9], g[10], g[11]};
    dJ[1][0] = e2[1] * n[2] - e2[2] * n[1];
    dJ[1][1] = e2[2] * n[0] - e2[0] * n[2];
    dJ[1][2] = e2[0] * n[1] - e2[1] * n[0];
    dJ[2][0] = n[1] * e1[2] - n[2] * e1[1];
    dJ[2][1] = n[2] * e1[0] - n[0] * e1[2];
    dJ[2][2] = n[0] * e1[1] - n[1] * e1[0];
    for (int d = 0; d < 3; ++d) dJ[0][d] = -(dJ[1][d] + dJ[2][d]);
}

}  // namespace bem
