// Regular pairs with long quadrature rules (more than BEM_QUAD_FAST_POINTS points: the
// reference's symmetric rules of orders 7-20, 13..79 points, bempp_cl/api/integration/
// triangle_gauss.py:143-173), FP64.
//
// The fast regular kernels keep a tile's quadrature points in LDS and registers, sized for at
// most 16 points. For the long rules every lane takes one (test element, trial element) pair and
// loops over all point pairs (pair_fp64_geom, the moment block S[a][b] of the FP64 paths), then
// forms the local block as every other kernel does (local_value) and scales it by the local
// multipliers (bempp_cl/core/numba_kernels.py:980-1074 scalar, 1853-1969 EFIE). The rule sits in
// the kernel arguments and is read with wavefront-uniform indices. A workgroup takes one test
// element and 256 trial elements, so the test geometry is uniform too.
#include "regular_impl.h"

namespace bem {

constexpr int kWideThreads = 256;

template <int OP, int NTS, int NSS, typename OutT, bool ATOMIC>
__global__ __launch_bounds__(kWideThreads) void regular_wide_kernel(const RegularWideArgs a) {
    constexpr bool CPLX = OpTraits<OP>::is_complex;
    const int sp = blockIdx.x * kWideThreads + threadIdx.x;
    const int tp = blockIdx.y;
    if (sp >= a.trial.n) return;
    const int e = a.test.elements[tp], f = a.trial.elements[sp];
    const double *gt = a.test.geom + (int64_t)e * BEM_GEOM_STRIDE;
    const double *gf = a.trial.geom + (int64_t)f * BEM_GEOM_STRIDE;
    bool adjacent = false;
    if (a.identical) {
        for (int i = 0; i < 3; ++i)
            for (int j = 0; j < 3; ++j) adjacent |= a.test.vids[4 * e + i] == a.trial.vids[4 * f + j];
    }
    const double tnm = a.test.nmult ? (double)a.test.nmult[e] : 1.0;
    const double snm = a.trial.nmult ? (double)a.trial.nmult[f] : 1.0;
    C2d S[NTS][NSS], S2[NTS][NSS], S3[NTS][NSS];
    if (!adjacent) pair_fp64_geom<green_op<OP>(), NTS, NSS, 0>(gt, gf, tnm, snm, a.rule, a.kr, a.ki, S, S2, S3);
    if (ATOMIC && adjacent) return;
    OutT *o = reinterpret_cast<OutT *>(a.out);
    for (int i = 0; i < NTS; ++i) {
        for (int j = 0; j < NSS; ++j) {
            const double m = a.test.mult[(int64_t)e * NTS + i] * a.trial.mult[(int64_t)f * NSS + j];
            C2d v = {0.0, 0.0};
            if (!adjacent && m != 0.0) {
                v = local_value<OP, NTS, NSS>(S, i, j, gt, gf, a.kr, a.ki, tnm * snm);
                v.re *= m;
                v.im *= m;
            }
            const int64_t idx = (int64_t)a.test.l2g[(int64_t)e * NTS + i] * a.ld + a.trial.l2g[(int64_t)f * NSS + j];
            if constexpr (ATOMIC) {
                if (m == 0.0) continue;
                if constexpr (CPLX) {
                    atomic_add(o + 2 * idx, (OutT)v.re);
                    atomic_add(o + 2 * idx + 1, (OutT)v.im);
                } else {
                    atomic_add(o + idx, (OutT)v.re);
                }
            } else if (a.accumulate) {
                if constexpr (CPLX) {
                    o[2 * idx] = (OutT)((double)o[2 * idx] + v.re);
                    o[2 * idx + 1] = (OutT)((double)o[2 * idx + 1] + v.im);
                } else {
                    o[idx] = (OutT)((double)o[idx] + v.re);
                }
            } else {
                if constexpr (CPLX) {
                    o[2 * idx] = (OutT)v.re;
                    o[2 * idx + 1] = (OutT)v.im;
                } else {
                    o[idx] = (OutT)v.re;
                }
            }
        }
    }
}

template <int OP, int NTS, int NSS, typename OutT>
static int wide_launch3(const RegularWideArgs &a, bool atomic, hipStream_t s) {
    const dim3 grid((a.trial.n + kWideThreads - 1) / kWideThreads, a.test.n);
    if (grid.x == 0 || grid.y == 0) return BEM_OK;
    if (atomic)
        hipLaunchKernelGGL((regular_wide_kernel<OP, NTS, NSS, OutT, true>), grid, dim3(kWideThreads), 0, s, a);
    else
        hipLaunchKernelGGL((regular_wide_kernel<OP, NTS, NSS, OutT, false>), grid, dim3(kWideThreads), 0, s, a);
    return check_launch("regular_wide_kernel");
}

template <int OP, typename OutT>
static int wide_launch2(const RegularWideArgs &a, bool atomic, hipStream_t s) {
    const int nt = nshape_of(a.test.shape), ns = nshape_of(a.trial.shape);
    if constexpr (OP == kMaxwellEFIE || is_hypersingular<OP>()) {
        if (nt != 3 || ns != 3) {
            set_error("operator %d needs three shape functions per element on both sides", OP);
            return BEM_ERR_INVALID_ARGUMENT;
        }
        return wide_launch3<OP, 3, 3, OutT>(a, atomic, s);
    } else {
        if (nt == 1 && ns == 1) return wide_launch3<OP, 1, 1, OutT>(a, atomic, s);
        if (nt == 1 && ns == 3) return wide_launch3<OP, 1, 3, OutT>(a, atomic, s);
        if (nt == 3 && ns == 1) return wide_launch3<OP, 3, 1, OutT>(a, atomic, s);
        return wide_launch3<OP, 3, 3, OutT>(a, atomic, s);
    }
}

template <int OP>
static int wide_launch1(const RegularWideArgs &a, bool atomic, int out_dtype, hipStream_t s) {
    constexpr bool CPLX = OpTraits<OP>::is_complex;
    if (CPLX ? out_dtype == BEM_C128 : out_dtype == BEM_F64) return wide_launch2<OP, double>(a, atomic, s);
    if (CPLX ? out_dtype == BEM_C64 : out_dtype == BEM_F32) return wide_launch2<OP, float>(a, atomic, s);
    set_error("output dtype %d does not match operator %d", out_dtype, OP);
    return BEM_ERR_INVALID_ARGUMENT;
}

int regular_wide_launch(int op, const RegularWideArgs &a, bool atomic, int out_dtype, hipStream_t s) {
    switch (op) {
    case BEM_LAPLACE_SINGLE_LAYER: return wide_launch1<kLaplaceSL>(a, atomic, out_dtype, s);
    case BEM_LAPLACE_DOUBLE_LAYER: return wide_launch1<kLaplaceDL>(a, atomic, out_dtype, s);
    case BEM_LAPLACE_ADJOINT_DOUBLE_LAYER: return wide_launch1<kLaplaceADL>(a, atomic, out_dtype, s);
    case BEM_HELMHOLTZ_SINGLE_LAYER: return wide_launch1<kHelmholtzSL>(a, atomic, out_dtype, s);
    case BEM_HELMHOLTZ_DOUBLE_LAYER: return wide_launch1<kHelmholtzDL>(a, atomic, out_dtype, s);
    case BEM_HELMHOLTZ_ADJOINT_DOUBLE_LAYER: return wide_launch1<kHelmholtzADL>(a, atomic, out_dtype, s);
    case BEM_MAXWELL_ELECTRIC_FIELD: return wide_launch1<kMaxwellEFIE>(a, atomic, out_dtype, s);
    case BEM_LAPLACE_HYPERSINGULAR: return wide_launch1<kLaplaceHS>(a, atomic, out_dtype, s);
    case BEM_HELMHOLTZ_HYPERSINGULAR: return wide_launch1<kHelmholtzHS>(a, atomic, out_dtype, s);
    default:
        set_error("operator %d takes at most %d quadrature points", op, BEM_QUAD_FAST_POINTS);
        return BEM_ERR_UNSUPPORTED;
    }
}

}  // namespace bem
