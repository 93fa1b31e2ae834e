// Near-field regular pairs of the mixed-precision mode, FP64.
//
// The FP32 regular kernel skips every non-adjacent pair whose centroids are closer than
// factor * mean diameter (pair_is_near); this kernel computes exactly those pairs in FP64
// from a host-built candidate list (a superset: pairs that fail the same test, or share a
// vertex, are ignored here). One lane per pair; the local block is multiplied by the local
// multipliers and added to the dense output with hardware atomics.
#include "regular_impl.h"

namespace bem {

struct NearArgs {
    SpaceArgs test, trial;
    RuleArgs rule;
    int n_pairs;
    const int32_t *pairs;
    double kr, ki, near_factor;
    void *out;
    void *out2;  // DUAL: double layer (out: single layer)
    int64_t ld;
    int identical;
};

// DUAL: OP is a single layer; the double layer of the same family goes to a.out2.
template <int OP, int NTS, int NSS, int QT, typename OutT, bool DUAL = false>
__global__ __launch_bounds__(256) void near_kernel(const NearArgs a) {
    constexpr bool CPLX = OpTraits<OP>::is_complex;
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= a.n_pairs) return;
    const int e = a.pairs[2 * (int64_t)p], f = a.pairs[2 * (int64_t)p + 1];
    const double *gt = a.test.geom + (int64_t)e * BEM_GEOM_STRIDE;
    const double *gf = a.trial.geom + (int64_t)f * BEM_GEOM_STRIDE;
    if (a.identical) {
        for (int i = 0; i < 3; ++i)
            for (int j = 0; j < 3; ++j)
                if (a.test.vids[4 * e + i] == a.trial.vids[4 * f + j]) return;
    }
    double ct[3], cf[3], dist2;
    centroid_rn(gt, ct);
    centroid_rn(gf, cf);
    if (!pair_is_near(ct, cf, gt[13], gf[13], a.near_factor, &dist2)) return;
    double tnm = a.test.nmult ? (double)a.test.nmult[e] : 1.0;
    double snm = a.trial.nmult ? (double)a.trial.nmult[f] : 1.0;
    C2d S[NTS][NSS], S2[NTS][NSS], S3[NTS][NSS];
    pair_fp64_geom<OP, NTS, NSS, QT, DUAL>(gt, gf, tnm, snm, a.rule, a.kr, a.ki, S, S2, S3);
    auto scatter = [&](const C2d(&Sm)[NTS][NSS], auto op_tag, void *out) {
        constexpr int OPX = decltype(op_tag)::value;
        OutT *o = reinterpret_cast<OutT *>(out);
        for (int i = 0; i < NTS; ++i) {
            for (int j = 0; j < NSS; ++j) {
                double m = a.test.mult[(int64_t)e * NTS + i] * a.trial.mult[(int64_t)f * NSS + j];
                if (m == 0.0) continue;
                C2d v = local_value<OPX, NTS, NSS>(Sm, i, j, gt, gf, a.kr, a.ki);
                int64_t idx = (int64_t)a.test.l2g[(int64_t)e * NTS + i] * a.ld + a.trial.l2g[(int64_t)f * NSS + j];
                if constexpr (CPLX) {
                    atomic_add(o + 2 * idx, (OutT)(v.re * m));
                    atomic_add(o + 2 * idx + 1, (OutT)(v.im * m));
                } else {
                    atomic_add(o + idx, (OutT)(v.re * m));
                }
            }
        }
    };
    scatter(S, std::integral_constant<int, OP>{}, a.out);
    if constexpr (DUAL) scatter(S2, std::integral_constant<int, OP + 1>{}, a.out2);
}

template <int OP, int NTS, int NSS, typename OutT, bool DUAL>
static int near_launch3(const NearArgs &a, hipStream_t s) {
    dim3 grid((a.n_pairs + 255) / 256);
    if (a.rule.n == 6) hipLaunchKernelGGL((near_kernel<OP, NTS, NSS, 6, OutT, DUAL>), grid, dim3(256), 0, s, a);
    else hipLaunchKernelGGL((near_kernel<OP, NTS, NSS, 0, OutT, DUAL>), grid, dim3(256), 0, s, a);
    return check_launch("near_kernel");
}

template <int OP, typename OutT, bool DUAL>
static int near_launch2(const NearArgs &a, hipStream_t s) {
    if constexpr (OP == kMaxwellEFIE) {
        return near_launch3<OP, 3, 3, OutT, false>(a, s);
    } else {
        int nt = nshape_of(a.test.shape), ns = nshape_of(a.trial.shape);
        if (nt == 1 && ns == 1) return near_launch3<OP, 1, 1, OutT, DUAL>(a, s);
        if (nt == 1 && ns == 3) return near_launch3<OP, 1, 3, OutT, DUAL>(a, s);
        if (nt == 3 && ns == 1) return near_launch3<OP, 3, 1, OutT, DUAL>(a, s);
        return near_launch3<OP, 3, 3, OutT, DUAL>(a, s);
    }
}

template <int OP, bool DUAL = false>
static int near_launch_op(const NearArgs &a, int out_dtype, hipStream_t s) {
    constexpr bool CPLX = OpTraits<OP>::is_complex;
    if (CPLX && out_dtype == BEM_C128) return near_launch2<OP, double, DUAL>(a, s);
    if (CPLX && out_dtype == BEM_C64) return near_launch2<OP, float, DUAL>(a, s);
    if (!CPLX && out_dtype == BEM_F64) return near_launch2<OP, double, DUAL>(a, s);
    if (!CPLX && out_dtype == BEM_F32) return near_launch2<OP, float, DUAL>(a, s);
    set_error("output dtype %d does not match operator %d", out_dtype, OP);
    return BEM_ERR_INVALID_ARGUMENT;
}

}  // namespace bem

using namespace bem;

static int near_entry(int op, const bem_space_t *test, const bem_space_t *trial, int n_pairs, const int32_t *pairs,
                      int grids_identical, const bem_quad_rule_t *rule, double k_re, double k_im, double near_factor,
                      void *out, void *out2, int64_t ld, int out_dtype, void *stream) {
    clear_error();
    if (!test || !trial || !rule || !out || (n_pairs > 0 && !pairs) || n_pairs < 0) {
        set_error("invalid argument");
        return BEM_ERR_INVALID_ARGUMENT;
    }
    int rc;
    if ((rc = check_space(test, "test")) || (rc = check_space(trial, "trial"))) return rc;
    if (rule->n_points < 1 || rule->n_points > 16) {
        set_error("quadrature rule must have 1..16 points");
        return BEM_ERR_INVALID_ARGUMENT;
    }
    if (n_pairs == 0) return BEM_OK;
    NearArgs a;
    a.test = to_args(test);
    a.trial = to_args(trial);
    a.rule = to_args(rule);
    a.n_pairs = n_pairs;
    a.pairs = pairs;
    a.kr = k_re;
    a.ki = k_im;
    a.near_factor = near_factor;
    a.out = out;
    a.out2 = out2;
    a.ld = ld;
    a.identical = grids_identical;
    hipStream_t s = (hipStream_t)stream;
    if (out2) {
        if (op == BEM_LAPLACE_SINGLE_LAYER) return near_launch_op<kLaplaceSL, true>(a, out_dtype, s);
        if (op == BEM_HELMHOLTZ_SINGLE_LAYER) return near_launch_op<kHelmholtzSL, true>(a, out_dtype, s);
        set_error("the single/double layer pair needs a single-layer operator, got %d", op);
        return BEM_ERR_UNSUPPORTED;
    }
    switch (op) {
    case BEM_LAPLACE_SINGLE_LAYER: return near_launch_op<kLaplaceSL>(a, out_dtype, s);
    case BEM_LAPLACE_DOUBLE_LAYER: return near_launch_op<kLaplaceDL>(a, out_dtype, s);
    case BEM_LAPLACE_ADJOINT_DOUBLE_LAYER: return near_launch_op<kLaplaceADL>(a, out_dtype, s);
    case BEM_HELMHOLTZ_SINGLE_LAYER: return near_launch_op<kHelmholtzSL>(a, out_dtype, s);
    case BEM_HELMHOLTZ_DOUBLE_LAYER: return near_launch_op<kHelmholtzDL>(a, out_dtype, s);
    case BEM_HELMHOLTZ_ADJOINT_DOUBLE_LAYER: return near_launch_op<kHelmholtzADL>(a, out_dtype, s);
    case BEM_MAXWELL_ELECTRIC_FIELD: return near_launch_op<kMaxwellEFIE>(a, out_dtype, s);
    default: set_error("unsupported operator %d", op); return BEM_ERR_UNSUPPORTED;
    }
}

extern "C" int bem_near_assemble(int op, const bem_space_t *test, const bem_space_t *trial, int n_pairs,
                                 const int32_t *pairs, int grids_identical, const bem_quad_rule_t *rule, double k_re,
                                 double k_im, double near_factor, void *out, int64_t ld, int out_dtype, void *stream) {
    return near_entry(op, test, trial, n_pairs, pairs, grids_identical, rule, k_re, k_im, near_factor, out, nullptr, ld,
                      out_dtype, stream);
}

extern "C" int bem_near_assemble_sl_dl(int op, const bem_space_t *test, const bem_space_t *trial, int n_pairs,
                                       const int32_t *pairs, int grids_identical, const bem_quad_rule_t *rule,
                                       double k_re, double k_im, double near_factor, void *out_single,
                                       void *out_double, int64_t ld, int out_dtype, void *stream) {
    if (!out_double) {
        clear_error();
        set_error("null argument");
        return BEM_ERR_INVALID_ARGUMENT;
    }
    return near_entry(op, test, trial, n_pairs, pairs, grids_identical, rule, k_re, k_im, near_factor, out_single,
                      out_double, ld, out_dtype, stream);
}
