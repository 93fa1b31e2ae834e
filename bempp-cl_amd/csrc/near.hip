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
// SYM (DUAL, one space): each pair stands for both orientations. The single-layer block goes to
// (e, f) and, transposed, to (f, e); the double layer to (e, f); the double layer of the reversed
// pair (same points, test normal) to (f, e). The tensor-product rule makes this exact.
template <int OP, int NTS, int NSS, int QT, typename OutT, bool DUAL = false, bool SYM = false>
__global__ __launch_bounds__(256) void near_kernel(const NearArgs a) {
    static_assert(!SYM || (DUAL && NTS == NSS), "the symmetric near pass pairs one space with itself");
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
    if (CPLX && fp64_phase_reduced(a.kr, a.ki, sqrt(dist2), gt[13] + gf[13]))
        pair_fp64_geom<green_op<OP>(), NTS, NSS, QT, DUAL, SYM, kPhReduced | kPhReal>(gt, gf, tnm, snm, a.rule, a.kr,
                                                                                      a.ki, S, S2, S3);
    else
        pair_fp64_geom<green_op<OP>(), NTS, NSS, QT, DUAL, SYM>(gt, gf, tnm, snm, a.rule, a.kr, a.ki, S, S2, S3);
    // to: 0 adds block (i, j) at (e_i, f_j), 1 at (f_j, e_i), 2 at both
    auto scatter = [&](const C2d(&Sm)[NTS][NSS], auto op_tag, void *out, int to) {
        constexpr int OPX = decltype(op_tag)::value;
        OutT *o = reinterpret_cast<OutT *>(out);
        auto add = [&](int64_t idx, const C2d &v, double m) {
            if constexpr (CPLX) {
                atomic_add(o + 2 * idx, (OutT)(v.re * m));
                atomic_add(o + 2 * idx + 1, (OutT)(v.im * m));
            } else {
                atomic_add(o + idx, (OutT)(v.re * m));
            }
        };
        for (int i = 0; i < NTS; ++i) {
            for (int j = 0; j < NSS; ++j) {
                double m = a.test.mult[(int64_t)e * NTS + i] * a.trial.mult[(int64_t)f * NSS + j];
                if (m == 0.0) continue;
                C2d v = local_value<OPX, NTS, NSS>(Sm, i, j, gt, gf, a.kr, a.ki, tnm * snm);
                const int64_t r = a.test.l2g[(int64_t)e * NTS + i], c = a.trial.l2g[(int64_t)f * NSS + j];
                if (to != 1) add(r * a.ld + c, v, m);
                if (to != 0) add(c * a.ld + r, v, m);
            }
        }
    };
    scatter(S, std::integral_constant<int, OP>{}, a.out, SYM ? 2 : 0);
    if constexpr (DUAL) scatter(S2, std::integral_constant<int, OP + 1>{}, a.out2, 0);
    if constexpr (SYM) scatter(S3, std::integral_constant<int, OP + 1>{}, a.out2, 1);
}

// The same pairs with a group of kNearGroup lanes per pair (6-point rule): lane g takes test point
// g with the trial points unrolled, so a lane carries only its point's trial moments and the
// kernel runs at several waves per SIMD instead of one. The moment blocks are summed over the
// group with xor shuffles, then the group's lanes share the entries of the scatter. When every
// pair of the wave has k r <= 1 at all its point pairs (bound: centroid distance plus both
// diameters), e^{ikr} comes from its Taylor series (sincos_small) without range reduction.
constexpr int kNearGroup = 8;

template <int OP, int NTS, int NSS, int QT, typename OutT, bool DUAL = false, bool SYM = false>
__global__ __launch_bounds__(256) void near_group_kernel(const NearArgs a) {
    static_assert(QT > 0 && QT <= kNearGroup, "one test point per lane of the group");
    static_assert(!SYM || (DUAL && NTS == NSS), "the symmetric near pass pairs one space with itself");
    constexpr bool CPLX = OpTraits<OP>::is_complex;
    constexpr int NB = NTS * NSS;
    constexpr int NOPS = SYM ? 3 : DUAL ? 2 : 1;
    const int g = threadIdx.x & (kNearGroup - 1);
    const int pid = (int)((blockIdx.x * (int64_t)blockDim.x + threadIdx.x) / kNearGroup);
    // every lane of a group sees the same pair and takes the same branches, so the shuffles
    // below only read lanes of active groups
    bool take = pid < a.n_pairs;
    int e = 0, f = 0;
    const double *gt = a.test.geom, *gf = a.trial.geom;
    double dist2 = 0.0;
    if (take) {
        e = a.pairs[2 * (int64_t)pid];
        f = a.pairs[2 * (int64_t)pid + 1];
        gt = a.test.geom + (int64_t)e * BEM_GEOM_STRIDE;
        gf = a.trial.geom + (int64_t)f * BEM_GEOM_STRIDE;
        if (a.identical) {
            for (int i = 0; i < 3; ++i)
                for (int j = 0; j < 3; ++j)
                    if (a.test.vids[4 * e + i] == a.trial.vids[4 * f + j]) take = false;
        }
        double ct[3], cf[3];
        centroid_rn(gt, ct);
        centroid_rn(gf, cf);
        if (take && !pair_is_near(ct, cf, gt[13], gf[13], a.near_factor, &dist2)) take = false;
    }
    const double rmax = sqrt(dist2) + gt[13] + gf[13];  // bounds every point distance of the pair
    const bool small = __all(!take || fabs(a.kr) * rmax <= 1.0);  // wave-uniform
    const bool reduced = a.ki == 0.0 && __all(!take || fabs(a.kr) * rmax < 0x1p25);
    if (!take) return;  // whole groups leave together

    const double tnm = a.test.nmult ? (double)a.test.nmult[e] : 1.0;
    const double snm = a.trial.nmult ? (double)a.trial.nmult[f] : 1.0;
    const double nx[3] = {gt[9] * tnm, gt[10] * tnm, gt[11] * tnm};
    const double ny[3] = {gf[9] * snm, gf[10] * snm, gf[11] * snm};
    // this lane's moment blocks: S[o][a][b] = u_a(xi_p) w_p J_e sum_q u_b(eta_q) w_q J_f G_o
    C2d S[NOPS][NTS][NSS];
#pragma unroll
    for (int o = 0; o < NOPS; ++o) zero_moments(S[o]);
    if (g < QT) {
        const int p = g;
        const double xi = a.rule.x[p], eta = a.rule.y[p];
        const double x0 = gt[0] + gt[3] * xi + gt[6] * eta;
        const double x1 = gt[1] + gt[4] * xi + gt[7] * eta;
        const double x2 = gt[2] + gt[5] * xi + gt[8] * eta;
        auto points = [&](auto phase_tag) {
            constexpr int SM = decltype(phase_tag)::value;
            C2d T[NOPS][NSS];
#pragma unroll
            for (int o = 0; o < NOPS; ++o)
#pragma unroll
                for (int b = 0; b < NSS; ++b) T[o][b].re = T[o][b].im = 0.0;
#pragma unroll
            for (int q = 0; q < QT; ++q) {
                const double s0 = a.rule.x[q], s1 = a.rule.y[q];
                const double dx = gf[0] + gf[3] * s0 + gf[6] * s1 - x0;
                const double dy = gf[1] + gf[4] * s0 + gf[7] * s1 - x1;
                const double dz = gf[2] + gf[5] * s0 + gf[8] * s1 - x2;
                const double r2 = dx * dx + dy * dy + dz * dz;
                double v[3][2];
                if constexpr (SYM)
                    green_sl_dl_rev<CPLX, double, SM>(dx, dy, dz, r2, nx, ny, a.kr, a.ki, v[0][0], v[0][1], v[1][0],
                                                      v[1][1], v[2][0], v[2][1]);
                else
                    green2<green_op<OP>(), DUAL, double, SM>(dx, dy, dz, r2, nx, ny, a.kr, a.ki, v[0][0], v[0][1], v[1][0],
                                                 v[1][1]);
                const double wq = a.rule.w[q] * gf[12];
#pragma unroll
                for (int o = 0; o < NOPS; ++o) add_trial_moments<NSS>(T[o], v[o][0] * wq, v[o][1] * wq, s0, s1);
            }
#pragma unroll
            for (int o = 0; o < NOPS; ++o) add_test_moments<NTS, NSS>(S[o], T[o], a.rule.w[p] * gt[12], xi, eta);
        };
        if (small && a.ki == 0.0) points(std::integral_constant<int, kPhSmall | kPhReal>{});
        else if (small) points(std::integral_constant<int, kPhSmall>{});
        else if (reduced) points(std::integral_constant<int, kPhReduced | kPhReal>{});
        else points(std::integral_constant<int, 0>{});
    }
    // sum over the group
#pragma unroll
    for (int o = 0; o < NOPS; ++o)
#pragma unroll
        for (int i = 0; i < NTS; ++i)
#pragma unroll
            for (int j = 0; j < NSS; ++j)
#pragma unroll
                for (int off = kNearGroup / 2; off >= 1; off >>= 1) {
                    S[o][i][j].re += __shfl_xor(S[o][i][j].re, off, kNearGroup);
                    if constexpr (CPLX) S[o][i][j].im += __shfl_xor(S[o][i][j].im, off, kNearGroup);
                }
    // lane g adds entries l = g, g + group, ... of the NOPS x NB blocks. SYM: op 0 (single layer)
    // at (e, f) and transposed at (f, e); op 1 at (e, f); op 2 (reversed pair) at (f, e).
    for (int l = g; l < NOPS * NB; l += kNearGroup) {
        const int o = l / NB, ij = l - o * NB, i = ij / NSS, j = ij - i * NSS;
        const double m = a.test.mult[(int64_t)e * NTS + i] * a.trial.mult[(int64_t)f * NSS + j];
        if (m == 0.0) continue;
        // compile-time block indices (a runtime index would put S in scratch)
        C2d v;
        if (o == 0) v = local_value<OP, NTS, NSS>(S[0], i, j, gt, gf, a.kr, a.ki, tnm * snm);
        else if (o == 1) v = local_value<(DUAL ? OP + 1 : OP), NTS, NSS>(S[DUAL ? 1 : 0], i, j, gt, gf, a.kr, a.ki);
        else v = local_value<(DUAL ? OP + 1 : OP), NTS, NSS>(S[NOPS - 1], i, j, gt, gf, a.kr, a.ki);
        const int64_t r = a.test.l2g[(int64_t)e * NTS + i], c = a.trial.l2g[(int64_t)f * NSS + j];
        OutT *out = reinterpret_cast<OutT *>(o == 0 ? a.out : a.out2);
        auto add = [&](int64_t idx) {
            if constexpr (CPLX) {
                atomic_add(out + 2 * idx, (OutT)(v.re * m));
                atomic_add(out + 2 * idx + 1, (OutT)(v.im * m));
            } else {
                atomic_add(out + idx, (OutT)(v.re * m));
            }
        };
        if (o != 2) add(r * a.ld + c);                     // (e, f)
        if (SYM && (o == 0 || o == 2)) add(c * a.ld + r);  // (f, e)
    }
}

template <int OP, int NTS, int NSS, typename OutT, bool DUAL, bool SYM = false>
static int near_launch3(const NearArgs &a, hipStream_t s) {
    if (a.rule.n == 6) {
        dim3 grid((unsigned)(((int64_t)a.n_pairs * kNearGroup + 255) / 256));
        hipLaunchKernelGGL((near_group_kernel<OP, NTS, NSS, 6, OutT, DUAL, SYM>), grid, dim3(256), 0, s, a);
        return check_launch("near_group_kernel");
    }
    dim3 grid((a.n_pairs + 255) / 256);
    hipLaunchKernelGGL((near_kernel<OP, NTS, NSS, 0, OutT, DUAL, SYM>), grid, dim3(256), 0, s, a);
    return check_launch("near_kernel");
}

template <int OP, typename OutT, bool DUAL, bool SYM = false>
static int near_launch2(const NearArgs &a, hipStream_t s) {
    if constexpr (OP == kMaxwellEFIE) {
        return near_launch3<OP, 3, 3, OutT, false>(a, s);
    } else if constexpr (is_hypersingular<OP>()) {
        if (nshape_of(a.test.shape) != 3 || nshape_of(a.trial.shape) != 3) {
            set_error("hypersingular operators need P1 shape functions on both sides");
            return BEM_ERR_INVALID_ARGUMENT;
        }
        return near_launch3<OP, 3, 3, OutT, false>(a, s);
    } else if constexpr (SYM) {
        const int nt = nshape_of(a.test.shape);
        if (nt == 1) return near_launch3<OP, 1, 1, OutT, true, true>(a, s);
        return near_launch3<OP, 3, 3, OutT, true, true>(a, s);
    } else {
        int nt = nshape_of(a.test.shape), ns = nshape_of(a.trial.shape);
        if (nt == 1 && ns == 1) return near_launch3<OP, 1, 1, OutT, DUAL>(a, s);
        if (nt == 1 && ns == 3) return near_launch3<OP, 1, 3, OutT, DUAL>(a, s);
        if (nt == 3 && ns == 1) return near_launch3<OP, 3, 1, OutT, DUAL>(a, s);
        return near_launch3<OP, 3, 3, OutT, DUAL>(a, s);
    }
}

template <int OP, bool DUAL = false, bool SYM = false>
static int near_launch_op(const NearArgs &a, int out_dtype, hipStream_t s) {
    constexpr bool CPLX = OpTraits<OP>::is_complex;
    if (CPLX && out_dtype == BEM_C128) return near_launch2<OP, double, DUAL, SYM>(a, s);
    if (CPLX && out_dtype == BEM_C64) return near_launch2<OP, float, DUAL, SYM>(a, s);
    if (!CPLX && out_dtype == BEM_F64) return near_launch2<OP, double, DUAL, SYM>(a, s);
    if (!CPLX && out_dtype == BEM_F32) return near_launch2<OP, float, DUAL, SYM>(a, s);
    set_error("output dtype %d does not match operator %d", out_dtype, OP);
    return BEM_ERR_INVALID_ARGUMENT;
}

}  // namespace bem

using namespace bem;

static int near_entry(int op, const bem_space_t *test, const bem_space_t *trial, int n_pairs, const int32_t *pairs,
                      int grids_identical, const bem_quad_rule_t *rule, double k_re, double k_im, double near_factor,
                      void *out, void *out2, int64_t ld, int out_dtype, void *stream, bool symmetric = false) {
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
    if (symmetric) {
        if (op == BEM_LAPLACE_SINGLE_LAYER) return near_launch_op<kLaplaceSL, true, true>(a, out_dtype, s);
        if (op == BEM_HELMHOLTZ_SINGLE_LAYER) return near_launch_op<kHelmholtzSL, true, true>(a, out_dtype, s);
        set_error("the symmetric single/double layer pair needs a single-layer operator, got %d", op);
        return BEM_ERR_UNSUPPORTED;
    }
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
    case BEM_LAPLACE_HYPERSINGULAR: return near_launch_op<kLaplaceHS>(a, out_dtype, s);
    case BEM_HELMHOLTZ_HYPERSINGULAR: return near_launch_op<kHelmholtzHS>(a, out_dtype, s);
    case BEM_MAXWELL_MAGNETIC_FIELD:
        return mfie_near_launch(a.test, a.trial, a.rule, k_re, k_im, grids_identical, near_factor, n_pairs, pairs, out,
                                ld, out_dtype, s);
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

extern "C" int bem_near_assemble_sl_dl_symmetric(int op, const bem_space_t *space, int n_pairs, const int32_t *pairs,
                                                 const bem_quad_rule_t *rule, double k_re, double k_im,
                                                 double mixed_near_factor, void *out_single, void *out_double,
                                                 int64_t ld, int out_dtype, void *stream) {
    if (!out_double) {
        clear_error();
        set_error("null argument");
        return BEM_ERR_INVALID_ARGUMENT;
    }
    return near_entry(op, space, space, n_pairs, pairs, 1, rule, k_re, k_im, mixed_near_factor, out_single,
                      out_double, ld, out_dtype, stream, true);
}
