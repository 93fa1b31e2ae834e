// Singular (Sauter-Schwab) pairs, FP64, for gfx950.
//
// A group of G lanes (8, or 32 for small launches) per (test element, trial element) pair that
// shares a vertex, an edge or the whole triangle (bempp_cl/core/numba_kernels.py:1471-1526 scalar,
// 1973-2061 EFIE); a wavefront takes 64 / G consecutive pairs, which are mostly of one kind (the
// pair table lists coincident, then edge, then vertex pairs). The pair's 512..1536 Duffy
// points are strided over the group's lanes; each lane accumulates the 3x3 moment block
// S[a][b] = sum_k G_k w_k u_a(xi_k) u_b(eta_k) with u = (1, xi_0, xi_1), the block is summed
// over the group with log2(G) shuffle steps (a reduction over the whole wavefront would cost 6
// steps for one pair), scaled by J_test J_trial and turned into the local matrix block exactly as
// for regular pairs. The group's lanes add the multiplied entries to the dense output with
// hardware atomics, which replaces the reference's host-side np.add.at scatter
// (bempp_cl/core/dense_assembler.py:78-98).
#include "bem_device.h"
#include "bem_internal.h"

namespace bem {

struct SingularArgs {
    SpaceArgs test, trial;
    int n_pairs;
    const int32_t *pairs;
    const double *tpts, *spts, *wts;
    double kr, ki;
    void *out;
    void *out2;  // DUAL: double layer (out: single layer)
    int64_t ld;
};

// lanes per pair: 8, or 32 for launches of fewer than kSingSmallPairs pairs, whose 8-lane groups
// would give fewer than four waves per SIMD on 256 CUs (1,280 triangles: 16,580 pairs, 2,073 waves;
// there 32 lanes per pair take 0.033 instead of 0.048 ms, while at 66,500 EFIE pairs they are 25 %
// slower than 8)
constexpr int kSingGroup = 8;
constexpr int kSingGroupSmall = 32;
constexpr int kSingSmallPairs = 32768;

template <int G, typename T>
__device__ __forceinline__ T group_sum(T v) {
#pragma unroll
    for (int off = G / 2; off >= 1; off >>= 1) v += __shfl_xor(v, off, G);
    return v;
}

// DUAL: OP is a single layer; the double layer of the same family (sharing r and e^{ikr} at
// every point) goes to a.out2.
template <int OP, int NTS, int NSS, typename OutT, bool DUAL = false, int G = kSingGroup>
__global__ __launch_bounds__(256) void singular_kernel(const SingularArgs a) {
    constexpr bool CPLX = OpTraits<OP>::is_complex;
    const int lane = threadIdx.x & (G - 1);  // lane within the pair's group
    const int pair = (int)(((int64_t)blockIdx.x * blockDim.x + threadIdx.x) / G);
    // the group's lanes share the pair, so whole groups leave together and the group shuffles
    // below only read lanes of groups that stay
    if (pair >= a.n_pairs) return;
    const int32_t *pd = a.pairs + 6 * (int64_t)pair;
    const int e = pd[0], f = pd[1], toff = pd[2], soff = pd[3], woff = pd[4], np = pd[5];
    const double *gt = a.test.geom + (int64_t)e * BEM_GEOM_STRIDE;
    const double *gf = a.trial.geom + (int64_t)f * BEM_GEOM_STRIDE;
    double tnm = a.test.nmult ? (double)a.test.nmult[e] : 1.0;
    double snm = a.trial.nmult ? (double)a.trial.nmult[f] : 1.0;
    double nx[3] = {gt[9] * tnm, gt[10] * tnm, gt[11] * tnm};
    double ny[3] = {gf[9] * snm, gf[10] * snm, gf[11] * snm};

    double Sr[NTS][NSS], Si[NTS][NSS], Sr2[DUAL ? NTS : 1][DUAL ? NSS : 1], Si2[DUAL ? NTS : 1][DUAL ? NSS : 1];
#pragma unroll
    for (int i = 0; i < NTS; ++i)
#pragma unroll
        for (int j = 0; j < NSS; ++j) {
            Sr[i][j] = Si[i][j] = 0.0;
            if constexpr (DUAL) Sr2[i][j] = Si2[i][j] = 0.0;
        }

    // every distance within the pair is at most diam_t + diam_f (triangles sharing a vertex), so a
    // wave-uniform bound selects the phase mode: the series sin/cos without range reduction when
    // k r <= 1, else (real wavenumbers) the branch-free reduction
    auto points = [&](auto phase_tag) {
        constexpr int PH = decltype(phase_tag)::value;
        for (int k = lane; k < np; k += G) {
            double xi0 = a.tpts[2 * (int64_t)(toff + k)], xi1 = a.tpts[2 * (int64_t)(toff + k) + 1];
            double et0 = a.spts[2 * (int64_t)(soff + k)], et1 = a.spts[2 * (int64_t)(soff + k) + 1];
            double w = a.wts[woff + k];
            double dx = (gf[0] + gf[3] * et0 + gf[6] * et1) - (gt[0] + gt[3] * xi0 + gt[6] * xi1);
            double dy = (gf[1] + gf[4] * et0 + gf[7] * et1) - (gt[1] + gt[4] * xi0 + gt[7] * xi1);
            double dz = (gf[2] + gf[5] * et0 + gf[8] * et1) - (gt[2] + gt[5] * xi0 + gt[8] * xi1);
            double r2 = dx * dx + dy * dy + dz * dz;
            double gre, gim, g2re, g2im;
            green2<green_op<OP>(), DUAL, double, PH>(dx, dy, dz, r2, nx, ny, a.kr, a.ki, gre, gim, g2re, g2im);
            gre *= w;
            gim *= w;
            g2re *= w;
            g2im *= w;
            double ut[3] = {1.0, xi0, xi1}, us[3] = {1.0, et0, et1};
#pragma unroll
            for (int i = 0; i < NTS; ++i)
#pragma unroll
                for (int j = 0; j < NSS; ++j) {
                    double u = ut[i] * us[j];
                    Sr[i][j] += gre * u;
                    Si[i][j] += gim * u;
                    if constexpr (DUAL) {
                        Sr2[i][j] += g2re * u;
                        Si2[i][j] += g2im * u;
                    }
                }
        }
    };
    const double kd = fabs(a.kr) * (gt[13] + gf[13]);
    if (kd <= 1.0 && a.ki == 0.0) points(std::integral_constant<int, kPhSmall | kPhReal>{});
    else if (kd <= 1.0) points(std::integral_constant<int, kPhSmall>{});
    else if (kd < 0x1p25 && a.ki == 0.0) points(std::integral_constant<int, kPhReduced | kPhReal>{});
    else points(std::integral_constant<int, 0>{});
    const double jj = gt[12] * gf[12];
    C2d S[NTS][NSS], S2[NTS][NSS];
#pragma unroll
    for (int i = 0; i < NTS; ++i)
#pragma unroll
        for (int j = 0; j < NSS; ++j) {
            S[i][j].re = group_sum<G>(Sr[i][j]) * jj;
            S[i][j].im = CPLX ? group_sum<G>(Si[i][j]) * jj : 0.0;
            if constexpr (DUAL) {
                S2[i][j].re = group_sum<G>(Sr2[i][j]) * jj;
                S2[i][j].im = CPLX ? group_sum<G>(Si2[i][j]) * jj : 0.0;
            }
        }
    // the group's lanes add entries l = lane, lane + G, ... (DUAL: l >= NTS*NSS are the double
    // layer's)
    const int nb = NTS * NSS;
    for (int ll = lane; ll < (DUAL ? 2 * nb : nb); ll += G) {
        const bool second = DUAL && ll >= nb;
        const int l = second ? ll - nb : ll;
        int i = l / NSS, j = l % NSS;
        double m = a.test.mult[(int64_t)e * NTS + i] * a.trial.mult[(int64_t)f * NSS + j];
        if (m != 0.0) {
            C2d v;
            if constexpr (DUAL) {
                v = second ? local_value<OP + 1, NTS, NSS>(S2, i, j, gt, gf, a.kr, a.ki)
                           : local_value<OP, NTS, NSS>(S, i, j, gt, gf, a.kr, a.ki);
            } else {
                v = local_value<OP, NTS, NSS>(S, i, j, gt, gf, a.kr, a.ki, tnm * snm);
            }
            int row = a.test.l2g[(int64_t)e * NTS + i];
            int col = a.trial.l2g[(int64_t)f * NSS + j];
            OutT *o = reinterpret_cast<OutT *>(second ? a.out2 : a.out);
            int64_t idx = (int64_t)row * a.ld + col;
            if constexpr (CPLX) {
                atomic_add(o + 2 * idx, (OutT)(v.re * m));
                atomic_add(o + 2 * idx + 1, (OutT)(v.im * m));
            } else {
                atomic_add(o + idx, (OutT)(v.re * m));
            }
        }
    }
}

template <int OP, typename OutT, bool DUAL, int G>
static int launch_sing_g(const SingularArgs &a, hipStream_t s) {
    dim3 grid((unsigned)(((int64_t)a.n_pairs * G + 255) / 256));
    if constexpr (OP == kMaxwellEFIE) {
        hipLaunchKernelGGL((singular_kernel<OP, 3, 3, OutT, false, G>), grid, dim3(256), 0, s, a);
    } else if constexpr (is_hypersingular<OP>()) {
        if (nshape_of(a.test.shape) != 3 || nshape_of(a.trial.shape) != 3) {
            set_error("hypersingular operators need P1 shape functions on both sides");
            return BEM_ERR_INVALID_ARGUMENT;
        }
        hipLaunchKernelGGL((singular_kernel<OP, 3, 3, OutT, false, G>), grid, dim3(256), 0, s, a);
    } else {
        int nt = nshape_of(a.test.shape), ns = nshape_of(a.trial.shape);
        if (nt == 1 && ns == 1)
            hipLaunchKernelGGL((singular_kernel<OP, 1, 1, OutT, DUAL, G>), grid, dim3(256), 0, s, a);
        else if (nt == 1 && ns == 3)
            hipLaunchKernelGGL((singular_kernel<OP, 1, 3, OutT, DUAL, G>), grid, dim3(256), 0, s, a);
        else if (nt == 3 && ns == 1)
            hipLaunchKernelGGL((singular_kernel<OP, 3, 1, OutT, DUAL, G>), grid, dim3(256), 0, s, a);
        else
            hipLaunchKernelGGL((singular_kernel<OP, 3, 3, OutT, DUAL, G>), grid, dim3(256), 0, s, a);
    }
    return check_launch("singular_kernel");
}

template <int OP, typename OutT, bool DUAL>
static int launch_sing(const SingularArgs &a, hipStream_t s) {
    if (a.n_pairs == 0) return BEM_OK;
    if (a.n_pairs < kSingSmallPairs) return launch_sing_g<OP, OutT, DUAL, kSingGroupSmall>(a, s);
    return launch_sing_g<OP, OutT, DUAL, kSingGroup>(a, s);
}

template <int OP, bool DUAL = false>
static int launch_sing_op(const SingularArgs &a, int out_dtype, hipStream_t s) {
    constexpr bool CPLX = OpTraits<OP>::is_complex;
    if (CPLX && out_dtype == BEM_C128) return launch_sing<OP, double, DUAL>(a, s);
    if (CPLX && out_dtype == BEM_C64) return launch_sing<OP, float, DUAL>(a, s);
    if (!CPLX && out_dtype == BEM_F64) return launch_sing<OP, double, DUAL>(a, s);
    if (!CPLX && out_dtype == BEM_F32) return launch_sing<OP, float, DUAL>(a, s);
    set_error("output dtype %d does not match operator %d", out_dtype, OP);
    return BEM_ERR_INVALID_ARGUMENT;
}

}  // namespace bem

using namespace bem;

static int singular_entry(int op, const bem_space_t *test, const bem_space_t *trial, int n_pairs,
                          const int32_t *pair_data, const double *test_points, const double *trial_points,
                          const double *weights, double k_re, double k_im, void *out, void *out2, int64_t ld,
                          int out_dtype, void *stream) {
    clear_error();
    if (!test || !trial || !out || (n_pairs > 0 && (!pair_data || !test_points || !trial_points || !weights))) {
        set_error("null argument");
        return BEM_ERR_INVALID_ARGUMENT;
    }
    if (n_pairs < 0) {
        set_error("negative pair count");
        return BEM_ERR_INVALID_ARGUMENT;
    }
    int rc;
    if ((rc = check_space(test, "test")) || (rc = check_space(trial, "trial"))) return rc;
    SingularArgs a;
    a.test = to_args(test);
    a.trial = to_args(trial);
    a.n_pairs = n_pairs;
    a.pairs = pair_data;
    a.tpts = test_points;
    a.spts = trial_points;
    a.wts = weights;
    a.kr = k_re;
    a.ki = k_im;
    a.out = out;
    a.out2 = out2;
    a.ld = ld;
    hipStream_t s = (hipStream_t)stream;
    if (out2) {
        if (op == BEM_LAPLACE_SINGLE_LAYER) return launch_sing_op<kLaplaceSL, true>(a, out_dtype, s);
        if (op == BEM_HELMHOLTZ_SINGLE_LAYER) return launch_sing_op<kHelmholtzSL, true>(a, out_dtype, s);
        set_error("the single/double layer pair needs a single-layer operator, got %d", op);
        return BEM_ERR_UNSUPPORTED;
    }
    switch (op) {
    case BEM_LAPLACE_SINGLE_LAYER: return launch_sing_op<kLaplaceSL>(a, out_dtype, s);
    case BEM_LAPLACE_DOUBLE_LAYER: return launch_sing_op<kLaplaceDL>(a, out_dtype, s);
    case BEM_LAPLACE_ADJOINT_DOUBLE_LAYER: return launch_sing_op<kLaplaceADL>(a, out_dtype, s);
    case BEM_HELMHOLTZ_SINGLE_LAYER: return launch_sing_op<kHelmholtzSL>(a, out_dtype, s);
    case BEM_HELMHOLTZ_DOUBLE_LAYER: return launch_sing_op<kHelmholtzDL>(a, out_dtype, s);
    case BEM_HELMHOLTZ_ADJOINT_DOUBLE_LAYER: return launch_sing_op<kHelmholtzADL>(a, out_dtype, s);
    case BEM_MAXWELL_ELECTRIC_FIELD: return launch_sing_op<kMaxwellEFIE>(a, out_dtype, s);
    case BEM_LAPLACE_HYPERSINGULAR: return launch_sing_op<kLaplaceHS>(a, out_dtype, s);
    case BEM_HELMHOLTZ_HYPERSINGULAR: return launch_sing_op<kHelmholtzHS>(a, out_dtype, s);
    case BEM_MAXWELL_MAGNETIC_FIELD:
        return mfie_singular_launch(a.test, a.trial, n_pairs, pair_data, test_points, trial_points, weights, k_re,
                                    k_im, out, ld, out_dtype, s);
    default: set_error("unsupported operator %d", op); return BEM_ERR_UNSUPPORTED;
    }
}

extern "C" int bem_singular_assemble(int op, const bem_space_t *test, const bem_space_t *trial, int n_pairs,
                                     const int32_t *pair_data, const double *test_points, const double *trial_points,
                                     const double *weights, double k_re, double k_im, void *out, int64_t ld,
                                     int out_dtype, void *stream) {
    return singular_entry(op, test, trial, n_pairs, pair_data, test_points, trial_points, weights, k_re, k_im, out,
                          nullptr, ld, out_dtype, stream);
}

extern "C" int bem_singular_assemble_sl_dl(int op, const bem_space_t *test, const bem_space_t *trial, int n_pairs,
                                           const int32_t *pair_data, const double *test_points,
                                           const double *trial_points, const double *weights, double k_re, double k_im,
                                           void *out_single, void *out_double, int64_t ld, int out_dtype,
                                           void *stream) {
    if (!out_double) {
        clear_error();
        set_error("null argument");
        return BEM_ERR_INVALID_ARGUMENT;
    }
    return singular_entry(op, test, trial, n_pairs, pair_data, test_points, trial_points, weights, k_re, k_im,
                          out_single, out_double, ld, out_dtype, stream);
}
