// Regular-pair dense Galerkin assembly for gfx950.
//
// Computes, for every (test element e, trial element f) pair that shares no vertex (or
// every pair when the grids differ),
//   A[g_e(i), g_f(j)] += m_e(i) m_f(j) sum_p sum_q w_p J_e w_q J_f phi_i(xi_p) psi_j(eta_q) G(x_p, y_q)
// (bempp_cl/core/numba_kernels.py:980-1074), and the Maxwell EFIE block
//   sum_pq G w w J J [ -ik phi_i.psi_j - 4 / (ik J_e J_f) ] l_i l_j m_i m_j
// (numba_kernels.py:1853-1969).
//
// Work decomposition: a 256-thread workgroup owns a tile of TM test elements x 256 trial
// elements. Each lane keeps one trial element (its quadrature points, weights, normal)
// in registers and walks the TM test elements staged in LDS, so every LDS read is a
// broadcast. Lanes of a wave hold consecutive trial elements, which makes the DP0 store
// of a matrix row coalesced.
//
// Per pair the kernel forms the 3x3 moment block
//   S[a][b] = sum_pq G'_pq u_a(xi_p) u_b(eta_q),  u = (1, xi_0, xi_1),  G' = G w_p J_e w_q J_f
// and derives every local block from it: DP0 uses S[0][0], P1 uses C^T S C with
// phi = (u0 - u1 - u2, u1, u2), and the RWG/SNC block uses the affine structure of the
// reference RWG values (xi - o_i) with o = (0,1), (1,0), (0,0).
//
// MIXED precision: pairs far apart (centroid distance >= factor * mean diameter) are
// evaluated in FP32 on coordinates relative to the element centroids, with the centroid
// difference formed in FP64, and summed with Kahan compensation; near pairs use FP64.
#include "bem_device.h"
#include "bem_internal.h"

namespace bem {

constexpr int kRegThreads = 256;
constexpr int kMaxQ = 16;

struct RegularArgs {
    SpaceArgs test, trial;
    RuleArgs rule;
    double kr, ki;
    double near_factor;
    void *out;
    int64_t ld;
    int identical;
    int atomic;
    int accumulate;
};

// ------------------------------------------------------------------ per-pair FP64 core
// Points are recomputed from the geometry rows; used by the FP64 mode (through the tile
// arrays) and by the near-pair path of the MIXED mode.
template <int OP, int NTS, int NSS, int QT>
__device__ __forceinline__ void pair_fp64_geom(const double *__restrict__ gt, const double *__restrict__ gf, double tnm,
                                               double snm, const RuleArgs &rule, double kr, double ki,
                                               C2d (&S)[NTS][NSS]) {
    const int nq = QT > 0 ? QT : rule.n;
    double nx[3] = {gt[9] * tnm, gt[10] * tnm, gt[11] * tnm};
    double ny[3] = {gf[9] * snm, gf[10] * snm, gf[11] * snm};
    double jt = gt[12], jf = gf[12];
#pragma unroll
    for (int a = 0; a < NTS; ++a)
#pragma unroll
        for (int b = 0; b < NSS; ++b) S[a][b].re = S[a][b].im = 0.0;
#pragma unroll(QT > 0 ? QT : 1)
    for (int p = 0; p < nq; ++p) {
        double xi = rule.x[p], eta = rule.y[p];
        double x0 = gt[0] + gt[3] * xi + gt[6] * eta;
        double x1 = gt[1] + gt[4] * xi + gt[7] * eta;
        double x2 = gt[2] + gt[5] * xi + gt[8] * eta;
        C2d T[NSS];
#pragma unroll
        for (int b = 0; b < NSS; ++b) T[b].re = T[b].im = 0.0;
#pragma unroll(QT > 0 ? QT : 1)
        for (int q = 0; q < nq; ++q) {
            double s0 = rule.x[q], s1 = rule.y[q];
            double dx = gf[0] + gf[3] * s0 + gf[6] * s1 - x0;
            double dy = gf[1] + gf[4] * s0 + gf[7] * s1 - x1;
            double dz = gf[2] + gf[5] * s0 + gf[8] * s1 - x2;
            double r2 = dx * dx + dy * dy + dz * dz;
            double gre, gim;
            green<OP, double>(dx, dy, dz, r2, nx, ny, kr, ki, gre, gim);
            double wq = rule.w[q] * jf;
            gre *= wq;
            gim *= wq;
            T[0].re += gre;
            T[0].im += gim;
            if constexpr (NSS == 3) {
                T[1].re += gre * s0;
                T[1].im += gim * s0;
                T[2].re += gre * s1;
                T[2].im += gim * s1;
            }
        }
        double wp = rule.w[p] * jt;
        double u[3] = {wp, wp * xi, wp * eta};
#pragma unroll
        for (int a = 0; a < NTS; ++a)
#pragma unroll
            for (int b = 0; b < NSS; ++b) {
                S[a][b].re += u[a] * T[b].re;
                S[a][b].im += u[a] * T[b].im;
            }
    }
}

template <typename OutT, bool CPLX>
__device__ __forceinline__ void store_entry(void *out, int64_t ld, int row, int col, C2d v, int atomic,
                                            int accumulate) {
    OutT *o = reinterpret_cast<OutT *>(out);
    int64_t idx = (int64_t)row * ld + col;
    if constexpr (CPLX) {
        OutT *p = o + 2 * idx;
        if (atomic) {
            atomic_add(p, (OutT)v.re);
            atomic_add(p + 1, (OutT)v.im);
        } else if (accumulate) {
            p[0] += (OutT)v.re;
            p[1] += (OutT)v.im;
        } else {
            // one 8- or 16-byte store of the complex pair
            using V2 = typename std::conditional<sizeof(OutT) == 8, double2, float2>::type;
            V2 val;
            val.x = (OutT)v.re;
            val.y = (OutT)v.im;
            *reinterpret_cast<V2 *>(p) = val;
        }
    } else {
        OutT *p = o + idx;
        if (atomic) atomic_add(p, (OutT)v.re);
        else if (accumulate) p[0] += (OutT)v.re;
        else p[0] = (OutT)v.re;
    }
}

// ------------------------------------------------------------------ kernel
template <int OP, int NTS, int NSS, int QT, int TM, bool MIXED, bool KAHAN, typename OutT>
__global__ __launch_bounds__(kRegThreads) void regular_kernel(const RegularArgs a) {
    constexpr bool CPLX = OpTraits<OP>::is_complex;
    constexpr int NT = NTS, NS = NSS;  // local block size equals the moment size
    const int nq = QT > 0 ? QT : a.rule.n;

    __shared__ int s_elem[TM];
    __shared__ int s_vid[TM][3];
    __shared__ float s_off[TM][kMaxQ][3];  // MIXED: test point offsets from the centroid
    __shared__ float s_wt[TM][kMaxQ];       // MIXED: w_p J_e
    __shared__ float s_nx[TM][3];
    __shared__ double s_c[TM][3];
    __shared__ double s_diam[TM];

    const int tid = threadIdx.x;
    const int tile0 = blockIdx.y * TM;
    const int jf = blockIdx.x * kRegThreads + tid;
    const bool valid = jf < a.trial.n;

    if (tid < TM && tile0 + tid < a.test.n) {
        int e = a.test.elements[tile0 + tid];
        const double *g = a.test.geom + (int64_t)e * BEM_GEOM_STRIDE;
        s_elem[tid] = e;
        s_vid[tid][0] = a.test.vids[4 * e];
        s_vid[tid][1] = a.test.vids[4 * e + 1];
        s_vid[tid][2] = a.test.vids[4 * e + 2];
        double c0 = g[0] + (g[3] + g[6]) / 3.0, c1 = g[1] + (g[4] + g[7]) / 3.0, c2 = g[2] + (g[5] + g[8]) / 3.0;
        s_c[tid][0] = c0;
        s_c[tid][1] = c1;
        s_c[tid][2] = c2;
        s_diam[tid] = g[13];
        if constexpr (MIXED) {
            double nm = a.test.nmult ? (double)a.test.nmult[e] : 1.0;
            s_nx[tid][0] = (float)(g[9] * nm);
            s_nx[tid][1] = (float)(g[10] * nm);
            s_nx[tid][2] = (float)(g[11] * nm);
            for (int p = 0; p < nq; ++p) {
                double xi = a.rule.x[p], eta = a.rule.y[p];
                s_off[tid][p][0] = (float)(g[3] * (xi - 1.0 / 3.0) + g[6] * (eta - 1.0 / 3.0));
                s_off[tid][p][1] = (float)(g[4] * (xi - 1.0 / 3.0) + g[7] * (eta - 1.0 / 3.0));
                s_off[tid][p][2] = (float)(g[5] * (xi - 1.0 / 3.0) + g[8] * (eta - 1.0 / 3.0));
                s_wt[tid][p] = (float)(a.rule.w[p] * g[12]);
            }
        }
    }
    __syncthreads();

    // trial element data in registers
    int f = 0;
    int fv0 = -1, fv1 = -1, fv2 = -1;
    double snm = 1.0;
    const double *gf = a.trial.geom;
    double cf0 = 0, cf1 = 0, cf2 = 0, diam_f = 0;
    float bq[MIXED ? kMaxQ : 1][3];
    float wq[MIXED ? kMaxQ : 1];
    float nyf[3] = {0.f, 0.f, 0.f};
    if (valid) {
        f = a.trial.elements[jf];
        gf = a.trial.geom + (int64_t)f * BEM_GEOM_STRIDE;
        fv0 = a.trial.vids[4 * f];
        fv1 = a.trial.vids[4 * f + 1];
        fv2 = a.trial.vids[4 * f + 2];
        snm = a.trial.nmult ? (double)a.trial.nmult[f] : 1.0;
        cf0 = gf[0] + (gf[3] + gf[6]) / 3.0;
        cf1 = gf[1] + (gf[4] + gf[7]) / 3.0;
        cf2 = gf[2] + (gf[5] + gf[8]) / 3.0;
        diam_f = gf[13];
        if constexpr (MIXED) {
            nyf[0] = (float)(gf[9] * snm);
            nyf[1] = (float)(gf[10] * snm);
            nyf[2] = (float)(gf[11] * snm);
#pragma unroll(QT > 0 ? QT : 1)
            for (int q = 0; q < nq; ++q) {
                double s0 = a.rule.x[q] - 1.0 / 3.0, s1 = a.rule.y[q] - 1.0 / 3.0;
                bq[q][0] = (float)(gf[3] * s0 + gf[6] * s1);
                bq[q][1] = (float)(gf[4] * s0 + gf[7] * s1);
                bq[q][2] = (float)(gf[5] * s0 + gf[8] * s1);
                wq[q] = (float)(a.rule.w[q] * gf[12]);
            }
        }
    }

    const int ntile = min(TM, a.test.n - tile0);
    for (int t = 0; t < ntile; ++t) {
        const int e = s_elem[t];
        const double *gt = a.test.geom + (int64_t)e * BEM_GEOM_STRIDE;
        bool adj = false;
        if (a.identical) {
            int v0 = s_vid[t][0], v1 = s_vid[t][1], v2 = s_vid[t][2];
            adj = v0 == fv0 || v0 == fv1 || v0 == fv2 || v1 == fv0 || v1 == fv1 || v1 == fv2 || v2 == fv0 ||
                  v2 == fv1 || v2 == fv2;
        }
        C2d S[NTS][NSS];
#pragma unroll
        for (int i = 0; i < NTS; ++i)
#pragma unroll
            for (int j = 0; j < NSS; ++j) S[i][j].re = S[i][j].im = 0.0;

        if (valid && !adj) {
            double tnm = a.test.nmult ? (double)a.test.nmult[e] : 1.0;
            if constexpr (!MIXED) {
                pair_fp64_geom<OP, NTS, NSS, QT>(gt, gf, tnm, snm, a.rule, a.kr, a.ki, S);
            } else {
                double d0 = cf0 - s_c[t][0], d1 = cf1 - s_c[t][1], d2 = cf2 - s_c[t][2];
                double dist2 = d0 * d0 + d1 * d1 + d2 * d2;
                double lim = a.near_factor * 0.5 * (s_diam[t] + diam_f);
                if (dist2 < lim * lim) {
                    pair_fp64_geom<OP, NTS, NSS, QT>(gt, gf, tnm, snm, a.rule, a.kr, a.ki, S);
                } else {
                    // FP32 far pair: d = c_f - c_e (FP64 -> FP32), y - x = d + b_q - a_p
                    const float fd0 = (float)d0, fd1 = (float)d1, fd2 = (float)d2;
                    const float kr = (float)a.kr, ki = (float)a.ki;
                    const float *nx = s_nx[t];
                    Acc<float, KAHAN> Sr[NTS][NSS], Si[NTS][NSS];
#pragma unroll
                    for (int i = 0; i < NTS; ++i)
#pragma unroll
                        for (int j = 0; j < NSS; ++j) {
                            Sr[i][j].zero();
                            Si[i][j].zero();
                        }
#pragma unroll(QT > 0 ? QT : 1)
                    for (int p = 0; p < nq; ++p) {
                        float ox = fd0 - s_off[t][p][0], oy = fd1 - s_off[t][p][1], oz = fd2 - s_off[t][p][2];
                        float Tr[NSS], Ti[NSS];
#pragma unroll
                        for (int b = 0; b < NSS; ++b) Tr[b] = Ti[b] = 0.f;
#pragma unroll(QT > 0 ? QT : 1)
                        for (int q = 0; q < nq; ++q) {
                            float dx = ox + bq[q][0], dy = oy + bq[q][1], dz = oz + bq[q][2];
                            float r2 = dx * dx + dy * dy + dz * dz;
                            float gre, gim;
                            green<OP, float>(dx, dy, dz, r2, nx, nyf, kr, ki, gre, gim);
                            gre *= wq[q];
                            gim *= wq[q];
                            Tr[0] += gre;
                            Ti[0] += gim;
                            if constexpr (NSS == 3) {
                                float s0 = (float)a.rule.x[q], s1 = (float)a.rule.y[q];
                                Tr[1] += gre * s0;
                                Ti[1] += gim * s0;
                                Tr[2] += gre * s1;
                                Ti[2] += gim * s1;
                            }
                        }
                        float wp = s_wt[t][p];
                        float u[3] = {wp, wp * (float)a.rule.x[p], wp * (float)a.rule.y[p]};
#pragma unroll
                        for (int i = 0; i < NTS; ++i)
#pragma unroll
                            for (int j = 0; j < NSS; ++j) {
                                Sr[i][j].add(u[i] * Tr[j]);
                                Si[i][j].add(u[i] * Ti[j]);
                            }
                    }
#pragma unroll
                    for (int i = 0; i < NTS; ++i)
#pragma unroll
                        for (int j = 0; j < NSS; ++j) {
                            S[i][j].re = (double)Sr[i][j].s - (double)Sr[i][j].c;
                            S[i][j].im = (double)Si[i][j].s - (double)Si[i][j].c;
                        }
                }
            }
        }

        if (valid) {
#pragma unroll
            for (int i = 0; i < NT; ++i) {
                double mi = a.test.mult[(int64_t)e * NT + i];
                int row = a.test.l2g[(int64_t)e * NT + i];
#pragma unroll
                for (int j = 0; j < NS; ++j) {
                    double mj = a.trial.mult[(int64_t)f * NS + j];
                    int col = a.trial.l2g[(int64_t)f * NS + j];
                    C2d v = local_value<OP, NTS, NSS>(S, i, j, gt, gf, a.kr, a.ki);
                    double m = mi * mj;
                    v.re *= m;
                    v.im *= m;
                    if (a.atomic) {
                        if (m != 0.0 && !adj) store_entry<OutT, CPLX>(a.out, a.ld, row, col, v, 1, 0);
                    } else {
                        store_entry<OutT, CPLX>(a.out, a.ld, row, col, v, 0, a.accumulate);
                    }
                }
            }
        }
    }
}

// ------------------------------------------------------------------ dispatch
template <int OP, int NTS, int NSS, int QT, bool MIXED, typename OutT>
static int launch3(const RegularArgs &a, int kahan, hipStream_t stream) {
    constexpr int TM = MIXED ? 16 : 8;
    dim3 grid((a.trial.n + kRegThreads - 1) / kRegThreads, (a.test.n + TM - 1) / TM);
    if (grid.x == 0 || grid.y == 0) return BEM_OK;
    if (grid.y > 65535u) {
        set_error("too many test elements for one launch (%d)", a.test.n);
        return BEM_ERR_INVALID_ARGUMENT;
    }
    if (MIXED && kahan)
        hipLaunchKernelGGL((regular_kernel<OP, NTS, NSS, QT, TM, MIXED, true, OutT>), grid, dim3(kRegThreads), 0, stream, a);
    else
        hipLaunchKernelGGL((regular_kernel<OP, NTS, NSS, QT, TM, MIXED, false, OutT>), grid, dim3(kRegThreads), 0, stream, a);
    return check_launch("regular_kernel");
}

template <int OP, int NTS, int NSS, bool MIXED, typename OutT>
static int launch2(const RegularArgs &a, int kahan, hipStream_t stream) {
    if (a.rule.n == 6) return launch3<OP, NTS, NSS, 6, MIXED, OutT>(a, kahan, stream);
    return launch3<OP, NTS, NSS, 0, MIXED, OutT>(a, kahan, stream);
}

template <int OP, bool MIXED, typename OutT>
static int launch1(const RegularArgs &a, int kahan, hipStream_t stream) {
    if constexpr (OP == kMaxwellEFIE) {
        return launch2<OP, 3, 3, MIXED, OutT>(a, kahan, stream);
    } else {
        int nt = nshape_of(a.test.shape), ns = nshape_of(a.trial.shape);
        if (nt == 1 && ns == 1) return launch2<OP, 1, 1, MIXED, OutT>(a, kahan, stream);
        if (nt == 1 && ns == 3) return launch2<OP, 1, 3, MIXED, OutT>(a, kahan, stream);
        if (nt == 3 && ns == 1) return launch2<OP, 3, 1, MIXED, OutT>(a, kahan, stream);
        return launch2<OP, 3, 3, MIXED, OutT>(a, kahan, stream);
    }
}

template <int OP>
static int launch_op(const RegularArgs &a, int precision, int kahan, int out_dtype, hipStream_t stream) {
    constexpr bool CPLX = OpTraits<OP>::is_complex;
    const bool mixed = precision == BEM_PRECISION_MIXED;
    if (CPLX) {
        if (out_dtype == BEM_C128) return mixed ? launch1<OP, true, double>(a, kahan, stream) : launch1<OP, false, double>(a, kahan, stream);
        if (out_dtype == BEM_C64) return mixed ? launch1<OP, true, float>(a, kahan, stream) : launch1<OP, false, float>(a, kahan, stream);
    } else {
        if (out_dtype == BEM_F64) return mixed ? launch1<OP, true, double>(a, kahan, stream) : launch1<OP, false, double>(a, kahan, stream);
        if (out_dtype == BEM_F32) return mixed ? launch1<OP, true, float>(a, kahan, stream) : launch1<OP, false, float>(a, kahan, stream);
    }
    set_error("output dtype %d does not match operator %d", out_dtype, OP);
    return BEM_ERR_INVALID_ARGUMENT;
}

}  // namespace bem

using namespace bem;

extern "C" int bem_regular_assemble(int op, const bem_space_t *test, const bem_space_t *trial, int grids_identical,
                                    const bem_quad_rule_t *rule, double k_re, double k_im, int precision,
                                    double mixed_near_factor, int kahan, void *out, int64_t ld, int out_dtype,
                                    int scatter, int accumulate, void *stream) {
    clear_error();
    if (!test || !trial || !rule || !out) {
        set_error("null argument");
        return BEM_ERR_INVALID_ARGUMENT;
    }
    int rc;
    if ((rc = check_space(test, "test")) || (rc = check_space(trial, "trial"))) return rc;
    if (rule->n_points < 1 || rule->n_points > 16) {
        set_error("quadrature rule must have 1..16 points, got %d", rule->n_points);
        return BEM_ERR_INVALID_ARGUMENT;
    }
    if (precision != BEM_PRECISION_FP64 && precision != BEM_PRECISION_MIXED) {
        set_error("unknown precision mode %d", precision);
        return BEM_ERR_INVALID_ARGUMENT;
    }
    if (op == BEM_MAXWELL_ELECTRIC_FIELD && (test->shapeset != BEM_SHAPE_RWG || trial->shapeset != BEM_SHAPE_RWG)) {
        set_error("the electric field operator needs RWG/SNC spaces on both sides");
        return BEM_ERR_INVALID_ARGUMENT;
    }
    if (op != BEM_MAXWELL_ELECTRIC_FIELD && (test->shapeset == BEM_SHAPE_RWG || trial->shapeset == BEM_SHAPE_RWG)) {
        set_error("scalar operators need scalar (P0/P1) spaces");
        return BEM_ERR_INVALID_ARGUMENT;
    }
    RegularArgs a;
    a.test = to_args(test);
    a.trial = to_args(trial);
    a.rule = to_args(rule);
    a.kr = k_re;
    a.ki = k_im;
    a.near_factor = mixed_near_factor;
    a.out = out;
    a.ld = ld;
    a.identical = grids_identical;
    a.atomic = scatter;
    a.accumulate = accumulate;
    hipStream_t s = (hipStream_t)stream;
    switch (op) {
    case BEM_LAPLACE_SINGLE_LAYER: return launch_op<kLaplaceSL>(a, precision, kahan, out_dtype, s);
    case BEM_LAPLACE_DOUBLE_LAYER: return launch_op<kLaplaceDL>(a, precision, kahan, out_dtype, s);
    case BEM_LAPLACE_ADJOINT_DOUBLE_LAYER: return launch_op<kLaplaceADL>(a, precision, kahan, out_dtype, s);
    case BEM_HELMHOLTZ_SINGLE_LAYER: return launch_op<kHelmholtzSL>(a, precision, kahan, out_dtype, s);
    case BEM_HELMHOLTZ_DOUBLE_LAYER: return launch_op<kHelmholtzDL>(a, precision, kahan, out_dtype, s);
    case BEM_HELMHOLTZ_ADJOINT_DOUBLE_LAYER: return launch_op<kHelmholtzADL>(a, precision, kahan, out_dtype, s);
    case BEM_MAXWELL_ELECTRIC_FIELD: return launch_op<kMaxwellEFIE>(a, precision, kahan, out_dtype, s);
    default: set_error("unsupported operator %d", op); return BEM_ERR_UNSUPPORTED;
    }
}
