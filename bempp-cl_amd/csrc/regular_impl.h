// Regular-pair dense Galerkin assembly for gfx950.
//
// Computes, for every (test element e, trial element f) pair that shares no vertex (or
// every pair when the grids differ),
//   A[g_e(i), g_f(j)] += m_e(i) m_f(j) sum_p sum_q w_p J_e w_q J_f phi_i(xi_p) psi_j(eta_q) G(x_p, y_q)
// (bempp_cl/core/numba_kernels.py:980-1074), and the Maxwell EFIE block
//   sum_pq G w w J J [ -ik phi_i.psi_j - 4 / (ik J_e J_f) ] l_i l_j m_i m_j
// (numba_kernels.py:1853-1969).
//
// Work decomposition: a 256-thread workgroup owns a tile of up to TM test elements and a
// block of up to 256 trial elements. Each lane keeps one trial element (quadrature point
// offsets, weights, normal) in registers and walks the test elements staged in LDS, so
// every LDS read is a broadcast.
//
// Per pair the kernel forms the 3x3 moment block
//   S[a][b] = sum_pq G'_pq u_a(xi_p) u_b(eta_q),  u = (1, xi_0, xi_1),  G' = G w_p J_e w_q J_f
// from which every local block follows (bem_device.h: local_value).
//
// Output scatter:
//  * direct (both spaces localised, DP0/DP1): every pair owns its entries; lanes hold
//    consecutive trial elements, so a DP0 row segment is one coalesced store.
//  * tile (P1, RWG/SNC: dofs shared between elements): the host splits the (locality-
//    ordered) element lists into test tiles and trial blocks and gives every block its
//    list of distinct dofs and every (element, shape function) its slot in that list.
//    The workgroup sums its contributions in an LDS tile [test slots x trial slots] with
//    LDS atomics, then adds the tile to HBM once per entry. This removes the ~20-fold
//    repetition of global atomics on the same entry (each P1 dof pair is touched by up to
//    36 element pairs).
//
// MIXED precision: pairs with centroid distance >= factor * mean diameter run in FP32.
// Per pair, the centroid offset d = c_f - c_e and the reference distance r0 = |d| (rounded
// to FP32) are formed in FP64, together with the phase k r0 reduced mod 2 pi. Per point
// pair only the small difference r - r0 = t / (r + r0) with t = |d + b_q - a_p|^2 - r0^2
// (a, b: offsets of the quadrature points from the centroids) is needed in FP32, so the
// FP32 phase error is ~k * eps * h instead of k * eps * r and the hardware v_sin/v_cos
// (argument in revolutions) see arguments of size k h. The pair's moments are multiplied
// by e^{ik r0} once. Sums over test points use Kahan compensation. Near pairs use FP64.
#pragma once

#include "bem_device.h"
#include "bem_internal.h"
#include "mfie_block.h"

namespace bem {

constexpr int kRegThreads = 256;
constexpr int kMaxQ = 16;
constexpr double kInv2Pi = 0.15915494309189533577;

struct BlockArgs {
    const int32_t *start;  // [nblocks + 1] offsets into the element list (NULL: fixed blocks)
    const int32_t *dofs;   // [nblocks * maxu] distinct global dofs per block
    const int32_t *ndofs;  // [nblocks]
    const int32_t *slots;  // [n_list * 3] slot of (list position, shape function) in its block
    const int16_t *inc;    // trial side: [nblocks * maxu * maxinc] incidences 3 * (pos - start) + j, -1 padded
    int nblocks;
    int maxu;
    int maxinc;
};

constexpr int kMaxInc = 12;  // incidences per dof and block held in registers by the owner lane
constexpr int kRwgMaxInc = 2;  // an RWG dof (edge) has two elements

struct RegularArgs {
    SpaceArgs test, trial;
    BlockArgs tb, sb;
    RuleArgs rule;
    double kr, ki;
    double near_factor;
    void *out;
    void *out2;  // DUAL kernels: the double layer (out holds the single layer)
    int64_t ld;
    int identical;
    int accumulate;
    // SYM kernels: work item w is trial block work[3w] with the test tiles [work[3w + 1], work[3w + 2])
    const int32_t *work;
    int nwork;
};


// NXP: 1 in MIXED mode, where the absolute FP64 test points are not needed; Q: points per element;
// EF: the FP32 EFIE epilogue's element data (edge vectors, edge lengths, 1 / J)
template <int TM, int NXP = TM, int Q = kMaxQ, bool EF = false>
struct TestTile {
    int elem[TM];
    int vid[TM][3];
    int slot[TM][3];
    double mult[TM][3];
    double c[TM][3];
    double diam[TM];
    double xp[NXP][Q][3];  // absolute test points (FP64 paths)
    float a[TM][Q][3];     // offsets from the centroid (FP32 path)
    float aa[TM][Q];       // |a|^2
    float wp[TM][Q];       // w_p J_e
    float nx[TM][3];       // test normal (multiplied)
    float ee[EF ? TM : 1][2][3];
    float el[EF ? TM : 1][3];
    float ej[EF ? TM : 1];
};

// test elements staged per SYM work item (a run of test tiles)
constexpr int kSymRunElements = BEM_TILE_SYM_RUN_ELEMENTS;
// below this many workgroups (8 per CU) the DP0 upper-triangle pass takes 4-element test tiles
constexpr int64_t kUpperSmallGrid = 2048;

// ------------------------------------------------------------------ FP64 pair
// T[b] += g u_b(eta_q) with u = (1, eta0, eta1)
template <int NSS>
__device__ __forceinline__ void add_trial_moments(C2d (&T)[NSS], double gre, double gim, double s0, double s1) {
    T[0].re += gre;
    T[0].im += gim;
    if constexpr (NSS == 3) {
        T[1].re += gre * s0;
        T[1].im += gim * s0;
        T[2].re += gre * s1;
        T[2].im += gim * s1;
    }
}

// S[a][b] += u_a(xi_p) T[b] with u = (wp, wp xi0, wp xi1)
template <int NTS, int NSS>
__device__ __forceinline__ void add_test_moments(C2d (&S)[NTS][NSS], const C2d (&T)[NSS], double wp, double xi,
                                                 double eta) {
    const double u[3] = {wp, wp * xi, wp * eta};
#pragma unroll
    for (int a = 0; a < NTS; ++a)
#pragma unroll
        for (int b = 0; b < NSS; ++b) {
            S[a][b].re += u[a] * T[b].re;
            S[a][b].im += u[a] * T[b].im;
        }
}

template <int NTS, int NSS>
__device__ __forceinline__ void zero_moments(C2d (&S)[NTS][NSS]) {
#pragma unroll
    for (int a = 0; a < NTS; ++a)
#pragma unroll
        for (int b = 0; b < NSS; ++b) S[a][b].re = S[a][b].im = 0.0;
}

// Green's function values of one point pair: OP (g), for DUAL the double layer (g2) and for
// SYM also the double layer of the reversed pair (g3, test normal nx).
template <int OP, bool DUAL, bool SYM, int PH = 0>
__device__ __forceinline__ void green3(double dx, double dy, double dz, double r2, const double *nx, const double *ny,
                                       double kr, double ki, double &gre, double &gim, double &g2re, double &g2im,
                                       double &g3re, double &g3im, double ph0 = 0.0) {
    if constexpr (SYM) {
        green_sl_dl_rev<OpTraits<OP>::is_complex, double, PH>(dx, dy, dz, r2, nx, ny, kr, ki, gre, gim, g2re, g2im,
                                                              g3re, g3im, ph0);
    } else {
        green2<OP, DUAL, double, PH>(dx, dy, dz, r2, nx, ny, kr, ki, gre, gim, g2re, g2im, ph0);
        g3re = g3im = 0.0;
    }
}

// FP64 pair from the geometry table (any rule). DUAL: S2 receives the double layer; SYM: S3
// receives the double layer of the reversed pair (indexed like S: test side first). PH: phase
// mode of the Green's functions.
template <int OP, int NTS, int NSS, int QT, bool DUAL = false, bool SYM = false, int PH = 0, typename R = RuleArgs>
__device__ __forceinline__ void pair_fp64_geom(const double *__restrict__ gt, const double *__restrict__ gf, double tnm,
                                               double snm, const R &rule, double kr, double ki,
                                               C2d (&S)[NTS][NSS], C2d (&S2)[NTS][NSS], C2d (&S3)[NTS][NSS]) {
    const int nq = QT > 0 ? QT : rule.n;
    double nx[3] = {gt[9] * tnm, gt[10] * tnm, gt[11] * tnm};
    double ny[3] = {gf[9] * snm, gf[10] * snm, gf[11] * snm};
    double jt = gt[12], jf = gf[12];
    zero_moments(S);
    if constexpr (DUAL) zero_moments(S2);
    if constexpr (SYM) zero_moments(S3);
#pragma unroll 1
    for (int p = 0; p < nq; ++p) {
        double xi = rule.x[p], eta = rule.y[p];
        double x0 = gt[0] + gt[3] * xi + gt[6] * eta;
        double x1 = gt[1] + gt[4] * xi + gt[7] * eta;
        double x2 = gt[2] + gt[5] * xi + gt[8] * eta;
        C2d T[NSS], T2[NSS], T3[NSS];
#pragma unroll
        for (int b = 0; b < NSS; ++b) T[b].re = T[b].im = T2[b].re = T2[b].im = T3[b].re = T3[b].im = 0.0;
#pragma unroll(QT > 0 ? QT : 1)
        for (int q = 0; q < nq; ++q) {
            double s0 = rule.x[q], s1 = rule.y[q];
            double dx = gf[0] + gf[3] * s0 + gf[6] * s1 - x0;
            double dy = gf[1] + gf[4] * s0 + gf[7] * s1 - x1;
            double dz = gf[2] + gf[5] * s0 + gf[8] * s1 - x2;
            double r2 = dx * dx + dy * dy + dz * dz;
            double gre, gim, g2re, g2im, g3re, g3im;
            green3<OP, DUAL, SYM, PH>(dx, dy, dz, r2, nx, ny, kr, ki, gre, gim, g2re, g2im, g3re, g3im);
            const double wq = rule.w[q] * jf;
            add_trial_moments<NSS>(T, gre * wq, gim * wq, s0, s1);
            if constexpr (DUAL) add_trial_moments<NSS>(T2, g2re * wq, g2im * wq, s0, s1);
            if constexpr (SYM) add_trial_moments<NSS>(T3, g3re * wq, g3im * wq, s0, s1);
        }
        add_test_moments<NTS, NSS>(S, T, rule.w[p] * jt, xi, eta);
        if constexpr (DUAL) add_test_moments<NTS, NSS>(S2, T2, rule.w[p] * jt, xi, eta);
        if constexpr (SYM) add_test_moments<NTS, NSS>(S3, T3, rule.w[p] * jt, xi, eta);
    }
}

// FP64 pair with test points from LDS and trial points from registers. DUAL: S2 = double
// layer; SYM: S3 = double layer of the reversed pair. PH: phase mode of the Green's functions.
template <int OP, int NTS, int NSS, int QT, bool DUAL = false, bool SYM = false, int PH = 0, typename TT>
__device__ __forceinline__ void pair_fp64_regs(const TT &T_, int t, const double *gt, double tnm,
                                               const double (&yq)[QT > 0 ? QT : 1][3], const double (&wq)[QT > 0 ? QT : 1],
                                               const double *ny, const RuleArgs &rule, double kr, double ki,
                                               C2d (&S)[NTS][NSS], C2d (&S2)[NTS][NSS], C2d (&S3)[NTS][NSS],
                                               double ph0 = 0.0) {
    double nx[3] = {gt[9] * tnm, gt[10] * tnm, gt[11] * tnm};
    double jt = gt[12];
    zero_moments(S);
    if constexpr (DUAL) zero_moments(S2);
    if constexpr (SYM) zero_moments(S3);
#pragma unroll 1
    for (int p = 0; p < QT; ++p) {
        double x0 = T_.xp[t][p][0], x1 = T_.xp[t][p][1], x2 = T_.xp[t][p][2];
        C2d T[NSS], T2[NSS], T3[NSS];
#pragma unroll
        for (int b = 0; b < NSS; ++b) T[b].re = T[b].im = T2[b].re = T2[b].im = T3[b].re = T3[b].im = 0.0;
#pragma unroll
        for (int q = 0; q < QT; ++q) {
            double dx = yq[q][0] - x0, dy = yq[q][1] - x1, dz = yq[q][2] - x2;
            double r2 = dx * dx + dy * dy + dz * dz;
            double gre, gim, g2re, g2im, g3re, g3im;
            green3<OP, DUAL, SYM, PH>(dx, dy, dz, r2, nx, ny, kr, ki, gre, gim, g2re, g2im, g3re, g3im, ph0);
            add_trial_moments<NSS>(T, gre * wq[q], gim * wq[q], rule.x[q], rule.y[q]);
            if constexpr (DUAL) add_trial_moments<NSS>(T2, g2re * wq[q], g2im * wq[q], rule.x[q], rule.y[q]);
            if constexpr (SYM) add_trial_moments<NSS>(T3, g3re * wq[q], g3im * wq[q], rule.x[q], rule.y[q]);
        }
        add_test_moments<NTS, NSS>(S, T, rule.w[p] * jt, rule.x[p], rule.y[p]);
        if constexpr (DUAL) add_test_moments<NTS, NSS>(S2, T2, rule.w[p] * jt, rule.x[p], rule.y[p]);
        if constexpr (SYM) add_test_moments<NTS, NSS>(S3, T3, rule.w[p] * jt, rule.x[p], rule.y[p]);
    }
}

// ------------------------------------------------------------------ FP32 far pair
// Trial element state of one lane (FP32 path)
template <int QT>
struct TrialF32 {
    static constexpr int Q = QT > 0 ? QT : kMaxQ;
    float b[Q][3];  // offsets of the quadrature points from the centroid
    float bb[Q];    // |b|^2
    float wq[Q];    // w_q J_f / (4 pi)
    float bn[Q];    // b . n_y
    float ny[3];
};

__device__ __forceinline__ float sin_rev(float x) { return __builtin_amdgcn_sinf(x); }
__device__ __forceinline__ float cos_rev(float x) { return __builtin_amdgcn_cosf(x); }

// DAMP: complex wavenumber (e^{-Im(k) r} factor); a template parameter so that real
// wavenumbers do not pay for an if-converted exponential per evaluation.
// DUAL (OP a single layer): S2 receives the double layer of the same family, sharing the
// distance, 1/r and e^{ik delta} of every point pair. SYM (with DUAL): S3 receives the double
// layer of the reversed pair (normal of the test element), from the same values again.
// ORB (6-point rule in the orbit order of orbit_rule, P1 trial side): the trial moments of every
// point value G come from the two orbit sums S_A, S_B: sum_q G_q = S_A + S_B and
//   sum_q G_q xi_q  = b_A S_A + b_B S_B + (a_A - b_A) G_A1 + (a_B - b_B) G_B1
//   sum_q G_q eta_q = b_A S_A + b_B S_B + (a_A - b_A) G_A2 + (a_B - b_B) G_B2,
// 11 operations per value and test point instead of 18.
template <int OP, int NTS, int NSS, int QT, bool KAHAN, bool DAMP, bool DUAL, bool SYM = false, bool ORB = false,
          typename TT>
__device__ __forceinline__ void pair_fp32(const TT &T_, int t, const TrialF32<QT> &F, double d0, double d1,
                                          double d2, double dist2, const RuleArgs &rule, double kr, double ki,
                                          C2f (&S)[NTS][NSS], C2f (&S2)[NTS][NSS], C2f (&S3)[NTS][NSS]) {
    static_assert(!ORB || (QT == 6 && NSS == 3), "the orbit moments need the 6-point rule and a P1 trial side");
    constexpr bool HELM = OP >= kHelmholtzSL;
    static_assert(!DUAL || OP == kLaplaceSL || OP == kHelmholtzSL, "DUAL pairs a single layer with its double layer");
    static_assert(!SYM || DUAL, "SYM extends the DUAL pass");
    constexpr bool NXDOT = OP == kLaplaceADL || OP == kHelmholtzADL || SYM;  // needs (x - y).n_x
    const int nq = QT > 0 ? QT : rule.n;
    const float fd0 = (float)d0, fd1 = (float)d1, fd2 = (float)d2;
    // reference distance r0 (FP32 value) and the FP64 residual |d|^2 - r0^2
    const float r0 = sqrtf((float)dist2);
    const float r0sq = r0 * r0;
    const float resid = (float)(dist2 - (double)r0 * (double)r0);
    const float kr_rev = (float)(kr * kInv2Pi);
    const float kr_f = (float)kr, ki_f = (float)ki;
    const float ki_log2e = (float)(ki * 1.4426950408889634);
    // K = 1 / kr_rev and r0 K (infinite for k = 0: the phase is then 0)
    const float rev_inv = 1.f / kr_rev, r0_rev_inv = r0 * rev_inv;

    // test normal projections of the trial offsets (ADL)
    float bnx[TrialF32<QT>::Q];
    if constexpr (NXDOT) {
#pragma unroll(QT > 0 ? QT : 1)
        for (int q = 0; q < nq; ++q)
            bnx[q] = F.b[q][0] * T_.nx[t][0] + F.b[q][1] * T_.nx[t][1] + F.b[q][2] * T_.nx[t][2];
    }

    Acc<float, KAHAN> Sr[NTS][NSS], Si[NTS][NSS];
    Acc<float, KAHAN> Sr2[DUAL ? NTS : 1][DUAL ? NSS : 1], Si2[DUAL ? NTS : 1][DUAL ? NSS : 1];
    Acc<float, KAHAN> Sr3[SYM ? NTS : 1][SYM ? NSS : 1], Si3[SYM ? NTS : 1][SYM ? NSS : 1];
#pragma unroll
    for (int i = 0; i < NTS; ++i)
#pragma unroll
        for (int j = 0; j < NSS; ++j) {
            Sr[i][j].zero();
            Si[i][j].zero();
            if constexpr (DUAL) {
                Sr2[i][j].zero();
                Si2[i][j].zero();
            }
            if constexpr (SYM) {
                Sr3[i][j].zero();
                Si3[i][j].zero();
            }
        }
    // the test-point loop stays rolled: unrolling it lets the scheduler keep all 36
    // evaluations in flight, which exhausts the VGPR file and drops occupancy to 1 wave
#pragma unroll 1
    for (int p = 0; p < nq; ++p) {
        // w = d - a_p: y - x = w + b_q
        const float w0 = fd0 - T_.a[t][p][0], w1 = fd1 - T_.a[t][p][1], w2 = fd2 - T_.a[t][p][2];
        // |w|^2 - r0^2 = resid - 2 d.a_p + |a_p|^2
        const float Ap = resid - 2.f * (fd0 * T_.a[t][p][0] + fd1 * T_.a[t][p][1] + fd2 * T_.a[t][p][2]) + T_.aa[t][p];
        const float tw0 = 2.f * w0, tw1 = 2.f * w1, tw2 = 2.f * w2;
        float wn = 0.f, wnx = 0.f;
        if constexpr (OP == kLaplaceDL || OP == kHelmholtzDL || DUAL) wn = w0 * F.ny[0] + w1 * F.ny[1] + w2 * F.ny[2];
        if constexpr (NXDOT) wnx = -(w0 * T_.nx[t][0] + w1 * T_.nx[t][1] + w2 * T_.nx[t][2]);
        float Tr[NSS], Ti[NSS], Tr2[DUAL ? NSS : 1], Ti2[DUAL ? NSS : 1], Tr3[SYM ? NSS : 1], Ti3[SYM ? NSS : 1];
#pragma unroll
        for (int b = 0; b < NSS; ++b) {
            Tr[b] = Ti[b] = 0.f;
            if constexpr (DUAL) Tr2[b] = Ti2[b] = 0.f;
            if constexpr (SYM) Tr3[b] = Ti3[b] = 0.f;
        }
        constexpr int QO = ORB ? 6 : 1;
        float Gv[6][QO];  // ORB: this test point's values at the 6 trial points (gre gim g2re g2im g3re g3im)
#pragma unroll(QT > 0 ? QT : 1)
        for (int q = 0; q < nq; ++q) {
            // t = |y - x|^2 - r0^2
            const float tt = Ap + F.bb[q] + tw0 * F.b[q][0] + tw1 * F.b[q][1] + tw2 * F.b[q][2];
            const float r2 = r0sq + tt;
            const float rinv = __builtin_amdgcn_rsqf(r2);
            float gre, gim, g2re = 0.f, g2im = 0.f, g3re = 0.f, g3im = 0.f;
            if constexpr (!HELM) {
                float g = rinv * F.wq[q];
                if constexpr (OP == kLaplaceSL) {
                    gre = g;
                    if constexpr (SYM) {
                        const float h = -g * rinv * rinv;
                        g2re = (wn + F.bn[q]) * h;
                        g3re = (wnx - bnx[q]) * h;
                    } else if constexpr (DUAL) {
                        g2re = -(wn + F.bn[q]) * g * rinv * rinv;
                    }
                } else if constexpr (OP == kLaplaceDL) {
                    gre = -(wn + F.bn[q]) * g * rinv * rinv;
                } else {  // (y - x).n_x = bnx - wnx
                    gre = (bnx[q] - wnx) * g * rinv * rinv;
                }
                gim = 0.f;
            } else {
                const float r = r2 * rinv;
                // phase in revolutions, k (r - r0) / (2 pi) = t / ((r + r0) K) with K = 2 pi / k
                float ph, delta = 0.f;
                if constexpr (DAMP) {
                    delta = tt * __builtin_amdgcn_rcpf(r + r0);
                    ph = kr_rev * delta;
                } else {
                    ph = tt * __builtin_amdgcn_rcpf(r * rev_inv + r0_rev_inv);
                }
                const float s = sin_rev(ph), c = cos_rev(ph);
                float g = rinv * F.wq[q];
                if constexpr (DAMP) g *= __builtin_amdgcn_exp2f(-ki_log2e * delta);
                if constexpr (OP == kHelmholtzSL || OP == kMaxwellEFIE) {
                    gre = c * g;
                    gim = s * g;
                    if constexpr (SYM) {
                        const float gd = g * rinv * rinv;
                        const float ar = DAMP ? -1.f - ki_f * r : -1.f, ai = kr_f * r;
                        const float hr = (ar * c - ai * s) * gd, hi = (ai * c + ar * s) * gd;
                        const float dot2 = wn + F.bn[q], dot3 = wnx - bnx[q];
                        g2re = hr * dot2;
                        g2im = hi * dot2;
                        g3re = hr * dot3;
                        g3im = hi * dot3;
                    } else if constexpr (DUAL) {
                        const float gd = g * (wn + F.bn[q]) * rinv * rinv;
                        const float ar = DAMP ? -1.f - ki_f * r : -1.f, ai = kr_f * r;
                        g2re = (ar * c - ai * s) * gd;
                        g2im = (ai * c + ar * s) * gd;
                    }
                } else {
                    float dot;
                    if constexpr (OP == kHelmholtzDL) dot = wn + F.bn[q];
                    else dot = wnx - bnx[q];
                    g *= dot * rinv * rinv;
                    const float ar = DAMP ? -1.f - ki_f * r : -1.f, ai = kr_f * r;
                    gre = (ar * c - ai * s) * g;
                    gim = (ai * c + ar * s) * g;
                }
            }
            if constexpr (ORB) {
                Gv[0][q] = gre;
                Gv[1][q] = gim;
                Gv[2][q] = g2re;
                Gv[3][q] = g2im;
                Gv[4][q] = g3re;
                Gv[5][q] = g3im;
                continue;
            }
            Tr[0] += gre;
            if constexpr (HELM) Ti[0] += gim;
            if constexpr (DUAL) {
                Tr2[0] += g2re;
                if constexpr (HELM) Ti2[0] += g2im;
            }
            if constexpr (SYM) {
                Tr3[0] += g3re;
                if constexpr (HELM) Ti3[0] += g3im;
            }
            if constexpr (NSS == 3) {
                const float s0 = (float)rule.x[q], s1 = (float)rule.y[q];
                Tr[1] += gre * s0;
                Tr[2] += gre * s1;
                if constexpr (HELM) {
                    Ti[1] += gim * s0;
                    Ti[2] += gim * s1;
                }
                if constexpr (DUAL) {
                    Tr2[1] += g2re * s0;
                    Tr2[2] += g2re * s1;
                    if constexpr (HELM) {
                        Ti2[1] += g2im * s0;
                        Ti2[2] += g2im * s1;
                    }
                }
                if constexpr (SYM) {
                    Tr3[1] += g3re * s0;
                    Tr3[2] += g3re * s1;
                    if constexpr (HELM) {
                        Ti3[1] += g3im * s0;
                        Ti3[2] += g3im * s1;
                    }
                }
            }
        }
        if constexpr (ORB) {
            auto orbit = [&](const float(&G)[QO], float(&T)[NSS]) {
                const float sa = G[0] + G[1] + G[2], sb = G[3] + G[4] + G[5];
                const float base = rule.orb[0] * sa + rule.orb[2] * sb;
                T[0] = sa + sb;
                T[1] = rule.orb[1] * G[1] + (rule.orb[3] * G[4] + base);
                T[2] = rule.orb[1] * G[2] + (rule.orb[3] * G[5] + base);
            };
            orbit(Gv[0], Tr);
            if constexpr (HELM) orbit(Gv[1], Ti);
            if constexpr (DUAL) {
                orbit(Gv[2], Tr2);
                if constexpr (HELM) orbit(Gv[3], Ti2);
            }
            if constexpr (SYM) {
                orbit(Gv[4], Tr3);
                if constexpr (HELM) orbit(Gv[5], Ti3);
            }
        }
        const float wp = T_.wp[t][p];
        const float u[3] = {wp, wp * (float)rule.x[p], wp * (float)rule.y[p]};
#pragma unroll
        for (int i = 0; i < NTS; ++i)
#pragma unroll
            for (int j = 0; j < NSS; ++j) {
                Sr[i][j].add(u[i] * Tr[j]);
                if constexpr (HELM) Si[i][j].add(u[i] * Ti[j]);
                if constexpr (DUAL) {
                    Sr2[i][j].add(u[i] * Tr2[j]);
                    if constexpr (HELM) Si2[i][j].add(u[i] * Ti2[j]);
                }
                if constexpr (SYM) {
                    Sr3[i][j].add(u[i] * Tr3[j]);
                    if constexpr (HELM) Si3[i][j].add(u[i] * Ti3[j]);
                }
            }
    }
    // multiply by e^{i k r0} e^{-ki r0} (FP64 phase reduced to one revolution)
    float pr = 1.f, pi = 0.f;
    if constexpr (HELM) {
        double rev = kr * kInv2Pi * (double)r0;
        rev -= rint(rev);
        float damp = ki != 0.0 ? __expf((float)(-ki * (double)r0)) : 1.f;
        pr = cos_rev((float)rev) * damp;
        pi = sin_rev((float)rev) * damp;
    }
#pragma unroll
    for (int i = 0; i < NTS; ++i)
#pragma unroll
        for (int j = 0; j < NSS; ++j) {
            float sr = Sr[i][j].s - Sr[i][j].c;
            float si = HELM ? Si[i][j].s - Si[i][j].c : 0.f;
            S[i][j].re = pr * sr - pi * si;
            S[i][j].im = pr * si + pi * sr;
            if constexpr (DUAL) {
                sr = Sr2[i][j].s - Sr2[i][j].c;
                si = HELM ? Si2[i][j].s - Si2[i][j].c : 0.f;
                S2[i][j].re = pr * sr - pi * si;
                S2[i][j].im = pr * si + pi * sr;
            }
            if constexpr (SYM) {
                sr = Sr3[i][j].s - Sr3[i][j].c;
                si = HELM ? Si3[i][j].s - Si3[i][j].c : 0.f;
                S3[i][j].re = pr * sr - pi * si;
                S3[i][j].im = pr * si + pi * sr;
            }
        }
}

// Scalar local block from FP32 moments: phi = (u0 - u1 - u2, u1, u2) on P1 sides, u0 on P0
// sides, i.e. V = C^T S C with the sparse 3x3 change of basis C (see local_value).
template <int NTS, int NSS>
__device__ __forceinline__ void scalar_block_f32(const C2f (&S)[NTS][NSS], C2f (&V)[NTS][NSS]) {
    C2f Y[NTS][NSS];
#pragma unroll
    for (int a = 0; a < NTS; ++a) {
        if constexpr (NSS == 3) {
            Y[a][0].re = S[a][0].re - S[a][1].re - S[a][2].re;
            Y[a][0].im = S[a][0].im - S[a][1].im - S[a][2].im;
            Y[a][1] = S[a][1];
            Y[a][2] = S[a][2];
        } else {
            Y[a][0] = S[a][0];
        }
    }
#pragma unroll
    for (int j = 0; j < NSS; ++j) {
        if constexpr (NTS == 3) {
            V[0][j].re = Y[0][j].re - Y[1][j].re - Y[2][j].re;
            V[0][j].im = Y[0][j].im - Y[1][j].im - Y[2][j].im;
            V[1][j] = Y[1][j];
            V[2][j] = Y[2][j];
        } else {
            V[0][j] = Y[0][j];
        }
    }
}

// Maxwell EFIE local block in FP32 from FP32 moments (the formula of local_value<kMaxwellEFIE>):
//   V[i][j] = l_i l_j [ -ik sum_ab M_ab (S[1+a][1+b] - o_i[a] S[0][1+b] - o_j[b] S[1+a][0] + o_i[a] o_j[b] S00)
//                       - 4 S00 / (ik) ] / (J_t J_f),   M = E_t^T E_f,  o = ((0, 1), (1, 0), (0, 0))
// et/ef: edge vectors v1 - v0, v2 - v0; lt/lf: edge lengths; jti/jfi: 1 / J; (ikr, iki) = ik and
// (invr, invi) = 1 / (ik). The moments already carry FP32 rounding; the combination has no
// cancellation beyond the factor ~3 of (xi - o) weights.
__device__ __forceinline__ void efie_block_f32(const C2f (&S)[3][3], const float (&et)[2][3], const float (&lt)[3],
                                               float jti, const float (&ef)[2][3], const float (&lf)[3], float jfi,
                                               float ikr, float iki, float invr, float invi, C2f (&V)[3][3]) {
    float M[2][2];
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
        for (int b = 0; b < 2; ++b) M[a][b] = et[a][0] * ef[b][0] + et[a][1] * ef[b][1] + et[a][2] * ef[b][2];
    const float jj = jti * jfi;
    const float s_re = 4.f * S[0][0].re * jj, s_im = 4.f * S[0][0].im * jj;
    const float c_re = -(invr * s_re - invi * s_im), c_im = -(invr * s_im + invi * s_re);
    constexpr float ox[3] = {0.f, 1.f, 0.f}, oy[3] = {1.f, 0.f, 0.f};
#pragma unroll
    for (int i = 0; i < 3; ++i) {
        // R[a][b'] = S[1 + a][b'] - o_i[a] S[0][b']
        const float oi[2] = {ox[i], oy[i]};
        C2f R[2][3];
#pragma unroll
        for (int a = 0; a < 2; ++a)
#pragma unroll
            for (int b = 0; b < 3; ++b) {
                R[a][b].re = S[1 + a][b].re - oi[a] * S[0][b].re;
                R[a][b].im = S[1 + a][b].im - oi[a] * S[0][b].im;
            }
#pragma unroll
        for (int j = 0; j < 3; ++j) {
            const float oj[2] = {ox[j], oy[j]};
            float tr = 0.f, ti = 0.f;
#pragma unroll
            for (int a = 0; a < 2; ++a)
#pragma unroll
                for (int b = 0; b < 2; ++b) {
                    tr += M[a][b] * (R[a][1 + b].re - oj[b] * R[a][0].re);
                    ti += M[a][b] * (R[a][1 + b].im - oj[b] * R[a][0].im);
                }
            tr *= jj;
            ti *= jj;
            const float l = lt[i] * lf[j];
            V[i][j].re = (-(ikr * tr - iki * ti) + c_re) * l;
            V[i][j].im = (-(ikr * ti + iki * tr) + c_im) * l;
        }
    }
}

// Hypersingular local block in FP32 from FP32 moments (the formula of local_value for
// kLaplaceHS / kHelmholtzHS): V[i][j] = C_ij S00 - k^2 (n_t . n_f) SL_ij with the curl products
//   C_ij = (nm_t / J_t) (nm_f / J_f) sum_ab c_i[a] c_j[b] (e_a^t . e_b^f),  c = curl_coef,
// SL the P1 single-layer block of the same moments. et/ef: edge vectors; sjt/sjf: nm / J;
// nt/nf: normals times nm; (k2r, k2i) = k^2 (zero for Laplace).
template <bool CPLX>
__device__ __forceinline__ void hs_block_f32(const C2f (&S)[3][3], const float (&et)[2][3], float sjt,
                                             const float (&nt)[3], const float (&ef)[2][3], float sjf,
                                             const float (&nf)[3], float k2r, float k2i, C2f (&V)[3][3]) {
    float M[2][2];
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
        for (int b = 0; b < 2; ++b) M[a][b] = et[a][0] * ef[b][0] + et[a][1] * ef[b][1] + et[a][2] * ef[b][2];
    const float jj = sjt * sjf;
    C2f SL[3][3];
    scalar_block_f32<3, 3>(S, SL);
    const float nn = nt[0] * nf[0] + nt[1] * nf[1] + nt[2] * nf[2];
    const float ar = k2r * nn, ai = k2i * nn;
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j) {
            float c = 0.f;
#pragma unroll
            for (int a = 0; a < 2; ++a)
#pragma unroll
                for (int b = 0; b < 2; ++b) c += (float)(curl_coef(i, a) * curl_coef(j, b)) * M[a][b];
            c *= jj;
            V[i][j].re = c * S[0][0].re;
            V[i][j].im = CPLX ? c * S[0][0].im : 0.f;
            if constexpr (CPLX) {
                V[i][j].re -= ar * SL[i][j].re - ai * SL[i][j].im;
                V[i][j].im -= ar * SL[i][j].im + ai * SL[i][j].re;
            }
        }
}

// Maxwell magnetic field (MFIE) block of a far pair in FP32 (mixed mode), from the moments of
// mfie_block.h: S[i][j] = sum_pq w_p J_t w_q J_f G (ikr - 1) / r^2 (x_p - y_q) . ((x_p - P_i) x (y_q - Q_j))
// with G = e^{ikr} / (4 pi r) (bempp_cl/core/numba_kernels.py:2151-2277); mfie_block_f32 scales it
// by l_i l_j / (J_t J_f). The distance uses the reference-distance split of pair_fp32: the phase
// e^{i k r0} is formed once per pair in FP64 and only r - r0 comes from the small FP32 offsets.
// kap: c_f - Q_j of the trial element.
template <int QT, typename TT>
__device__ __forceinline__ void mfie_pair_fp32(const TT &T_, int t, const TrialF32<QT> &F, const float (&kap)[3][3],
                                               double d0, double d1, double d2, double dist2, const RuleArgs &rule,
                                               double kr, double ki, C2f (&S)[3][3]) {
    const int nq = QT > 0 ? QT : rule.n;
    const float fd[3] = {(float)d0, (float)d1, (float)d2};
    const float r0 = sqrtf((float)dist2);
    const float r0sq = r0 * r0;
    const float resid = (float)(dist2 - (double)r0 * (double)r0);
    const float kr_rev = (float)(kr * kInv2Pi), kr_f = (float)kr, ki_f = (float)ki;
    const float ki_log2e = (float)(ki * 1.4426950408889634);
    // e^{i k r0} (with e^{-Im k r0}), the phase reduced to one revolution in FP64
    double rev = kr * (double)r0 * kInv2Pi;
    rev -= rint(rev);
    float c0 = cos_rev((float)rev), s0 = sin_rev((float)rev);
    if (ki != 0.0) {
        const float dm = (float)exp(-ki * (double)r0);
        c0 *= dm;
        s0 *= dm;
    }
    mfie::PairMoments<float> M;
    M.clear();
#pragma unroll 1
    for (int p = 0; p < nq; ++p) {
        const float ap[3] = {T_.a[t][p][0], T_.a[t][p][1], T_.a[t][p][2]};
        // x_p - y_q = w - b_q with w = a_p - d; |x - y|^2 - r0^2 = Ap - 2 w.b_q + |b_q|^2
        const float w0 = ap[0] - fd[0], w1 = ap[1] - fd[1], w2 = ap[2] - fd[2];
        const float Ap = resid - 2.f * (fd[0] * ap[0] + fd[1] * ap[1] + fd[2] * ap[2]) + T_.aa[t][p];
        const float tw0 = 2.f * w0, tw1 = 2.f * w1, tw2 = 2.f * w2;
        const float wp = T_.wp[t][p];
        float Rp[2] = {0.f, 0.f}, Yp[2][3] = {{0.f, 0.f, 0.f}, {0.f, 0.f, 0.f}};
#pragma unroll(QT > 0 ? QT : 1)
        for (int q = 0; q < nq; ++q) {
            const float tt = Ap + F.bb[q] - (tw0 * F.b[q][0] + tw1 * F.b[q][1] + tw2 * F.b[q][2]);
            const float r2 = r0sq + tt;
            const float rinv = __builtin_amdgcn_rsqf(r2);
            const float r = r2 * rinv;
            const float delta = tt * __builtin_amdgcn_rcpf(r + r0);  // r - r0
            const float ph = kr_rev * delta;
            const float c = cos_rev(ph), sn = sin_rev(ph);
            float er = c0 * c - s0 * sn, ei = c0 * sn + s0 * c;
            if (ki_f != 0.f) {
                const float dm = __builtin_amdgcn_exp2f(-ki_log2e * delta);
                er *= dm;
                ei *= dm;
            }
            // e^{ikr} ((-Im k r - 1) + i Re k r) w_p J_t w_q J_f / (4 pi r^3)
            const float mag = wp * F.wq[q] * rinv * rinv * rinv;
            const float fr = -ki_f * r - 1.f, fi = kr_f * r;
            const float vr = mag * (er * fr - ei * fi), vi = mag * (er * fi + ei * fr);
            Rp[0] += vr;
            Rp[1] += vi;
#pragma unroll
            for (int d = 0; d < 3; ++d) {
                Yp[0][d] += vr * F.b[q][d];
                Yp[1][d] += vi * F.b[q][d];
            }
        }
        M.add_point(Rp, Yp, ap, fd);
    }
    float pi[3][3], ar[3][3], ai[3][3];
    mfie::rwg_offsets(T_.ee[t][0], T_.ee[t][1], pi);
    M.block(pi, kap, fd, ar, ai);
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j) {
            S[i][j].re = ar[i][j];
            S[i][j].im = ai[i][j];
        }
}

// V_ij = S_ij l_i l_j / (J_t J_f): lt, lf edge lengths, jti, jfi the inverse Jacobians
__device__ __forceinline__ void mfie_block_f32(const C2f (&S)[3][3], const float (&lt)[3], float jti,
                                               const float (&lf)[3], float jfi, C2f (&V)[3][3]) {
    const float jj = jti * jfi;
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j) {
            const float m = lt[i] * lf[j] * jj;
            V[i][j].re = S[i][j].re * m;
            V[i][j].im = S[i][j].im * m;
        }
}

// Edge vectors, edge lengths (|v1 - v0|, |v2 - v0|, |v1 - v2|) and 1 / J of a geometry row, in FP32.
__device__ __forceinline__ void efie_geom_f32(const double *g, float (&e)[2][3], float (&l)[3], float &jinv) {
#pragma unroll
    for (int d = 0; d < 3; ++d) {
        e[0][d] = (float)g[3 + d];
        e[1][d] = (float)g[6 + d];
    }
    l[0] = (float)sqrt(g[3] * g[3] + g[4] * g[4] + g[5] * g[5]);
    l[1] = (float)sqrt(g[6] * g[6] + g[7] * g[7] + g[8] * g[8]);
    const double d0 = g[3] - g[6], d1 = g[4] - g[7], d2 = g[5] - g[8];
    l[2] = (float)sqrt(d0 * d0 + d1 * d1 + d2 * d2);
    jinv = (float)(1.0 / g[12]);
}

template <typename OutT, bool CPLX>
__device__ __forceinline__ void store_direct(void *out, int64_t ld, int row, int col, C2d v, int accumulate) {
    OutT *o = reinterpret_cast<OutT *>(out);
    int64_t idx = (int64_t)row * ld + col;
    if constexpr (CPLX) {
        using V2 = typename std::conditional<sizeof(OutT) == 8, double2, float2>::type;
        V2 *p = reinterpret_cast<V2 *>(o + 2 * idx);
        V2 val;
        if (accumulate) {
            val = *p;
            val.x += (OutT)v.re;
            val.y += (OutT)v.im;
        } else {
            val.x = (OutT)v.re;
            val.y = (OutT)v.im;
        }
        *p = val;
    } else {
        OutT *p = o + idx;
        if (accumulate) p[0] += (OutT)v.re;
        else p[0] = (OutT)v.re;
    }
}

// ------------------------------------------------------------------ kernel
// DUAL: OP is a single layer and the kernel also assembles the double layer of the same family
// into a.out2 from the same point-pair geometry and phase factors.
// SYM (TILE, test space == trial space on one grid): every unordered element pair is
// evaluated once. The test tiles are cut inside the trial blocks of one element order; a work
// item is a trial block with a run of test tiles, all of them in this block or in earlier ones
// (the workgroup keeps its trial elements in registers across the run). In the tile's own
// block only the pairs with list position of the trial element > that of the test element
// count. Each pair (e, f) adds its single-layer block S_ef to a.out at (e, f) only (the caller
// adds the transpose afterwards: S_fe = S_ef^T); with DUAL also its double-layer block to a.out2
// at (e, f) and the double layer of the reversed pair (test normal) to a.out2 at (f, e) (REV).
// waves per SIMD the launch bound asks for: the FP32 fused pass and the symmetric single layer
// fit 3; the symmetric single + double layer pass and the mixed EFIE passes (ordered and
// symmetric) keep more state (2);
// the single-operator FP64 passes, ordered or symmetric, 2 (their few spills lie outside the
// point-pair loops); the others leave it to the compiler
template <int OP, bool MIXED, bool DUAL, bool SYM>
constexpr int regular_waves() {
    return !MIXED ? (DUAL ? 1 : 2) : SYM ? (DUAL || is_rwg_op<OP>() ? 2 : 3) : DUAL ? 3 : is_rwg_op<OP>() ? 2 : 1;
}

// UPPER (direct DP0 stores, a symmetric operator on one space): only the pairs whose trial element
// comes later in the list than the test element are evaluated; each is stored at (i, j) and at its
// mirror image (j, i), and the diagonal gets zeros (bem_regular_assemble_symmetric).
// ORB: a.rule is a 6-point rule in orbit order (orbit_rule); the FP32 pairs use the orbit moments.
template <int OP, int NTS, int NSS, int QT, int TM, bool MIXED, bool KAHAN, bool TILE, int MAXU_T, int MAXU_S,
          typename OutT, bool DUAL = false, bool SYM = false, bool UPPER = false, bool ORB = false>
__global__ __launch_bounds__(kRegThreads, (regular_waves<OP, MIXED, DUAL, SYM>())) void regular_kernel(
    const RegularArgs a) {
    static_assert(!SYM || TILE, "the symmetric pass is tiled");
    // REV: the double layer of the reversed pair (symmetric single + double layer pass)
    constexpr bool REV = SYM && DUAL;
    static_assert(!UPPER || (!TILE && !DUAL && NTS == 1 && NSS == 1), "UPPER stores DP0 blocks directly");
    constexpr bool CPLX = OpTraits<OP>::is_complex;
    constexpr int NT = NTS, NS = NSS;
    constexpr int NC = CPLX ? 2 : 1;
    constexpr int OP2 = OP + 1;  // the double layer of the family (DUAL only)
    constexpr int GOP = green_op<OP>();  // the point-pair sums use this operator's Green's function
    // OWNER: the trial-side reduction of the tile runs without atomics. For every test
    // element each lane stores its local block to LDS; after a barrier the lane owning trial
    // slot s sums the block entries of the (element, j) incidences of s and updates its own
    // column of the tile. The FP64 mode, which is compute bound, adds with LDS atomics.
    constexpr bool OWNER = TILE && MIXED;
    using AccT = OutT;  // LDS accumulation type: float in MIXED (complex64), double in FP64 mode
    const int nq = QT > 0 ? QT : a.rule.n;

    // test elements in LDS: one tile, or for SYM the whole run of tiles of the work item
    constexpr int TMC = SYM ? kSymRunElements : TM;
    constexpr int QC = QT > 0 ? QT : kMaxQ;
    // EFX: the MIXED EFIE block is formed in FP32 from staged edge data (efie_block_f32)
    constexpr bool EFX = OP == kMaxwellEFIE && MIXED;
    // HSX: the MIXED hypersingular block is formed in FP32 from staged edge vectors and nm / J
    constexpr bool HSX = is_hypersingular<OP>() && MIXED;
    // MFX: the MIXED magnetic field block (mfie_pair_fp32 / mfie_block_f32); there is no FP64 or untiled mode
    constexpr bool MFX = OP == kMaxwellMFIE;
    static_assert(!MFX || (MIXED && TILE && !DUAL && NTS == 3 && NSS == 3), "the tiled MFIE pass is mixed precision");
    __shared__ TestTile<TMC, MIXED ? 1 : TMC, QC, EFX || HSX || MFX> T_;
    __shared__ AccT acc[TILE ? MAXU_T * MAXU_S * NC : 1];
    __shared__ AccT acc2[TILE && DUAL ? MAXU_T * MAXU_S * NC : 1];
    __shared__ AccT acc3[TILE && REV ? MAXU_T * MAXU_S * NC : 1];
    // owner-pass stage, [(j * NT + i)][lane][re, im]; column kRegThreads of the rows j = 0 stays
    // zero and stands in for missing incidences, so the gather below needs no branches
    constexpr int SROW = kRegThreads + 1;
    __shared__ __align__(16) float stage[OWNER ? NT * NS * SROW * NC : 1];
    // TILE: the trial block's dofs and, double-buffered, the current and the next test tile's dofs
    // and (start, end, dof count). The tile loop reads them from LDS only: on gfx9 a wait for a
    // global load also waits for every older global atomic, so a global load between two tile
    // flushes would stall the wave until the previous flush's atomics have completed.
    __shared__ int32_t sh_sdofs[TILE ? MAXU_S : 1];
    __shared__ int32_t sh_tdofs[TILE ? 2 : 1][TILE ? MAXU_T : 1];
    __shared__ int32_t sh_tmeta[TILE ? 2 : 1][3];
    // UPPER: the tile's mirror images, [trial lane][test element] (padded row), written out after
    // the tile as whole row segments; 16-byte entries (complex128) are stored directly instead
    using MirT = typename std::conditional<
        CPLX, typename std::conditional<sizeof(OutT) == 8, double2, float2>::type, OutT>::type;
    constexpr bool MSTAGE = UPPER && sizeof(MirT) <= 8;
    __shared__ MirT mir[MSTAGE ? kRegThreads * (TM + 1) : 1];

    const int tid = threadIdx.x;
    int block_id = blockIdx.x, tile_begin = blockIdx.y, tile_end = blockIdx.y + 1;
    if constexpr (SYM) {
        block_id = a.work[3 * blockIdx.x];
        tile_begin = a.work[3 * blockIdx.x + 1];
        tile_end = a.work[3 * blockIdx.x + 2];
        // malformed work items (the host never produces them); uniform per workgroup, before any barrier
        if (block_id < 0 || block_id >= a.sb.nblocks || tile_begin < 0 || tile_end > a.tb.nblocks ||
            tile_begin > tile_end)
            return;
    }
    int s0, s1;
    if constexpr (TILE) {
        s0 = a.sb.start[block_id];
        s1 = a.sb.start[block_id + 1];
    } else {
        s0 = blockIdx.x * kRegThreads;
        s1 = min(s0 + kRegThreads, a.trial.n);
    }
    // guards against malformed block descriptions (the host never produces them)
    s1 = min(s1, s0 + kRegThreads);
    const int jf = s0 + tid;
    const bool valid = jf < s1;
    // UPPER: a trial block that ends at or before the test tile holds no pair of the upper
    // triangle (uniform per workgroup, before any barrier)
    if constexpr (UPPER) {
        if (s1 <= (int)blockIdx.y * TM + 1) return;
    }

    // stage the test elements at list positions [pos0, pos0 + n) into T_[0, n): all lanes share
    // the element tasks and the (element, quadrature point) tasks
    auto stage_tests = [&](int pos0, int n) {
        for (int task = tid; task < n * (1 + nq); task += kRegThreads) {
            const int el = task < n ? task : (task - n) / nq;
            const int pos = pos0 + el;
            const int e = a.test.elements[pos];
            const double *g = a.test.geom + (int64_t)e * BEM_GEOM_STRIDE;
            if (task < n) {
                T_.elem[el] = e;
                for (int i = 0; i < 3; ++i) T_.vid[el][i] = a.test.vids[4 * e + i];
                for (int i = 0; i < NT; ++i) {
                    T_.mult[el][i] = a.test.mult[(int64_t)e * NT + i];
                    if constexpr (TILE) T_.slot[el][i] = min(max(a.tb.slots[3 * pos + i], 0), MAXU_T - 1);
                }
                centroid_rn(g, T_.c[el]);
                T_.diam[el] = g[13];
                const double nm = a.test.nmult ? (double)a.test.nmult[e] : 1.0;
                T_.nx[el][0] = (float)(g[9] * nm);
                T_.nx[el][1] = (float)(g[10] * nm);
                T_.nx[el][2] = (float)(g[11] * nm);
                if constexpr (EFX || HSX || MFX) efie_geom_f32(g, T_.ee[el], T_.el[el], T_.ej[el]);
                if constexpr (HSX) T_.ej[el] *= (float)nm;
            } else {
                const int p = (task - n) - el * nq;
                const double xi = a.rule.x[p], eta = a.rule.y[p];
                float off[3];
                for (int d = 0; d < 3; ++d) {
                    off[d] = (float)(g[3 + d] * (xi - 1.0 / 3.0) + g[6 + d] * (eta - 1.0 / 3.0));
                    if constexpr (!MIXED) T_.xp[el][p][d] = g[d] + g[3 + d] * xi + g[6 + d] * eta;
                    T_.a[el][p][d] = off[d];
                }
                T_.aa[el][p] = off[0] * off[0] + off[1] * off[1] + off[2] * off[2];
                T_.wp[el][p] = (float)(a.rule.w[p] * g[12]);
            }
        }
    };
    int run0 = 0;
    if constexpr (SYM) {
        run0 = a.tb.start[tile_begin];
        const int run_n = a.tb.start[tile_end] - run0;
        if (run_n < 0 || run_n > TMC) return;  // malformed work item; uniform, before any barrier
        stage_tests(run0, run_n);  // the first tile's barrier publishes it
    }

    // trial element of this lane
    int f = 0;
    int fv0 = -1, fv1 = -1, fv2 = -1;
    double snm = 1.0, cf0 = 0, cf1 = 0, cf2 = 0, diam_f = 0;
    const double *gf = a.trial.geom;
    double mf[NS];
    int slot_f[NS];
    TrialF32<QT> F32;
    double yq[QT > 0 ? QT : 1][3], wq64[QT > 0 ? QT : 1], ny64[3] = {0, 0, 0};
    float ef[2][3] = {{0.f, 0.f, 0.f}, {0.f, 0.f, 0.f}}, lf[3] = {0.f, 0.f, 0.f}, jfi = 0.f;  // EFX
    float ikr = 0.f, iki = 0.f, invr = 0.f, invi = 0.f;  // EFX: ik = -Im k + i Re k and 1 / (ik)
    float kapf[3][3] = {{0.f, 0.f, 0.f}, {0.f, 0.f, 0.f}, {0.f, 0.f, 0.f}};  // MFX: c_f - Q_j
    const float k2r = (float)(a.kr * a.kr - a.ki * a.ki), k2i = (float)(2.0 * a.kr * a.ki);  // HSX: k^2
    if constexpr (EFX) {
        const double den = a.ki * a.ki + a.kr * a.kr;
        ikr = (float)(-a.ki);
        iki = (float)a.kr;
        invr = (float)(-a.ki / den);
        invi = (float)(-a.kr / den);
    }
#pragma unroll
    for (int j = 0; j < NS; ++j) {
        mf[j] = 0.0;
        slot_f[j] = 0;
    }
    if (valid) {
        f = a.trial.elements[jf];
        gf = a.trial.geom + (int64_t)f * BEM_GEOM_STRIDE;
        fv0 = a.trial.vids[4 * f];
        fv1 = a.trial.vids[4 * f + 1];
        fv2 = a.trial.vids[4 * f + 2];
        snm = a.trial.nmult ? (double)a.trial.nmult[f] : 1.0;
        double cfv[3];
        centroid_rn(gf, cfv);
        cf0 = cfv[0];
        cf1 = cfv[1];
        cf2 = cfv[2];
        diam_f = gf[13];
#pragma unroll
        for (int j = 0; j < NS; ++j) {
            mf[j] = a.trial.mult[(int64_t)f * NS + j];
            if constexpr (TILE) slot_f[j] = min(max(a.sb.slots[3 * jf + j], 0), MAXU_S - 1);
        }
        ny64[0] = gf[9] * snm;
        ny64[1] = gf[10] * snm;
        ny64[2] = gf[11] * snm;
        if constexpr (EFX || HSX || MFX) efie_geom_f32(gf, ef, lf, jfi);
        if constexpr (HSX) jfi *= (float)snm;
        if constexpr (MFX) mfie::rwg_offsets(ef[0], ef[1], kapf);
        if constexpr (MIXED) {
            F32.ny[0] = (float)ny64[0];
            F32.ny[1] = (float)ny64[1];
            F32.ny[2] = (float)ny64[2];
#pragma unroll(QT > 0 ? QT : 1)
            for (int q = 0; q < nq; ++q) {
                double s0q = a.rule.x[q] - 1.0 / 3.0, s1q = a.rule.y[q] - 1.0 / 3.0;
                float b0 = (float)(gf[3] * s0q + gf[6] * s1q);
                float b1 = (float)(gf[4] * s0q + gf[7] * s1q);
                float b2 = (float)(gf[5] * s0q + gf[8] * s1q);
                F32.b[q][0] = b0;
                F32.b[q][1] = b1;
                F32.b[q][2] = b2;
                F32.bb[q] = b0 * b0 + b1 * b1 + b2 * b2;
                F32.wq[q] = (float)(a.rule.w[q] * gf[12] * kInv4Pi);
                F32.bn[q] = b0 * F32.ny[0] + b1 * F32.ny[1] + b2 * F32.ny[2];
            }
        } else if constexpr (QT > 0) {
#pragma unroll
            for (int q = 0; q < QT; ++q) {
                double s0q = a.rule.x[q], s1q = a.rule.y[q];
                for (int d = 0; d < 3; ++d) yq[q][d] = gf[d] + gf[3 + d] * s0q + gf[6 + d] * s1q;
                wq64[q] = a.rule.w[q] * gf[12];
            }
        }
    }

    // owner lane state: the incidences (lane p, shape function j) of trial slots tid (and
    // tid + 256 when a block has more slots than lanes) in this block, as stage offsets of row
    // (j, i = 0) and column p; missing ones point at the zero column
    constexpr int OWNS = OWNER && MAXU_S > kRegThreads ? 2 : 1;
    constexpr int KINC = is_rwg_op<OP>() ? kRwgMaxInc : kMaxInc;
    int own_off[OWNER ? OWNS : 1][OWNER ? KINC : 1];
    int own_nds = 0;
    if constexpr (OWNER) {
        own_nds = min(a.sb.ndofs[block_id], MAXU_S);
#pragma unroll
        for (int o = 0; o < OWNS; ++o) {
            const int sl = tid + o * kRegThreads;
            const int16_t *src =
                a.sb.inc + ((int64_t)block_id * a.sb.maxu + min(sl, a.sb.maxu - 1)) * a.sb.maxinc;
#pragma unroll
            for (int k = 0; k < KINC; ++k) {
                const int code = (sl < own_nds && k < a.sb.maxinc) ? (int)src[k] : -1;
                const int p = code / 3, j = code - 3 * p;
                own_off[o][k] = (code >= 0 && code < 3 * kRegThreads) ? (j * NT * SROW + p) * NC : kRegThreads * NC;
            }
        }
        if (tid < NT * NC) stage[(tid / NC) * SROW * NC + kRegThreads * NC + tid % NC] = 0.f;  // zero column
    }

    // load test tile `tl`'s dofs and (start, end, dof count) into LDS buffer `buf` (no atomics of this
    // workgroup are outstanding when it runs: before the tile loop, or before a flush's atomics)
    auto load_tile_meta = [&](int tl, int buf) {
        if constexpr (TILE) {
            const int nmax = min(a.tb.maxu, MAXU_T);
            if (tid < nmax) sh_tdofs[buf][tid] = a.tb.dofs[(int64_t)tl * a.tb.maxu + tid];
            if (tid == kRegThreads - 1) {
                sh_tmeta[buf][0] = a.tb.start[tl];
                sh_tmeta[buf][1] = a.tb.start[tl + 1];
                sh_tmeta[buf][2] = min(max(a.tb.ndofs[tl], 0), nmax);
            }
        }
    };
    int ndt = 0, nds = 0;
    if constexpr (TILE) {
        nds = min(max(a.sb.ndofs[block_id], 0), min(a.sb.maxu, MAXU_S));
        for (int s = tid; s < nds; s += kRegThreads) sh_sdofs[s] = a.sb.dofs[(int64_t)block_id * a.sb.maxu + s];
        load_tile_meta(tile_begin, 0);  // published by the first tile's barrier
    }

    for (int tile_id = tile_begin; tile_id < tile_end; ++tile_id) {
        int t0, t1;
        const int buf = (tile_id - tile_begin) & 1;
        if constexpr (TILE) {
            if (tile_id == tile_begin) {  // the LDS copy is published by the barrier below
                t0 = a.tb.start[tile_id];
                t1 = a.tb.start[tile_id + 1];
                ndt = min(max(a.tb.ndofs[tile_id], 0), min(a.tb.maxu, MAXU_T));
            } else {  // loaded during the previous tile's flush, published by its closing barrier
                t0 = __builtin_amdgcn_readfirstlane(sh_tmeta[buf][0]);
                t1 = __builtin_amdgcn_readfirstlane(sh_tmeta[buf][1]);
                ndt = __builtin_amdgcn_readfirstlane(sh_tmeta[buf][2]);
            }
        } else {
            t0 = tile_id * TM;
            t1 = min(t0 + TM, a.test.n);
        }
        const int ntile = max(0, min(t1 - t0, TM));
        // SYM: the test tile lies in this trial block; only pairs with jf > test position count
        const bool diag = SYM && t0 >= s0 && t0 < s1;
        const int toff = SYM ? t0 - run0 : 0;  // T_ index of the tile's first element

        if constexpr (TILE) {
            for (int i = tid; i < MAXU_T * MAXU_S * (CPLX ? 2 : 1); i += kRegThreads) {
                acc[i] = (AccT)0;
                if constexpr (DUAL) acc2[i] = (AccT)0;
                if constexpr (REV) acc3[i] = (AccT)0;
            }
        }
        if constexpr (!SYM) stage_tests(t0, ntile);
        __syncthreads();

        for (int t = 0; t < ntile; ++t) {
            const int te = toff + t;
            const int e = T_.elem[te];
            const double *gt = a.test.geom + (int64_t)e * BEM_GEOM_STRIDE;
            bool adj = false;
            if (a.identical) {
                int v0 = T_.vid[te][0], v1 = T_.vid[te][1], v2 = T_.vid[te][2];
                adj = v0 == fv0 || v0 == fv1 || v0 == fv2 || v1 == fv0 || v1 == fv1 || v1 == fv2 || v2 == fv0 ||
                      v2 == fv1 || v2 == fv2;
            }
            if constexpr (SYM) adj = adj || (diag && jf <= t0 + t);  // the pair belongs to the other orientation
            if constexpr (UPPER) adj = adj || jf <= t0 + t;            // lower triangle: other pairs' mirror images
            C2d S[NTS][NSS], S2[NTS][NSS], S3[NTS][NSS];     // FP64 moments (FP64 mode; MIXED paths that need the FP64 epilogue)
            C2f Sf[NTS][NSS], Sf2[NTS][NSS], Sf3[NTS][NSS];  // FP32 moments (MIXED)
    #pragma unroll
            for (int i = 0; i < NTS; ++i)
    #pragma unroll
                for (int j = 0; j < NSS; ++j) {
                    S[i][j].re = S[i][j].im = 0.0;
                    Sf[i][j].re = Sf[i][j].im = 0.f;
                    S2[i][j].re = S2[i][j].im = 0.0;
                    Sf2[i][j].re = Sf2[i][j].im = 0.f;
                    S3[i][j].re = S3[i][j].im = 0.0;
                    Sf3[i][j].re = Sf3[i][j].im = 0.f;
                }

            // MIXED: near pairs are left to the FP64 near-field kernel (bem_near_assemble)
            bool skip = adj;
            // normal multipliers of the pair (the hypersingular epilogue)
            const double nmp = (a.test.nmult ? (double)a.test.nmult[e] : 1.0) * snm;
            if (valid && !adj) {
                if constexpr (!MIXED) {
                    double tnm = a.test.nmult ? (double)a.test.nmult[e] : 1.0;
                    bool fast = false, shift = false;
                    double R0 = 0.0;
                    if constexpr (OpTraits<GOP>::is_complex) {
                        const double c0 = cf0 - T_.c[te][0], c1 = cf1 - T_.c[te][1], c2 = cf2 - T_.c[te][2];
                        R0 = sqrt(c0 * c0 + c1 * c1 + c2 * c2);
                        fast = fp64_phase_reduced(a.kr, a.ki, R0, T_.diam[te] + diam_f);
                        // |k (r - R0)| <= k (diam_e + diam_f) <= 1 on every lane: e^{ikr} = e^{ik R0} e^{ik (r - R0)}
                        // with the second factor from its Taylor series, the first once per pair
                        shift = fast && __all(fabs(a.kr) * (T_.diam[te] + diam_f) <= 1.0);
                    }
                    if constexpr (QT > 0 && OpTraits<GOP>::is_complex) {
                        if (shift) {
                            const double ph0 = a.kr * R0;
                            pair_fp64_regs<GOP, NTS, NSS, QT, DUAL, REV, kPhShift | kPhReal>(
                                T_, te, gt, tnm, yq, wq64, ny64, a.rule, a.kr, a.ki, S, S2, S3, ph0);
                            double ps, pc;
                            sincos_reduced(ph0, &ps, &pc);
                            auto rotate = [&](C2d(&Sm)[NTS][NSS]) {
#pragma unroll
                                for (int i = 0; i < NTS; ++i)
#pragma unroll
                                    for (int j = 0; j < NSS; ++j) {
                                        const double re = Sm[i][j].re, im = Sm[i][j].im;
                                        Sm[i][j].re = re * pc - im * ps;
                                        Sm[i][j].im = re * ps + im * pc;
                                    }
                            };
                            rotate(S);
                            if constexpr (DUAL) rotate(S2);
                            if constexpr (REV) rotate(S3);
                        } else if (fast)
                            pair_fp64_regs<GOP, NTS, NSS, QT, DUAL, REV, kPhReduced | kPhReal>(
                                T_, te, gt, tnm, yq, wq64, ny64, a.rule, a.kr, a.ki, S, S2, S3);
                        else
                            pair_fp64_regs<GOP, NTS, NSS, QT, DUAL, REV>(T_, te, gt, tnm, yq, wq64, ny64, a.rule, a.kr,
                                                                         a.ki, S, S2, S3);
                    } else if constexpr (QT > 0) {
                        if (fast)
                            pair_fp64_regs<GOP, NTS, NSS, QT, DUAL, REV, kPhReduced | kPhReal>(
                                T_, te, gt, tnm, yq, wq64, ny64, a.rule, a.kr, a.ki, S, S2, S3);
                        else
                            pair_fp64_regs<GOP, NTS, NSS, QT, DUAL, REV>(T_, te, gt, tnm, yq, wq64, ny64, a.rule, a.kr,
                                                                         a.ki, S, S2, S3);
                    } else {
                        if (fast)
                            pair_fp64_geom<GOP, NTS, NSS, QT, DUAL, REV, kPhReduced | kPhReal>(
                                gt, gf, tnm, snm, a.rule, a.kr, a.ki, S, S2, S3);
                        else
                            pair_fp64_geom<GOP, NTS, NSS, QT, DUAL, REV>(gt, gf, tnm, snm, a.rule, a.kr, a.ki, S, S2,
                                                                         S3);
                    }
                } else {
                    const double cfv[3] = {cf0, cf1, cf2};
                    double dist2;
                    skip = pair_is_near(T_.c[te], cfv, T_.diam[te], diam_f, a.near_factor, &dist2);
                    if (!skip) {
                        double d0 = cf0 - T_.c[te][0], d1 = cf1 - T_.c[te][1], d2 = cf2 - T_.c[te][2];
                        if constexpr (MFX)
                            mfie_pair_fp32<QT>(T_, te, F32, kapf, d0, d1, d2, dist2, a.rule, a.kr, a.ki, Sf);
                        else if (OpTraits<OP>::is_complex && a.ki != 0.0)
                            pair_fp32<GOP, NTS, NSS, QT, KAHAN, true, DUAL, REV, ORB>(T_, te, F32, d0, d1, d2, dist2, a.rule,
                                                                                       a.kr, a.ki, Sf, Sf2, Sf3);
                        else
                            pair_fp32<GOP, NTS, NSS, QT, KAHAN, false, DUAL, REV, ORB>(T_, te, F32, d0, d1, d2, dist2,
                                                                                        a.rule, a.kr, a.ki, Sf, Sf2, Sf3);
                    }
                }
            }
            if constexpr (MIXED && !OWNER) {
    #pragma unroll
                for (int i = 0; i < NTS; ++i)
    #pragma unroll
                    for (int j = 0; j < NSS; ++j) {
                        S[i][j].re = (double)Sf[i][j].re;
                        S[i][j].im = (double)Sf[i][j].im;
                        if constexpr (DUAL) {
                            S2[i][j].re = (double)Sf2[i][j].re;
                            S2[i][j].im = (double)Sf2[i][j].im;
                        }
                        if constexpr (REV) {
                            S3[i][j].re = (double)Sf3[i][j].re;
                            S3[i][j].im = (double)Sf3[i][j].im;
                        }
                    }
            }

            if constexpr (OWNER) {
                // stage this lane's block (zero for idle lanes, skipped pairs, zero multipliers), then
                // the owner lane of every trial slot adds the entries of its incidences to its tile column
                auto owner_pass = [&](const C2f(&Sm)[NTS][NSS], AccT *tile) {
                    C2f V[NTS][NSS];
                    if constexpr (OP == kMaxwellEFIE)
                        efie_block_f32(Sm, T_.ee[te], T_.el[te], T_.ej[te], ef, lf, jfi, ikr, iki, invr, invi, V);
                    else if constexpr (HSX)
                        hs_block_f32<CPLX>(Sm, T_.ee[te], T_.ej[te], T_.nx[te], ef, jfi, F32.ny, k2r, k2i, V);
                    else if constexpr (MFX)
                        mfie_block_f32(Sm, T_.el[te], T_.ej[te], lf, jfi, V);
                    else
                        scalar_block_f32<NTS, NSS>(Sm, V);
    #pragma unroll
                    for (int i = 0; i < NT; ++i)
    #pragma unroll
                        for (int j = 0; j < NS; ++j) {
                            float vr = 0.f, vi = 0.f;
                            const float m = (float)T_.mult[te][i] * (float)mf[j];
                            if (valid && !skip && m != 0.f) {
                                vr = V[i][j].re * m;
                                vi = V[i][j].im * m;
                            }
                            float *dst = stage + ((j * NT + i) * SROW + tid) * NC;
                            if constexpr (CPLX) *reinterpret_cast<float2 *>(dst) = make_float2(vr, vi);
                            else dst[0] = vr;
                        }
                    __syncthreads();
    #pragma unroll
                    for (int o = 0; o < OWNS; ++o) {
                        const int sl = tid + o * kRegThreads;
                        if (sl >= own_nds) break;
                        // all reads of a row are issued before any sum: one LDS round trip per row
                        auto gather = [&](int i, int k0, int k1, float &sr, float &si) {
    #pragma unroll
                            for (int k = k0; k < k1; ++k) {
                                const float *src = stage + own_off[o][k] + i * SROW * NC;
                                if constexpr (CPLX) {
                                    const float2 v = *reinterpret_cast<const float2 *>(src);
                                    sr += v.x;
                                    si += v.y;
                                } else {
                                    sr += src[0];
                                }
                            }
                        };
                        constexpr int KLO = KINC < 6 ? KINC : 6;
                        const bool more = a.sb.maxinc > KLO;  // uniform: vertices of valence > 6
    #pragma unroll
                        for (int i = 0; i < NT; ++i) {
                            if (T_.mult[te][i] == 0.0) continue;
                            float sr = 0.f, si = 0.f;
                            gather(i, 0, KLO, sr, si);
                            if (more) gather(i, KLO, KINC, sr, si);
                            AccT *dst = tile + NC * (T_.slot[te][i] * MAXU_S + sl);
                            dst[0] += (AccT)sr;
                            if constexpr (CPLX) dst[1] += (AccT)si;
                        }
                    }
                    __syncthreads();
                };
                owner_pass(Sf, acc);
                if constexpr (DUAL) owner_pass(Sf2, acc2);
                if constexpr (REV) owner_pass(Sf3, acc3);
            } else if (valid) {
    #pragma unroll
                for (int i = 0; i < NT; ++i) {
                    const double mi = T_.mult[te][i];
    #pragma unroll
                    for (int j = 0; j < NS; ++j) {
                        const double m = mi * mf[j];
                        if constexpr (TILE) {
                            if (m != 0.0 && !skip) {
                                C2d v = local_value<OP, NTS, NSS>(S, i, j, gt, gf, a.kr, a.ki, nmp);
                                const int off = (CPLX ? 2 : 1) * (T_.slot[te][i] * MAXU_S + slot_f[j]);
                                unsafeAtomicAdd(acc + off, (AccT)(v.re * m));
                                if constexpr (CPLX) unsafeAtomicAdd(acc + off + 1, (AccT)(v.im * m));
                                if constexpr (DUAL) {
                                    C2d v2 = local_value<OP2, NTS, NSS>(S2, i, j, gt, gf, a.kr, a.ki);
                                    unsafeAtomicAdd(acc2 + off, (AccT)(v2.re * m));
                                    if constexpr (CPLX) unsafeAtomicAdd(acc2 + off + 1, (AccT)(v2.im * m));
                                }
                                if constexpr (REV) {
                                    C2d v3 = local_value<OP2, NTS, NSS>(S3, i, j, gt, gf, a.kr, a.ki);
                                    unsafeAtomicAdd(acc3 + off, (AccT)(v3.re * m));
                                    if constexpr (CPLX) unsafeAtomicAdd(acc3 + off + 1, (AccT)(v3.im * m));
                                }
                            }
                        } else {
                            C2d v = local_value<OP, NTS, NSS>(S, i, j, gt, gf, a.kr, a.ki, nmp);
                            v.re *= m;
                            v.im *= m;
                            int row = a.test.l2g[(int64_t)e * NT + i];
                            int col = a.trial.l2g[(int64_t)f * NS + j];
                            if constexpr (UPPER) {
                                // the entry and its mirror image (dof = list position): a lane
                                // whose trial element comes before the test element stores
                                // nothing (its entries are other pairs' mirror images), the
                                // diagonal stores the zero of its coincident pair
                                if (jf >= t0 + t) store_direct<OutT, CPLX>(a.out, a.ld, row, col, v, 0);
                                if constexpr (MSTAGE) {
                                    MirT mv;
                                    if constexpr (CPLX) {
                                        mv.x = (OutT)v.re;
                                        mv.y = (OutT)v.im;
                                    } else {
                                        mv = (OutT)v.re;
                                    }
                                    mir[tid * (TM + 1) + t] = mv;
                                } else {
                                    if (jf > t0 + t) store_direct<OutT, CPLX>(a.out, a.ld, col, row, v, 0);
                                }
                            } else {
                                store_direct<OutT, CPLX>(a.out, a.ld, row, col, v, a.accumulate);
                            }
                            if constexpr (DUAL) {
                                C2d v2 = local_value<OP2, NTS, NSS>(S2, i, j, gt, gf, a.kr, a.ki);
                                v2.re *= m;
                                v2.im *= m;
                                store_direct<OutT, CPLX>(a.out2, a.ld, row, col, v2, a.accumulate);
                            }
                        }
                    }
                }
            }
        }

        if constexpr (MSTAGE) {
            // the mirror images of the tile: row jr (a trial position) gets columns t0 .. t0 + ntile,
            // TM consecutive lanes per row, so a wave stores whole row segments (a 128-byte line for
            // 16 complex64 entries); entries left of the diagonal are other tiles', the diagonal is zero
            __syncthreads();
            MirT *o = reinterpret_cast<MirT *>(a.out);
            const int c = tid % TM, ic = t0 + c;
            for (int r = tid / TM; r < s1 - s0; r += kRegThreads / TM) {
                const int jr = s0 + r;
                if (c < ntile && jr >= ic) {
                    MirT mv;
                    if (jr > ic) {
                        mv = mir[r * (TM + 1) + c];
                    } else if constexpr (CPLX) {
                        mv.x = mv.y = (OutT)0;
                    } else {
                        mv = (OutT)0;
                    }
                    o[(int64_t)jr * a.ld + ic] = mv;
                }
            }
        }

        if constexpr (TILE) {
            __syncthreads();
            // the next tile's dofs and bounds, before this flush issues its atomics
            if (tile_id + 1 < tile_end) load_tile_meta(tile_id + 1, buf ^ 1);
            const int32_t *dt = sh_tdofs[buf];
            const int32_t *ds = sh_sdofs;
            // transposed: entry (test slot ia, trial slot is) goes to (ds[is], dt[ia]); consecutive
            // lanes then walk one output row. Only LDS reads and global atomics: nothing waits for
            // the atomics, which complete while the next tile is computed.
            auto flush = [&](const AccT *tile, void *out, bool transposed) {
                OutT *o = reinterpret_cast<OutT *>(out);
                for (int idx = tid; idx < ndt * nds; idx += kRegThreads) {
                    int ia, is;
                    if (transposed) {
                        is = idx / ndt;
                        ia = idx - is * ndt;
                    } else {
                        ia = idx / nds;
                        is = idx - ia * nds;
                    }
                    const AccT *src = tile + (CPLX ? 2 : 1) * (ia * MAXU_S + is);
                    const int64_t pos = transposed ? (int64_t)ds[is] * a.ld + dt[ia] : (int64_t)dt[ia] * a.ld + ds[is];
                    if constexpr (CPLX) {
                        if (src[0] != (AccT)0) unsafeAtomicAdd(o + 2 * pos, src[0]);
                        if (src[1] != (AccT)0) unsafeAtomicAdd(o + 2 * pos + 1, src[1]);
                    } else {
                        if (src[0] != (AccT)0) unsafeAtomicAdd(o + pos, src[0]);
                    }
                }
            };
            flush(acc, a.out, false);
            if constexpr (DUAL) flush(acc2, a.out2, false);
            if constexpr (REV) flush(acc3, a.out2, true);
        }
        // the next tile reuses the test tile and the LDS tiles
        if (tile_id + 1 < tile_end) __syncthreads();
    }  // tile loop
}

// ------------------------------------------------------------------ dispatch
template <int OP, int NTS, int NSS, int QT, bool MIXED, typename OutT, bool DUAL = false, bool SYM = false,
          bool UPPER = false>
static int launch3(const RegularArgs &a, int kahan, bool tile, hipStream_t stream) {
    // DUAL keeps two LDS tiles and SYM three; their smaller test tiles keep the LDS footprint
    // at 3 workgroups per CU
    // mixed EFIE: trial blocks of up to BEM_TILE_RWG_MIXED_TRIAL_DOFS edges (a 256-element patch),
    // owner lanes with two slots each, smaller test tiles to stay at 2 workgroups per CU
    constexpr bool RWGM = is_rwg_op<OP>() && MIXED;
    constexpr int TM = SYM ? BEM_TILE_SYM_TEST_ELEMENTS
                     : DUAL ? BEM_TILE_PAIR_TEST_ELEMENTS
                     : RWGM ? BEM_TILE_RWG_MIXED_TEST_ELEMENTS
                            : BEM_TILE_TEST_ELEMENTS;
    constexpr bool CPLX = OpTraits<OP>::is_complex;
    // LDS tile of distinct dofs: test slots x trial slots (BEM_TILE_TEST_DOFS x BEM_TILE_TRIAL_DOFS)
    constexpr int MT = SYM ? BEM_TILE_SYM_TEST_DOFS
                     : DUAL ? BEM_TILE_PAIR_TEST_DOFS
                     : RWGM ? BEM_TILE_RWG_MIXED_TEST_DOFS
                            : BEM_TILE_TEST_DOFS;
    constexpr int MS = RWGM ? BEM_TILE_RWG_MIXED_TRIAL_DOFS : BEM_TILE_TRIAL_DOFS;
    constexpr int KMAX = is_rwg_op<OP>() ? kRwgMaxInc : kMaxInc;
    (void)CPLX;
    dim3 grid;
    if (tile) {
        if (a.tb.maxu > MT || a.sb.maxu > MS || a.tb.maxu < 1 || a.sb.maxu < 1) {
            set_error("dof blocks larger than the LDS tile (%d x %d > %d x %d)", a.tb.maxu, a.sb.maxu, MT, MS);
            return BEM_ERR_INVALID_ARGUMENT;
        }
        if (MIXED && (a.sb.inc == nullptr || a.sb.maxinc < 1 || a.sb.maxinc > KMAX)) {
            set_error("mixed precision needs trial incidence lists with 1..%d entries per dof", KMAX);
            return BEM_ERR_INVALID_ARGUMENT;
        }
        grid = SYM ? dim3(a.nwork) : dim3(a.sb.nblocks, a.tb.nblocks);
    } else {
        if (SYM) {
            set_error("the symmetric pass needs dof blocks");
            return BEM_ERR_INVALID_ARGUMENT;
        }
        if (UPPER && a.test.n != a.trial.n) {
            set_error("the symmetric DP0 pass needs one space on both sides");
            return BEM_ERR_INVALID_ARGUMENT;
        }
        grid = dim3((a.trial.n + kRegThreads - 1) / kRegThreads, (a.test.n + TM - 1) / TM);
    }
    if (grid.x == 0 || grid.y == 0) return BEM_OK;
    if (grid.y > 65535u) {
        set_error("too many test tiles for one launch (%u)", grid.y);
        return BEM_ERR_INVALID_ARGUMENT;
    }
#define BEM_LAUNCH(K, T)                                                                                              \
    hipLaunchKernelGGL((regular_kernel<OP, NTS, NSS, QT, TM, MIXED, K, T, (T ? MT : 1), (T ? MS : 1), OutT, DUAL, \
                                       SYM && T, UPPER && !T>),                                                       \
                       grid, dim3(kRegThreads), 0, stream, a)
    if (tile) {
        if constexpr (UPPER) {
            set_error("the symmetric DP0 pass stores directly (no dof blocks)");
            return BEM_ERR_INVALID_ARGUMENT;
        }
        if constexpr (SYM && MIXED && QT == 6 && NSS == 3 && (OP == kLaplaceSL || OP == kHelmholtzSL)) {
            // the symmetric single layer (and double layer) pass with the orbit moments when the rule allows
            RegularArgs b = a;
            if (!kahan && orbit_rule(b.rule)) {
                hipLaunchKernelGGL((regular_kernel<OP, NTS, NSS, QT, TM, MIXED, false, true, MT, MS, OutT, DUAL, true,
                                                   false, true>),
                                   grid, dim3(kRegThreads), 0, stream, b);
                return check_launch("regular_kernel");
            }
        }
        if constexpr (OP == kMaxwellMFIE) {
            BEM_LAUNCH(false, true);
        } else {
            if (MIXED && kahan) BEM_LAUNCH(true, true);
            else BEM_LAUNCH(false, true);
        }
    } else if constexpr (UPPER) {
        // a small DP0 space (1,280 elements: 400 workgroups of 16-element test tiles, half of them
        // below the diagonal) leaves most SIMDs without a wave: 4-element test tiles give four
        // times the workgroups
        constexpr int TMS = 4;
        if ((int64_t)grid.x * grid.y < kUpperSmallGrid) {
            const dim3 gs(grid.x, (a.test.n + TMS - 1) / TMS);
            if (MIXED && kahan)
                hipLaunchKernelGGL((regular_kernel<OP, NTS, NSS, QT, TMS, MIXED, true, false, 1, 1, OutT, false, false,
                                                   true>),
                                   gs, dim3(kRegThreads), 0, stream, a);
            else
                hipLaunchKernelGGL((regular_kernel<OP, NTS, NSS, QT, TMS, MIXED, false, false, 1, 1, OutT, false, false,
                                                   true>),
                                   gs, dim3(kRegThreads), 0, stream, a);
        } else {
            if (MIXED && kahan) BEM_LAUNCH(true, false);
            else BEM_LAUNCH(false, false);
        }
    } else if constexpr (!SYM && OP != kMaxwellMFIE) {
        if (MIXED && kahan) BEM_LAUNCH(true, false);
        else BEM_LAUNCH(false, false);
    }
#undef BEM_LAUNCH
    return check_launch("regular_kernel");
}

template <int OP, int NTS, int NSS, bool MIXED, typename OutT, bool DUAL = false, bool SYM = false, bool UPPER = false>
static int launch2(const RegularArgs &a, int kahan, bool tile, hipStream_t stream) {
    if (a.rule.n == 6) return launch3<OP, NTS, NSS, 6, MIXED, OutT, DUAL, SYM, UPPER>(a, kahan, tile, stream);
    return launch3<OP, NTS, NSS, 0, MIXED, OutT, DUAL, SYM, UPPER>(a, kahan, tile, stream);
}

template <int OP, bool MIXED, typename OutT, bool DUAL = false, bool SYM = false, bool UPPER = false>
static int launch1(const RegularArgs &a, int kahan, bool tile, hipStream_t stream) {
    if constexpr (OP == kMaxwellEFIE) {
        return launch2<OP, 3, 3, MIXED, OutT, false, SYM>(a, kahan, tile, stream);
    } else if constexpr (is_hypersingular<OP>()) {
        if (nshape_of(a.test.shape) != 3 || nshape_of(a.trial.shape) != 3) {
            set_error("hypersingular operators need P1 shape functions on both sides");
            return BEM_ERR_INVALID_ARGUMENT;
        }
        return launch2<OP, 3, 3, MIXED, OutT, false, SYM>(a, kahan, tile, stream);
    } else if constexpr (UPPER) {
        if (nshape_of(a.test.shape) != 1 || nshape_of(a.trial.shape) != 1) {
            set_error("the symmetric DP0 pass needs piecewise constant spaces");
            return BEM_ERR_INVALID_ARGUMENT;
        }
        return launch2<OP, 1, 1, MIXED, OutT, false, false, true>(a, kahan, tile, stream);
    } else if constexpr (SYM) {
        // shared dofs (the tiled path) on one space: P1
        if (nshape_of(a.test.shape) != 3 || nshape_of(a.trial.shape) != 3) {
            set_error("the symmetric pass needs a space with three shape functions per element (P1)");
            return BEM_ERR_INVALID_ARGUMENT;
        }
        return launch2<OP, 3, 3, MIXED, OutT, DUAL, true>(a, kahan, tile, stream);
    } else {
        int nt = nshape_of(a.test.shape), ns = nshape_of(a.trial.shape);
        if (nt == 1 && ns == 1) return launch2<OP, 1, 1, MIXED, OutT, DUAL>(a, kahan, tile, stream);
        if (nt == 1 && ns == 3) return launch2<OP, 1, 3, MIXED, OutT, DUAL>(a, kahan, tile, stream);
        if (nt == 3 && ns == 1) return launch2<OP, 3, 1, MIXED, OutT, DUAL>(a, kahan, tile, stream);
        return launch2<OP, 3, 3, MIXED, OutT, DUAL>(a, kahan, tile, stream);
    }
}

template <int OP, bool DUAL = false, bool SYM = false, bool UPPER = false>
static int launch_op(const RegularArgs &a, int precision, int kahan, bool tile, int out_dtype, hipStream_t stream) {
    constexpr bool CPLX = OpTraits<OP>::is_complex;
    const bool mixed = precision == BEM_PRECISION_MIXED;
    if (CPLX) {
        if (out_dtype == BEM_C128)
            return mixed ? launch1<OP, true, double, DUAL, SYM, UPPER>(a, kahan, tile, stream)
                         : launch1<OP, false, double, DUAL, SYM, UPPER>(a, kahan, tile, stream);
        if (out_dtype == BEM_C64)
            return mixed ? launch1<OP, true, float, DUAL, SYM, UPPER>(a, kahan, tile, stream)
                         : launch1<OP, false, float, DUAL, SYM, UPPER>(a, kahan, tile, stream);
    } else {
        if (out_dtype == BEM_F64)
            return mixed ? launch1<OP, true, double, DUAL, SYM, UPPER>(a, kahan, tile, stream)
                         : launch1<OP, false, double, DUAL, SYM, UPPER>(a, kahan, tile, stream);
        if (out_dtype == BEM_F32)
            return mixed ? launch1<OP, true, float, DUAL, SYM, UPPER>(a, kahan, tile, stream)
                         : launch1<OP, false, float, DUAL, SYM, UPPER>(a, kahan, tile, stream);
    }
    set_error("output dtype %d does not match operator %d", out_dtype, OP);
    return BEM_ERR_INVALID_ARGUMENT;
}

// One translation unit per operator instantiates launch_op<OP> (regular_op<k>.hip).
int regular_launch_op(int op, const RegularArgs &a, int precision, int kahan, bool tile, int out_dtype, hipStream_t s);
#define BEM_DECLARE_REGULAR_OP(N) \
    int regular_launch_##N(const RegularArgs &a, int precision, int kahan, bool tile, int out_dtype, hipStream_t s);
BEM_DECLARE_REGULAR_OP(0)
BEM_DECLARE_REGULAR_OP(1)
BEM_DECLARE_REGULAR_OP(2)
BEM_DECLARE_REGULAR_OP(3)
BEM_DECLARE_REGULAR_OP(4)
BEM_DECLARE_REGULAR_OP(5)
BEM_DECLARE_REGULAR_OP(6)
BEM_DECLARE_REGULAR_OP(7)
BEM_DECLARE_REGULAR_OP(8)
#undef BEM_DECLARE_REGULAR_OP
// single layer + double layer in one pass (regular_dual<k>.hip, k = the single layer)
// far pairs of the magnetic field operator in mixed precision (tiled, owner reduction; regular_mfie.hip)
int regular_launch_mfie(const RegularArgs &a, int out_dtype, hipStream_t s);
int regular_launch_dual_0(const RegularArgs &a, int precision, int kahan, bool tile, int out_dtype, hipStream_t s);
int regular_launch_dual_3(const RegularArgs &a, int precision, int kahan, bool tile, int out_dtype, hipStream_t s);
// symmetric single + double layer pass (regular_sym<k>.hip)
int regular_launch_sym_0(const RegularArgs &a, int precision, int kahan, int out_dtype, hipStream_t s);
int regular_launch_sym_3(const RegularArgs &a, int precision, int kahan, int out_dtype, hipStream_t s);
// the symmetric pass of one symmetric operator alone (single layers, hypersingular operators)
int regular_launch_symsl_0(const RegularArgs &a, int precision, int kahan, int out_dtype, hipStream_t s);
int regular_launch_symsl_3(const RegularArgs &a, int precision, int kahan, int out_dtype, hipStream_t s);
int regular_launch_symsl_7(const RegularArgs &a, int precision, int kahan, int out_dtype, hipStream_t s);
int regular_launch_symsl_8(const RegularArgs &a, int precision, int kahan, int out_dtype, hipStream_t s);
int regular_launch_symsl_6(const RegularArgs &a, int precision, int kahan, int out_dtype, hipStream_t s);
// symmetric DP0 single layer, upper triangle (regular_upper.hip)
int regular_launch_upper_0(const RegularArgs &a, int precision, int kahan, int out_dtype, hipStream_t s);
int regular_launch_upper_3(const RegularArgs &a, int precision, int kahan, int out_dtype, hipStream_t s);

}  // namespace bem
