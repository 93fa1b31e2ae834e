// Adjoint (vertex and density gradient) kernels of the Helmholtz single layer, DP0, for gfx950.
//
// Galerkin operator. For V_ef = sum_pq w_p J_e w_q J_f G(|y_q - x_p|), G(r) = e^{ikr} / (4 pi r)
// (bempp_cl/core/numba_kernels.py:298-320, 980-1074; singular pairs with the Sauter-Schwab
// points of bempp_cl/core/singular_assembler.py), these kernels add to grad[v] the derivative
// with respect to the vertex coordinates of
//   Re sum_ef conj(x_e) V_ef y_f
// for two dof vectors x, y (x = the cotangent of V c, y = c gives the vertex gradient of any
// loss of V c). With x_q = v0 + (v1 - v0) xi + (v2 - v0) eta linear in the vertices,
//   dV_ef/dy_q = w_p J_e w_q J_f G'(r) (y_q - x_p) / r,   G'(r) = e^{ikr} (ikr - 1) / (4 pi r^2),
//   dV_ef/dJ_f = sum_pq w_p J_e w_q G,   dJ/dv1 = e2 x n, dJ/dv2 = n x e1, dJ/dv0 = -(dJ/dv1 + dJ/dv2).
// The regular part uses the symmetry V_ef = V_fe of the single layer on one space: the
// derivative with respect to the test element's vertices equals the trial-side derivative of
// the transposed pair, so both orientations carry the weight conj(x_e) y_f + conj(x_f) y_e.
// FP64 mode: one lane per trial element f accumulates, over all test elements e, the
// trial-side derivative, and no cross-lane reductions are needed. MIXED mode: far pairs in
// FP32 with the reference-distance split (far_split.h), each unordered pair once, both sides'
// derivatives from one evaluation (adjoint_far_kernel<TM, true>); near pairs (the ones the
// mixed forward kernels compute in FP64) in FP64, all ordered pairs. Singular pairs (one
// wavefront per pair, FP64) differentiate both elements with the weight conj(x_e) y_f, since
// their quadrature is not symmetric.
//
// Potential. For u_p = sum_e m_e c_e sum_q w_q J_e G(|x_p - y_q|) (bempp_cl/core/numba_kernels.py:
// 1788-1849) the potential adjoint adds (P^H g)_e to grad_c and the vertex derivative of
// Re sum_pe conj(g_p) P_pe c_e to grad (FP64, one lane per element).
#include "far_split.h"
#include "regular_impl.h"

namespace bem {
namespace adj {

constexpr int kQ = 6;
constexpr int kThreads = 256;

struct Args {
    SpaceArgs sp;
    RuleArgs rule;
    double k;
    double near_factor;
    const double *x;  // complex128 [ndofs]
    const double *y;  // complex128 [ndofs]
    double *grad;     // [n_vertices * 3]
    int range;        // test elements per workgroup
    int nty;          // number of test ranges
};

__device__ __forceinline__ void atomic_vertex_grad(double *grad, const int32_t *vids, int e, const double (&G)[3][3]) {
#pragma unroll
    for (int j = 0; j < 3; ++j) {
        const int v = vids[4 * e + j];
#pragma unroll
        for (int d = 0; d < 3; ++d) atomic_add(grad + 3 * (int64_t)v + d, G[j][d]);
    }
}

// trial-side vertex gradient of element f from the accumulated point sums:
//   G_j = J_f sum_q phi_j(eta_q) A_q + S dJ_f/dv_j,  phi = (1 - eta0 - eta1, eta0, eta1)
__device__ __forceinline__ void fold_trial(const double *gf, const RuleArgs &rule, const double (&A)[kQ][3], double S,
                                           double (&G)[3][3]) {
    double dJ[3][3];
    jacobian_gradient(gf, dJ);
#pragma unroll
    for (int j = 0; j < 3; ++j)
#pragma unroll
        for (int d = 0; d < 3; ++d) G[j][d] = S * dJ[j][d];
#pragma unroll
    for (int q = 0; q < kQ; ++q) {
        const double phi[3] = {1.0 - rule.x[q] - rule.y[q], rule.x[q], rule.y[q]};
#pragma unroll
        for (int j = 0; j < 3; ++j)
#pragma unroll
            for (int d = 0; d < 3; ++d) G[j][d] += gf[12] * phi[j] * A[q][d];
    }
}

// conj(x_a) y_b for complex128 arrays
__device__ __forceinline__ void cmul_conj(const double *x, int a, const double *y, int b, double &re, double &im) {
    const double xr = x[2 * a], xi = x[2 * a + 1], yr = y[2 * b], yi = y[2 * b + 1];
    re = xr * yr + xi * yi;
    im = xr * yi - xi * yr;
}

// ------------------------------------------------------------------ regular pairs, FP32 (MIXED)
template <int TM>
struct TileF32 {
    int elem[TM];
    int vid[TM][3];
    double c[TM][3];
    double diam[TM];
    float a[TM][kQ][3];
    float aa[TM][kQ];
    float wp[TM][kQ];  // w_p J_e
    float xr[TM], xi[TM], yr[TM], yi[TM];  // m_e x_e, m_e y_e
};

// UPPER: every unordered pair once. The lane of trial element f takes only the test elements
// earlier in the list; besides f's trial-side derivative it forms the test-side derivative of e
// from the same point pairs (d/dx_p = -d/dy_q of G(|y_q - x_p|), the Jacobian term J_f / J_e
// times the pair sum). Since phi_0 = 1 - xi - eta, e's vertex gradient needs only three moments
// of the point derivatives P_p = sum_q d/dy_q: sum_p P_p (= the pair's sum over q), sum_p xi_p P_p
// and sum_p eta_p P_p. Those 9 values and the pair sum are summed over the wave (DPP), over the
// workgroup in LDS, and folded into e's vertex gradient once per test tile.
template <int TM, bool UPPER = false>
__global__ __launch_bounds__(kThreads, UPPER ? 3 : 1) void adjoint_far_kernel(const Args a) {
    __shared__ TileF32<TM> T_;
    __shared__ float tB[UPPER ? TM : 1][3][3];  // UPPER: J_f * (sum_p, sum_p xi_p, sum_p eta_p) P_p per tile element
    __shared__ float tS[UPPER ? TM : 1];        // UPPER: J_f * (pair sum) per tile element
    const int tid = threadIdx.x;
    const int jf = blockIdx.x * kThreads + tid;
    const bool valid = jf < a.sp.n;
    const int e0 = blockIdx.y * a.range;
    int e1 = min(e0 + a.range, a.sp.n);
    if constexpr (UPPER) {
        // test positions must lie below the largest trial position of the block (uniform, before any barrier)
        const int last = min((int)(blockIdx.x * kThreads + kThreads), a.sp.n) - 1;
        if (e0 >= last) return;
        e1 = min(e1, last);
    }
    const float kf = (float)a.k;
    const float krev = (float)(a.k * kInv2Pi);

    int f = 0, fv0 = -1, fv1 = -1, fv2 = -1;
    double cf[3] = {0.0, 0.0, 0.0}, diam_f = 0.0;
    float b[kQ][3], bb[kQ], wq[kQ];  // offsets from the centroid, w_q / (4 pi)
    float xfr = 0.f, xfi = 0.f, yfr = 0.f, yfi = 0.f;
    const double *gf = a.sp.geom;
    if (valid) {
        f = a.sp.elements[jf];
        gf = a.sp.geom + (int64_t)f * BEM_GEOM_STRIDE;
        fv0 = a.sp.vids[4 * f];
        fv1 = a.sp.vids[4 * f + 1];
        fv2 = a.sp.vids[4 * f + 2];
        centroid_rn(gf, cf);
        diam_f = gf[13];
        const double m = a.sp.mult[f];
        const int dof = a.sp.l2g[f];
        xfr = (float)(m * a.x[2 * dof]);
        xfi = (float)(m * a.x[2 * dof + 1]);
        yfr = (float)(m * a.y[2 * dof]);
        yfi = (float)(m * a.y[2 * dof + 1]);
    }
#pragma unroll
    for (int q = 0; q < kQ; ++q) {
        const double s0 = a.rule.x[q] - 1.0 / 3.0, s1 = a.rule.y[q] - 1.0 / 3.0;
        for (int d = 0; d < 3; ++d) b[q][d] = valid ? (float)(gf[3 + d] * s0 + gf[6 + d] * s1) : 0.f;
        bb[q] = b[q][0] * b[q][0] + b[q][1] * b[q][1] + b[q][2] * b[q][2];
        wq[q] = (float)(a.rule.w[q] * kInv4Pi);
    }
    double A[kQ][3], S = 0.0;
#pragma unroll
    for (int q = 0; q < kQ; ++q) A[q][0] = A[q][1] = A[q][2] = 0.0;

    const float jf32 = (float)gf[12];  // UPPER: J_f (test-side weights)
    const int lane = tid & 63;
    for (int t0 = e0; t0 < e1; t0 += TM) {
        const int ntile = min(TM, e1 - t0);
        __syncthreads();
        if (tid < ntile) {
            const int e = a.sp.elements[t0 + tid];
            const double *g = a.sp.geom + (int64_t)e * BEM_GEOM_STRIDE;
            T_.elem[tid] = e;
            for (int i = 0; i < 3; ++i) T_.vid[tid][i] = a.sp.vids[4 * e + i];
            centroid_rn(g, T_.c[tid]);
            T_.diam[tid] = g[13];
            for (int p = 0; p < kQ; ++p) {
                const double xi = a.rule.x[p] - 1.0 / 3.0, eta = a.rule.y[p] - 1.0 / 3.0;
                for (int d = 0; d < 3; ++d) T_.a[tid][p][d] = (float)(g[3 + d] * xi + g[6 + d] * eta);
                T_.aa[tid][p] = T_.a[tid][p][0] * T_.a[tid][p][0] + T_.a[tid][p][1] * T_.a[tid][p][1] +
                                T_.a[tid][p][2] * T_.a[tid][p][2];
                T_.wp[tid][p] = (float)(a.rule.w[p] * g[12]);
            }
            const double m = a.sp.mult[e];
            const int dof = a.sp.l2g[e];
            T_.xr[tid] = (float)(m * a.x[2 * dof]);
            T_.xi[tid] = (float)(m * a.x[2 * dof + 1]);
            T_.yr[tid] = (float)(m * a.y[2 * dof]);
            T_.yi[tid] = (float)(m * a.y[2 * dof + 1]);
        }
        if constexpr (UPPER) {
            for (int i = tid; i < TM * 10; i += kThreads) {
                if (i < TM * 9) (&tB[0][0][0])[i] = 0.f;
                else tS[i - TM * 9] = 0.f;
            }
        }
        __syncthreads();
        for (int t = 0; t < ntile; ++t) {
            bool skip = !valid;
            if (valid) {
                skip = T_.vid[t][0] == fv0 || T_.vid[t][0] == fv1 || T_.vid[t][0] == fv2 || T_.vid[t][1] == fv0 ||
                       T_.vid[t][1] == fv1 || T_.vid[t][1] == fv2 || T_.vid[t][2] == fv0 || T_.vid[t][2] == fv1 ||
                       T_.vid[t][2] == fv2;
                double dist2;
                if (!skip) skip = pair_is_near(T_.c[t], cf, T_.diam[t], diam_f, a.near_factor, &dist2);
            }
            if constexpr (UPPER) {
                skip = skip || t0 + t >= jf;
                // the test-side sums are reduced over the wave, so a lane that skips stays converged
                // and computes a harmless far pair with zero weight, unless the whole wave skips
                if (__all(skip)) continue;
            } else {
                if (skip) continue;
            }
            // W = conj(x_e) y_f + conj(x_f) y_e
            float Wr = T_.xr[t] * yfr + T_.xi[t] * yfi + xfr * T_.yr[t] + xfi * T_.yi[t];
            float Wi = T_.xr[t] * yfi - T_.xi[t] * yfr + xfr * T_.yi[t] - xfi * T_.yr[t];
            double cpair[3] = {cf[0], cf[1], cf[2]};
            if (UPPER && skip) {
                Wr = Wi = 0.f;
                cpair[0] = T_.c[t][0] + 1.0e3 * (1.0 + T_.diam[t]);  // far from every point of e
            }
            const FarPair P = far_pair_setup(T_.c[t], cpair);
            float cs, sn;
            phase_rev_f32(a.k, (double)P.r0 * kInv2Pi, cs, sn);
            const float Er = Wr * cs - Wi * sn, Ei = Wr * sn + Wi * cs;  // W e^{i k r0}
            float Aq[kQ][3], sj = 0.f;
            float M1[3] = {0.f, 0.f, 0.f}, M2[3] = {0.f, 0.f, 0.f};  // UPPER: sum_p xi_p P_p, sum_p eta_p P_p
#pragma unroll
            for (int q = 0; q < kQ; ++q) Aq[q][0] = Aq[q][1] = Aq[q][2] = 0.f;
#pragma unroll 1
            for (int p = 0; p < kQ; ++p) {
                const float w0 = P.fd0 - T_.a[t][p][0], w1 = P.fd1 - T_.a[t][p][1], w2 = P.fd2 - T_.a[t][p][2];
                const float Ap = P.resid - 2.f * (P.fd0 * T_.a[t][p][0] + P.fd1 * T_.a[t][p][1] + P.fd2 * T_.a[t][p][2]) +
                                 T_.aa[t][p];
                const float wp = T_.wp[t][p];
                float P0 = 0.f, P1 = 0.f, P2 = 0.f;  // UPPER: P_p = sum_q tq (y_q - x_p)
#pragma unroll
                for (int q = 0; q < kQ; ++q) {
                    const float tt = Ap + bb[q] + 2.f * (w0 * b[q][0] + w1 * b[q][1] + w2 * b[q][2]);
                    const float r2 = P.r0sq + tt;
                    const float rinv = __builtin_amdgcn_rsqf(r2);
                    const float r = r2 * rinv;
                    const float ph = krev * (tt * __builtin_amdgcn_rcpf(r + P.r0));
                    const float c = __builtin_amdgcn_cosf(ph), s = __builtin_amdgcn_sinf(ph);
                    const float Zr = Er * c - Ei * s, Zi = Er * s + Ei * c;
                    const float amp = rinv * wq[q] * wp;
                    sj += amp * Zr;
                    const float tq = amp * (-Zr - kf * r * Zi) * rinv * rinv;
                    const float d0 = tq * (w0 + b[q][0]), d1 = tq * (w1 + b[q][1]), d2 = tq * (w2 + b[q][2]);
                    Aq[q][0] += d0;
                    Aq[q][1] += d1;
                    Aq[q][2] += d2;
                    if constexpr (UPPER) {
                        P0 += d0;
                        P1 += d1;
                        P2 += d2;
                    }
                }
                if constexpr (UPPER) {
                    const float xi = (float)a.rule.x[p], eta = (float)a.rule.y[p];
                    M1[0] += xi * P0;
                    M1[1] += xi * P1;
                    M1[2] += xi * P2;
                    M2[0] += eta * P0;
                    M2[1] += eta * P1;
                    M2[2] += eta * P2;
                }
            }
#pragma unroll
            for (int q = 0; q < kQ; ++q)
                for (int d = 0; d < 3; ++d) A[q][d] += (double)Aq[q][d];
            S += (double)sj;
            if constexpr (UPPER) {
                // test side: d/dx_p = -P_p J_f (tq lacks J_f); J_e S_e = J_f * (pair sum)
                float M0[3];
#pragma unroll
                for (int d = 0; d < 3; ++d) {
                    M0[d] = Aq[0][d];
#pragma unroll
                    for (int q = 1; q < kQ; ++q) M0[d] += Aq[q][d];
                }
                float R[10];
                wave_total2_dpp(M0[0] * jf32, M0[1] * jf32, R[0], R[1]);
                wave_total2_dpp(M0[2] * jf32, M1[0] * jf32, R[2], R[3]);
                wave_total2_dpp(M1[1] * jf32, M1[2] * jf32, R[4], R[5]);
                wave_total2_dpp(M2[0] * jf32, M2[1] * jf32, R[6], R[7]);
                wave_total2_dpp(M2[2] * jf32, sj * jf32, R[8], R[9]);
                if (lane == 0) {
#pragma unroll
                    for (int m = 0; m < 9; ++m) atomicAdd(&tB[t][m / 3][m % 3], R[m]);
                    atomicAdd(&tS[t], R[9]);
                }
            }
        }
        if constexpr (UPPER) {
            // fold the test side of this tile: G_e[j] = S_e dJ_e/dv_j - sum_p phi_j(xi_p) J_f P_p
            __syncthreads();
            if (tid < ntile) {
                const int e = T_.elem[tid];
                const double *g = a.sp.geom + (int64_t)e * BEM_GEOM_STRIDE;
                double dJ[3][3], G[3][3];
                jacobian_gradient(g, dJ);
                const double Se = (double)tS[tid] / g[12];
#pragma unroll
                for (int d = 0; d < 3; ++d) {
                    const double m0 = tB[tid][0][d], m1 = tB[tid][1][d], m2 = tB[tid][2][d];
                    G[0][d] = Se * dJ[0][d] - (m0 - m1 - m2);
                    G[1][d] = Se * dJ[1][d] - m1;
                    G[2][d] = Se * dJ[2][d] - m2;
                }
                atomic_vertex_grad(a.grad, a.sp.vids, e, G);
            }
        }
    }
    if (valid) {
        double G[3][3];
        fold_trial(gf, a.rule, A, S, G);
        atomic_vertex_grad(a.grad, a.sp.vids, f, G);
    }
}

// ------------------------------------------------------------------ regular pairs, FP64
template <int TM>
struct TileF64 {
    int elem[TM];
    int vid[TM][3];
    double xp[TM][kQ][3];
    double wp[TM][kQ];
    double xr[TM], xi[TM], yr[TM], yi[TM];
};

// point-pair sums of one pair in FP64: A_q += t (y_q - x_p), S += w w J_e Re(W G)
__device__ __forceinline__ void pair_terms_fp64(double k, double Wr, double Wi, const double (&xp)[3], double wpJ,
                                                const double (&yq)[3], double wq, double (&Aq)[3], double &S) {
    const double dx = yq[0] - xp[0], dy = yq[1] - xp[1], dz = yq[2] - xp[2];
    const double r = sqrt(dx * dx + dy * dy + dz * dz);
    double s, c;
    sincos_fast(k * r, &s, &c);
    const double Zr = Wr * c - Wi * s, Zi = Wr * s + Wi * c;
    const double amp = wpJ * wq * kInv4Pi / r;
    S += amp * Zr;
    const double t = amp * (-Zr - k * r * Zi) / (r * r);
    Aq[0] += t * dx;
    Aq[1] += t * dy;
    Aq[2] += t * dz;
}

template <int TM>
__global__ __launch_bounds__(kThreads) void adjoint_fp64_kernel(const Args a) {
    __shared__ TileF64<TM> T_;
    const int tid = threadIdx.x;
    const int jf = blockIdx.x * kThreads + tid;
    const bool valid = jf < a.sp.n;
    const int e0 = blockIdx.y * a.range, e1 = min(e0 + a.range, a.sp.n);

    int f = 0, fv0 = -1, fv1 = -1, fv2 = -1;
    double yq[kQ][3], xfr = 0, xfi = 0, yfr = 0, yfi = 0;
    const double *gf = a.sp.geom;
    if (valid) {
        f = a.sp.elements[jf];
        gf = a.sp.geom + (int64_t)f * BEM_GEOM_STRIDE;
        fv0 = a.sp.vids[4 * f];
        fv1 = a.sp.vids[4 * f + 1];
        fv2 = a.sp.vids[4 * f + 2];
        const double m = a.sp.mult[f];
        const int dof = a.sp.l2g[f];
        xfr = m * a.x[2 * dof];
        xfi = m * a.x[2 * dof + 1];
        yfr = m * a.y[2 * dof];
        yfi = m * a.y[2 * dof + 1];
    }
#pragma unroll
    for (int q = 0; q < kQ; ++q)
        for (int d = 0; d < 3; ++d) yq[q][d] = valid ? gf[d] + gf[3 + d] * a.rule.x[q] + gf[6 + d] * a.rule.y[q] : 0.0;
    double A[kQ][3], S = 0.0;
#pragma unroll
    for (int q = 0; q < kQ; ++q) A[q][0] = A[q][1] = A[q][2] = 0.0;

    for (int t0 = e0; t0 < e1; t0 += TM) {
        const int ntile = min(TM, e1 - t0);
        __syncthreads();
        if (tid < ntile) {
            const int e = a.sp.elements[t0 + tid];
            const double *g = a.sp.geom + (int64_t)e * BEM_GEOM_STRIDE;
            T_.elem[tid] = e;
            for (int i = 0; i < 3; ++i) T_.vid[tid][i] = a.sp.vids[4 * e + i];
            for (int p = 0; p < kQ; ++p) {
                for (int d = 0; d < 3; ++d) T_.xp[tid][p][d] = g[d] + g[3 + d] * a.rule.x[p] + g[6 + d] * a.rule.y[p];
                T_.wp[tid][p] = a.rule.w[p] * g[12];
            }
            const double m = a.sp.mult[e];
            const int dof = a.sp.l2g[e];
            T_.xr[tid] = m * a.x[2 * dof];
            T_.xi[tid] = m * a.x[2 * dof + 1];
            T_.yr[tid] = m * a.y[2 * dof];
            T_.yi[tid] = m * a.y[2 * dof + 1];
        }
        __syncthreads();
        for (int t = 0; t < ntile; ++t) {
            if (!valid) continue;
            if (T_.vid[t][0] == fv0 || T_.vid[t][0] == fv1 || T_.vid[t][0] == fv2 || T_.vid[t][1] == fv0 ||
                T_.vid[t][1] == fv1 || T_.vid[t][1] == fv2 || T_.vid[t][2] == fv0 || T_.vid[t][2] == fv1 ||
                T_.vid[t][2] == fv2)
                continue;
            const double Wr = T_.xr[t] * yfr + T_.xi[t] * yfi + xfr * T_.yr[t] + xfi * T_.yi[t];
            const double Wi = T_.xr[t] * yfi - T_.xi[t] * yfr + xfr * T_.yi[t] - xfi * T_.yr[t];
#pragma unroll 1
            for (int p = 0; p < kQ; ++p) {
                const double xp[3] = {T_.xp[t][p][0], T_.xp[t][p][1], T_.xp[t][p][2]};
                const double wpJ = T_.wp[t][p];
#pragma unroll
                for (int q = 0; q < kQ; ++q) pair_terms_fp64(a.k, Wr, Wi, xp, wpJ, yq[q], a.rule.w[q], A[q], S);
            }
        }
    }
    if (valid) {
        double G[3][3];
        fold_trial(gf, a.rule, A, S, G);
        atomic_vertex_grad(a.grad, a.sp.vids, f, G);
    }
}

// ------------------------------------------------------------------ near pairs (MIXED), FP64
__global__ __launch_bounds__(256) void adjoint_near_kernel(const Args a, int n_pairs, const int32_t *pairs) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= n_pairs) return;
    const int e = pairs[2 * (int64_t)p], f = pairs[2 * (int64_t)p + 1];
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j)
            if (a.sp.vids[4 * e + i] == a.sp.vids[4 * f + j]) return;
    const double *gt = a.sp.geom + (int64_t)e * BEM_GEOM_STRIDE;
    const double *gf = a.sp.geom + (int64_t)f * BEM_GEOM_STRIDE;
    double ct[3], cf[3], dist2;
    centroid_rn(gt, ct);
    centroid_rn(gf, cf);
    if (!pair_is_near(ct, cf, gt[13], gf[13], a.near_factor, &dist2)) return;
    const double me = a.sp.mult[e], mf = a.sp.mult[f];
    const int de = a.sp.l2g[e], df = a.sp.l2g[f];
    double r1, i1, r2, i2;
    cmul_conj(a.x, de, a.y, df, r1, i1);
    cmul_conj(a.x, df, a.y, de, r2, i2);
    const double Wr = me * mf * (r1 + r2), Wi = me * mf * (i1 + i2);
    double A[kQ][3], S = 0.0;
    double yq[kQ][3];
#pragma unroll
    for (int q = 0; q < kQ; ++q) {
        A[q][0] = A[q][1] = A[q][2] = 0.0;
#pragma unroll
        for (int d = 0; d < 3; ++d) yq[q][d] = gf[d] + gf[3 + d] * a.rule.x[q] + gf[6 + d] * a.rule.y[q];
    }
#pragma unroll 1
    for (int pp = 0; pp < kQ; ++pp) {
        const double xp[3] = {gt[0] + gt[3] * a.rule.x[pp] + gt[6] * a.rule.y[pp],
                              gt[1] + gt[4] * a.rule.x[pp] + gt[7] * a.rule.y[pp],
                              gt[2] + gt[5] * a.rule.x[pp] + gt[8] * a.rule.y[pp]};
        const double wpJ = a.rule.w[pp] * gt[12];
#pragma unroll
        for (int q = 0; q < kQ; ++q) pair_terms_fp64(a.k, Wr, Wi, xp, wpJ, yq[q], a.rule.w[q], A[q], S);
    }
    double G[3][3];
    fold_trial(gf, a.rule, A, S, G);
    atomic_vertex_grad(a.grad, a.sp.vids, f, G);
}

// ------------------------------------------------------------------ singular pairs, FP64, both elements
// kSingLanes lanes per pair (8 pairs per wavefront): the 512..1536 Duffy points are strided over
// them and the 19 sums reduced in 3 shuffle steps (a wavefront per pair left most lanes of the
// vertex pairs with 8 points and spent 6 steps per sum)
constexpr int kSingLanes = 8;

template <typename T>
__device__ __forceinline__ T sing_group_sum(T v) {
#pragma unroll
    for (int off = kSingLanes / 2; off >= 1; off >>= 1) v += __shfl_xor(v, off, kSingLanes);
    return v;
}

__global__ __launch_bounds__(256) void adjoint_singular_kernel(const Args a, int n_pairs, const int32_t *pd_all,
                                                               const double *tpts, const double *spts,
                                                               const double *wts) {
    const int lane = threadIdx.x & (kSingLanes - 1);
    const int pair = (int)(((int64_t)blockIdx.x * blockDim.x + threadIdx.x) / kSingLanes);
    // the group's lanes share the pair, so whole groups leave together and the group shuffles
    // below only read lanes of groups that stay
    if (pair >= n_pairs) return;
    const int32_t *pd = pd_all + 6 * (int64_t)pair;
    const int e = pd[0], f = pd[1], np = pd[5];
    const double *tp = tpts + 2 * (int64_t)pd[2], *sp = spts + 2 * (int64_t)pd[3], *w = wts + pd[4];
    const double *gt = a.sp.geom + (int64_t)e * BEM_GEOM_STRIDE;
    const double *gf = a.sp.geom + (int64_t)f * BEM_GEOM_STRIDE;
    double Wr, Wi;
    cmul_conj(a.x, a.sp.l2g[e], a.y, a.sp.l2g[f], Wr, Wi);
    const double m = a.sp.mult[e] * a.sp.mult[f];
    Wr *= m;
    Wi *= m;
    // Ay[j]: trial vertex j, Ax[j]: test vertex j (both without the J_e J_f factor); S: sum w Re(W G)
    double Ay[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}}, Ax[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}}, S = 0.0;
    auto points = [&](auto phase_tag) {
        constexpr int PH = decltype(phase_tag)::value;
        for (int kk = lane; kk < np; kk += kSingLanes) {
            const double xi0 = tp[2 * kk], xi1 = tp[2 * kk + 1], et0 = sp[2 * kk], et1 = sp[2 * kk + 1];
            double dv[3];
#pragma unroll
            for (int d = 0; d < 3; ++d)
                dv[d] = (gf[d] + gf[3 + d] * et0 + gf[6 + d] * et1) - (gt[d] + gt[3 + d] * xi0 + gt[6 + d] * xi1);
            const double r2 = dv[0] * dv[0] + dv[1] * dv[1] + dv[2] * dv[2];
            const double rinv = rsqrt_t(r2), r = r2 * rinv;
            double s, c;
            sincos_t<PH>(a.k * r, &s, &c);
            const double Zr = Wr * c - Wi * s, Zi = Wr * s + Wi * c;
            const double amp = w[kk] * kInv4Pi * rinv;
            S += amp * Zr;
            const double t = amp * (-Zr - a.k * r * Zi) * (rinv * rinv);
            const double py[3] = {1.0 - et0 - et1, et0, et1}, px[3] = {1.0 - xi0 - xi1, xi0, xi1};
#pragma unroll
            for (int j = 0; j < 3; ++j)
#pragma unroll
                for (int d = 0; d < 3; ++d) {
                    Ay[j][d] += py[j] * t * dv[d];
                    Ax[j][d] -= px[j] * t * dv[d];
                }
        }
    };
    // every distance within the pair is at most diam_e + diam_f (triangles sharing a vertex): the
    // series sin/cos without range reduction when |k| r <= 1, else the branch-free reduction
    const double kd = fabs(a.k) * (gt[13] + gf[13]);
    if (kd <= 1.0) points(std::integral_constant<int, kPhSmall | kPhReal>{});
    else if (kd < 0x1p25) points(std::integral_constant<int, kPhReduced | kPhReal>{});
    else points(std::integral_constant<int, kPhReal>{});
#pragma unroll
    for (int j = 0; j < 3; ++j)
#pragma unroll
        for (int d = 0; d < 3; ++d) {
            Ay[j][d] = sing_group_sum(Ay[j][d]);
            Ax[j][d] = sing_group_sum(Ax[j][d]);
        }
    S = sing_group_sum(S);
    for (int o = lane; o < 18; o += kSingLanes) {
        const bool trial_side = o < 9;
        const int j = (o % 9) / 3, d = o % 3;
        const double *g = trial_side ? gf : gt;
        const double other_J = trial_side ? gt[12] : gf[12];
        double dJ[3][3];
        jacobian_gradient(g, dJ);
        // select entry (j, d) with compile-time indices (keeps the arrays in registers)
        double acc = 0.0, dj = 0.0;
#pragma unroll
        for (int jj = 0; jj < 3; ++jj)
#pragma unroll
            for (int dd = 0; dd < 3; ++dd)
                if (jj == j && dd == d) {
                    acc = trial_side ? Ay[jj][dd] : Ax[jj][dd];
                    dj = dJ[jj][dd];
                }
        const double val = gt[12] * gf[12] * acc + S * other_J * dj;
        const int v = a.sp.vids[4 * (trial_side ? f : e) + j];
        atomic_add(a.grad + 3 * (int64_t)v + d, val);
    }
}

// ------------------------------------------------------------------ potential adjoint, FP64
struct PotArgs {
    SpaceArgs sp;
    RuleArgs rule;
    double k;
    const double *c;      // complex128 [ndofs]
    const double *pts;    // [3 * n_points]
    const double *g;      // complex128 [n_points] cotangent of u
    int n_points;
    int chunk;            // points per workgroup
    double *grad_c;       // complex128 [ndofs] or NULL
    double *grad;         // [n_vertices * 3] or NULL
};

constexpr int kPotTile = 256;

__global__ __launch_bounds__(kThreads, 2) void potential_adjoint_kernel(const PotArgs a) {
    __shared__ double P_[kPotTile][3];
    __shared__ double G_[kPotTile][2];
    // running max of |x| over the staged points (a bound for the current tile)
    __shared__ unsigned int pmax_bits;
    const int tid = threadIdx.x;
    if (tid == 0) pmax_bits = 0u;  // published by the first tile's barrier
    const int je = blockIdx.x * kThreads + tid;
    const bool valid = je < a.sp.n;
    const int p0 = blockIdx.y * a.chunk, p1 = min(p0 + a.chunk, a.n_points);
    int e = 0;
    const double *ge = a.sp.geom;
    double cr = 0.0, ci = 0.0, m = 0.0;
    if (valid) {
        e = a.sp.elements[je];
        ge = a.sp.geom + (int64_t)e * BEM_GEOM_STRIDE;
        m = a.sp.mult[e];
        const int dof = a.sp.l2g[e];
        cr = m * a.c[2 * dof];
        ci = m * a.c[2 * dof + 1];
    }
    double yq[kQ][3], ynorm = 0.0;  // ynorm: max |y| over the element's quadrature points
#pragma unroll
    for (int q = 0; q < kQ; ++q) {
        for (int d = 0; d < 3; ++d) yq[q][d] = valid ? ge[d] + ge[3 + d] * a.rule.x[q] + ge[6 + d] * a.rule.y[q] : 0.0;
        ynorm = fmax(ynorm, sqrt(yq[q][0] * yq[q][0] + yq[q][1] * yq[q][1] + yq[q][2] * yq[q][2]));
    }
    double A[kQ][3], S = 0.0, hr = 0.0, hi = 0.0;  // hr + i hi = sum_p conj(P_pe / m_e) g_p
#pragma unroll
    for (int q = 0; q < kQ; ++q) A[q][0] = A[q][1] = A[q][2] = 0.0;
    for (int t0 = p0; t0 < p1; t0 += kPotTile) {
        const int nt = min(kPotTile, p1 - t0);
        __syncthreads();
        if (tid < nt) {
            double x2 = 0.0;
            for (int d = 0; d < 3; ++d) {
                const double x = a.pts[3 * (int64_t)(t0 + tid) + d];
                P_[tid][d] = x;
                x2 += x * x;
            }
            G_[tid][0] = a.g[2 * (int64_t)(t0 + tid)];
            G_[tid][1] = a.g[2 * (int64_t)(t0 + tid) + 1];
            // non-negative floats order like their bit patterns
            atomicMax(&pmax_bits, __float_as_uint((float)(sqrt(x2) * (1.0 + 1e-6))));
        }
        __syncthreads();
        if (!valid) continue;
        auto points = [&](auto phase_tag) {
            constexpr int PH = decltype(phase_tag)::value;
            for (int t = 0; t < nt; ++t) {
                const double gr = G_[t][0], gi = G_[t][1];
                // W = conj(g_p) m_e c_e
                const double Wr = gr * cr + gi * ci, Wi = gr * ci - gi * cr;
#pragma unroll
                for (int q = 0; q < kQ; ++q) {
                    const double dx = yq[q][0] - P_[t][0], dy = yq[q][1] - P_[t][1], dz = yq[q][2] - P_[t][2];
                    const double r2 = dx * dx + dy * dy + dz * dz;
                    const double rinv = rsqrt_t(r2), r = r2 * rinv;
                    double s, c;
                    sincos_t<PH>(a.k * r, &s, &c);
                    const double amp = a.rule.w[q] * kInv4Pi * rinv;  // without J_e
                    // grad_c: conj(G) g
                    hr += amp * (c * gr + s * gi);
                    hi += amp * (c * gi - s * gr);
                    const double Zr = Wr * c - Wi * s, Zi = Wr * s + Wi * c;
                    S += amp * Zr;
                    const double tq = amp * (-Zr - a.k * r * Zi) * (rinv * rinv);
                    A[q][0] += tq * dx;
                    A[q][1] += tq * dy;
                    A[q][2] += tq * dz;
                }
            }
        };
        // |k| r <= |k| (|x| + |y|) < 2^25 on every pair of the wave: the branch-free phase
        // reduction, which lets the compiler interleave the quadrature points
        const double pmax = (double)__uint_as_float(pmax_bits);
        if (__all(fabs(a.k) * (pmax + ynorm) < 0x1p25)) points(std::integral_constant<int, kPhReduced | kPhReal>{});
        else points(std::integral_constant<int, kPhReal>{});
    }
    if (!valid) return;
    if (a.grad_c) {
        const int dof = a.sp.l2g[e];
        atomic_add(a.grad_c + 2 * dof, m * ge[12] * hr);
        atomic_add(a.grad_c + 2 * dof + 1, m * ge[12] * hi);
    }
    if (a.grad) {
        double G[3][3];
        fold_trial(ge, a.rule, A, S, G);
        atomic_vertex_grad(a.grad, a.sp.vids, e, G);
    }
}

}  // namespace adj
}  // namespace bem

using namespace bem;

static int check_dp0_space(const bem_space_t *s, const bem_quad_rule_t *rule) {
    int rc = check_space(s, "space");
    if (rc) return rc;
    if (s->shapeset != BEM_SHAPE_P0) {
        set_error("the adjoint kernels support DP0 spaces");
        return BEM_ERR_UNSUPPORTED;
    }
    if (!rule || rule->n_points != adj::kQ) {
        set_error("the adjoint kernels use the 6-point (order 4) regular rule");
        return BEM_ERR_UNSUPPORTED;
    }
    return BEM_OK;
}

extern "C" int bem_adjoint_single_layer(const bem_space_t *space, const bem_quad_rule_t *rule, double k, int precision,
                                        double mixed_near_factor, const void *x, const void *y, int n_near,
                                        const int32_t *near_pairs, int n_singular, const int32_t *singular_pairs,
                                        const double *test_points, const double *trial_points, const double *weights,
                                        double *grad_vertices, void *stream) {
    clear_error();
    if (!space || !x || !y || !grad_vertices) {
        set_error("null argument");
        return BEM_ERR_INVALID_ARGUMENT;
    }
    int rc = check_dp0_space(space, rule);
    if (rc) return rc;
    if (precision != BEM_PRECISION_FP64 && precision != BEM_PRECISION_MIXED) {
        set_error("unknown precision mode %d", precision);
        return BEM_ERR_INVALID_ARGUMENT;
    }
    if (n_near < 0 || n_singular < 0 || (n_near > 0 && !near_pairs) ||
        (n_singular > 0 && (!singular_pairs || !test_points || !trial_points || !weights))) {
        set_error("missing near-field or singular pair arrays");
        return BEM_ERR_INVALID_ARGUMENT;
    }
    if (!(fabs(k) < 1e300)) {
        set_error("wavenumber must be finite");
        return BEM_ERR_INVALID_ARGUMENT;
    }
    const bool mixed = precision == BEM_PRECISION_MIXED;
    adj::Args a;
    a.sp = to_args(space);
    a.rule = to_args(rule);
    a.k = k;
    a.near_factor = mixed_near_factor;
    a.x = (const double *)x;
    a.y = (const double *)y;
    a.grad = grad_vertices;
    hipStream_t s = (hipStream_t)stream;
    if (a.sp.n > 0) {
        // test ranges sized so that the grid holds >= ~8 workgroups per CU
        const int ntx = (a.sp.n + adj::kThreads - 1) / adj::kThreads;
        int range = 2048;
        while (range > 128 && (long)ntx * ((a.sp.n + range - 1) / range) < 2048) range /= 2;
        a.range = range;
        a.nty = (a.sp.n + range - 1) / range;
        const dim3 grid(ntx, a.nty);
        // mixed: each unordered pair once (both sides' derivatives from one evaluation)
        if (mixed) hipLaunchKernelGGL((adj::adjoint_far_kernel<32, true>), grid, dim3(adj::kThreads), 0, s, a);
        else hipLaunchKernelGGL((adj::adjoint_fp64_kernel<32>), grid, dim3(adj::kThreads), 0, s, a);
        if ((rc = check_launch("adjoint_regular_kernel"))) return rc;
    }
    if (mixed && n_near > 0) {
        hipLaunchKernelGGL(adj::adjoint_near_kernel, dim3((n_near + 255) / 256), dim3(256), 0, s, a, n_near, near_pairs);
        if ((rc = check_launch("adjoint_near_kernel"))) return rc;
    }
    if (n_singular > 0) {
        hipLaunchKernelGGL(adj::adjoint_singular_kernel,
                           dim3((unsigned)(((int64_t)n_singular * adj::kSingLanes + 255) / 256)), dim3(256), 0, s, a,
                           n_singular, singular_pairs, test_points, trial_points, weights);
        if ((rc = check_launch("adjoint_singular_kernel"))) return rc;
    }
    return BEM_OK;
}

extern "C" int bem_adjoint_potential_single_layer(const bem_space_t *space, const bem_quad_rule_t *rule, double k,
                                                  const void *coefficients, int n_points, const double *points,
                                                  const void *cotangent, void *grad_coefficients,
                                                  double *grad_vertices, void *stream) {
    clear_error();
    if (!space || !coefficients || (n_points > 0 && (!points || !cotangent)) || (!grad_coefficients && !grad_vertices)) {
        set_error("null argument");
        return BEM_ERR_INVALID_ARGUMENT;
    }
    int rc = check_dp0_space(space, rule);
    if (rc) return rc;
    if (n_points < 0 || !(fabs(k) < 1e300)) {
        set_error("invalid number of points or wavenumber");
        return BEM_ERR_INVALID_ARGUMENT;
    }
    if (n_points == 0 || space->n_elements == 0) return BEM_OK;
    adj::PotArgs a;
    a.sp = to_args(space);
    a.rule = to_args(rule);
    a.k = k;
    a.c = (const double *)coefficients;
    a.pts = points;
    a.g = (const double *)cotangent;
    a.n_points = n_points;
    a.grad_c = (double *)grad_coefficients;
    a.grad = grad_vertices;
    const int nbx = (a.sp.n + adj::kThreads - 1) / adj::kThreads;
    int chunk = adj::kPotTile * 8;
    while (chunk > adj::kPotTile && (long)nbx * ((n_points + chunk - 1) / chunk) < 2048) chunk /= 2;
    a.chunk = chunk;
    const dim3 grid(nbx, (n_points + chunk - 1) / chunk);
    hipLaunchKernelGGL(adj::potential_adjoint_kernel, grid, dim3(adj::kThreads), 0, (hipStream_t)stream, a);
    return check_launch("potential_adjoint_kernel");
}
