// The 3 x 3 block of a regular pair of the Maxwell magnetic field operator (MFIE) from moments of
// the kernel values, shared by the MFIE kernels (mfie.hip) and the tiled regular kernel
// (regular_impl.h). The RWG function i of an element is A_i(x) = x - P_i (times l_i / J), P_i the
// vertex opposite edge i: with the reference coordinates o = ((0, 1), (1, 0), (0, 0)),
// P_i = v0 + e1 o_i0 + e2 o_i1 (bempp_cl/core/numba_kernels.py:2151-2277).
#pragma once

#include <hip/hip_runtime.h>

#include "bem_device.h"
#include "bem_internal.h"

namespace bem {
namespace mfie {

constexpr double kOx[3] = {0.0, 1.0, 0.0}, kOy[3] = {1.0, 0.0, 0.0};

// c - P_i for the three RWG functions (c the centroid, P_i the vertex opposite edge i: A_i(x) = x - P_i)
template <typename T, typename E>
__device__ __forceinline__ void rwg_offsets(const E *e1, const E *e2, T (&pi)[3][3]) {
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int d = 0; d < 3; ++d)
            pi[i][d] = (T)(e1[d] * (E)(1.0 / 3.0 - kOx[i]) + e2[d] * (E)(1.0 / 3.0 - kOy[i]));
}

template <typename T>
__device__ __forceinline__ void cross3(const T *u, const T *v, T (&w)[3]) {
    w[0] = u[1] * v[2] - u[2] * v[1];
    w[1] = u[2] * v[0] - u[0] * v[2];
    w[2] = u[0] * v[1] - u[1] * v[0];
}

template <typename T>
__device__ __forceinline__ T dot3(const T *u, const T *v) {
    return u[0] * v[0] + u[1] * v[1] + u[2] * v[2];
}

// The 3 x 3 block of a regular pair from moments of the kernel values. With x_p = c_e + a_p,
// y_q = c_f + b_q, fd = c_f - c_e, pi_i = c_e - P_i, kappa_j = c_f - Q_j and G_pq the weighted
// values w_p w_q G (ikr - 1) / r^2,
//   V_ij = sum_pq G_pq (x_p - y_q) . ((a_p + pi_i) x (b_q + kappa_j))
//        = S + pi_i . U - kappa_j . V + pi_i . (kappa_j x W)
// since a . (a x b) = b . (a x b) = 0, where
//   S = -sum G fd . (a x b),  U = sum G b x (a - fd),  V = sum G (b + fd) x a,  W = sum G (a - b - fd).
// All four are linear in b_q: per test point only R_p = sum_q G_pq and Y_p = sum_q G_pq b_q are
// summed over the trial points (8 multiply-adds per point pair instead of the 54 of the direct
// 3 x 3 contraction), and add_point folds them with a_p. Index [2] is real / imaginary part.
template <typename T>
struct PairMoments {
    T R[2], Ra[2][3], Rh[2][3], YA[2][3], Ys[2][3], Sh[2];

    __device__ __forceinline__ void clear() {
#pragma unroll
        for (int c = 0; c < 2; ++c) {
            R[c] = Sh[c] = T(0);
#pragma unroll
            for (int d = 0; d < 3; ++d) Ra[c][d] = Rh[c][d] = YA[c][d] = Ys[c][d] = T(0);
        }
    }

    __device__ __forceinline__ void add_point(const T (&Rp)[2], const T (&Yp)[2][3], const T (&ap)[3],
                                              const T (&fd)[3]) {
        T h[3];
        cross3(fd, ap, h);
#pragma unroll
        for (int c = 0; c < 2; ++c) {
            T ya[3];
            cross3(Yp[c], ap, ya);
            R[c] += Rp[c];
            Sh[c] += dot3(Yp[c], h);
#pragma unroll
            for (int d = 0; d < 3; ++d) {
                Ra[c][d] += Rp[c] * ap[d];
                Rh[c][d] += Rp[c] * h[d];
                YA[c][d] += ya[d];
                Ys[c][d] += Yp[c][d];
            }
        }
    }

    __device__ __forceinline__ void block(const T (&pi)[3][3], const T (&kap)[3][3], const T (&fd)[3], T (&ar)[3][3],
                                          T (&ai)[3][3]) const {
        T PU[2][3], KV[2][3], KW[2][3][3];
#pragma unroll
        for (int c = 0; c < 2; ++c) {
            T yf[3], U[3], V[3], W[3];
            cross3(Ys[c], fd, yf);
#pragma unroll
            for (int d = 0; d < 3; ++d) {
                U[d] = YA[c][d] - yf[d];
                V[d] = YA[c][d] + Rh[c][d];
                W[d] = Ra[c][d] - Ys[c][d] - R[c] * fd[d];
            }
#pragma unroll
            for (int i = 0; i < 3; ++i) {
                PU[c][i] = dot3(pi[i], U);
                KV[c][i] = dot3(kap[i], V);
                cross3(kap[i], W, KW[c][i]);
            }
        }
#pragma unroll
        for (int i = 0; i < 3; ++i)
#pragma unroll
            for (int j = 0; j < 3; ++j) {
                ar[i][j] = PU[0][i] - KV[0][j] - Sh[0] + dot3(pi[i], KW[0][j]);
                ai[i][j] = PU[1][i] - KV[1][j] - Sh[1] + dot3(pi[i], KW[1][j]);
            }
    }
};

// edge lengths |v1 - v0|, |v2 - v0|, |v1 - v2| of a geometry row
__device__ __forceinline__ void edge_lengths(const double *g, double (&l)[3]) {
    l[0] = sqrt(g[3] * g[3] + g[4] * g[4] + g[5] * g[5]);
    l[1] = sqrt(g[6] * g[6] + g[7] * g[7] + g[8] * g[8]);
    const double d0 = g[3] - g[6], d1 = g[4] - g[7], d2 = g[5] - g[8];
    l[2] = sqrt(d0 * d0 + d1 * d1 + d2 * d2);
}

// The 3 x 3 block V_ij / (l_i l_j) of a regular pair in FP64 with the plain rule weights (the
// Jacobians cancel the Piola maps): ap(p) the offsets of the test points from the test centroid,
// e1 / e2 the test edge vectors, b the trial offsets, kap c_f - Q_j of the trial element,
// fd = c_f - c_e.
// PH: phase mode of e^{ikr} (kPhReduced | kPhReal when the caller has bounded |k r| on the wave's
// pairs and the wavenumber is real).
template <int PH = 0, int Q, typename AP>
__device__ __forceinline__ void pair_block_fp64(AP ap_of, const double *e1, const double *e2, const double (&b)[Q][3],
                                                const double (&kap)[3][3], const double (&fd)[3], const RuleArgs &rule,
                                                double kr, double ki, double (&ar)[3][3], double (&ai)[3][3]) {
    PairMoments<double> M;
    M.clear();
#pragma unroll 1
    for (int p = 0; p < Q; ++p) {
        double ap[3];
        ap_of(p, ap);
        const double d0 = ap[0] - fd[0], d1 = ap[1] - fd[1], d2 = ap[2] - fd[2];
        double Rp[2] = {0.0, 0.0}, Yp[2][3] = {{0.0, 0.0, 0.0}, {0.0, 0.0, 0.0}};
#pragma unroll
        for (int q = 0; q < Q; ++q) {
            const double u0 = d0 - b[q][0], u1 = d1 - b[q][1], u2 = d2 - b[q][2];  // x_p - y_q
            const double r2 = u0 * u0 + u1 * u1 + u2 * u2;
            const double rinv = rsqrt_t(r2), r = r2 * rinv;
            double sn, cs;
            sincos_t<PH>(kr * r, &sn, &cs);
            const double damp = (PH & kPhReal) ? 1.0 : ki != 0.0 ? exp(-ki * r) : 1.0;
            const double g = rule.w[p] * rule.w[q] * kInv4Pi * damp * (rinv * rinv * rinv);
            const double fr = (PH & kPhReal) ? -1.0 : -ki * r - 1.0, fi = kr * r;
            const double vr = g * (cs * fr - sn * fi), vi = g * (cs * fi + sn * fr);
            Rp[0] += vr;
            Rp[1] += vi;
#pragma unroll
            for (int d = 0; d < 3; ++d) {
                Yp[0][d] += vr * b[q][d];
                Yp[1][d] += vi * b[q][d];
            }
        }
        M.add_point(Rp, Yp, ap, fd);
    }
    double pi[3][3];
    rwg_offsets(e1, e2, pi);
    M.block(pi, kap, fd, ar, ai);
}

}  // namespace mfie
}  // namespace bem
