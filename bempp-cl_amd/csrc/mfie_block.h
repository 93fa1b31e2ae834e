// The 3 x 3 block of a regular pair of the Maxwell magnetic field operator (MFIE) from moments of
// the kernel values, shared by the MFIE kernels (mfie.hip) and the tiled regular kernel
// (regular_impl.h). The RWG function i of an element is A_i(x) = x - P_i (times l_i / J), P_i the
// vertex opposite edge i: with the reference coordinates o = ((0, 1), (1, 0), (0, 0)),
// P_i = v0 + e1 o_i0 + e2 o_i1 (bempp_cl/core/numba_kernels.py:2151-2277).
#pragma once

#include <hip/hip_runtime.h>

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

}  // namespace mfie
}  // namespace bem
