// Maxwell magnetic field boundary operator (MFIE), RWG x SNC, for gfx950.
//
// V_ij = l_i l_j sum_pq w_p w_q G(r) (ik r - 1) / r^2 (x_p - y_q) . (A_i(p) x B_j(q))
// (bempp_cl/core/numba_kernels.py:2151-2277 regular, 2065-2147 singular), with G the Helmholtz
// single-layer kernel e^{ikr} / (4 pi r), A_i = E_t (xi_p - o_i) and B_j = E_f (eta_q - o_j) the
// Piola-mapped RWG functions times their Jacobians (the J_t J_f of the quadrature cancels the
// 1/J of the two Piola maps), o = ((0, 1), (1, 0), (0, 0)) and l the edge lengths.
// The kernels add to the output with atomics:
//  * regular, FP64: one lane per trial element, test elements staged in LDS in tiles of 16;
//    pairs that share a vertex are left to the singular kernel (grids identical);
//  * regular, mixed, without dof blocks: far pairs in FP32 with the pair's reference distance
//    r0 = |c_f - c_e| and e^{i k r0} formed in FP64 and only r - r0 (from the small offsets of the
//    points from the centroids) in FP32, as the other mixed kernels (far_split.h); with dof blocks
//    the far pairs go through the tiled regular kernel (regular_mfie.hip) instead;
//  * near pairs of the mixed mode (pair_is_near), FP64: one lane per candidate pair
//    (bem_near_assemble);
//  * singular, FP64: eight lanes per Sauter-Schwab pair, its Duffy points strided over the
//    group, the 3 x 3 block reduced across the group.
#include "bem_internal.h"
#include "far_split.h"
#include "mfie_block.h"

namespace bem {
namespace mfie {

constexpr int kThreads = 256;
constexpr int kTile = 16;  // test elements per workgroup
constexpr int kQ = 6;      // quadrature points of the regular kernel (the order-4 rule)
constexpr double kInv2Pi = 0.15915494309189533577;  // phases in revolutions

struct Args {
    SpaceArgs test, trial;
    RuleArgs rule;
    double kr, ki;
    double near_factor;  // mixed mode: the pair_is_near split
    int identical;
    void *out;
    int64_t ld;
    // singular pairs
    int n_pairs;
    const int32_t *pairs;
    const double *tpts, *spts, *wts;
};


// E (xi - o_i, eta - o'_i) for the three RWG functions of an element (edge vectors e1, e2)
__device__ __forceinline__ void piola_j(const double *e1, const double *e2, double xi, double eta, double (&A)[3][3]) {
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int d = 0; d < 3; ++d) A[i][d] = e1[d] * (xi - kOx[i]) + e2[d] * (eta - kOy[i]);
}


// acc[i][j] += wgt G (ik r - 1) / r^2 (x - y) . (A_i x B_j), complex wavenumber k = kr + i ki
__device__ __forceinline__ void accumulate(const double (&diff)[3], const double (&A)[3][3], const double (&B)[3][3],
                                           double wgt, double kr, double ki, double (&ar)[3][3], double (&ai)[3][3]) {
    const double r2 = diff[0] * diff[0] + diff[1] * diff[1] + diff[2] * diff[2];
    const double r = sqrt(r2);
    double sn, cs;
    sincos_fast(kr * r, &sn, &cs);
    const double damp = ki != 0.0 ? exp(-ki * r) : 1.0;
    const double g = wgt * kInv4Pi * damp / (r * r2);  // |G| / r^2 times the weight
    // G (ik r - 1) / r^2 = g e^{i kr r} ((-ki r - 1) + i kr r)
    const double fr = -ki * r - 1.0, fi = kr * r;
    const double vr = g * (cs * fr - sn * fi), vi = g * (cs * fi + sn * fr);
#pragma unroll
    for (int i = 0; i < 3; ++i) {
        // D = diff x A_i, then D . B_j = diff . (A_i x B_j)
        const double D0 = diff[1] * A[i][2] - diff[2] * A[i][1];
        const double D1 = diff[2] * A[i][0] - diff[0] * A[i][2];
        const double D2 = diff[0] * A[i][1] - diff[1] * A[i][0];
#pragma unroll
        for (int j = 0; j < 3; ++j) {
            const double s = D0 * B[j][0] + D1 * B[j][1] + D2 * B[j][2];
            ar[i][j] += vr * s;
            ai[i][j] += vi * s;
        }
    }
}

// dofs and multiplier x edge length of the three RWG functions of element e of a space
__device__ __forceinline__ void dof_factors(const SpaceArgs &sp, int e, int (&dof)[3], double (&ml)[3]) {
    double l[3];
    edge_lengths(sp.geom + (int64_t)e * BEM_GEOM_STRIDE, l);
#pragma unroll
    for (int i = 0; i < 3; ++i) {
        dof[i] = sp.l2g[3 * (int64_t)e + i];
        ml[i] = sp.mult[3 * (int64_t)e + i] * l[i];
    }
}

// out[dof_t[i], dof_f[j]] += (ar + i ai)[i][j] ml_t[i] ml_f[j] for the 9 entries of a pair
template <typename OutT, typename AccT>
__device__ __forceinline__ void scatter_block(const Args &a, const int *dt, const double *mt, const int (&df)[3],
                                              const double (&mf)[3], const AccT (&ar)[3][3], const AccT (&ai)[3][3]) {
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j) {
            const double m = mt[i] * mf[j];
            if (m == 0.0) continue;
            OutT *o = reinterpret_cast<OutT *>(a.out) + 2 * ((int64_t)dt[i] * a.ld + df[j]);
            atomic_add(o, (OutT)(ar[i][j] * m));
            atomic_add(o + 1, (OutT)(ai[i][j] * m));
        }
}


// offsets of the quadrature points from the centroid, b_q = e1 (x_q - 1/3) + e2 (y_q - 1/3)
__device__ __forceinline__ void centroid_offsets(const double *g, const RuleArgs &rule, double (&b)[kQ][3]) {
#pragma unroll
    for (int q = 0; q < kQ; ++q)
#pragma unroll
        for (int d = 0; d < 3; ++d) b[q][d] = g[3 + d] * (rule.x[q] - 1.0 / 3.0) + g[6 + d] * (rule.y[q] - 1.0 / 3.0);
}

struct Tile {
    int elem[kTile];
    int vid[kTile][3];
    double a[kTile][kQ][3];  // x_p - c_e
    double e1[kTile][3], e2[kTile][3];
    double c[kTile][3];  // centroids
    double diam[kTile];
    int dof[kTile][3];
    double ml[kTile][3];  // multiplier x edge length
};

template <typename OutT>
__global__ __launch_bounds__(kThreads) void mfie_regular_kernel(const Args a) {
    __shared__ Tile T_;
    const int tid = threadIdx.x;
    const int jf = blockIdx.x * kThreads + tid;
    const bool valid = jf < a.trial.n;
    const int t0 = blockIdx.y * kTile;
    const int ntile = min(kTile, a.test.n - t0);
    if (tid < ntile) {
        const int e = a.test.elements[t0 + tid];
        const double *g = a.test.geom + (int64_t)e * BEM_GEOM_STRIDE;
        T_.elem[tid] = e;
        for (int i = 0; i < 3; ++i) T_.vid[tid][i] = a.test.vids[4 * e + i];
        for (int d = 0; d < 3; ++d) {
            T_.e1[tid][d] = g[3 + d];
            T_.e2[tid][d] = g[6 + d];
        }
        for (int p = 0; p < kQ; ++p)
            for (int d = 0; d < 3; ++d)
                T_.a[tid][p][d] = g[3 + d] * (a.rule.x[p] - 1.0 / 3.0) + g[6 + d] * (a.rule.y[p] - 1.0 / 3.0);
        centroid_rn(g, T_.c[tid]);
        T_.diam[tid] = g[13];
        dof_factors(a.test, e, T_.dof[tid], T_.ml[tid]);
    }
    __syncthreads();
    if (!valid) return;
    const int f = a.trial.elements[jf];
    const double *gf = a.trial.geom + (int64_t)f * BEM_GEOM_STRIDE;
    const int fv0 = a.trial.vids[4 * f], fv1 = a.trial.vids[4 * f + 1], fv2 = a.trial.vids[4 * f + 2];
    double cf[3];
    centroid_rn(gf, cf);
    int fdof[3];
    double fml[3];
    dof_factors(a.trial, f, fdof, fml);
    // trial points as offsets b_q from the centroid
    double b[kQ][3], kap[3][3];
    centroid_offsets(gf, a.rule, b);
    rwg_offsets(gf + 3, gf + 6, kap);
    for (int t = 0; t < ntile; ++t) {
        if (a.identical) {
            bool adj = false;
            for (int i = 0; i < 3; ++i) {
                const int v = T_.vid[t][i];
                adj = adj || v == fv0 || v == fv1 || v == fv2;
            }
            if (adj) continue;
        }
        const double fd[3] = {cf[0] - T_.c[t][0], cf[1] - T_.c[t][1], cf[2] - T_.c[t][2]};
        double ar[3][3], ai[3][3];
        auto ap_of = [&](int p, double (&ap)[3]) {
            for (int d = 0; d < 3; ++d) ap[d] = T_.a[t][p][d];
        };
        const double fdn = sqrt(fd[0] * fd[0] + fd[1] * fd[1] + fd[2] * fd[2]);
        if (fp64_phase_reduced(a.kr, a.ki, fdn, T_.diam[t] + gf[13]))
            pair_block_fp64<kPhReduced | kPhReal>(ap_of, T_.e1[t], T_.e2[t], b, kap, fd, a.rule, a.kr, a.ki, ar, ai);
        else
            pair_block_fp64(ap_of, T_.e1[t], T_.e2[t], b, kap, fd, a.rule, a.kr, a.ki, ar, ai);
        scatter_block<OutT>(a, T_.dof[t], T_.ml[t], fdof, fml, ar, ai);
    }
}

// ------------------------------------------------------------------ near pairs (mixed mode), FP64
// one lane per candidate pair (e, f) of a.pairs ((n, 2) rows); pairs that share a vertex (grids
// identical) or lie outside the near factor are skipped
template <typename OutT>
__global__ __launch_bounds__(kThreads) void mfie_near_pairs_kernel(const Args a) {
    const int p = blockIdx.x * kThreads + threadIdx.x;
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
    double at[kQ][3], b[kQ][3], kap[3][3];
    centroid_offsets(gt, a.rule, at);
    centroid_offsets(gf, a.rule, b);
    rwg_offsets(gf + 3, gf + 6, kap);
    const double fd[3] = {cf[0] - ct[0], cf[1] - ct[1], cf[2] - ct[2]};
    double ar[3][3], ai[3][3];
    auto ap_of = [&](int q, double (&ap)[3]) {
        for (int d = 0; d < 3; ++d) ap[d] = at[q][d];
    };
    if (fp64_phase_reduced(a.kr, a.ki, sqrt(dist2), gt[13] + gf[13]))
        pair_block_fp64<kPhReduced | kPhReal>(ap_of, gt + 3, gt + 6, b, kap, fd, a.rule, a.kr, a.ki, ar, ai);
    else
        pair_block_fp64(ap_of, gt + 3, gt + 6, b, kap, fd, a.rule, a.kr, a.ki, ar, ai);
    int tdof[3], fdof[3];
    double tml[3], fml[3];
    dof_factors(a.test, e, tdof, tml);
    dof_factors(a.trial, f, fdof, fml);
    scatter_block<OutT>(a, tdof, tml, fdof, fml, ar, ai);
}

// ------------------------------------------------------------------ mixed: far pairs in FP32
struct TileF32 {
    int elem[kTile];
    int vid[kTile][3];
    double c[kTile][3];
    double diam[kTile];
    float a[kTile][kQ][3];  // x_p - c_e
    float e1[kTile][3], e2[kTile][3];
    int dof[kTile][3];
    double ml[kTile][3];  // multiplier x edge length
};

template <typename OutT>
__global__ __launch_bounds__(kThreads) void mfie_regular_mixed_kernel(const Args a) {
    __shared__ TileF32 T_;
    const int tid = threadIdx.x;
    const int jf = blockIdx.x * kThreads + tid;
    const bool valid = jf < a.trial.n;
    const int t0 = blockIdx.y * kTile;
    const int ntile = min(kTile, a.test.n - t0);
    if (tid < ntile) {
        const int e = a.test.elements[t0 + tid];
        const double *g = a.test.geom + (int64_t)e * BEM_GEOM_STRIDE;
        T_.elem[tid] = e;
        for (int i = 0; i < 3; ++i) T_.vid[tid][i] = a.test.vids[4 * e + i];
        centroid_rn(g, T_.c[tid]);
        T_.diam[tid] = g[13];
        for (int d = 0; d < 3; ++d) {
            T_.e1[tid][d] = (float)g[3 + d];
            T_.e2[tid][d] = (float)g[6 + d];
        }
        for (int p = 0; p < kQ; ++p)
            for (int d = 0; d < 3; ++d)
                T_.a[tid][p][d] = (float)(g[3 + d] * (a.rule.x[p] - 1.0 / 3.0) + g[6 + d] * (a.rule.y[p] - 1.0 / 3.0));
        dof_factors(a.test, e, T_.dof[tid], T_.ml[tid]);
    }
    __syncthreads();
    if (!valid) return;
    const int f = a.trial.elements[jf];
    const double *gf = a.trial.geom + (int64_t)f * BEM_GEOM_STRIDE;
    const int fv0 = a.trial.vids[4 * f], fv1 = a.trial.vids[4 * f + 1], fv2 = a.trial.vids[4 * f + 2];
    double cf[3];
    centroid_rn(gf, cf);
    int fdof[3];
    double fml[3];
    dof_factors(a.trial, f, fdof, fml);
    // trial points as offsets b_q from the centroid, FP32
    float b[kQ][3], kap[3][3], wq[kQ];
#pragma unroll
    for (int q = 0; q < kQ; ++q) {
        const double s0 = a.rule.x[q], s1 = a.rule.y[q];
#pragma unroll
        for (int d = 0; d < 3; ++d) b[q][d] = (float)(gf[3 + d] * (s0 - 1.0 / 3.0) + gf[6 + d] * (s1 - 1.0 / 3.0));
        wq[q] = (float)(a.rule.w[q] * kInv4Pi);
    }
    rwg_offsets(gf + 3, gf + 6, kap);
    const float kr = (float)a.kr, ki = (float)a.ki, krev = (float)(a.kr * kInv2Pi);
    for (int t = 0; t < ntile; ++t) {
        if (a.identical) {
            bool adj = false;
            for (int i = 0; i < 3; ++i) {
                const int v = T_.vid[t][i];
                adj = adj || v == fv0 || v == fv1 || v == fv2;
            }
            if (adj) continue;
        }
        double dist2;
        if (pair_is_near(T_.c[t], cf, T_.diam[t], gf[13], a.near_factor, &dist2)) continue;  // near pass
        const FarPair P = far_pair_setup(T_.c[t], cf);  // fd = c_f - c_e
        float cs0, sn0;
        phase_rev_f32(a.kr, (double)P.r0 * kInv2Pi, cs0, sn0);
        const float damp0 = a.ki != 0.0 ? (float)exp(-a.ki * (double)P.r0) : 1.f;
        cs0 *= damp0;
        sn0 *= damp0;
        const float fd[3] = {P.fd0, P.fd1, P.fd2};
        PairMoments<float> M;
        M.clear();
#pragma unroll 1
        for (int p = 0; p < kQ; ++p) {
            const float wp = (float)a.rule.w[p];
            const float ap0 = T_.a[t][p][0], ap1 = T_.a[t][p][1], ap2 = T_.a[t][p][2];
            float Rp[2] = {0.f, 0.f}, Yp[2][3] = {{0.f, 0.f, 0.f}, {0.f, 0.f, 0.f}};
#pragma unroll
            for (int q = 0; q < kQ; ++q) {
                // x - y = (c_e + a_p) - (c_f + b_q) = u - fd with u = a_p - b_q small
                const float u0 = ap0 - b[q][0], u1 = ap1 - b[q][1], u2 = ap2 - b[q][2];
                const float tt = P.resid - 2.f * (P.fd0 * u0 + P.fd1 * u1 + P.fd2 * u2) + (u0 * u0 + u1 * u1 + u2 * u2);
                const float r2 = P.r0sq + tt;
                const float rinv = __builtin_amdgcn_rsqf(r2);
                const float r = r2 * rinv;
                const float delta = tt * __builtin_amdgcn_rcpf(r + P.r0);  // r - r0
                const float c = __builtin_amdgcn_cosf(krev * delta), s = __builtin_amdgcn_sinf(krev * delta);
                float er = cs0 * c - sn0 * s, ei = cs0 * s + sn0 * c;  // e^{ik r} (with e^{-Im k r0})
                if (ki != 0.f) {
                    const float dm = __expf(-ki * delta);
                    er *= dm;
                    ei *= dm;
                }
                // G (ik r - 1) / r^2 times the weights: e^{ikr} ((-Im k r - 1) + i Re k r) w_p w_q / (4 pi r^3)
                const float mag = wp * wq[q] * rinv * rinv * rinv;
                const float fr = -ki * r - 1.f, fi = kr * r;
                const float vr = mag * (er * fr - ei * fi), vi = mag * (er * fi + ei * fr);
                Rp[0] += vr;
                Rp[1] += vi;
#pragma unroll
                for (int d = 0; d < 3; ++d) {
                    Yp[0][d] += vr * b[q][d];
                    Yp[1][d] += vi * b[q][d];
                }
            }
            const float ap[3] = {ap0, ap1, ap2};
            M.add_point(Rp, Yp, ap, fd);
        }
        float pi[3][3], ar[3][3], ai[3][3];
        rwg_offsets(T_.e1[t], T_.e2[t], pi);
        M.block(pi, kap, fd, ar, ai);
        scatter_block<OutT>(a, T_.dof[t], T_.ml[t], fdof, fml, ar, ai);
    }
}

constexpr int kSingGroup = 8;  // lanes per singular pair

template <typename T>
__device__ __forceinline__ T group_sum(T v) {
#pragma unroll
    for (int off = kSingGroup / 2; off >= 1; off >>= 1) v += __shfl_xor(v, off, kSingGroup);
    return v;
}

// kSingGroup lanes per Sauter-Schwab pair (eight pairs per wavefront): lane g of the group takes
// the Duffy points g, g + 8, ...; the 3 x 3 block is summed over the group with xor shuffles and
// its entries are scattered by the group's lanes (lane g entry g, lane 0 also entry 8).
template <typename OutT>
__global__ __launch_bounds__(kThreads) void mfie_singular_kernel(const Args a) {
    const int g = threadIdx.x & (kSingGroup - 1);
    const int pair = (blockIdx.x * kThreads + threadIdx.x) / kSingGroup;
    if (pair >= a.n_pairs) return;  // uniform per group
    const int32_t *pd = a.pairs + 6 * (int64_t)pair;
    const int e = pd[0], f = pd[1], toff = pd[2], soff = pd[3], woff = pd[4], np = pd[5];
    const double *gt = a.test.geom + (int64_t)e * BEM_GEOM_STRIDE;
    const double *gf = a.trial.geom + (int64_t)f * BEM_GEOM_STRIDE;
    double ar[3][3], ai[3][3];
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j) ar[i][j] = ai[i][j] = 0.0;
    for (int k = g; k < np; k += kSingGroup) {
        const double xi0 = a.tpts[2 * (int64_t)(toff + k)], xi1 = a.tpts[2 * (int64_t)(toff + k) + 1];
        const double et0 = a.spts[2 * (int64_t)(soff + k)], et1 = a.spts[2 * (int64_t)(soff + k) + 1];
        double A[3][3], B[3][3];
        piola_j(gt + 3, gt + 6, xi0, xi1, A);
        piola_j(gf + 3, gf + 6, et0, et1, B);
        double diff[3];
#pragma unroll
        for (int d = 0; d < 3; ++d)
            diff[d] = (gt[d] + gt[3 + d] * xi0 + gt[6 + d] * xi1) - (gf[d] + gf[3 + d] * et0 + gf[6 + d] * et1);
        accumulate(diff, A, B, a.wts[woff + k], a.kr, a.ki, ar, ai);
    }
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j) {
            ar[i][j] = group_sum(ar[i][j]);
            ai[i][j] = group_sum(ai[i][j]);
        }
    int tdof[3], fdof[3];
    double tml[3], fml[3];
    dof_factors(a.test, e, tdof, tml);
    dof_factors(a.trial, f, fdof, fml);
    // compile-time indices into the register blocks
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j) {
            const int entry = 3 * i + j;
            if (g != entry % kSingGroup) continue;
            const double m = tml[i] * fml[j];
            if (m == 0.0) continue;
            OutT *o = reinterpret_cast<OutT *>(a.out) + 2 * ((int64_t)tdof[i] * a.ld + fdof[j]);
            atomic_add(o, (OutT)(ar[i][j] * m));
            atomic_add(o + 1, (OutT)(ai[i][j] * m));
        }
}

}  // namespace mfie

// regular pairs of the MFIE (out_dtype BEM_C64 or BEM_C128), added to the output: FP64, or
// MIXED (far pairs only, FP32; the near pairs are left to mfie_near_launch)
int mfie_regular_launch(const SpaceArgs &test, const SpaceArgs &trial, const RuleArgs &rule, double kr, double ki,
                        int identical, int precision, double near_factor, void *out, int64_t ld, int out_dtype,
                        hipStream_t s) {
    if (rule.n != mfie::kQ) {
        set_error("the magnetic field operator uses the 6-point regular rule (order 4), got %d points", rule.n);
        return BEM_ERR_INVALID_ARGUMENT;
    }
    if (out_dtype != BEM_C128 && out_dtype != BEM_C64) {
        set_error("the magnetic field operator needs complex output");
        return BEM_ERR_INVALID_ARGUMENT;
    }
    if (test.n == 0 || trial.n == 0) return BEM_OK;
    const int64_t ntiles = (test.n + mfie::kTile - 1) / mfie::kTile;
    if (ntiles > 65535) {
        set_error("too many test tiles for one launch (%lld)", (long long)ntiles);
        return BEM_ERR_INVALID_ARGUMENT;
    }
    mfie::Args a{};
    a.test = test;
    a.trial = trial;
    a.rule = rule;
    a.kr = kr;
    a.ki = ki;
    a.near_factor = near_factor;
    a.identical = identical;
    a.out = out;
    a.ld = ld;
    const dim3 grid((trial.n + mfie::kThreads - 1) / mfie::kThreads, (unsigned)ntiles);
    const bool c128 = out_dtype == BEM_C128;
    if (precision == BEM_PRECISION_MIXED) {
        if (c128) hipLaunchKernelGGL(mfie::mfie_regular_mixed_kernel<double>, grid, dim3(mfie::kThreads), 0, s, a);
        else hipLaunchKernelGGL(mfie::mfie_regular_mixed_kernel<float>, grid, dim3(mfie::kThreads), 0, s, a);
        return check_launch("mfie_regular_mixed_kernel");
    }
    if (c128) hipLaunchKernelGGL(mfie::mfie_regular_kernel<double>, grid, dim3(mfie::kThreads), 0, s, a);
    else hipLaunchKernelGGL(mfie::mfie_regular_kernel<float>, grid, dim3(mfie::kThreads), 0, s, a);
    return check_launch("mfie_regular_kernel");
}

// near pairs of the mixed mode among n_pairs candidates ((n, 2) int32 rows), FP64, added to the output
int mfie_near_launch(const SpaceArgs &test, const SpaceArgs &trial, const RuleArgs &rule, double kr, double ki,
                     int identical, double near_factor, int n_pairs, const int32_t *pairs, void *out, int64_t ld,
                     int out_dtype, hipStream_t s) {
    if (rule.n != mfie::kQ) {
        set_error("the magnetic field operator uses the 6-point regular rule (order 4), got %d points", rule.n);
        return BEM_ERR_INVALID_ARGUMENT;
    }
    if (out_dtype != BEM_C128 && out_dtype != BEM_C64) {
        set_error("the magnetic field operator needs complex output");
        return BEM_ERR_INVALID_ARGUMENT;
    }
    if (n_pairs == 0) return BEM_OK;
    mfie::Args a{};
    a.test = test;
    a.trial = trial;
    a.rule = rule;
    a.kr = kr;
    a.ki = ki;
    a.near_factor = near_factor;
    a.identical = identical;
    a.out = out;
    a.ld = ld;
    a.n_pairs = n_pairs;
    a.pairs = pairs;
    const dim3 grid((unsigned)((n_pairs + mfie::kThreads - 1) / mfie::kThreads));
    if (out_dtype == BEM_C128) hipLaunchKernelGGL(mfie::mfie_near_pairs_kernel<double>, grid, dim3(mfie::kThreads), 0, s, a);
    else hipLaunchKernelGGL(mfie::mfie_near_pairs_kernel<float>, grid, dim3(mfie::kThreads), 0, s, a);
    return check_launch("mfie_near_pairs_kernel");
}

// singular pairs of the MFIE, added to the output
int mfie_singular_launch(const SpaceArgs &test, const SpaceArgs &trial, int n_pairs, const int32_t *pairs,
                         const double *tpts, const double *spts, const double *wts, double kr, double ki, void *out,
                         int64_t ld, int out_dtype, hipStream_t s) {
    if (n_pairs == 0) return BEM_OK;
    mfie::Args a{};
    a.test = test;
    a.trial = trial;
    a.kr = kr;
    a.ki = ki;
    a.out = out;
    a.ld = ld;
    a.n_pairs = n_pairs;
    a.pairs = pairs;
    a.tpts = tpts;
    a.spts = spts;
    a.wts = wts;
    const dim3 grid((unsigned)(((int64_t)n_pairs * mfie::kSingGroup + mfie::kThreads - 1) / mfie::kThreads));
    if (out_dtype == BEM_C128)
        hipLaunchKernelGGL(mfie::mfie_singular_kernel<double>, grid, dim3(mfie::kThreads), 0, s, a);
    else if (out_dtype == BEM_C64)
        hipLaunchKernelGGL(mfie::mfie_singular_kernel<float>, grid, dim3(mfie::kThreads), 0, s, a);
    else {
        set_error("the magnetic field operator needs complex output");
        return BEM_ERR_INVALID_ARGUMENT;
    }
    return check_launch("mfie_singular_kernel");
}

}  // namespace bem
