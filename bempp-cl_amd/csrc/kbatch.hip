// Frequency-batched Helmholtz single layer, DP0 x DP0, for gfx950.
//
// Assembles the Helmholtz single-layer Galerkin matrices of nk real wavenumbers k_m at once:
//   out[m][e][f] = sum_pq w_p J_e w_q J_f e^{i k_m r_pq} / (4 pi r_pq)
// over regular pairs, plus the near-field (mixed mode) and singular (Sauter-Schwab) pairs
// in FP64. The reference evaluates one wavenumber per assembly
// (bempp_cl/core/numba_kernels.py:298-320 inside the regular loop at 980-1074).
//
// Wavenumber expansion (mixed mode). For a pair with reference distance r0 and point
// offsets delta_pq = r_pq - r0,
//   sum_pq a_pq e^{i k r_pq} = e^{i k r0} sum_n (i k)^n / n! M_n,   M_n = sum_pq a_pq delta_pq^n,
// truncated after the order N at which the remainder bound x^{N+1} / (N+1)! (x = k_max
// max |delta|, the Lagrange bound for e^{i theta}) drops below a tolerance (1e-9 of
// sum |a| for FP32 far pairs, 1e-14 for FP64 near and singular pairs). The moments depend on
// the geometry only and are formed once per pair; every wavenumber then costs a polynomial
// in k^2 (Horner) and one phase factor e^{i k r0}, instead of 36 sin/cos evaluations. Since
// |delta_pq| <= |a_p| + |b_q| (offsets of the quadrature points from the centroids), the
// order follows from the element sizes; the order is made uniform per wavefront and pairs
// needing more than the largest compiled order fall back to direct evaluation.
// FP64 mode evaluates every point pair and wavenumber directly (it is the reference the
// mixed sweep is checked against).
//
// Regular pairs run in one persistent launch: gridDim.x workgroups walk the
// (test tile x trial block) space with a static stride. Each lane owns one trial element;
// the test elements of the tile are staged in LDS and read as broadcasts.
#include "far_split.h"
#include "regular_impl.h"

namespace bem {
namespace kb {

constexpr int kKMax = BEM_KBATCH_MAX_WAVENUMBERS;
constexpr int kQ = 6;
constexpr float kTol32 = 1e-9f;
// FP64 moment expansions of the mixed mode's near and singular pairs: their sums are stored in
// complex64 (relative rounding 6e-8), so a series tail of 1e-10 of the leading moment is exact
// enough (at k = 16 on 5,120 triangles the singular pairs then take 16 instead of 24 moments)
constexpr double kTolMoments = 1e-10;

struct Args {
    SpaceArgs test, trial;
    RuleArgs rule;
    int nk;
    double k[kKMax];
    double kmax;
    double near_factor;
    void *out;
    int64_t ld;
    int64_t kstride;
    int identical;
    int ntx, nty;
};

// Smallest N with x^{N+1} / (N+1)! <= tol (x >= 0); nmax + 1 if that is above nmax.
template <typename T>
__device__ __forceinline__ int taylor_order(T x, T tol, int nmax) {
    T term = x;
    int n = 0;
    while (!(term <= tol)) {
        ++n;
        if (n > nmax) return nmax + 1;
        term *= x / (T)(n + 1);
    }
    return n;
}

// sum_n (i k)^n c_n (c_n already divided by n!): even and odd parts by Horner in k^2
template <int NM, typename T>
__device__ __forceinline__ void taylor_eval(const T (&c)[NM + 1], T k, T &re, T &im) {
    const T kap = k * k;
    constexpr int NE = NM / 2;        // highest even index / 2
    constexpr int NO = (NM - 1) / 2;  // highest odd index / 2
    T e = ((NE & 1) ? -c[2 * NE] : c[2 * NE]);
#pragma unroll
    for (int j = NE - 1; j >= 0; --j) e = e * kap + ((j & 1) ? -c[2 * j] : c[2 * j]);
    T o = ((NO & 1) ? -c[2 * NO + 1] : c[2 * NO + 1]);
#pragma unroll
    for (int j = NO - 1; j >= 0; --j) o = o * kap + ((j & 1) ? -c[2 * j + 1] : c[2 * j + 1]);
    re = e;
    im = o * k;
}

// divide the moments by n! (constants folded at compile time) and scale by s
template <int NM, typename T>
__device__ __forceinline__ void scale_factorials(T (&c)[NM + 1], T s) {
    double inv = 1.0;
#pragma unroll
    for (int n = 0; n <= NM; ++n) {
        if (n > 0) inv /= n;
        c[n] *= s * (T)inv;
    }
}

// ------------------------------------------------------------------ FP32 far pairs
template <int TM>
struct TileF32 {
    int elem[TM];
    int vid[TM][3];
    double mult[TM];
    double c[TM][3];
    double diam[TM];
    float a[TM][kQ][3];  // quadrature point offsets from the centroid
    float aa[TM][kQ];    // |a|^2
    float wp[TM][kQ];    // w_p J_e
    float rho[TM];       // max |a_p|
};

struct TrialF32Lane {
    float b[kQ][3];
    float bb[kQ];
    float wq[kQ];  // w_q J_f / (4 pi)
};

// visit the 36 point pairs: f(amplitude w w J J / (4 pi r), delta = r - r0)
template <int TM, typename F>
__device__ __forceinline__ void far_points(const TileF32<TM> &T_, int t, const TrialF32Lane &L, const FarPair &P, F &&f) {
#pragma unroll 1
    for (int p = 0; p < kQ; ++p) {
        const float w0 = P.fd0 - T_.a[t][p][0], w1 = P.fd1 - T_.a[t][p][1], w2 = P.fd2 - T_.a[t][p][2];
        const float Ap =
            P.resid - 2.f * (P.fd0 * T_.a[t][p][0] + P.fd1 * T_.a[t][p][1] + P.fd2 * T_.a[t][p][2]) + T_.aa[t][p];
        const float wp = T_.wp[t][p];
#pragma unroll
        for (int q = 0; q < kQ; ++q) {
            const float tt = Ap + L.bb[q] + 2.f * (w0 * L.b[q][0] + w1 * L.b[q][1] + w2 * L.b[q][2]);
            const float r2 = P.r0sq + tt;
            const float rinv = __builtin_amdgcn_rsqf(r2);
            const float r = r2 * rinv;
            f(rinv * L.wq[q] * wp, tt * __builtin_amdgcn_rcpf(r + P.r0));
        }
    }
}

__device__ __forceinline__ void store_c64(void *out, int64_t idx, float re, float im) {
    float2 v;
    v.x = re;
    v.y = im;
    *reinterpret_cast<float2 *>(reinterpret_cast<float *>(out) + 2 * idx) = v;
}

template <int NM, int TM>
__device__ __forceinline__ void far_pair_moments(const Args &a, const TileF32<TM> &T_, int t, const TrialF32Lane &L,
                                                 const FarPair &P, bool skip, float m, bool valid, int64_t base) {
    float c[NM + 1];
#pragma unroll
    for (int n = 0; n <= NM; ++n) c[n] = 0.f;
    if (!skip) {
        far_points(T_, t, L, P, [&](float amp, float dl) {
            float pw = amp;
#pragma unroll
            for (int n = 0; n <= NM; ++n) {
                c[n] += pw;
                if (n < NM) pw *= dl;
            }
        });
    }
    scale_factorials<NM>(c, m);
    const double r0rev = (double)P.r0 * kInv2Pi;
    for (int km = 0; km < a.nk; ++km) {
        const double kd = a.k[km];
        float re, im, cs, sn;
        taylor_eval<NM>(c, (float)kd, re, im);
        phase_rev_f32(kd, r0rev, cs, sn);
        if (valid) store_c64(a.out, km * a.kstride + base, cs * re - sn * im, cs * im + sn * re);
    }
}

// direct evaluation for pairs too large for the compiled expansion orders
template <int TM>
__device__ __forceinline__ void far_pair_direct(const Args &a, const TileF32<TM> &T_, int t, const TrialF32Lane &L,
                                                const FarPair &P, bool skip, float m, bool valid, int64_t base) {
    const double r0rev = (double)P.r0 * kInv2Pi;
    for (int km = 0; km < a.nk; ++km) {
        const double kd = a.k[km];
        const float krev = (float)(kd * kInv2Pi);
        float sr = 0.f, si = 0.f;
        if (!skip) {
            far_points(T_, t, L, P, [&](float amp, float dl) {
                const float ph = krev * dl;
                sr += amp * __builtin_amdgcn_cosf(ph);
                si += amp * __builtin_amdgcn_sinf(ph);
            });
        }
        float cs, sn;
        phase_rev_f32(kd, r0rev, cs, sn);
        if (valid) store_c64(a.out, km * a.kstride + base, m * (cs * sr - sn * si), m * (cs * si + sn * sr));
    }
}

template <int TM>
__global__ __launch_bounds__(kRegThreads) void kbatch_far_kernel(const Args a) {
    __shared__ TileF32<TM> T_;
    const int tid = threadIdx.x;
    const float kmax = (float)a.kmax;
    const int ntiles = a.ntx * a.nty;
    for (int tile = blockIdx.x; tile < ntiles; tile += gridDim.x) {
        const int bx = tile % a.ntx, by = tile / a.ntx;
        const int t0 = by * TM, ntile = min(TM, a.test.n - t0);
        const int jf = bx * kRegThreads + tid;
        const bool valid = jf < a.trial.n;
        __syncthreads();  // the previous tile's readers are done with T_
        if (tid < ntile) {
            const int e = a.test.elements[t0 + tid];
            const double *g = a.test.geom + (int64_t)e * BEM_GEOM_STRIDE;
            T_.elem[tid] = e;
            for (int i = 0; i < 3; ++i) T_.vid[tid][i] = a.test.vids[4 * e + i];
            T_.mult[tid] = a.test.mult[e];
            centroid_rn(g, T_.c[tid]);
            T_.diam[tid] = g[13];
            float rho2 = 0.f;
            for (int p = 0; p < kQ; ++p) {
                const double xi = a.rule.x[p] - 1.0 / 3.0, eta = a.rule.y[p] - 1.0 / 3.0;
                for (int d = 0; d < 3; ++d) T_.a[tid][p][d] = (float)(g[3 + d] * xi + g[6 + d] * eta);
                T_.aa[tid][p] = T_.a[tid][p][0] * T_.a[tid][p][0] + T_.a[tid][p][1] * T_.a[tid][p][1] +
                                T_.a[tid][p][2] * T_.a[tid][p][2];
                T_.wp[tid][p] = (float)(a.rule.w[p] * g[12]);
                rho2 = fmaxf(rho2, T_.aa[tid][p]);
            }
            T_.rho[tid] = sqrtf(rho2);
        }
        __syncthreads();

        int fv0 = -1, fv1 = -1, fv2 = -1, f = 0;
        double cf[3] = {0.0, 0.0, 0.0}, diam_f = 0.0, mf = 0.0;
        TrialF32Lane L;
        float rho2_f = 0.f;
        if (valid) {
            f = a.trial.elements[jf];
            const double *gf = a.trial.geom + (int64_t)f * BEM_GEOM_STRIDE;
            fv0 = a.trial.vids[4 * f];
            fv1 = a.trial.vids[4 * f + 1];
            fv2 = a.trial.vids[4 * f + 2];
            centroid_rn(gf, cf);
            diam_f = gf[13];
            mf = a.trial.mult[f];
#pragma unroll
            for (int q = 0; q < kQ; ++q) {
                const double s0 = a.rule.x[q] - 1.0 / 3.0, s1 = a.rule.y[q] - 1.0 / 3.0;
                L.b[q][0] = (float)(gf[3] * s0 + gf[6] * s1);
                L.b[q][1] = (float)(gf[4] * s0 + gf[7] * s1);
                L.b[q][2] = (float)(gf[5] * s0 + gf[8] * s1);
                L.bb[q] = L.b[q][0] * L.b[q][0] + L.b[q][1] * L.b[q][1] + L.b[q][2] * L.b[q][2];
                L.wq[q] = (float)(a.rule.w[q] * gf[12] * kInv4Pi);
                rho2_f = fmaxf(rho2_f, L.bb[q]);
            }
        } else {
#pragma unroll
            for (int q = 0; q < kQ; ++q) {
                L.b[q][0] = L.b[q][1] = L.b[q][2] = 0.f;
                L.bb[q] = 0.f;
                L.wq[q] = 0.f;
            }
        }
        const float rho_fw = wave_max64(sqrtf(rho2_f));
        const int64_t col = valid ? a.trial.l2g[f] : 0;

        for (int t = 0; t < ntile; ++t) {
            const int e = T_.elem[t];
            bool skip = !valid;
            if (valid) {
                if (a.identical)
                    skip = T_.vid[t][0] == fv0 || T_.vid[t][0] == fv1 || T_.vid[t][0] == fv2 || T_.vid[t][1] == fv0 ||
                           T_.vid[t][1] == fv1 || T_.vid[t][1] == fv2 || T_.vid[t][2] == fv0 ||
                           T_.vid[t][2] == fv1 || T_.vid[t][2] == fv2;
                double dist2;
                if (!skip) skip = pair_is_near(T_.c[t], cf, T_.diam[t], diam_f, a.near_factor, &dist2);
            }
            const int64_t base = valid ? (int64_t)a.test.l2g[e] * a.ld + col : 0;
            const float m = (float)(T_.mult[t] * mf);
            const FarPair P = far_pair_setup(T_.c[t], cf);
            // expansion order for the largest pair of the wavefront (x is wave-uniform)
            const float x = kmax * (T_.rho[t] + rho_fw) * 1.0001f;
            const int N = taylor_order(x, kTol32, 16);
            const int bucket = __builtin_amdgcn_readfirstlane(N <= 4 ? 0 : N <= 6 ? 1 : N <= 8 ? 2 : N <= 10 ? 3 : N <= 12 ? 4
                                                             : N <= 16 ? 5 : 6);
            switch (bucket) {
            case 0: far_pair_moments<4>(a, T_, t, L, P, skip, m, valid, base); break;
            case 1: far_pair_moments<6>(a, T_, t, L, P, skip, m, valid, base); break;
            case 2: far_pair_moments<8>(a, T_, t, L, P, skip, m, valid, base); break;
            case 3: far_pair_moments<10>(a, T_, t, L, P, skip, m, valid, base); break;
            case 4: far_pair_moments<12>(a, T_, t, L, P, skip, m, valid, base); break;
            case 5: far_pair_moments<16>(a, T_, t, L, P, skip, m, valid, base); break;
            default: far_pair_direct(a, T_, t, L, P, skip, m, valid, base); break;
            }
        }
    }
}

// ------------------------------------------------------------------ FP64 regular pairs
template <int TM>
struct TileF64 {
    int elem[TM];
    int vid[TM][3];
    double mult[TM];
    double xp[TM][kQ][3];
    double wp[TM][kQ];
};

template <int TM>
__global__ __launch_bounds__(kRegThreads) void kbatch_fp64_kernel(const Args a) {
    __shared__ TileF64<TM> T_;
    const int tid = threadIdx.x;
    const int ntiles = a.ntx * a.nty;
    for (int tile = blockIdx.x; tile < ntiles; tile += gridDim.x) {
        const int bx = tile % a.ntx, by = tile / a.ntx;
        const int t0 = by * TM, ntile = min(TM, a.test.n - t0);
        const int jf = bx * kRegThreads + tid;
        const bool valid = jf < a.trial.n;
        __syncthreads();
        if (tid < ntile) {
            const int e = a.test.elements[t0 + tid];
            const double *g = a.test.geom + (int64_t)e * BEM_GEOM_STRIDE;
            T_.elem[tid] = e;
            for (int i = 0; i < 3; ++i) T_.vid[tid][i] = a.test.vids[4 * e + i];
            T_.mult[tid] = a.test.mult[e];
            for (int p = 0; p < kQ; ++p) {
                for (int d = 0; d < 3; ++d) T_.xp[tid][p][d] = g[d] + g[3 + d] * a.rule.x[p] + g[6 + d] * a.rule.y[p];
                T_.wp[tid][p] = a.rule.w[p] * g[12];
            }
        }
        __syncthreads();

        int f = 0, fv0 = -1, fv1 = -1, fv2 = -1;
        double mf = 0.0, yq[kQ][3], wq[kQ];
        if (valid) {
            f = a.trial.elements[jf];
            const double *gf = a.trial.geom + (int64_t)f * BEM_GEOM_STRIDE;
            fv0 = a.trial.vids[4 * f];
            fv1 = a.trial.vids[4 * f + 1];
            fv2 = a.trial.vids[4 * f + 2];
            mf = a.trial.mult[f];
#pragma unroll
            for (int q = 0; q < kQ; ++q) {
                for (int d = 0; d < 3; ++d) yq[q][d] = gf[d] + gf[3 + d] * a.rule.x[q] + gf[6 + d] * a.rule.y[q];
                wq[q] = a.rule.w[q] * gf[12] * kInv4Pi;
            }
        } else {
#pragma unroll
            for (int q = 0; q < kQ; ++q) {
                yq[q][0] = yq[q][1] = yq[q][2] = 0.0;
                wq[q] = 0.0;
            }
        }
        const int64_t col = valid ? a.trial.l2g[f] : 0;
        for (int t = 0; t < ntile; ++t) {
            const int e = T_.elem[t];
            bool skip = !valid;
            if (valid && a.identical)
                skip = T_.vid[t][0] == fv0 || T_.vid[t][0] == fv1 || T_.vid[t][0] == fv2 || T_.vid[t][1] == fv0 ||
                       T_.vid[t][1] == fv1 || T_.vid[t][1] == fv2 || T_.vid[t][2] == fv0 || T_.vid[t][2] == fv1 ||
                       T_.vid[t][2] == fv2;
            const int64_t base = valid ? (int64_t)a.test.l2g[e] * a.ld + col : 0;
            const double m = T_.mult[t] * mf;
            for (int km = 0; km < a.nk; ++km) {
                const double k = a.k[km];
                double sr = 0.0, si = 0.0;
                if (!skip) {
#pragma unroll 1
                    for (int p = 0; p < kQ; ++p) {
                        const double wp = T_.wp[t][p];
#pragma unroll
                        for (int q = 0; q < kQ; ++q) {
                            const double dx = yq[q][0] - T_.xp[t][p][0], dy = yq[q][1] - T_.xp[t][p][1],
                                         dz = yq[q][2] - T_.xp[t][p][2];
                            const double r = sqrt(dx * dx + dy * dy + dz * dz);
                            double s, c;
                            sincos_fast(k * r, &s, &c);
                            const double g = wq[q] * wp / r;
                            sr += c * g;
                            si += s * g;
                        }
                    }
                }
                if (valid) {
                    double2 v;
                    v.x = sr * m;
                    v.y = si * m;
                    *reinterpret_cast<double2 *>(reinterpret_cast<double *>(a.out) + 2 * (km * a.kstride + base)) = v;
                }
            }
        }
    }
}

// ------------------------------------------------------------------ FP64 near-field pairs (mixed mode)
__device__ __forceinline__ void add_c(void *out, int64_t idx, float re, float im) {
    float *o = reinterpret_cast<float *>(out) + 2 * idx;
    atomic_add(o, re);
    atomic_add(o + 1, im);
}

__device__ __forceinline__ void add_c(void *out, int64_t idx, double re, double im) {
    double *o = reinterpret_cast<double *>(out) + 2 * idx;
    atomic_add(o, re);
    atomic_add(o + 1, im);
}

struct NearPair {
    const double *gt, *gf;
    double r0;
    double m;
    int64_t base;
};

// visit the 36 point pairs of a near pair: f(amplitude w w J J / (4 pi r), r)
template <typename F>
__device__ __forceinline__ void near_points(const Args &a, const NearPair &P, F &&f) {
    const double *gt = P.gt, *gf = P.gf;
#pragma unroll 1
    for (int p = 0; p < kQ; ++p) {
        const double xi = a.rule.x[p], eta = a.rule.y[p];
        const double x0 = gt[0] + gt[3] * xi + gt[6] * eta, x1 = gt[1] + gt[4] * xi + gt[7] * eta,
                     x2 = gt[2] + gt[5] * xi + gt[8] * eta;
        const double wp = a.rule.w[p] * gt[12] * gf[12] * kInv4Pi;
#pragma unroll
        for (int q = 0; q < kQ; ++q) {
            const double s0 = a.rule.x[q], s1 = a.rule.y[q];
            const double dx = gf[0] + gf[3] * s0 + gf[6] * s1 - x0, dy = gf[1] + gf[4] * s0 + gf[7] * s1 - x1,
                         dz = gf[2] + gf[5] * s0 + gf[8] * s1 - x2;
            const double r = sqrt(dx * dx + dy * dy + dz * dz);
            f(wp * a.rule.w[q] / r, r);
        }
    }
}

template <int NM>
__device__ __forceinline__ void near_pair_moments(const Args &a, const NearPair &P, bool skip) {
    double c[NM + 1];
#pragma unroll
    for (int n = 0; n <= NM; ++n) c[n] = 0.0;
    if (skip) return;
    near_points(a, P, [&](double amp, double r) {
        const double dl = r - P.r0;
        double pw = amp;
#pragma unroll
        for (int n = 0; n <= NM; ++n) {
            c[n] += pw;
            if (n < NM) pw *= dl;
        }
    });
    scale_factorials<NM>(c, P.m);
    for (int km = 0; km < a.nk; ++km) {
        const double k = a.k[km];
        double re, im, s, cs;
        taylor_eval<NM>(c, k, re, im);
        sincos_fast(k * P.r0, &s, &cs);
        add_c(a.out, km * a.kstride + P.base, (float)(cs * re - s * im), (float)(cs * im + s * re));
    }
}

__device__ __forceinline__ void near_pair_direct(const Args &a, const NearPair &P, bool skip) {
    if (skip) return;
    for (int km = 0; km < a.nk; ++km) {
        const double k = a.k[km];
        double sr = 0.0, si = 0.0;
        near_points(a, P, [&](double amp, double r) {
            double s, c;
            sincos_fast(k * r, &s, &c);
            sr += c * amp;
            si += s * amp;
        });
        add_c(a.out, km * a.kstride + P.base, (float)(sr * P.m), (float)(si * P.m));
    }
}

// max distance of the 6 quadrature points from the centroid
__device__ __forceinline__ double quad_radius(const double *g, const RuleArgs &rule) {
    double r2 = 0.0;
    for (int p = 0; p < kQ; ++p) {
        const double xi = rule.x[p] - 1.0 / 3.0, eta = rule.y[p] - 1.0 / 3.0;
        const double o0 = g[3] * xi + g[6] * eta, o1 = g[4] * xi + g[7] * eta, o2 = g[5] * xi + g[8] * eta;
        r2 = fmax(r2, o0 * o0 + o1 * o1 + o2 * o2);
    }
    return sqrt(r2);
}

__global__ __launch_bounds__(256) void kbatch_near_kernel(const Args a, int n_pairs, const int32_t *pairs) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    // no early return: the expansion order is reduced over the wavefront
    bool skip = p >= n_pairs;
    NearPair P;
    P.gt = a.test.geom;
    P.gf = a.trial.geom;
    P.r0 = 1.0;
    P.m = 0.0;
    P.base = 0;
    double x = 0.0;
    if (!skip) {
        const int e = pairs[2 * (int64_t)p], f = pairs[2 * (int64_t)p + 1];
        P.gt = a.test.geom + (int64_t)e * BEM_GEOM_STRIDE;
        P.gf = a.trial.geom + (int64_t)f * BEM_GEOM_STRIDE;
        if (a.identical) {
            for (int i = 0; i < 3; ++i)
                for (int j = 0; j < 3; ++j)
                    if (a.test.vids[4 * e + i] == a.trial.vids[4 * f + j]) skip = true;
        }
        double ct[3], cf[3], dist2;
        centroid_rn(P.gt, ct);
        centroid_rn(P.gf, cf);
        if (!skip) skip = !pair_is_near(ct, cf, P.gt[13], P.gf[13], a.near_factor, &dist2);
        if (!skip) {
            const double d0 = cf[0] - ct[0], d1 = cf[1] - ct[1], d2 = cf[2] - ct[2];
            P.r0 = sqrt(d0 * d0 + d1 * d1 + d2 * d2);
            P.m = a.test.mult[e] * a.trial.mult[f];
            P.base = (int64_t)a.test.l2g[e] * a.ld + a.trial.l2g[f];
            x = a.kmax * (quad_radius(P.gt, a.rule) + quad_radius(P.gf, a.rule)) * 1.0001;
        }
    }
    const int N = wave_max64(skip ? 0 : taylor_order(x, kTolMoments, 32));
    const int bucket = __builtin_amdgcn_readfirstlane(N <= 8 ? 0 : N <= 16 ? 1 : N <= 24 ? 2 : N <= 32 ? 3 : 4);
    switch (bucket) {
    case 0: near_pair_moments<8>(a, P, skip); break;
    case 1: near_pair_moments<16>(a, P, skip); break;
    case 2: near_pair_moments<24>(a, P, skip); break;
    case 3: near_pair_moments<32>(a, P, skip); break;
    default: near_pair_direct(a, P, skip); break;
    }
}

// ------------------------------------------------------------------ FP64 singular pairs
struct SingularPair {
    const double *gt, *gf;
    const double *tp, *sp, *w;  // this pair's points and weights
    int np;
    double scale;  // J_e J_f / (4 pi) times the multipliers
    int64_t base;
};

// visit this lane's points: f(weight / r, r)
template <typename F>
__device__ __forceinline__ void singular_points(const SingularPair &P, int lane, F &&f) {
    const double *gt = P.gt, *gf = P.gf;
    for (int kk = lane; kk < P.np; kk += 64) {
        const double xi0 = P.tp[2 * kk], xi1 = P.tp[2 * kk + 1];
        const double et0 = P.sp[2 * kk], et1 = P.sp[2 * kk + 1];
        const double dx = (gf[0] + gf[3] * et0 + gf[6] * et1) - (gt[0] + gt[3] * xi0 + gt[6] * xi1);
        const double dy = (gf[1] + gf[4] * et0 + gf[7] * et1) - (gt[1] + gt[4] * xi0 + gt[7] * xi1);
        const double dz = (gf[2] + gf[5] * et0 + gf[8] * et1) - (gt[2] + gt[5] * xi0 + gt[8] * xi1);
        const double r = sqrt(dx * dx + dy * dy + dz * dz);
        f(P.w[kk] / r, r);
    }
}

// mixed mode: moments about rc = (diam_e + diam_f) / 2, the centre of the distance range
template <int NM, typename OutT>
__device__ __forceinline__ void singular_pair_moments(const Args &a, const SingularPair &P, int lane, double rc) {
    double c[NM + 1];
#pragma unroll
    for (int n = 0; n <= NM; ++n) c[n] = 0.0;
    singular_points(P, lane, [&](double g, double r) {
        const double dl = r - rc;
        double pw = g;
#pragma unroll
        for (int n = 0; n <= NM; ++n) {
            c[n] += pw;
            if (n < NM) pw *= dl;
        }
    });
#pragma unroll
    for (int n = 0; n <= NM; ++n) c[n] = wave_sum64(c[n]);
    scale_factorials<NM>(c, P.scale);
    // lane m evaluates wavenumber m
    for (int km = lane; km < a.nk; km += 64) {
        const double k = a.k[km];
        double re, im, s, cs;
        taylor_eval<NM>(c, k, re, im);
        sincos_fast(k * rc, &s, &cs);
        add_c(a.out, km * a.kstride + P.base, (OutT)(cs * re - s * im), (OutT)(cs * im + s * re));
    }
}

template <typename OutT>
__device__ __forceinline__ void singular_pair_direct(const Args &a, const SingularPair &P, int lane) {
    for (int km = 0; km < a.nk; ++km) {
        const double k = a.k[km];
        double sr = 0.0, si = 0.0;
        singular_points(P, lane, [&](double g, double r) {
            double s, c;
            sincos_fast(k * r, &s, &c);
            sr += c * g;
            si += s * g;
        });
        sr = wave_sum64(sr);
        si = wave_sum64(si);
        if (lane == 0) add_c(a.out, km * a.kstride + P.base, (OutT)(sr * P.scale), (OutT)(si * P.scale));
    }
}

// one wavefront per pair; MOMENTS: mixed-mode expansion, otherwise direct FP64 evaluation
template <bool MOMENTS, typename OutT>
__global__ __launch_bounds__(256) void kbatch_singular_kernel(const Args a, int n_pairs, const int32_t *pd_all,
                                                              const double *tpts, const double *spts,
                                                              const double *wts) {
    const int lane = threadIdx.x & 63;
    const int pair = blockIdx.x * 4 + (threadIdx.x >> 6);
    if (pair >= n_pairs) return;  // uniform per wavefront
    const int32_t *pd = pd_all + 6 * (int64_t)pair;
    const int e = pd[0], f = pd[1];
    SingularPair P;
    P.gt = a.test.geom + (int64_t)e * BEM_GEOM_STRIDE;
    P.gf = a.trial.geom + (int64_t)f * BEM_GEOM_STRIDE;
    P.tp = tpts + 2 * (int64_t)pd[2];
    P.sp = spts + 2 * (int64_t)pd[3];
    P.w = wts + pd[4];
    P.np = pd[5];
    P.scale = P.gt[12] * P.gf[12] * kInv4Pi * a.test.mult[e] * a.trial.mult[f];
    P.base = (int64_t)a.test.l2g[e] * a.ld + a.trial.l2g[f];
    if constexpr (MOMENTS) {
        // any two points of triangles sharing a vertex are at most diam_e + diam_f apart
        const double rc = 0.5 * (P.gt[13] + P.gf[13]);
        const int N = taylor_order(a.kmax * rc * 1.0001, kTolMoments, 32);
        const int bucket = __builtin_amdgcn_readfirstlane(N <= 8 ? 0 : N <= 16 ? 1 : N <= 24 ? 2 : N <= 32 ? 3 : 4);
        switch (bucket) {
        case 0: singular_pair_moments<8, OutT>(a, P, lane, rc); break;
        case 1: singular_pair_moments<16, OutT>(a, P, lane, rc); break;
        case 2: singular_pair_moments<24, OutT>(a, P, lane, rc); break;
        case 3: singular_pair_moments<32, OutT>(a, P, lane, rc); break;
        default: singular_pair_direct<OutT>(a, P, lane); break;
        }
    } else {
        singular_pair_direct<OutT>(a, P, lane);
    }
}

}  // namespace kb
}  // namespace bem

using namespace bem;

extern "C" int bem_kbatch_assemble(const bem_space_t *test, const bem_space_t *trial, int grids_identical,
                                   const bem_quad_rule_t *rule, int nk, const double *wavenumbers, int precision,
                                   int parts, double mixed_near_factor, const bem_kbatch_config_t *config, void *out,
                                   int64_t ld, int64_t matrix_stride, int out_dtype, int n_near,
                                   const int32_t *near_pairs, int n_singular, const int32_t *singular_pairs,
                                   const double *test_points, const double *trial_points, const double *weights,
                                   void *stream) {
    clear_error();
    if (!test || !trial || !rule || !out || !wavenumbers || !config) {
        set_error("null argument");
        return BEM_ERR_INVALID_ARGUMENT;
    }
    if (nk < 1 || nk > kb::kKMax) {
        set_error("number of wavenumbers must be 1..%d", kb::kKMax);
        return BEM_ERR_INVALID_ARGUMENT;
    }
    int rc;
    if ((rc = check_space(test, "test")) || (rc = check_space(trial, "trial"))) return rc;
    if (test->shapeset != BEM_SHAPE_P0 || trial->shapeset != BEM_SHAPE_P0) {
        set_error("the frequency-batched assembler supports DP0 x DP0 spaces");
        return BEM_ERR_UNSUPPORTED;
    }
    if (rule->n_points != kb::kQ) {
        set_error("the frequency-batched assembler uses the 6-point (order 4) regular rule");
        return BEM_ERR_UNSUPPORTED;
    }
    if (precision != BEM_PRECISION_FP64 && precision != BEM_PRECISION_MIXED) {
        set_error("unknown precision mode %d", precision);
        return BEM_ERR_INVALID_ARGUMENT;
    }
    const bool mixed = precision == BEM_PRECISION_MIXED;
    if ((mixed && out_dtype != BEM_C64) || (!mixed && out_dtype != BEM_C128)) {
        set_error("output must be complex64 in mixed mode and complex128 in FP64 mode");
        return BEM_ERR_INVALID_ARGUMENT;
    }
    if (n_near < 0 || n_singular < 0 || (n_near > 0 && !near_pairs) ||
        (n_singular > 0 && (!singular_pairs || !test_points || !trial_points || !weights))) {
        set_error("missing near-field or singular pair arrays");
        return BEM_ERR_INVALID_ARGUMENT;
    }
    if (config->test_tile != 8 && config->test_tile != 16 && config->test_tile != 32) {
        set_error("test_tile must be 8, 16 or 32");
        return BEM_ERR_INVALID_ARGUMENT;
    }
    if (config->workgroups_per_cu < 1 || config->workgroups_per_cu > 16) {
        set_error("workgroups_per_cu must be 1..16");
        return BEM_ERR_INVALID_ARGUMENT;
    }
    if (ld < 1 || matrix_stride < 1) {
        set_error("ld and matrix_stride must be positive");
        return BEM_ERR_INVALID_ARGUMENT;
    }
    kb::Args a;
    a.test = to_args(test);
    a.trial = to_args(trial);
    a.rule = to_args(rule);
    a.nk = nk;
    a.kmax = 0.0;
    for (int i = 0; i < kb::kKMax; ++i) {
        a.k[i] = i < nk ? wavenumbers[i] : 0.0;
        if (i < nk) {
            if (!(fabs(wavenumbers[i]) < 1e300)) {
                set_error("wavenumbers must be finite");
                return BEM_ERR_INVALID_ARGUMENT;
            }
            a.kmax = fmax(a.kmax, fabs(wavenumbers[i]));
        }
    }
    a.near_factor = mixed_near_factor;
    a.out = out;
    a.ld = ld;
    a.kstride = matrix_stride;
    a.identical = grids_identical;
    const int TM = config->test_tile;
    a.ntx = (a.trial.n + kRegThreads - 1) / kRegThreads;
    a.nty = (a.test.n + TM - 1) / TM;
    hipStream_t s = (hipStream_t)stream;
    const long ntiles = (long)a.ntx * a.nty;
    if ((parts & BEM_KBATCH_REGULAR) && ntiles > 0) {
        int dev = 0, ncu = 0;
        if (hipGetDevice(&dev) != hipSuccess ||
            hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess) {
            set_error("cannot query the device");
            return BEM_ERR_DEVICE;
        }
        const long want = (long)ncu * config->workgroups_per_cu;
        const dim3 grid((unsigned)(want < ntiles ? want : ntiles));
#define KB_LAUNCH(TMV)                                                                                 \
    do {                                                                                               \
        if (mixed) hipLaunchKernelGGL((kb::kbatch_far_kernel<TMV>), grid, dim3(kRegThreads), 0, s, a); \
        else hipLaunchKernelGGL((kb::kbatch_fp64_kernel<TMV>), grid, dim3(kRegThreads), 0, s, a);      \
    } while (0)
        if (TM == 8) KB_LAUNCH(8);
        else if (TM == 16) KB_LAUNCH(16);
        else KB_LAUNCH(32);
#undef KB_LAUNCH
        if ((rc = check_launch("kbatch_regular_kernel"))) return rc;
    }
    if ((parts & BEM_KBATCH_NEAR) && mixed && n_near > 0) {
        hipLaunchKernelGGL(kb::kbatch_near_kernel, dim3((n_near + 255) / 256), dim3(256), 0, s, a, n_near, near_pairs);
        if ((rc = check_launch("kbatch_near_kernel"))) return rc;
    }
    if ((parts & BEM_KBATCH_SINGULAR) && n_singular > 0) {
        const dim3 grid((n_singular + 3) / 4);
        if (mixed)
            hipLaunchKernelGGL((kb::kbatch_singular_kernel<true, float>), grid, dim3(256), 0, s, a, n_singular,
                               singular_pairs, test_points, trial_points, weights);
        else
            hipLaunchKernelGGL((kb::kbatch_singular_kernel<false, double>), grid, dim3(256), 0, s, a, n_singular,
                               singular_pairs, test_points, trial_points, weights);
        if ((rc = check_launch("kbatch_singular_kernel"))) return rc;
    }
    return BEM_OK;
}
