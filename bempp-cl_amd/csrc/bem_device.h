// Device-side helpers shared by the dense BEM kernels (gfx950).
//
// Green's functions follow bempp_cl/core/numba_kernels.py:176-398 (regular) and
// 453-551 (singular); they are the textbook Laplace/Helmholtz kernels
//   SL  G = e^{ikr} / (4 pi r)
//   DL  (ikr - 1) e^{ikr} (y - x).n_y / (4 pi r^3)       (Laplace: k = 0)
//   ADL (ikr - 1) e^{ikr} (x - y).n_x / (4 pi r^3)
// with complex k = kr + i ki entering as e^{ikr} = e^{i kr r} e^{-ki r}.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <type_traits>

#include "bempp_hip.h"

namespace bem {

constexpr double kInv4Pi = 0.079577471545947667884;

enum Op {
    kLaplaceSL = BEM_LAPLACE_SINGLE_LAYER,
    kLaplaceDL = BEM_LAPLACE_DOUBLE_LAYER,
    kLaplaceADL = BEM_LAPLACE_ADJOINT_DOUBLE_LAYER,
    kHelmholtzSL = BEM_HELMHOLTZ_SINGLE_LAYER,
    kHelmholtzDL = BEM_HELMHOLTZ_DOUBLE_LAYER,
    kHelmholtzADL = BEM_HELMHOLTZ_ADJOINT_DOUBLE_LAYER,
    kMaxwellEFIE = BEM_MAXWELL_ELECTRIC_FIELD,
    kLaplaceHS = BEM_LAPLACE_HYPERSINGULAR,
    kHelmholtzHS = BEM_HELMHOLTZ_HYPERSINGULAR,
    kMaxwellMFIE = BEM_MAXWELL_MAGNETIC_FIELD,
};

// operators on RWG/SNC spaces (edge dofs: at most two elements per dof)
template <int OP>
constexpr bool is_rwg_op() {
    return OP == kMaxwellEFIE || OP == kMaxwellMFIE;
}

// hypersingular operators: moments of their single layer, combined with surface curls in the epilogue
template <int OP>
constexpr bool is_hypersingular() {
    return OP == kLaplaceHS || OP == kHelmholtzHS;
}

// the operator whose Green's function the point-pair sums of OP evaluate
template <int OP>
constexpr int green_op() {
    return OP == kLaplaceHS ? kLaplaceSL : OP == kHelmholtzHS ? kHelmholtzSL : OP;
}

template <int OP>
struct OpTraits {
    static constexpr bool is_complex = OP >= kHelmholtzSL && OP != kLaplaceHS;
    static constexpr bool needs_trial_normal = OP == kLaplaceDL || OP == kHelmholtzDL;
    static constexpr bool needs_test_normal = OP == kLaplaceADL || OP == kHelmholtzADL;
};

// 1 / sqrt(x) for positive normal x: the hardware estimate (v_rsq_f64) refined by two Newton steps
// y <- y + y (1 - x y^2) / 2; about 1e-16 relative error (checked against long double on the host),
// a third of the instructions of 1.0 / sqrt(x).
__device__ __forceinline__ double rsqrt_t(double x) {
    double y = __builtin_amdgcn_rsq(x);
    double e = fma(-x * y, y, 1.0);
    y = fma(0.5 * y, e, y);
    e = fma(-x * y, y, 1.0);
    return fma(0.5 * y, e, y);
}
__device__ __forceinline__ float rsqrt_t(float x) { return __builtin_amdgcn_rsqf(x); }

// sin and cos of |x| <= 1 from their Taylor series through x^17 and x^18 (remainders
// |x|^19/19! and |x|^20/20!, below 1e-17), with no range reduction.
__device__ __forceinline__ void sincos_small(double x, double *s, double *c) {
    const double x2 = x * x;
    double ps = 1.0 / 355687428096000.0;  // 1/17!
    ps = ps * x2 - 1.0 / 1307674368000.0;  // 1/15!
    ps = ps * x2 + 1.0 / 6227020800.0;     // 1/13!
    ps = ps * x2 - 1.0 / 39916800.0;       // 1/11!
    ps = ps * x2 + 1.0 / 362880.0;         // 1/9!
    ps = ps * x2 - 1.0 / 5040.0;
    ps = ps * x2 + 1.0 / 120.0;
    ps = ps * x2 - 1.0 / 6.0;
    *s = x + x * x2 * ps;
    double pc = 1.0 / 6402373705728000.0;  // 1/18!
    pc = pc * x2 - 1.0 / 20922789888000.0;  // 1/16!
    pc = pc * x2 + 1.0 / 87178291200.0;     // 1/14!
    pc = pc * x2 - 1.0 / 479001600.0;       // 1/12!
    pc = pc * x2 + 1.0 / 3628800.0;         // 1/10!
    pc = pc * x2 - 1.0 / 40320.0;
    pc = pc * x2 + 1.0 / 720.0;
    pc = pc * x2 - 1.0 / 24.0;
    pc = pc * x2 + 0.5;
    *c = 1.0 - x2 * pc;
}

// sin and cos of x (radians), |x| < 2^26: x = n pi / 2 + y with n = rint(2 x / pi) and |y| <= pi / 4
// from a three-part Cody-Waite subtraction (pi / 2 split into three doubles; FMA keeps each step to
// one rounding), the Taylor polynomials of sin y through y^15 and of cos y through y^16 (remainders
// below 3e-17 at |y| = pi / 4), then the quadrant's swap and signs. Absolute error about 1.4e-16 for
// |x| <= 1e7 (checked against long double on the host), with half the instructions of the library
// sincos and no branch.
__device__ __forceinline__ void sincos_reduced(double x, double *sp, double *cp) {
    const double n = rint(x * 0.63661977236758134308);
    double y = fma(-n, 1.5707963267948966, x);
    y = fma(-n, 6.123233995736766e-17, y);
    y = fma(-n, -1.4973849048591698e-33, y);
    const double z = y * y;
    double ps = -1.0 / 1307674368000.0;  // -1/15!
    ps = fma(ps, z, 1.0 / 6227020800.0);
    ps = fma(ps, z, -1.0 / 39916800.0);
    ps = fma(ps, z, 1.0 / 362880.0);
    ps = fma(ps, z, -1.0 / 5040.0);
    ps = fma(ps, z, 1.0 / 120.0);
    ps = fma(ps, z, -1.0 / 6.0);
    const double s = fma(y * z, ps, y);
    double pc = 1.0 / 20922789888000.0;  // 1/16!
    pc = fma(pc, z, -1.0 / 87178291200.0);
    pc = fma(pc, z, 1.0 / 479001600.0);
    pc = fma(pc, z, -1.0 / 3628800.0);
    pc = fma(pc, z, 1.0 / 40320.0);
    pc = fma(pc, z, -1.0 / 720.0);
    pc = fma(pc, z, 1.0 / 24.0);
    pc = fma(pc, z, -0.5);
    const double c = fma(z, pc, 1.0);
    const int q = (int)n & 3;
    const double S = (q & 1) ? c : s, C = (q & 1) ? s : c;
    *sp = (q & 2) ? -S : S;
    *cp = ((q + 1) & 2) ? -C : C;
}

// the library sincos out of line: its large-argument reduction needs many registers, which an
// inlined copy would claim in every kernel that merely may take it
__device__ __noinline__ void sincos_large(double x, double *sp, double *cp) { sincos(x, sp, cp); }

// sincos_reduced, with the library routine for |x| >= 2^26
__device__ __forceinline__ void sincos_fast(double x, double *sp, double *cp) {
    if (!(fabs(x) < 67108864.0)) {
        sincos_large(x, sp, cp);
        return;
    }
    sincos_reduced(x, sp, cp);
}

// Phase modes of the FP64 Green's functions (template argument PH): kPhSmall, the caller
// guarantees |k r| <= 1 (Taylor series, sincos_small); kPhReduced, the caller guarantees
// |k r| < 2^26 (sincos_reduced, no large-argument branch); otherwise sincos_fast. kPhReal may be
// or-ed in when the wavenumber is real: the e^{-Im(k) r} factor and its branch are left out.
// kPhShift: the Green's functions return e^{-i ph0} times their value, with the phase k r - ph0
// (|k r - ph0| <= 1 guaranteed by the caller) from sincos_small; the caller multiplies by e^{i ph0}.
constexpr int kPhSmall = 1, kPhReduced = 2, kPhReal = 4, kPhShift = 8;

// Wave-uniform: the wavenumber is real and |k r| < 2^25 on every point pair of every active lane's
// element pair (r <= centroid distance + both diameters), so the FP64 Green's functions may take
// the branch-free phase (kPhReduced | kPhReal), which lets the compiler interleave the point pairs.
__device__ __forceinline__ bool fp64_phase_reduced(double kr, double ki, double centroid_dist, double diam_sum) {
    return ki == 0.0 && __all(fabs(kr) * (centroid_dist + diam_sum) < 0x1p25);
}

template <int PH = 0>
__device__ __forceinline__ void sincos_t(double x, double *s, double *c) {
    if constexpr ((PH & 3) == kPhSmall) sincos_small(x, s, c);
    else if constexpr ((PH & 3) == kPhReduced) sincos_reduced(x, s, c);
    else sincos_fast(x, s, c);
}
template <int PH = 0>
__device__ __forceinline__ void sincos_t(float x, float *s, float *c) { sincosf(x, s, c); }
// sin and cos of the phase k r (k r - ph0 with kPhShift)
template <int PH, typename T>
__device__ __forceinline__ void phase_t(T kr, T r, T ph0, T *s, T *c) {
    if constexpr ((PH & kPhShift) && sizeof(T) == 8) sincos_small(fma(kr, r, -ph0), s, c);
    else sincos_t<PH>(kr * r, s, c);
}
__device__ __forceinline__ double exp_t(double x) { return exp(x); }
__device__ __forceinline__ float exp_t(float x) { return __expf(x); }

// Green's function of operator OP for d = y - x, r2 = |d|^2.
// nx: test normal (ADL), ny: trial normal (DL). Returns (re, im).
template <int OP, typename T, int PH = 0>
__device__ __forceinline__ void green(T dx, T dy, T dz, T r2, const T *nx, const T *ny, T kr, T ki, T &gre, T &gim,
                                      T ph0 = T(0)) {
    T rinv = rsqrt_t(r2);
    if constexpr (OP == kLaplaceSL) {
        gre = T(kInv4Pi) * rinv;
        gim = T(0);
    } else if constexpr (OP == kLaplaceDL) {
        T dot = dx * ny[0] + dy * ny[1] + dz * ny[2];
        gre = -T(kInv4Pi) * dot * rinv * rinv * rinv;
        gim = T(0);
    } else if constexpr (OP == kLaplaceADL) {
        T dot = dx * nx[0] + dy * nx[1] + dz * nx[2];
        gre = T(kInv4Pi) * dot * rinv * rinv * rinv;
        gim = T(0);
    } else if constexpr (OP == kHelmholtzSL || OP == kMaxwellEFIE) {
        T r = r2 * rinv;
        T s, c;
        phase_t<PH>(kr, r, ph0, &s, &c);
        T g = T(kInv4Pi) * rinv;
        if constexpr (!(PH & kPhReal)) { if (ki != T(0)) g *= exp_t(-ki * r); }
        gre = c * g;
        gim = s * g;
    } else {  // Helmholtz DL / ADL
        T dot;
        if constexpr (OP == kHelmholtzDL) dot = dx * ny[0] + dy * ny[1] + dz * ny[2];
        else dot = -(dx * nx[0] + dy * nx[1] + dz * nx[2]);
        T r = r2 * rinv;
        T s, c;
        phase_t<PH>(kr, r, ph0, &s, &c);
        T g = T(kInv4Pi) * dot * rinv * rinv * rinv;
        if constexpr (!(PH & kPhReal)) { if (ki != T(0)) g *= exp_t(-ki * r); }
        T fr = c * g, fi = s * g;
        T a = T(-1) - ki * r, b = kr * r;  // i k r - 1 = a + i b
        gre = a * fr - b * fi;
        gim = b * fr + a * fi;
    }
}

// Single and double layer of one family at the same point pair, sharing r, 1/r and e^{ikr}:
// (g1re, g1im) = SL, (g2re, g2im) = DL with the trial normal ny.
template <bool HELM, typename T, int PH = 0>
__device__ __forceinline__ void green_sl_dl(T dx, T dy, T dz, T r2, const T *ny, T kr, T ki, T &g1re, T &g1im,
                                            T &g2re, T &g2im, T ph0 = T(0)) {
    const T rinv = rsqrt_t(r2);
    const T dot = dx * ny[0] + dy * ny[1] + dz * ny[2];
    if constexpr (!HELM) {
        g1re = T(kInv4Pi) * rinv;
        g1im = T(0);
        g2re = -g1re * dot * rinv * rinv;
        g2im = T(0);
    } else {
        const T r = r2 * rinv;
        T s, c;
        phase_t<PH>(kr, r, ph0, &s, &c);
        T g = T(kInv4Pi) * rinv;
        if constexpr (!(PH & kPhReal)) { if (ki != T(0)) g *= exp_t(-ki * r); }
        g1re = c * g;
        g1im = s * g;
        const T gd = g * dot * rinv * rinv;
        const T fr = c * gd, fi = s * gd;
        const T a = (PH & kPhReal) ? T(-1) : T(-1) - ki * r, b = kr * r;
        g2re = a * fr - b * fi;
        g2im = b * fr + a * fi;
    }
}

// Single layer, double layer (normal n_y) and the double layer of the reversed pair (x on the
// trial side, y on the test side: kernel G'(r) (x - y).n_x / r), all from one distance and one
// e^{ikr}; (dx, dy, dz) = y - x. The third value is the (f, e) double-layer entry that the
// symmetric pass computes together with the (e, f) entries.
template <bool HELM, typename T, int PH = 0>
__device__ __forceinline__ void green_sl_dl_rev(T dx, T dy, T dz, T r2, const T *nx, const T *ny, T kr, T ki, T &g1re,
                                                T &g1im, T &g2re, T &g2im, T &g3re, T &g3im, T ph0 = T(0)) {
    const T rinv = rsqrt_t(r2);
    const T dot2 = dx * ny[0] + dy * ny[1] + dz * ny[2];
    const T dot3 = -(dx * nx[0] + dy * nx[1] + dz * nx[2]);
    if constexpr (!HELM) {
        g1re = T(kInv4Pi) * rinv;
        g1im = T(0);
        const T h = -g1re * rinv * rinv;
        g2re = h * dot2;
        g3re = h * dot3;
        g2im = g3im = T(0);
    } else {
        const T r = r2 * rinv;
        T s, c;
        phase_t<PH>(kr, r, ph0, &s, &c);
        T g = T(kInv4Pi) * rinv;
        if constexpr (!(PH & kPhReal)) { if (ki != T(0)) g *= exp_t(-ki * r); }
        g1re = c * g;
        g1im = s * g;
        const T gd = g * rinv * rinv;
        const T fr = c * gd, fi = s * gd;
        const T a = (PH & kPhReal) ? T(-1) : T(-1) - ki * r, b = kr * r;
        const T hr = a * fr - b * fi, hi = b * fr + a * fi;
        g2re = hr * dot2;
        g2im = hi * dot2;
        g3re = hr * dot3;
        g3im = hi * dot3;
    }
}

// Green's function of OP and, for DUAL (OP a single layer), also the double layer of the same
// family at the same point pair (shared r and e^{ikr}).
template <int OP, bool DUAL, typename T, int PH = 0>
__device__ __forceinline__ void green2(T dx, T dy, T dz, T r2, const T *nx, const T *ny, T kr, T ki, T &g1re, T &g1im,
                                       T &g2re, T &g2im, T ph0 = T(0)) {
    if constexpr (DUAL) {
        green_sl_dl<OpTraits<OP>::is_complex, T, PH>(dx, dy, dz, r2, ny, kr, ki, g1re, g1im, g2re, g2im, ph0);
    } else {
        green<OP, T, PH>(dx, dy, dz, r2, nx, ny, kr, ki, g1re, g1im, ph0);
        g2re = g2im = T(0);
    }
}

// Compensated (Kahan) accumulator; plain sum when K = false.
template <typename T, bool K>
struct Acc {
    T s, c;
    __device__ __forceinline__ void zero() { s = T(0); c = T(0); }
    __device__ __forceinline__ void add(T x) {
        if constexpr (K) {
            T y = x - c;
            T t = s + y;
            c = (t - s) - y;
            s = t;
        } else {
            s += x;
        }
    }
};

// Atomic adds to HBM (hardware float atomics on gfx950).
__device__ __forceinline__ void atomic_add(double *p, double v) { unsafeAtomicAdd(p, v); }
__device__ __forceinline__ void atomic_add(float *p, float v) { unsafeAtomicAdd(p, v); }

struct C2d {
    double re, im;
};

struct C2f {
    float re, im;
};

// Element centroid v0 + (e1 + e2) / 3 from a geometry row, with explicitly rounded
// operations so that every kernel evaluating it gets bit-identical results.
__device__ __forceinline__ void centroid_rn(const double *g, double *c) {
#pragma unroll
    for (int d = 0; d < 3; ++d) c[d] = __dadd_rn(g[d], __dmul_rn(__dadd_rn(g[3 + d], g[6 + d]), 1.0 / 3.0));
}

// Near-field test of the mixed-precision mode: |c_t - c_f| < factor * (diam_t + diam_f) / 2.
// The FP32 regular kernel skips exactly the pairs for which this returns true and the FP64
// near-field kernel computes exactly those, so it is evaluated without contraction.
__device__ __forceinline__ bool pair_is_near(const double *ct, const double *cf, double diam_t, double diam_f,
                                             double factor, double *dist2_out) {
    double d0 = __dadd_rn(cf[0], -ct[0]), d1 = __dadd_rn(cf[1], -ct[1]), d2 = __dadd_rn(cf[2], -ct[2]);
    double dist2 = __dadd_rn(__dadd_rn(__dmul_rn(d0, d0), __dmul_rn(d1, d1)), __dmul_rn(d2, d2));
    double lim = __dmul_rn(__dmul_rn(factor, 0.5), __dadd_rn(diam_t, diam_f));
    *dist2_out = dist2;
    return dist2 < __dmul_rn(lim, lim);
}

// ------------------------------------------------------------------ epilogue
// Surface curl of the P1 shape function i of a flat triangle times its Jacobian J:
// curl phi = n x grad phi gives curl phi_1 = e2 / J, curl phi_2 = -e1 / J and
// curl phi_0 = (e1 - e2) / J (e1 = v1 - v0, e2 = v2 - v0), i.e. the coefficients of (e1, e2).
__host__ __device__ constexpr double curl_coef(int i, int a) {
    return i == 0 ? (a == 0 ? 1.0 : -1.0) : i == 1 ? (a == 0 ? 0.0 : 1.0) : (a == 0 ? -1.0 : 0.0);
}

// Local block value (i, j) from the moment block S, for scalar shapesets, EFIE and the
// hypersingular operators. nmp: product of the two elements' normal multipliers (used by the
// hypersingular operators, whose curls and normals both carry them).
template <int OP, int NTS, int NSS>
__device__ __forceinline__ C2d local_value(const C2d (&S)[NTS][NSS], int i, int j, const double *gt,
                                                  const double *gf, double kr, double ki, double nmp = 1.0) {
    C2d v;
    if constexpr (is_hypersingular<OP>()) {
        // curl_i . curl_j S00 - k^2 (n_t . n_f) phi_i psi_j-moment (bempp_cl/core/numba_kernels.py:1284-1299)
        static_assert(NTS == 3 && NSS == 3, "hypersingular operators need P1 shape functions on both sides");
        const C2d sl = local_value<kLaplaceSL, NTS, NSS>(S, i, j, gt, gf, kr, ki);
        double c = 0.0;
#pragma unroll
        for (int a = 0; a < 2; ++a)
#pragma unroll
            for (int b = 0; b < 2; ++b) {
                const double m = gt[3 + 3 * a] * gf[3 + 3 * b] + gt[4 + 3 * a] * gf[4 + 3 * b] + gt[5 + 3 * a] * gf[5 + 3 * b];
                c += curl_coef(i, a) * curl_coef(j, b) * m;
            }
        c *= nmp / (gt[12] * gf[12]);
        v.re = c * S[0][0].re;
        v.im = c * S[0][0].im;
        if constexpr (OP == kHelmholtzHS) {
            const double nn = nmp * (gt[9] * gf[9] + gt[10] * gf[10] + gt[11] * gf[11]);
            const double k2r = (kr * kr - ki * ki) * nn, k2i = 2.0 * kr * ki * nn;
            v.re -= k2r * sl.re - k2i * sl.im;
            v.im -= k2r * sl.im + k2i * sl.re;
        }
    } else if constexpr (OP == kMaxwellEFIE) {
        // M = E_t^T E_f, E = [v1 - v0, v2 - v0]
        double M[2][2];
#pragma unroll
        for (int a = 0; a < 2; ++a)
#pragma unroll
            for (int b = 0; b < 2; ++b)
                M[a][b] = gt[3 + 3 * a] * gf[3 + 3 * b] + gt[4 + 3 * a] * gf[4 + 3 * b] + gt[5 + 3 * a] * gf[5 + 3 * b];
        const double ox[3] = {0.0, 1.0, 0.0}, oy[3] = {1.0, 0.0, 0.0};
        double oi[2] = {ox[i], oy[i]}, oj[2] = {ox[j], oy[j]};
        // sum (xi - o_i)^T M (eta - o_j) G'
        C2d term = {0.0, 0.0};
#pragma unroll
        for (int a = 0; a < 2; ++a)
#pragma unroll
            for (int b = 0; b < 2; ++b) {
                double mab = M[a][b];
                term.re += mab * (S[1 + a][1 + b].re - oi[a] * S[0][1 + b].re - oj[b] * S[1 + a][0].re + oi[a] * oj[b] * S[0][0].re);
                term.im += mab * (S[1 + a][1 + b].im - oi[a] * S[0][1 + b].im - oj[b] * S[1 + a][0].im + oi[a] * oj[b] * S[0][0].im);
            }
        double jj = 1.0 / (gt[12] * gf[12]);
        // -ik term / (J J) - 4 S00 / (ik J J), ik = -ki + i kr
        double ikr = -ki, iki = kr;
        double den = ikr * ikr + iki * iki;
        double inv_re = ikr / den, inv_im = -iki / den;
        double t_re = term.re * jj, t_im = term.im * jj;
        double s_re = 4.0 * S[0][0].re * jj, s_im = 4.0 * S[0][0].im * jj;
        v.re = -(ikr * t_re - iki * t_im) - (inv_re * s_re - inv_im * s_im);
        v.im = -(ikr * t_im + iki * t_re) - (inv_re * s_im + inv_im * s_re);
        // edge lengths l0 = |v1 - v0|, l1 = |v2 - v0|, l2 = |v1 - v2|
        auto elen = [](const double *g, int k) {
            double a0, a1, a2;
            if (k == 0) { a0 = g[3]; a1 = g[4]; a2 = g[5]; }
            else if (k == 1) { a0 = g[6]; a1 = g[7]; a2 = g[8]; }
            else { a0 = g[3] - g[6]; a1 = g[4] - g[7]; a2 = g[5] - g[8]; }
            return sqrt(a0 * a0 + a1 * a1 + a2 * a2);
        };
        double l = elen(gt, i) * elen(gf, j);
        v.re *= l;
        v.im *= l;
    } else {
        // phi_i = sum_a C[a][i] u_a with phi = (u0 - u1 - u2, u1, u2); DP0 uses u0 only
        auto coef = [](int a, int i) -> double {
            if (i == 0) return a == 0 ? 1.0 : -1.0;
            return a == i ? 1.0 : 0.0;
        };
        v.re = v.im = 0.0;
#pragma unroll
        for (int a = 0; a < NTS; ++a)
#pragma unroll
            for (int b = 0; b < NSS; ++b) {
                double c = (NTS == 1 ? 1.0 : coef(a, i)) * (NSS == 1 ? 1.0 : coef(b, j));
                v.re += c * S[a][b].re;
                v.im += c * S[a][b].im;
            }
    }
    return v;
}

}  // namespace bem
