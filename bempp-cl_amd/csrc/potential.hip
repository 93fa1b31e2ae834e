// Dense potential evaluation for gfx950.
//
//   u(x_m) = sum_e sum_q G(x_m, y_eq) c_eq,  c_eq = J_e w_q sum_i phi_i(q) x_{e,i}
// (bempp_cl/core/numba_kernels.py:1788-1849) and the EFIE potential
//   E(x_m) = sum_eq G(x_m, y_eq) [ ik t1_eq - (x_m - y_eq) (ik r - 1) t2_eq / (ik r^2) ]
// (numba_kernels.py:2281-2359). Three launches:
//   1. prep: one lane per (support element, quadrature point) writes the source position,
//      normal and source weights (c, or t1 and t2) to the workspace;
//   2. eval: a 256-lane workgroup owns 256 targets and one chunk of sources; sources are
//      staged through LDS in tiles of 256 and read as broadcasts; partial sums per chunk go
//      to the workspace;
//   3. reduce: sums the chunk partials in a fixed order (deterministic result).
// Far fields (numba_kernels.py:402-449, 2437-2570): the targets are unit directions xh and
// the kernel is K = e^{-i Re(k) xh.y} / (4 pi) (the reference puts Re k in the exponent); they
// are evaluated in FP64 (far_field_eval_fp64) or, MIXED, with FP32 offsets within groups of
// sources (far_field_eval_mixed).
// MIXED: per (target, tile) the distance r0 = |x - c| to the tile centre c and e^{ik r0}
// are formed in FP64; per source only r - r0 = u.(2v + u) / (r + r0), with v = x - c and
// u = c - y ... small vectors, is needed, so the FP32 phase error stays ~k * eps * tile size
// instead of k * eps * r.
#include "bem_device.h"
#include "bem_internal.h"

#include <mutex>
#include <type_traits>

namespace bem {

constexpr int kPotThreads = 256;
constexpr int kPotTile = 256;
constexpr int kSrcFields = 14;  // px py pz nx ny nz c_re c_im t1 (6): EFIE uses c as t2

// Taylor coefficients of e^{ik delta} - 1 in powers of delta for a real wavenumber k (the EFIE
// potential's series, see efie_em1).
struct EfieSeries {
    float c[5];  // (-1)^n k^2n / (2n)!, n = 1..5
    float s[4];  // (-1)^n k^(2n+1) / (2n+1)!, n = 1..4
    float k;
};

static EfieSeries efie_series(double k) {
    EfieSeries e;
    double cn = 1.0, sn = k;
    for (int n = 1; n <= 5; ++n) {
        cn *= -k * k / ((2.0 * n - 1.0) * (2.0 * n));
        e.c[n - 1] = (float)cn;
        if (n <= 4) {
            sn *= -k * k / ((2.0 * n) * (2.0 * n + 1.0));
            e.s[n - 1] = (float)sn;
        }
    }
    e.k = (float)k;
    return e;
}

struct PotArgs {
    SpaceArgs sp;
    RuleArgsWide rule;  // read by the prep kernel only
    const double *coeffs;  // complex128 per (element, shape)
    int npts;
    const double *pts;
    double kr, ki;
    double *src;  // [kSrcFields][nsrc_pad]
    int nsrc, nsrc_pad;
    double *partial;  // [nchunk][kdim][npts][2]
    int nchunk, chunk_len;
    void *out;
    EfieSeries ser;  // series of e^{i Re(k) delta} - 1 (EFIE potential)
};

template <int OP>
__global__ __launch_bounds__(256) void potential_prep_kernel(const PotArgs a) {
    int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= a.nsrc_pad) return;
    double *src = a.src;
    const int P = a.nsrc_pad;
    if (s >= a.nsrc) {  // padding sources contribute nothing
        for (int k = 0; k < kSrcFields; ++k) src[k * P + s] = 0.0;
        src[0 * P + s] = 1e30;  // far away: G ~ 0 and no division by zero
        return;
    }
    const int nq = a.rule.n;
    int ei = s / nq, q = s % nq;
    int e = a.sp.elements[ei];
    const double *g = a.sp.geom + (int64_t)e * BEM_GEOM_STRIDE;
    double xi = a.rule.x[q], eta = a.rule.y[q];
    for (int d = 0; d < 3; ++d) src[d * P + s] = g[d] + g[3 + d] * xi + g[6 + d] * eta;
    double nm = a.sp.nmult ? (double)a.sp.nmult[e] : 1.0;
    for (int d = 0; d < 3; ++d) src[(3 + d) * P + s] = g[9 + d] * nm;
    const double *c = a.coeffs;
    if constexpr (OP == kMaxwellEFIE) {
        double t2r = 0, t2i = 0, t1r[3] = {0, 0, 0}, t1i[3] = {0, 0, 0};
        const double ox[3] = {0.0, 1.0, 0.0}, oy[3] = {1.0, 0.0, 0.0};
        for (int i = 0; i < 3; ++i) {
            double a0, a1, a2;
            if (i == 0) { a0 = g[3]; a1 = g[4]; a2 = g[5]; }
            else if (i == 1) { a0 = g[6]; a1 = g[7]; a2 = g[8]; }
            else { a0 = g[3] - g[6]; a1 = g[4] - g[7]; a2 = g[5] - g[8]; }
            double len = sqrt(a0 * a0 + a1 * a1 + a2 * a2);
            double fr = a.rule.w[q] * len * c[2 * (3 * (int64_t)e + i)];
            double fi = a.rule.w[q] * len * c[2 * (3 * (int64_t)e + i) + 1];
            // Piola-mapped RWG value times J: E (xi - o_i)
            double r0 = xi - ox[i], r1 = eta - oy[i];
            for (int d = 0; d < 3; ++d) {
                double b = g[3 + d] * r0 + g[6 + d] * r1;
                t1r[d] += fr * b;
                t1i[d] += fi * b;
            }
            t2r += 2.0 * fr;
            t2i += 2.0 * fi;
        }
        src[6 * P + s] = t2r;
        src[7 * P + s] = t2i;
        for (int d = 0; d < 3; ++d) {
            src[(8 + 2 * d) * P + s] = t1r[d];
            src[(9 + 2 * d) * P + s] = t1i[d];
        }
    } else {
        int nsh = a.sp.shape == BEM_SHAPE_P0 ? 1 : 3;
        double phi[3] = {1.0 - xi - eta, xi, eta};
        if (nsh == 1) phi[0] = 1.0;
        double cr = 0, ci = 0;
        for (int i = 0; i < nsh; ++i) {
            double f = g[12] * a.rule.w[q] * phi[i];
            cr += f * c[2 * ((int64_t)e * nsh + i)];
            ci += f * c[2 * ((int64_t)e * nsh + i) + 1];
        }
        src[6 * P + s] = cr;
        src[7 * P + s] = ci;
    }
}

// Largest |y|^2 over the staged sources of a tile (0 for lanes without a source): per-wave
// maximum by shuffles into wmax[wave]; the caller's barrier publishes it.
__device__ __forceinline__ void tile_extent_stage(double y2, double *wmax) {
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) y2 = fmax(y2, __shfl_xor(y2, off, 64));
    if ((threadIdx.x & 63) == 0) wmax[threadIdx.x >> 6] = y2;
}

// Wave-uniform: a real wavenumber and |k| (|x| + max |y|) < 2^25 for every active lane's target x
// over the tile's sources, so every phase |k r| (r <= |x| + |y|) or |k xh.y| takes the branch-free
// reduction (kPhReduced | kPhReal).
__device__ __forceinline__ bool tile_phase_reduced(double kr, double ki, double xnorm, const double *wmax) {
    double y2 = wmax[0];
#pragma unroll
    for (int w = 1; w < kPotThreads / 64; ++w) y2 = fmax(y2, wmax[w]);
    return ki == 0.0 && __all(fabs(kr) * (xnorm + sqrt(y2)) < 0x1p25);
}

// One source's FP64 contribution to the potential at x (the body of potential_eval_fp64, also the
// mixed kernels' path for targets close to a tile's sources). field(k) returns field k of the
// source record (px py pz nx ny nz, c or t2, EFIE t1). EFIE: term_d = ik t1_d - (x - y)_d
// (ik r - 1) t2 / (ik r^2) times G, with (ik r - 1) / (ik r^2) = 1/r - alpha / r^2, alpha = 1 / (ik).
template <int OP, int PH, typename F, int KD>
__device__ __forceinline__ void fp64_source(F field, double x0, double x1, double x2, double kr, double ki, double alr,
                                            double ali, double (&accr)[KD], double (&acci)[KD]) {
    const double zero3[3] = {0.0, 0.0, 0.0};
    const double dx = field(0) - x0, dy = field(1) - x1, dz = field(2) - x2;
    const double r2 = dx * dx + dy * dy + dz * dz;
    const double ny[3] = {field(3), field(4), field(5)};
    double gre, gim;
    green<OP == kMaxwellEFIE ? kHelmholtzSL : OP, double, PH>(dx, dy, dz, r2, zero3, ny, kr, ki, gre, gim);
    if constexpr (OP == kMaxwellEFIE) {
        const double rinv = rsqrt_t(r2), rinv2 = rinv * rinv;
        const double qr = rinv - alr * rinv2, qi = -ali * rinv2;  // (ik r - 1) / (ik r^2)
        const double t2r = field(6), t2i = field(7);
        const double qtr = qr * t2r - qi * t2i, qti = qr * t2i + qi * t2r;
        const double diffv[3] = {-dx, -dy, -dz};
        const double ikr_re = -ki, ikr_im = kr;  // ik
#pragma unroll
        for (int d = 0; d < 3; ++d) {
            const double t1r = field(8 + 2 * d), t1i = field(9 + 2 * d);
            const double tr = (ikr_re * t1r - ikr_im * t1i) - diffv[d] * qtr;
            const double ti = (ikr_re * t1i + ikr_im * t1r) - diffv[d] * qti;
            accr[d] += gre * tr - gim * ti;
            acci[d] += gre * ti + gim * tr;
        }
    } else {
        const double cr = field(6), ci = field(7);
        accr[0] += gre * cr - gim * ci;
        acci[0] += gre * ci + gim * cr;
    }
}

// One source's FP64 contribution to the magnetic field potential at x: (x - y) x (G'(r) / r) t1 with
// G'(r) / r = e^{ikr} (ik r - 1) / (4 pi r^3) (the body of far_field_eval_fp64<kMagneticPotential>,
// also the mixed kernel's path for targets close to a group's sources).
template <int PH, typename F>
__device__ __forceinline__ void fp64_magnetic_source(F field, double x0, double x1, double x2, double kr, double ki,
                                                     double (&accr)[3], double (&acci)[3]) {
    const double d0 = x0 - field(0), d1 = x1 - field(1), d2 = x2 - field(2);
    const double r2 = d0 * d0 + d1 * d1 + d2 * d2, rinv = rsqrt_t(r2), r = r2 * rinv;
    double sn, cs;
    sincos_t<PH>(kr * r, &sn, &cs);
    const double damp = (PH & kPhReal) ? 1.0 : ki != 0.0 ? exp(-ki * r) : 1.0;
    const double g = kInv4Pi * damp * (rinv * rinv * rinv);
    const double fr = (PH & kPhReal) ? -1.0 : -ki * r - 1.0, fi = kr * r;  // ik r - 1
    const double vr = g * (cs * fr - sn * fi), vi = g * (cs * fi + sn * fr);
    double wr[3], wi[3];
#pragma unroll
    for (int d = 0; d < 3; ++d) {
        const double t1r = field(8 + 2 * d), t1i = field(9 + 2 * d);
        wr[d] = vr * t1r - vi * t1i;
        wi[d] = vr * t1i + vi * t1r;
    }
    accr[0] += d1 * wr[2] - d2 * wr[1];
    acci[0] += d1 * wi[2] - d2 * wi[1];
    accr[1] += d2 * wr[0] - d0 * wr[2];
    acci[1] += d2 * wi[0] - d0 * wi[2];
    accr[2] += d0 * wr[1] - d1 * wr[0];
    acci[2] += d0 * wi[1] - d1 * wi[0];
}

// Mixed kernels: a (target, tile or group) pair whose reference distance r0 is below
// kNearTileFactor times the radius R of the tile's sources around the reference takes the FP64
// path: the FP32 offsets lose relative accuracy once the target comes within a fraction of R of
// the sources (targets just off the surface), while targets beyond r0 >= 1.5 R stay within 1e-6
// of FP64 (measured for the single, double layer and EFIE potentials on icosphere levels 2 and 4).
constexpr double kNearTileFactor = 1.5;

// The FP64 path of the mixed kernels for the sources [s_begin, s_end) of the workspace.
template <int OP, int KD>
__device__ __forceinline__ void sources_fp64_near(const PotArgs &a, int s_begin, int s_end, double x0, double x1,
                                               double x2, double alr, double ali, double *accr, double *acci) {
    double ar[KD], ai[KD];
#pragma unroll
    for (int d = 0; d < KD; ++d) ar[d] = ai[d] = 0.0;
    const int64_t P = a.nsrc_pad;
    for (int s = s_begin; s < s_end; ++s)
        fp64_source<OP, 0>([&](int k) { return a.src[k * P + s]; }, x0, x1, x2, a.kr, a.ki, alr, ali, ar, ai);
#pragma unroll
    for (int d = 0; d < KD; ++d) {
        accr[d] += ar[d];
        acci[d] += ai[d];
    }
}

// FP64 evaluation (fp64_source per source).
template <int OP>
__global__ __launch_bounds__(kPotThreads) void potential_eval_fp64(const PotArgs a) {
    constexpr int KD = OP == kMaxwellEFIE ? 3 : 1;
    __shared__ double sh[kSrcFields][kPotTile];
    __shared__ double wmax[kPotThreads / 64];
    const int m = blockIdx.x * kPotThreads + threadIdx.x;
    const int chunk = blockIdx.y;
    const int s0 = chunk * a.chunk_len;
    const int s1 = min(s0 + a.chunk_len, a.nsrc);
    double x0 = 0, x1 = 0, x2 = 0;
    if (m < a.npts) {
        x0 = a.pts[3 * (int64_t)m];
        x1 = a.pts[3 * (int64_t)m + 1];
        x2 = a.pts[3 * (int64_t)m + 2];
    }
    const double xnorm = sqrt(x0 * x0 + x1 * x1 + x2 * x2);
    double accr[KD], acci[KD];
    for (int d = 0; d < KD; ++d) accr[d] = acci[d] = 0.0;
    const double kr = a.kr, ki = a.ki;
    const double kk = kr * kr + ki * ki;
    const double alr = kk > 0.0 ? -ki / kk : 0.0, ali = kk > 0.0 ? -kr / kk : 0.0;  // alpha = 1 / (ik)
    for (int t0 = s0; t0 < s1; t0 += kPotTile) {
        __syncthreads();
        {
            const int s = t0 + threadIdx.x;
            const int nf = OP == kMaxwellEFIE ? kSrcFields : 8;
            for (int k = 0; k < nf; ++k) sh[k][threadIdx.x] = s < s1 ? a.src[(int64_t)k * a.nsrc_pad + s] : 0.0;
            const double y0 = sh[0][threadIdx.x], y1 = sh[1][threadIdx.x], y2 = sh[2][threadIdx.x];
            tile_extent_stage(y0 * y0 + y1 * y1 + y2 * y2, wmax);
        }
        __syncthreads();
        const int nt = min(kPotTile, s1 - t0);
        auto sources = [&](auto ph_tag) {
            constexpr int PH = decltype(ph_tag)::value;
            for (int j = 0; j < nt; ++j)
                fp64_source<OP, PH>([&](int k) { return sh[k][j]; }, x0, x1, x2, kr, ki, alr, ali, accr, acci);
        };
        if (OpTraits<OP>::is_complex && tile_phase_reduced(kr, ki, xnorm, wmax))
            sources(std::integral_constant<int, kPhReduced | kPhReal>{});
        else
            sources(std::integral_constant<int, 0>{});
    }
    if (m < a.npts) {
        for (int d = 0; d < KD; ++d) {
            double *p = a.partial + 2 * (((int64_t)chunk * KD + d) * a.npts + m);
            p[0] = accr[d];
            p[1] = acci[d];
        }
    }
}

// sin and cos of a half angle x (radians) with relative accuracy, for e^{i 2x} - 1 = -2 sin^2 x +
// 2i sin x cos x, a small value that must not inherit the absolute error of the hardware sine:
// Taylor series for |x| <= 0.5 (remainders x^11 / 11! and x^10 / 10! < 3e-10), sincosf beyond.
__device__ __forceinline__ void half_angle_sincos(float x, float &s, float &c) {
    if (fabsf(x) <= 0.5f) {
        const float x2 = x * x;
        float p = fmaf(x2, 1.f / 362880.f, -1.f / 5040.f);
        p = fmaf(x2, p, 1.f / 120.f);
        p = fmaf(x2, p, -1.f / 6.f);
        s = fmaf(x * x2, p, x);
        float q = fmaf(x2, 1.f / 40320.f, -1.f / 720.f);
        q = fmaf(x2, q, 1.f / 24.f);
        q = fmaf(x2, q, -0.5f);
        c = fmaf(x2, q, 1.f);
    } else {
        sincosf(x, &s, &c);
    }
}

// e^{ik delta} - 1 = (cos y - 1) + i sin y, y = k delta (k real), with relative accuracy from the
// Taylor series in delta with the powers of k folded into the coefficients:
//   cos y - 1 = d2 (c2 + d2 (c4 + ...)),  sin y = delta s,  s = k + d2 (s3 + d2 (s5 + ...)),
// d2 = delta^2. MODE 0 (|y| <= 0.25): through y^6 and y^5; MODE 1 (|y| <= 0.5): through y^8 and
// y^7; MODE 2 (|y| <= 1): through y^10 and y^9 (remainders below 5e-8 of the leading terms
// y^2 / 2 and y). MODE 3, any y: the series of MODE 2 in y where |y| <= 1, half-angle sines beyond
// (si is then not used).
template <int MODE>
__device__ __forceinline__ void efie_em1(float delta, const EfieSeries &e, float &em1r, float &em1i, float &si) {
    const float d2 = delta * delta;
    if constexpr (MODE < 3) {
        constexpr int NC = MODE + 3, NS = MODE + 2;  // terms of the cosine and sine series
        float p = e.c[NC - 1];
#pragma unroll
        for (int n = NC - 2; n >= 0; --n) p = fmaf(d2, p, e.c[n]);
        em1r = d2 * p;
        float q = e.s[NS - 1];
#pragma unroll
        for (int n = NS - 2; n >= 0; --n) q = fmaf(d2, q, e.s[n]);
        si = fmaf(d2, q, e.k);
        em1i = delta * si;
    } else {
        const float y = e.k * delta;
        if (fabsf(y) <= 1.f) {
            const float y2 = y * y;
            float p = fmaf(y2, -1.f / 3628800.f, 1.f / 40320.f);
            p = fmaf(y2, p, -1.f / 720.f);
            p = fmaf(y2, p, 1.f / 24.f);
            p = fmaf(y2, p, -0.5f);
            em1r = y2 * p;
            float q = fmaf(y2, 1.f / 362880.f, -1.f / 5040.f);
            q = fmaf(y2, q, 1.f / 120.f);
            q = fmaf(y2, q, -1.f / 6.f);
            em1i = fmaf(y * y2, q, y);
        } else {
            float hs, hc;
            sincosf(0.5f * y, &hs, &hc);
            em1r = -2.f * hs * hs;
            em1i = 2.f * hs * hc;
        }
        si = 0.f;
    }
}

// MIXED EFIE potential: FP32 per source, FP64 per (target, group of 32 sources, a few
// neighbouring elements). With c the group's middle source, v = x - c, u = c - y
// (so x - y = v + u), r0 = |v|, delta = r - r0, E = e^{ik delta}, h(r) = 1/r^2 - alpha/r^3
// (alpha = 1/(ik)), one source contributes e^{ik r0} / (4 pi) [ ik (E / r) t1 - E h(r) (v + u) t2 ].
// It is split into
//   the group-centre term  ik t1 / r0 - h(r0) v t2   (group sums of t1, t2 in FP64), and
//   the remainder          f t1' - (b v + c u) t2'   (FP32, t1' = ik t1 / 4 pi, t2' = t2 / 4 pi),
// f = E/r - 1/r0, b = E h(r) - h(r0), c = E h(r). f and b are small differences formed without
// cancellation from E - 1 (series in delta) and q = 1/r0 - 1/r = delta / (r r0):
//   f = (E - 1) / r - q,   dh = h(r) - h(r0) = -q (1/r + 1/r0) + alpha q (1/r^2 + 1/(r r0) + 1/r0^2),
//   b = dh + (E - 1) h(r),  c = h(r) + (E - 1) h(r).
// b v t2' is summed per group as v B, B = sum b t2'. Sources are staged per tile of 256 as three
// FP32 records (u, |u|^2), (t2', t1'_x), (t1'_y, t1'_z), read as LDS broadcasts. The group radius
// max |u| bounds |delta| and selects the shortest series for e^{ik delta} - 1 (wave-uniform).
// DAMP: complex wavenumber, E = e^{-Im(k) delta} e^{i Re(k) delta}.
template <bool DAMP>
__global__ __launch_bounds__(kPotThreads) void potential_eval_mixed_efie(const PotArgs a) {
    constexpr int GS = 32;  // sources per reference group
    constexpr int NG = kPotTile / GS;
    __shared__ float4 shu[kPotTile];  // u, |u|^2
    __shared__ float4 shq[kPotTile];  // t2', t1'_x
    __shared__ float4 shp[kPotTile];  // t1'_y, t1'_z
    __shared__ double sc[NG][3];
    __shared__ double wsum[NG][8];    // group sums of t2 and t1 (FP64)
    __shared__ float srad[NG];        // group radius max |u| (bounds |delta|)
    const int m = blockIdx.x * kPotThreads + threadIdx.x;
    const int chunk = blockIdx.y;
    const int s0 = chunk * a.chunk_len;
    const int s1 = min(s0 + a.chunk_len, a.nsrc);
    const int P = a.nsrc_pad;
    double x0 = 0, x1 = 0, x2 = 0;
    if (m < a.npts) {
        x0 = a.pts[3 * (int64_t)m];
        x1 = a.pts[3 * (int64_t)m + 1];
        x2 = a.pts[3 * (int64_t)m + 2];
    }
    double accr[3] = {0.0, 0.0, 0.0}, acci[3] = {0.0, 0.0, 0.0};
    const float ki = (float)a.ki;
    const EfieSeries &ser = a.ser;
    // alpha = 1/(ik) (0 for k = 0, where the t2 term has no 1/r^3 part)
    const double kkd = a.kr * a.kr + a.ki * a.ki;
    const double alr_d = kkd > 0.0 ? -a.ki / kkd : 0.0, ali_d = kkd > 0.0 ? -a.kr / kkd : 0.0;
    const float al_re = (float)alr_d, al_im = (float)ali_d;
    const float kabs = fabsf((float)a.kr);
    for (int t0 = s0; t0 < s1; t0 += kPotTile) {
        const int nt = min(kPotTile, s1 - t0);
        __syncthreads();
        {
            const int j = threadIdx.x, g = j / GS, s = t0 + j;
            const bool ok = j < nt;
            // reference point of the group: its middle source (a real source, so r0 > 0 for every
            // target off the sources)
            const int cidx = t0 + min(g * GS + GS / 2, nt - 1);
            double y[3], c[3];
#pragma unroll
            for (int d = 0; d < 3; ++d) {
                y[d] = ok ? a.src[(int64_t)d * P + s] : 0.0;
                c[d] = a.src[(int64_t)d * P + cidx];
            }
            if (j % GS == 0)
                for (int d = 0; d < 3; ++d) sc[g][d] = c[d];
            double w[8];
#pragma unroll
            for (int k = 0; k < 8; ++k) w[k] = ok ? a.src[(int64_t)(6 + k) * P + s] : 0.0;
            float4 U = make_float4(0.f, 0.f, 0.f, 0.f), Q = U, S = U;
            double uu = 0.0;
            if (ok) {
                const double u0 = c[0] - y[0], u1 = c[1] - y[1], u2 = c[2] - y[2];
                uu = u0 * u0 + u1 * u1 + u2 * u2;
                U = make_float4((float)u0, (float)u1, (float)u2, (float)uu);
                const double ikr = -a.ki, iki = a.kr;
                float t1[6];
#pragma unroll
                for (int d = 0; d < 3; ++d) {
                    const double tr = w[2 + 2 * d], ti = w[3 + 2 * d];
                    t1[2 * d] = (float)(kInv4Pi * (ikr * tr - iki * ti));
                    t1[2 * d + 1] = (float)(kInv4Pi * (ikr * ti + iki * tr));
                }
                Q = make_float4((float)(kInv4Pi * w[0]), (float)(kInv4Pi * w[1]), t1[0], t1[1]);
                S = make_float4(t1[2], t1[3], t1[4], t1[5]);
            }
            shu[j] = U;
            shq[j] = Q;
            shp[j] = S;
#pragma unroll
            for (int k = 0; k < 8; ++k) {
#pragma unroll
                for (int off = GS / 2; off >= 1; off >>= 1) w[k] += __shfl_xor(w[k], off, 64);
            }
#pragma unroll
            for (int off = GS / 2; off >= 1; off >>= 1) uu = fmax(uu, __shfl_xor(uu, off, 64));
            if (j % GS == 0) {
#pragma unroll
                for (int k = 0; k < 8; ++k) wsum[g][k] = w[k];
                srad[g] = (float)(sqrt(uu) * (1.0 + 1e-6));
            }
        }
        __syncthreads();
        for (int g = 0; g < NG && g * GS < nt; ++g) {
            const int j0 = g * GS, j1 = min(j0 + GS, nt);
            // |Re(k) delta| <= |Re(k)| max |u| (wave-uniform)
            const float ymax = __int_as_float(__builtin_amdgcn_readfirstlane(__float_as_int(kabs * srad[g])));
            // per (target, group) FP64 reference distance and phase
            const double v0d = x0 - sc[g][0], v1d = x1 - sc[g][1], v2d = x2 - sc[g][2];
            const double r0sq_d = v0d * v0d + v1d * v1d + v2d * v2d;
            // targets close to the group's sources: the group in FP64 for the whole wave
            const double rn = kNearTileFactor * (double)srad[g];
            if (__any(m < a.npts && r0sq_d < rn * rn)) {
                sources_fp64_near<kMaxwellEFIE, 3>(a, t0 + j0, t0 + j1, x0, x1, x2, alr_d, ali_d, accr, acci);
                continue;
            }
            const double r0d = sqrt(r0sq_d);
            const double i0 = 1.0 / r0d, i2 = i0 * i0, i3 = i2 * i0;
            const float tv0 = (float)(2.0 * v0d), tv1 = (float)(2.0 * v1d), tv2 = (float)(2.0 * v2d);
            const float r0 = (float)r0d, r0sq = (float)r0sq_d;
            const float r0inv = (float)i0, r0inv2 = (float)i2;
            const float h0r = (float)(i2 - alr_d * i3), h0i = (float)(-ali_d * i3);
            float Ar[3] = {0.f, 0.f, 0.f}, Ai[3] = {0.f, 0.f, 0.f}, Br = 0.f, Bi = 0.f;
            auto sources = [&](auto mode) {
                constexpr int MODE = decltype(mode)::value;
#pragma unroll 2
                for (int j = j0; j < j1; ++j) {
                    const float4 U = shu[j];
                    const float tt = fmaf(tv0, U.x, fmaf(tv1, U.y, fmaf(tv2, U.z, U.w)));  // r^2 - r0^2
                    const float r2 = r0sq + tt;
                    const float rinv = rsqrt_t(r2);
                    const float delta = tt * __builtin_amdgcn_rcpf(fmaf(r2, rinv, r0));
                    float em1r, em1i, si;  // E - 1, and (E - 1) / delta's imaginary part (series)
                    efie_em1<MODE>(delta, ser, em1r, em1i, si);
                    const float D = delta * rinv;
                    const float q = D * r0inv;  // 1/r0 - 1/r
                    float fi;
                    if constexpr (DAMP) {
                        const float dmp = __expf(-ki * delta), dm1 = expm1f(-ki * delta);
                        em1r = fmaf(dmp, em1r, dm1);
                        em1i = dmp * em1i;
                        fi = em1i * rinv;
                    } else {
                        fi = MODE < 3 ? D * si : em1i * rinv;
                    }
                    const float fr = fmaf(em1r, rinv, -q);
                    const float s1 = rinv + r0inv;
                    const float q1 = q * s1, q3 = q * fmaf(s1, rinv, r0inv2);
                    // h(r) - h(r0); alpha is imaginary for a real wavenumber
                    const float dhr = DAMP ? fmaf(al_re, q3, -q1) : -q1, dhi = al_im * q3;
                    const float hr = h0r + dhr, hi = h0i + dhi;
                    const float pr = em1r * hr - em1i * hi, pi = em1r * hi + em1i * hr;  // (E - 1) h(r)
                    const float br = dhr + pr, bi = dhi + pi, cr = hr + pr, ci = hi + pi;
                    const float4 Q = shq[j], S = shp[j];
                    Br = fmaf(br, Q.x, fmaf(-bi, Q.y, Br));
                    Bi = fmaf(br, Q.y, fmaf(bi, Q.x, Bi));
                    const float ctr = cr * Q.x - ci * Q.y, cti = cr * Q.y + ci * Q.x;
                    Ar[0] = fmaf(fr, Q.z, fmaf(-fi, Q.w, fmaf(-ctr, U.x, Ar[0])));
                    Ai[0] = fmaf(fr, Q.w, fmaf(fi, Q.z, fmaf(-cti, U.x, Ai[0])));
                    Ar[1] = fmaf(fr, S.x, fmaf(-fi, S.y, fmaf(-ctr, U.y, Ar[1])));
                    Ai[1] = fmaf(fr, S.y, fmaf(fi, S.x, fmaf(-cti, U.y, Ai[1])));
                    Ar[2] = fmaf(fr, S.z, fmaf(-fi, S.w, fmaf(-ctr, U.z, Ar[2])));
                    Ai[2] = fmaf(fr, S.w, fmaf(fi, S.z, fmaf(-cti, U.z, Ai[2])));
                }
            };
            // the folded series need k^10 / 10! in FP32 range (|k| < 1000)
            if (ymax <= 0.24f && kabs < 1000.f)
                sources(std::integral_constant<int, 0>{});
            else if (ymax <= 0.49f && kabs < 1000.f)
                sources(std::integral_constant<int, 1>{});
            else if (ymax <= 0.98f && kabs < 1000.f)
                sources(std::integral_constant<int, 2>{});
            else
                sources(std::integral_constant<int, 3>{});
            // group-centre term in FP64, ik T1 / r0 - h(r0) v T2 (times 1/(4 pi)), plus the
            // FP32 remainder with its v B part, all times e^{ik r0}
            double ph_s, ph_c;
            sincos_fast(a.kr * r0d, &ph_s, &ph_c);
            const double damp0 = DAMP ? exp(-a.ki * r0d) : 1.0;
            const double pr = ph_c * damp0, pi = ph_s * damp0;
            const double ikr_d = -a.ki, iki_d = a.kr;
            const double hr0 = i2 - alr_d * i3, hi0 = -ali_d * i3;
            const double vd[3] = {v0d, v1d, v2d};
#pragma unroll
            for (int d = 0; d < 3; ++d) {
                const double t1r = wsum[g][2 + 2 * d], t1i = wsum[g][3 + 2 * d];
                const double gr = hr0 * vd[d], gi = hi0 * vd[d];
                const float vf = (float)vd[d];
                const double cr = kInv4Pi * ((ikr_d * t1r - iki_d * t1i) * i0 - (gr * wsum[g][0] - gi * wsum[g][1])) +
                                  (double)(Ar[d] - vf * Br);
                const double ci = kInv4Pi * ((ikr_d * t1i + iki_d * t1r) * i0 - (gr * wsum[g][1] + gi * wsum[g][0])) +
                                  (double)(Ai[d] - vf * Bi);
                accr[d] += pr * cr - pi * ci;
                acci[d] += pr * ci + pi * cr;
            }
        }  // groups
    }
    if (m < a.npts) {
        for (int d = 0; d < 3; ++d) {
            double *p = a.partial + 2 * (((int64_t)chunk * 3 + d) * a.npts + m);
            p[0] = accr[d];
            p[1] = acci[d];
        }
    }
}

// MIXED scalar potentials (Laplace and Helmholtz single and double layer). The sources of a tile
// of 256 are staged once as FP32 records relative to the tile's reference source c:
//   u = c - y, |u|^2, the weights c_j / (4 pi) and, for the double layer, n_y and u.n_y,
// read by every lane as LDS broadcasts. Per (target, tile) the FP64 part forms v = x - c,
// r0^2 = |v|^2, r0 and e^{i k r0}; per source only
//   t = r^2 - r0^2 = 2 v.u + |u|^2,  r = sqrt(r0^2 + t),  k (r - r0) / 2 pi = t / ((r + r0) K)
// (K = 2 pi / k) are needed in FP32, with the hardware rsq, rcp, sin and cos (revolutions). DAMP:
// complex wavenumber, e^{-Im(k) (r - r0)} per source (exp2), e^{-Im(k) r0} per tile.
template <int OP, bool DAMP>
__global__ __launch_bounds__(kPotThreads) void potential_eval_mixed_scalar(const PotArgs a) {
    constexpr bool HELM = OP == kHelmholtzSL || OP == kHelmholtzDL;
    constexpr bool DL = OP == kLaplaceDL || OP == kHelmholtzDL;
    static_assert(!DAMP || HELM, "only Helmholtz wavenumbers are complex");
    __shared__ float4 shu[kPotTile];           // u, |u|^2
    __shared__ float4 shw[kPotTile];           // Re c / 4 pi, Im c / 4 pi, u.n_y, -
    __shared__ float4 shn[DL ? kPotTile : 1];  // n_y, -
    __shared__ double sc[3];
    __shared__ float srad2[kPotThreads / 64];  // per wave: max |u|^2 of its staged sources
    const int m = blockIdx.x * kPotThreads + threadIdx.x;
    const int chunk = blockIdx.y;
    const int s0 = chunk * a.chunk_len;
    const int s1 = min(s0 + a.chunk_len, a.nsrc);
    double x0 = 0, x1 = 0, x2 = 0;
    if (m < a.npts) {
        x0 = a.pts[3 * (int64_t)m];
        x1 = a.pts[3 * (int64_t)m + 1];
        x2 = a.pts[3 * (int64_t)m + 2];
    }
    double accr = 0.0, acci = 0.0;
    const float kr_f = (float)a.kr, ki_f = (float)a.ki;
    const float kr_rev = (float)(a.kr * 0.15915494309189533577);
    const float ki_log2e = (float)(a.ki * 1.4426950408889634);
    const float kinv = 1.f / kr_rev;  // K (infinite for k = 0: the phase is then 0)
    const int P = a.nsrc_pad;
    for (int t0 = s0; t0 < s1; t0 += kPotTile) {
        const int nt = min(kPotTile, s1 - t0);
        __syncthreads();
        {
            const int cidx = t0 + min(kPotTile / 2, nt - 1);  // reference: the tile's middle source
            double c[3];
            for (int d = 0; d < 3; ++d) c[d] = a.src[(int64_t)d * P + cidx];
            if (threadIdx.x == 0)
                for (int d = 0; d < 3; ++d) sc[d] = c[d];
            const int j = threadIdx.x, s = t0 + j;
            float4 U = make_float4(0.f, 0.f, 0.f, 0.f), W = make_float4(0.f, 0.f, 0.f, 0.f);
            float4 N = make_float4(0.f, 0.f, 0.f, 0.f);
            if (j < nt) {
                const double u0 = c[0] - a.src[s], u1 = c[1] - a.src[(int64_t)P + s], u2 = c[2] - a.src[2 * (int64_t)P + s];
                U = make_float4((float)u0, (float)u1, (float)u2, (float)(u0 * u0 + u1 * u1 + u2 * u2));
                W.x = (float)(kInv4Pi * a.src[6 * (int64_t)P + s]);
                W.y = (float)(kInv4Pi * a.src[7 * (int64_t)P + s]);
                if constexpr (DL) {
                    const double n0 = a.src[3 * (int64_t)P + s], n1 = a.src[4 * (int64_t)P + s],
                                 n2 = a.src[5 * (int64_t)P + s];
                    N = make_float4((float)n0, (float)n1, (float)n2, 0.f);
                    W.z = (float)(u0 * n0 + u1 * n1 + u2 * n2);
                }
            }
            shu[j] = U;
            shw[j] = W;
            if constexpr (DL) shn[j] = N;
            float w2 = U.w;
#pragma unroll
            for (int off = 32; off >= 1; off >>= 1) w2 = fmaxf(w2, __shfl_xor(w2, off, 64));
            if ((j & 63) == 0) srad2[j >> 6] = w2;
        }
        __syncthreads();
        // per (target, tile): FP64 reference distance and phase
        const double v0d = x0 - sc[0], v1d = x1 - sc[1], v2d = x2 - sc[2];
        const double r0sq_d = v0d * v0d + v1d * v1d + v2d * v2d;
        {
            // targets close to the tile's sources: this tile in FP64 for the whole wave
            float R2 = srad2[0];
#pragma unroll
            for (int w = 1; w < kPotThreads / 64; ++w) R2 = fmaxf(R2, srad2[w]);
            if (__any(m < a.npts && r0sq_d < kNearTileFactor * kNearTileFactor * (double)R2 * (1.0 + 1e-6))) {
                sources_fp64_near<OP, 1>(a, t0, t0 + nt, x0, x1, x2, 0.0, 0.0, &accr, &acci);
                continue;
            }
        }
        const double r0d = sqrt(r0sq_d);
        const float tv0 = (float)(2.0 * v0d), tv1 = (float)(2.0 * v1d), tv2 = (float)(2.0 * v2d);
        const float v0 = (float)v0d, v1 = (float)v1d, v2 = (float)v2d;
        const float r0sq = (float)r0sq_d, r0 = (float)r0d;
        const float r0k = r0 * kinv;
        float Ar = 0.f, Ai = 0.f;
#pragma unroll 4
        for (int j = 0; j < nt; ++j) {
            const float4 U = shu[j];
            const float4 W = shw[j];
            const float tt = tv0 * U.x + (tv1 * U.y + (tv2 * U.z + U.w));  // r^2 - r0^2
            const float r2 = r0sq + tt;
            const float rinv = __builtin_amdgcn_rsqf(r2);
            float er = rinv, ei = 0.f;  // e^{ik(r - r0)} / r (Laplace: 1 / r)
            float r = 0.f;
            if constexpr (HELM) {
                r = r2 * rinv;
                float ph, amp = rinv;
                if constexpr (DAMP) {
                    const float delta = tt * __builtin_amdgcn_rcpf(r + r0);
                    ph = kr_rev * delta;
                    amp *= __builtin_amdgcn_exp2f(-ki_log2e * delta);
                } else {
                    ph = tt * __builtin_amdgcn_rcpf(r * kinv + r0k);
                }
                const float cs = __builtin_amdgcn_cosf(ph), sn = __builtin_amdgcn_sinf(ph);
                if constexpr (!DL) {
                    // single layer: the weight times e^{ik(r - r0)} first, then one FMA by the amplitude
                    Ar += amp * (cs * W.x - sn * W.y);
                    Ai += amp * (cs * W.y + sn * W.x);
                    continue;
                }
                er = cs * amp;
                ei = sn * amp;
            }
            if constexpr (DL) {
                // (i k r - 1) (y - x).n_y / r^2 with y - x = -(v + u)
                const float4 N = shn[j];
                const float dot = -(v0 * N.x + (v1 * N.y + (v2 * N.z + W.z)));
                const float f = dot * rinv * rinv;
                if constexpr (HELM) {
                    const float ar = DAMP ? -1.f - ki_f * r : -1.f, ai = kr_f * r;
                    const float pr = (ar * er - ai * ei) * f, pi = (ar * ei + ai * er) * f;
                    er = pr;
                    ei = pi;
                } else {
                    er = -er * f;
                }
            }
            if constexpr (HELM) {
                Ar += er * W.x - ei * W.y;
                Ai += er * W.y + ei * W.x;
            } else {
                Ar += er * W.x;
                Ai += er * W.y;
            }
        }
        // multiply the tile sum by e^{i k r0} (e^{-Im(k) r0}) in FP64
        double pr = 1.0, pi = 0.0;
        if constexpr (HELM) {
            double sn, cs;
            sincos_fast(a.kr * r0d, &sn, &cs);
            const double damp0 = DAMP ? exp(-a.ki * r0d) : 1.0;
            pr = cs * damp0;
            pi = sn * damp0;
        }
        accr += pr * Ar - pi * Ai;
        acci += pr * Ai + pi * Ar;
    }
    if (m < a.npts) {
        double *p = a.partial + 2 * ((int64_t)chunk * a.npts + m);
        p[0] = accr;
        p[1] = acci;
    }
}

// Far-field integrands per source (sources from potential_prep_kernel: c for scalar spaces,
// t1 and t2 of the EFIE for RWG):
//   Helmholtz single layer  K c
//   Helmholtz double layer  -i Re(k) (xh . n_y) K c
//   Maxwell electric field  K (ik t1 - xh t2)
//   Maxwell magnetic field  xh x (K ik t1), the cross product taken once per direction
// and, with the same sources, the magnetic field potential (numba_kernels.py:2363-2433) at points x:
//   sum (x - y) x (G (ik r - 1) t1 / r^2),  G = e^{ikr} / (4 pi r), r = |x - y|
enum FarField { kFarSL = 0, kFarDL = 1, kFarE = 2, kFarH = 3, kMagneticPotential = 4 };

template <int FF>
__global__ __launch_bounds__(kPotThreads) void far_field_eval_fp64(const PotArgs a) {
    constexpr bool MAXWELL = FF == kFarE || FF == kFarH || FF == kMagneticPotential;
    constexpr int KD = MAXWELL ? 3 : 1;
    constexpr int NF = MAXWELL ? kSrcFields : 8;
    __shared__ double sh[NF][kPotTile];
    __shared__ double wmax[kPotThreads / 64];
    const int m = blockIdx.x * kPotThreads + threadIdx.x;
    const int chunk = blockIdx.y;
    const int s0 = chunk * a.chunk_len;
    const int s1 = min(s0 + a.chunk_len, a.nsrc);
    double x0 = 0, x1 = 0, x2 = 0;
    if (m < a.npts) {
        x0 = a.pts[3 * (int64_t)m];
        x1 = a.pts[3 * (int64_t)m + 1];
        x2 = a.pts[3 * (int64_t)m + 2];
    }
    const double xh[3] = {x0, x1, x2};
    const double xnorm = sqrt(x0 * x0 + x1 * x1 + x2 * x2);
    double accr[KD], acci[KD];
    for (int d = 0; d < KD; ++d) accr[d] = acci[d] = 0.0;
    const double kr = a.kr, ki = a.ki;
    for (int t0 = s0; t0 < s1; t0 += kPotTile) {
        __syncthreads();
        {
            const int s = t0 + threadIdx.x;
            for (int k = 0; k < NF; ++k) sh[k][threadIdx.x] = s < s1 ? a.src[(int64_t)k * a.nsrc_pad + s] : 0.0;
            const double y0 = sh[0][threadIdx.x], y1 = sh[1][threadIdx.x], y2 = sh[2][threadIdx.x];
            tile_extent_stage(y0 * y0 + y1 * y1 + y2 * y2, wmax);
        }
        __syncthreads();
        const int nt = min(kPotTile, s1 - t0);
        auto sources = [&](auto ph_tag) {
        constexpr int PH = decltype(ph_tag)::value;
        for (int j = 0; j < nt; ++j) {
            if constexpr (FF == kMagneticPotential) {
                fp64_magnetic_source<PH>([&](int k) { return sh[k][j]; }, x0, x1, x2, kr, ki, accr, acci);
                continue;
            }
            const double dot = x0 * sh[0][j] + x1 * sh[1][j] + x2 * sh[2][j];
            double sn, cs;
            sincos_t<PH>(-kr * dot, &sn, &cs);
            const double gre = kInv4Pi * cs, gim = kInv4Pi * sn;
            if constexpr (FF == kFarSL || FF == kFarDL) {
                double vr = gre, vi = gim;
                if constexpr (FF == kFarDL) {
                    const double f = -kr * (x0 * sh[3][j] + x1 * sh[4][j] + x2 * sh[5][j]);
                    vr = -f * gim;  // i f K
                    vi = f * gre;
                }
                const double cr = sh[6][j], ci = sh[7][j];
                accr[0] += vr * cr - vi * ci;
                acci[0] += vr * ci + vi * cr;
            } else {
                const double t2r = sh[6][j], t2i = sh[7][j];
#pragma unroll
                for (int d = 0; d < 3; ++d) {
                    const double t1r = sh[8 + 2 * d][j], t1i = sh[9 + 2 * d][j];
                    // ik t1 with ik = -ki + i kr
                    double tr = -ki * t1r - kr * t1i, ti = -ki * t1i + kr * t1r;
                    if constexpr (FF == kFarE) {
                        tr -= xh[d] * t2r;
                        ti -= xh[d] * t2i;
                    }
                    accr[d] += gre * tr - gim * ti;
                    acci[d] += gre * ti + gim * tr;
                }
            }
        }
        };
        if (tile_phase_reduced(kr, ki, xnorm, wmax))
            sources(std::integral_constant<int, kPhReduced | kPhReal>{});
        else
            sources(std::integral_constant<int, 0>{});
    }
    if (m < a.npts) {
        if constexpr (FF == kFarH) {
            const double vr[3] = {accr[0], accr[1], accr[2]}, vi[3] = {acci[0], acci[1], acci[2]};
#pragma unroll
            for (int d = 0; d < 3; ++d) {
                const int d1 = (d + 1) % 3, d2 = (d + 2) % 3;
                accr[d] = xh[d1] * vr[d2] - xh[d2] * vr[d1];
                acci[d] = xh[d1] * vi[d2] - xh[d2] * vi[d1];
            }
        }
        for (int d = 0; d < KD; ++d) {
            double *p = a.partial + 2 * (((int64_t)chunk * KD + d) * a.npts + m);
            p[0] = accr[d];
            p[1] = acci[d];
        }
    }
}

// MIXED far fields and magnetic field potential. Sources are staged in FP32 as offsets u = y - c
// from a reference source c (FP64) of their group of 32.
// Far fields: with q_j the source's integrand weight (c, -i Re(k) (xh.n) c, ik t1 - xh t2 or
// ik t1) and theta_j = -Re(k) xh.u_j,
//   sum_j K_j q_j = K_c [ sum_j q_j + sum_j (e^{i theta_j} - 1) q_j ],  K_c = e^{-i Re(k) xh.c} / (4 pi).
// K_c and the group sums of the weights (hence sum_j q_j) are FP64; only the second sum, small
// and free of cancellation (e^{i theta} - 1 from half-angle sines), is FP32. The weights cancel
// over the surface (RWG divergences sum to zero per basis function, sum n c dS over a closed
// surface vanishes for constant c), so the leading term must not carry FP32 rounding.
// Magnetic potential: the reference-distance split of potential_eval_mixed_efie, r0 = |x - c| and
// e^{ik r0} in FP64 per (target, group), r - r0 in FP32 from the small offsets.
template <int FF>
__global__ __launch_bounds__(kPotThreads) void far_field_eval_mixed(const PotArgs a) {
    constexpr bool POT = FF == kMagneticPotential;
    constexpr bool MAXWELL = FF == kFarE || FF == kFarH || POT;
    constexpr int KD = MAXWELL ? 3 : 1;
    constexpr int NF = MAXWELL ? kSrcFields : 8;
    constexpr int NW = FF == kFarSL ? 2 : FF == kFarDL ? 6 : 8;  // FP64 group-summed weights
    constexpr int GS = 32;                                          // sources per reference group
    constexpr int NG = kPotTile / GS;
    constexpr double kInv2PiD = 0.15915494309189533577;
    __shared__ float sh[NF][kPotTile];
    __shared__ double sc[NG][3];
    __shared__ double wsum[NG][POT ? 1 : NW];
    __shared__ float grad2[POT ? NG : 1];  // magnetic potential: the group's max |u|^2
    const int m = blockIdx.x * kPotThreads + threadIdx.x;
    const int chunk = blockIdx.y;
    const int s0 = chunk * a.chunk_len;
    const int s1 = min(s0 + a.chunk_len, a.nsrc);
    double x0 = 0, x1 = 0, x2 = 0;
    if (m < a.npts) {
        x0 = a.pts[3 * (int64_t)m];
        x1 = a.pts[3 * (int64_t)m + 1];
        x2 = a.pts[3 * (int64_t)m + 2];
    }
    const double xh[3] = {x0, x1, x2};
    const float xf0 = (float)x0, xf1 = (float)x1, xf2 = (float)x2;
    double accr[KD], acci[KD];
    for (int d = 0; d < KD; ++d) accr[d] = acci[d] = 0.0;
    const float kr = (float)a.kr, ki = (float)a.ki;
    const float kr_rev = (float)(a.kr * kInv2PiD);
    const float ki_log2e = (float)(a.ki * 1.4426950408889634);
    for (int t0 = s0; t0 < s1; t0 += kPotTile) {
        const int nt = min(kPotTile, s1 - t0);
        __syncthreads();
        {
            const int g = threadIdx.x / GS;
            const int cidx = t0 + min(g * GS + GS / 2, nt - 1);
            double c[3];
            for (int d = 0; d < 3; ++d) c[d] = a.src[(int64_t)d * a.nsrc_pad + cidx];
            if (threadIdx.x % GS == 0)
                for (int d = 0; d < 3; ++d) sc[g][d] = c[d];
            const int s = t0 + threadIdx.x;
            const bool ok = threadIdx.x < nt;
            for (int d = 0; d < 3; ++d) sh[d][threadIdx.x] = ok ? (float)(a.src[(int64_t)d * a.nsrc_pad + s] - c[d]) : 0.f;
            for (int k = 3; k < NF; ++k) sh[k][threadIdx.x] = ok ? (float)a.src[(int64_t)k * a.nsrc_pad + s] : 0.f;
            if constexpr (POT) {
                float u2 = sh[0][threadIdx.x] * sh[0][threadIdx.x] + sh[1][threadIdx.x] * sh[1][threadIdx.x] +
                           sh[2][threadIdx.x] * sh[2][threadIdx.x];
#pragma unroll
                for (int off = GS / 2; off >= 1; off >>= 1) u2 = fmaxf(u2, __shfl_xor(u2, off, 64));
                if (threadIdx.x % GS == 0) grad2[g] = u2;
            }
            if constexpr (!POT) {
                double v[NW];
                if constexpr (FF == kFarDL) {
                    // n_j c_j: [d][re, im]
                    const double cr = ok ? a.src[(int64_t)6 * a.nsrc_pad + s] : 0.0;
                    const double ci = ok ? a.src[(int64_t)7 * a.nsrc_pad + s] : 0.0;
#pragma unroll
                    for (int d = 0; d < 3; ++d) {
                        const double n = ok ? a.src[(int64_t)(3 + d) * a.nsrc_pad + s] : 0.0;
                        v[2 * d] = n * cr;
                        v[2 * d + 1] = n * ci;
                    }
                } else {
#pragma unroll
                    for (int k = 0; k < NW; ++k) v[k] = ok ? a.src[(int64_t)(6 + k) * a.nsrc_pad + s] : 0.0;
                }
#pragma unroll
                for (int k = 0; k < NW; ++k) {
#pragma unroll
                    for (int off = GS / 2; off >= 1; off >>= 1) v[k] += __shfl_xor(v[k], off, 64);
                }
                if (threadIdx.x % GS == 0)
#pragma unroll
                    for (int k = 0; k < NW; ++k) wsum[g][k] = v[k];
            }
        }
        __syncthreads();
        for (int g = 0; g < NG && g * GS < nt; ++g) {
            const int j0 = g * GS, j1 = min(j0 + GS, nt);
            float Er[KD], Ei[KD];
            for (int d = 0; d < KD; ++d) Er[d] = Ei[d] = 0.f;
            double pr, pi;  // the group's FP64 factor
            if constexpr (POT) {
                const double v0d = x0 - sc[g][0], v1d = x1 - sc[g][1], v2d = x2 - sc[g][2];
                const double r0sq_d = v0d * v0d + v1d * v1d + v2d * v2d;
                // targets close to the group's sources: the group in FP64 for the whole wave
                if (__any(m < a.npts && r0sq_d < kNearTileFactor * kNearTileFactor * (double)grad2[g] * (1.0 + 1e-6))) {
                    const int64_t P = a.nsrc_pad;
                    for (int jj = t0 + j0; jj < t0 + j1; ++jj)
                        fp64_magnetic_source<0>([&](int k) { return a.src[k * P + jj]; }, x0, x1, x2, a.kr, a.ki,
                                                accr, acci);
                    continue;
                }
                const double r0d = sqrt(r0sq_d);
                double sn, cs;
                sincos_fast(a.kr * r0d, &sn, &cs);
                const double dmp = a.ki != 0.0 ? exp(-a.ki * r0d) : 1.0;
                pr = cs * dmp;
                pi = sn * dmp;
                const float v0 = (float)v0d, v1 = (float)v1d, v2 = (float)v2d;
                const float r0 = (float)r0d, r0sq = (float)r0sq_d;
                for (int j = j0; j < j1; ++j) {
                    const float u0 = sh[0][j], u1 = sh[1][j], u2 = sh[2][j];
                    const float w0 = v0 - u0, w1 = v1 - u1, w2 = v2 - u2;  // x - y
                    const float tt = u0 * (u0 - 2.f * v0) + u1 * (u1 - 2.f * v1) + u2 * (u2 - 2.f * v2);
                    const float r2 = r0sq + tt;
                    const float rinv = __builtin_amdgcn_rsqf(r2);
                    const float r = r2 * rinv;
                    const float delta = tt * __builtin_amdgcn_rcpf(r + r0);  // r - r0
                    const float ph = kr_rev * delta;
                    float c = __builtin_amdgcn_cosf(ph), sn1 = __builtin_amdgcn_sinf(ph);
                    if (ki != 0.f) {
                        const float dm = __builtin_amdgcn_exp2f(-ki_log2e * delta);
                        c *= dm;
                        sn1 *= dm;
                    }
                    // e^{ik delta} (ik r - 1) / (4 pi r^3)
                    const float gg = float(kInv4Pi) * rinv * rinv * rinv;
                    const float fr = -ki * r - 1.f, fi = kr * r;
                    const float vr = gg * (c * fr - sn1 * fi), vi = gg * (c * fi + sn1 * fr);
                    float wr[3], wi[3];
#pragma unroll
                    for (int d = 0; d < 3; ++d) {
                        const float t1r = sh[8 + 2 * d][j], t1i = sh[9 + 2 * d][j];
                        wr[d] = vr * t1r - vi * t1i;
                        wi[d] = vr * t1i + vi * t1r;
                    }
                    Er[0] += w1 * wr[2] - w2 * wr[1];
                    Ei[0] += w1 * wi[2] - w2 * wi[1];
                    Er[1] += w2 * wr[0] - w0 * wr[2];
                    Ei[1] += w2 * wi[0] - w0 * wi[2];
                    Er[2] += w0 * wr[1] - w1 * wr[0];
                    Ei[2] += w0 * wi[1] - w1 * wi[0];
                }
            } else {
                double sn, cs;
                sincos_fast(-a.kr * (x0 * sc[g][0] + x1 * sc[g][1] + x2 * sc[g][2]), &sn, &cs);
                pr = kInv4Pi * cs;
                pi = kInv4Pi * sn;
                for (int j = j0; j < j1; ++j) {
                    // e^{i theta} - 1 = -2 sin^2(theta / 2) + 2 i sin(theta / 2) cos(theta / 2)
                    float hs, hc;
                    half_angle_sincos(-0.5f * kr * (xf0 * sh[0][j] + xf1 * sh[1][j] + xf2 * sh[2][j]), hs, hc);
                    const float em1r = -2.f * hs * hs, em1i = 2.f * hs * hc;
                    if constexpr (FF == kFarSL || FF == kFarDL) {
                        float qr = sh[6][j], qi = sh[7][j];
                        if constexpr (FF == kFarDL) {
                            // -i Re(k) (xh.n) c
                            const float f = -kr * (xf0 * sh[3][j] + xf1 * sh[4][j] + xf2 * sh[5][j]);
                            const float cr = qr;
                            qr = -f * qi;
                            qi = f * cr;
                        }
                        Er[0] += em1r * qr - em1i * qi;
                        Ei[0] += em1r * qi + em1i * qr;
                    } else {
                        const float xfd[3] = {xf0, xf1, xf2};
                        const float t2r = sh[6][j], t2i = sh[7][j];
#pragma unroll
                        for (int d = 0; d < 3; ++d) {
                            const float t1r = sh[8 + 2 * d][j], t1i = sh[9 + 2 * d][j];
                            float qr = -ki * t1r - kr * t1i, qi = -ki * t1i + kr * t1r;  // ik t1
                            if constexpr (FF == kFarE) {
                                qr -= xfd[d] * t2r;
                                qi -= xfd[d] * t2i;
                            }
                            Er[d] += em1r * qr - em1i * qi;
                            Ei[d] += em1r * qi + em1i * qr;
                        }
                    }
                }
                // the group sums: sum_j q_j in FP64, added to the FP32 remainder
                double Qr[KD], Qi[KD];
                if constexpr (FF == kFarSL) {
                    Qr[0] = wsum[g][0];
                    Qi[0] = wsum[g][1];
                } else if constexpr (FF == kFarDL) {
                    const double nr = x0 * wsum[g][0] + x1 * wsum[g][2] + x2 * wsum[g][4];
                    const double ni = x0 * wsum[g][1] + x1 * wsum[g][3] + x2 * wsum[g][5];
                    Qr[0] = a.kr * ni;  // -i Re(k) (nr + i ni)
                    Qi[0] = -a.kr * nr;
                } else {
#pragma unroll
                    for (int d = 0; d < 3; ++d) {
                        const double t1r = wsum[g][2 + 2 * d], t1i = wsum[g][3 + 2 * d];
                        Qr[d] = -a.ki * t1r - a.kr * t1i;
                        Qi[d] = -a.ki * t1i + a.kr * t1r;
                        if constexpr (FF == kFarE) {
                            Qr[d] -= xh[d] * wsum[g][0];
                            Qi[d] -= xh[d] * wsum[g][1];
                        }
                    }
                }
                for (int d = 0; d < KD; ++d) {
                    const double er = Qr[d] + (double)Er[d], ei = Qi[d] + (double)Ei[d];
                    accr[d] += pr * er - pi * ei;
                    acci[d] += pr * ei + pi * er;
                }
                continue;
            }
            for (int d = 0; d < KD; ++d) {
                accr[d] += pr * Er[d] - pi * Ei[d];
                acci[d] += pr * Ei[d] + pi * Er[d];
            }
        }
    }
    if (m < a.npts) {
        if constexpr (FF == kFarH) {
            const double vr[3] = {accr[0], accr[1], accr[2]}, vi[3] = {acci[0], acci[1], acci[2]};
#pragma unroll
            for (int d = 0; d < 3; ++d) {
                const int d1 = (d + 1) % 3, d2 = (d + 2) % 3;
                accr[d] = xh[d1] * vr[d2] - xh[d2] * vr[d1];
                acci[d] = xh[d1] * vi[d2] - xh[d2] * vi[d1];
            }
        }
        for (int d = 0; d < KD; ++d) {
            double *p = a.partial + 2 * (((int64_t)chunk * KD + d) * a.npts + m);
            p[0] = accr[d];
            p[1] = acci[d];
        }
    }
}

template <typename OutT>
__global__ void potential_reduce_kernel(const PotArgs a, int kd) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= kd * a.npts) return;
    double sr = 0, si = 0;
    for (int c = 0; c < a.nchunk; ++c) {
        const double *p = a.partial + 2 * ((int64_t)c * kd * a.npts + i);
        sr += p[0];
        si += p[1];
    }
    OutT *o = reinterpret_cast<OutT *>(a.out);
    o[2 * (int64_t)i] = (OutT)sr;
    o[2 * (int64_t)i + 1] = (OutT)si;
}

static void plan(int npts, int nsrc, int *nchunk, int *chunk_len, int *nsrc_pad) {
    *nsrc_pad = ((nsrc + kPotTile - 1) / kPotTile) * kPotTile;
    const int tblocks = max(1, (npts + kPotThreads - 1) / kPotThreads);  // no points: one block's plan
    const int ntiles = max(1, *nsrc_pad / kPotTile);                     // no sources: one empty chunk
    int want = (4096 + tblocks - 1) / tblocks;  // aim at >= 4096 workgroups
    if (want > ntiles) want = ntiles;
    if (want > 65535) want = 65535;
    if (want < 1) want = 1;
    int tiles_per_chunk = (ntiles + want - 1) / want;
    *chunk_len = tiles_per_chunk * kPotTile;
    *nchunk = (*nsrc_pad + *chunk_len - 1) / *chunk_len;
}

static bool is_far_field_op(int op) {
    return op == BEM_HELMHOLTZ_FAR_FIELD_SINGLE_LAYER || op == BEM_HELMHOLTZ_FAR_FIELD_DOUBLE_LAYER ||
           op == BEM_MAXWELL_ELECTRIC_FAR_FIELD || op == BEM_MAXWELL_MAGNETIC_FAR_FIELD;
}

static bool is_potential_op(int op) {
    return op == BEM_LAPLACE_SINGLE_LAYER || op == BEM_LAPLACE_DOUBLE_LAYER || op == BEM_HELMHOLTZ_SINGLE_LAYER ||
           op == BEM_HELMHOLTZ_DOUBLE_LAYER || op == BEM_MAXWELL_ELECTRIC_FIELD || op == BEM_MAXWELL_MAGNETIC_FIELD ||
           is_far_field_op(op);
}

// the operators whose sources are RWG fields (t1, t2) and whose values are 3-vectors
static bool is_vector_potential_op(int op) {
    return op == BEM_MAXWELL_ELECTRIC_FIELD || op == BEM_MAXWELL_ELECTRIC_FAR_FIELD ||
           op == BEM_MAXWELL_MAGNETIC_FAR_FIELD || op == BEM_MAXWELL_MAGNETIC_FIELD;
}

// FF < 0: the potential of OP; FF >= 0: far field FF (sources prepared as for OP)
template <int OP, int FF = -1>
static int run_potential(PotArgs &a, int precision, int out_dtype, hipStream_t s) {
    int kd = OP == kMaxwellEFIE ? 3 : 1;
    int rc;
    // an empty space has no sources: only the reduction runs (it writes zeros)
    if (a.nsrc > 0) {
        hipLaunchKernelGGL((potential_prep_kernel<OP>), dim3((a.nsrc_pad + 255) / 256), dim3(256), 0, s, a);
        if ((rc = check_launch("potential_prep_kernel"))) return rc;
        dim3 grid((a.npts + kPotThreads - 1) / kPotThreads, a.nchunk);
        if constexpr (FF >= 0) {
            if (precision == BEM_PRECISION_MIXED)
                hipLaunchKernelGGL((far_field_eval_mixed<FF>), grid, dim3(kPotThreads), 0, s, a);
            else
                hipLaunchKernelGGL((far_field_eval_fp64<FF>), grid, dim3(kPotThreads), 0, s, a);
        }
        else if (precision == BEM_PRECISION_MIXED && OP != kMaxwellEFIE) {
            if constexpr (OP == kHelmholtzSL || OP == kHelmholtzDL) {
                if (a.ki != 0.0)
                    hipLaunchKernelGGL((potential_eval_mixed_scalar<OP, true>), grid, dim3(kPotThreads), 0, s, a);
                else
                    hipLaunchKernelGGL((potential_eval_mixed_scalar<OP, false>), grid, dim3(kPotThreads), 0, s, a);
            } else if constexpr (OP != kMaxwellEFIE) {
                hipLaunchKernelGGL((potential_eval_mixed_scalar<OP, false>), grid, dim3(kPotThreads), 0, s, a);
            }
        } else if (precision == BEM_PRECISION_MIXED) {
            if (a.ki != 0.0)
                hipLaunchKernelGGL((potential_eval_mixed_efie<true>), grid, dim3(kPotThreads), 0, s, a);
            else
                hipLaunchKernelGGL((potential_eval_mixed_efie<false>), grid, dim3(kPotThreads), 0, s, a);
        }
        else
            hipLaunchKernelGGL((potential_eval_fp64<OP>), grid, dim3(kPotThreads), 0, s, a);
        if ((rc = check_launch("potential_eval"))) return rc;
    } else {
        a.nchunk = 0;
    }
    int n = kd * a.npts;
    if (out_dtype == BEM_C128)
        hipLaunchKernelGGL((potential_reduce_kernel<double>), dim3((n + 255) / 256), dim3(256), 0, s, a, kd);
    else
        hipLaunchKernelGGL((potential_reduce_kernel<float>), dim3((n + 255) / 256), dim3(256), 0, s, a, kd);
    return check_launch("potential_reduce_kernel");
}

}  // namespace bem

using namespace bem;

extern "C" size_t bem_potential_workspace_size(int op, const bem_space_t *space, const bem_quad_rule_t *rule,
                                               int n_points, int precision) {
    (void)precision;
    if (!space || !rule || n_points < 0) return 0;
    int nsrc = space->n_elements * rule->n_points, nchunk, chunk_len, nsrc_pad;
    plan(n_points, nsrc, &nchunk, &chunk_len, &nsrc_pad);
    int kd = is_vector_potential_op(op) ? 3 : 1;
    return sizeof(double) * ((size_t)kSrcFields * nsrc_pad + (size_t)2 * nchunk * kd * n_points) + 256;
}

// Checks, workspace layout and launches of bem_potential_evaluate.
static int potential_evaluate(int op, const bem_space_t *space, const void *coefficients, int n_points,
                              const double *points, const bem_quad_rule_t *rule, double k_re, double k_im,
                              int precision, void *out, int out_dtype, void *workspace, size_t workspace_bytes,
                              hipStream_t s) {
    // no points: points and out may be null (an empty tensor has no storage)
    if (!space || !rule || !coefficients || (n_points > 0 && (!points || !out)) || !workspace) {
        set_error("null argument");
        return BEM_ERR_INVALID_ARGUMENT;
    }
    if (!is_potential_op(op)) {
        set_error("operator %d has no potential", op);
        return BEM_ERR_UNSUPPORTED;
    }
    int rc;
    if ((rc = check_space(space, "potential"))) return rc;
    if (is_vector_potential_op(op) != (space->shapeset == BEM_SHAPE_RWG)) {
        set_error("the Maxwell potentials and far fields need an RWG space, scalar ones a scalar space");
        return BEM_ERR_INVALID_ARGUMENT;
    }
    if (out_dtype != BEM_C128 && out_dtype != BEM_C64) {
        set_error("potential output must be complex (BEM_C128 or BEM_C64)");
        return BEM_ERR_INVALID_ARGUMENT;
    }
    if (rule->n_points < 1 || rule->n_points > BEM_QUAD_MAX_POINTS) {
        set_error("quadrature rule must have 1..%d points", BEM_QUAD_MAX_POINTS);
        return BEM_ERR_INVALID_ARGUMENT;
    }
    size_t need = bem_potential_workspace_size(op, space, rule, n_points, precision);
    if (workspace_bytes < need) {
        set_error("workspace too small: %zu < %zu bytes", workspace_bytes, need);
        return BEM_ERR_WORKSPACE;
    }
    if (n_points == 0) return BEM_OK;
    PotArgs a;
    a.sp = to_args(space);
    a.rule = to_args_wide(rule);
    a.coeffs = (const double *)coefficients;
    a.npts = n_points;
    a.pts = points;
    a.kr = (op == BEM_LAPLACE_SINGLE_LAYER || op == BEM_LAPLACE_DOUBLE_LAYER) ? 0.0 : k_re;
    a.ki = (op == BEM_LAPLACE_SINGLE_LAYER || op == BEM_LAPLACE_DOUBLE_LAYER) ? 0.0 : k_im;
    a.nsrc = space->n_elements * rule->n_points;
    plan(n_points, a.nsrc, &a.nchunk, &a.chunk_len, &a.nsrc_pad);
    uintptr_t base = ((uintptr_t)workspace + 255) & ~(uintptr_t)255;
    a.src = (double *)base;
    a.partial = a.src + (size_t)kSrcFields * a.nsrc_pad;
    a.out = out;
    a.ser = efie_series(a.kr);
    switch (op) {
    case BEM_LAPLACE_SINGLE_LAYER: return run_potential<kLaplaceSL>(a, precision, out_dtype, s);
    case BEM_LAPLACE_DOUBLE_LAYER: return run_potential<kLaplaceDL>(a, precision, out_dtype, s);
    case BEM_HELMHOLTZ_SINGLE_LAYER: return run_potential<kHelmholtzSL>(a, precision, out_dtype, s);
    case BEM_HELMHOLTZ_DOUBLE_LAYER: return run_potential<kHelmholtzDL>(a, precision, out_dtype, s);
    case BEM_HELMHOLTZ_FAR_FIELD_SINGLE_LAYER: return run_potential<kHelmholtzSL, kFarSL>(a, precision, out_dtype, s);
    case BEM_HELMHOLTZ_FAR_FIELD_DOUBLE_LAYER: return run_potential<kHelmholtzSL, kFarDL>(a, precision, out_dtype, s);
    case BEM_MAXWELL_ELECTRIC_FAR_FIELD: return run_potential<kMaxwellEFIE, kFarE>(a, precision, out_dtype, s);
    case BEM_MAXWELL_MAGNETIC_FAR_FIELD: return run_potential<kMaxwellEFIE, kFarH>(a, precision, out_dtype, s);
    case BEM_MAXWELL_MAGNETIC_FIELD: return run_potential<kMaxwellEFIE, kMagneticPotential>(a, precision, out_dtype, s);
    default: return run_potential<kMaxwellEFIE>(a, precision, out_dtype, s);
    }
}

extern "C" int bem_potential_evaluate(int op, const bem_space_t *space, const void *coefficients, int n_points,
                                      const double *points, const bem_quad_rule_t *rule, double k_re, double k_im,
                                      int precision, void *out, int out_dtype, void *workspace, size_t workspace_bytes,
                                      void *stream) {
    clear_error();
    return potential_evaluate(op, space, coefficients, n_points, points, rule, k_re, k_im, precision, out, out_dtype,
                              workspace, workspace_bytes, (hipStream_t)stream);
}

// ------------------------------------------------------------------ allocation-free variant
// bem_potential_evaluate_noalloc: the workspace of bem_potential_evaluate lives in this
// translation unit's device array (one copy per GPU, held until the process ends). Its device
// address is looked up once per device with hipGetSymbolAddress; host code never passes the
// array's own name to a kernel (on the host that name is a stand-in, not a device address).
// Nothing in the workspace needs to start at zero: the prep kernel writes every source slot
// (padding included) and the eval kernels write every (chunk, target) partial. Calls that
// overlap on one GPU share the array and race.
namespace {

static __device__ __align__(256) unsigned char scratch[BEM_POTENTIAL_SCRATCH_BYTES];

constexpr int kMaxScratchDevices = 64;
void *g_scratch_ptr[kMaxScratchDevices];
std::mutex g_scratch_mutex;

// device address of `scratch` on the current device (looked up on the first call per device)
int scratch_pointer(void **out) {
    int dev = -1;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess || dev < 0 || dev >= kMaxScratchDevices) {
        set_error("no usable current device for the potential scratch buffer (%s)", hipGetErrorString(e));
        return BEM_ERR_DEVICE;
    }
    std::lock_guard<std::mutex> lock(g_scratch_mutex);
    if (!g_scratch_ptr[dev]) {
        void *p = nullptr;
        e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p) {
            set_error("hipGetSymbolAddress(scratch) failed on device %d: %s", dev, hipGetErrorString(e));
            return BEM_ERR_DEVICE;
        }
        g_scratch_ptr[dev] = p;
    }
    *out = g_scratch_ptr[dev];
    return BEM_OK;
}

}  // namespace

extern "C" int bem_potential_evaluate_noalloc(int op, const bem_space_t *space, const void *coefficients,
                                              int n_points, const double *points, const bem_quad_rule_t *rule,
                                              double k_re, double k_im, int precision, void *out, int out_dtype,
                                              void *stream) {
    clear_error();
    if (!space || !rule) {
        set_error("null argument");
        return BEM_ERR_INVALID_ARGUMENT;
    }
    size_t need = bem_potential_workspace_size(op, space, rule, n_points, precision);
    if (need > (size_t)BEM_POTENTIAL_SCRATCH_BYTES) {
        set_error("potential needs %zu bytes of workspace, more than the %zu-byte scratch buffer of "
                  "bem_potential_evaluate_noalloc; use bem_potential_evaluate with a caller workspace",
                  need, (size_t)BEM_POTENTIAL_SCRATCH_BYTES);
        return BEM_ERR_WORKSPACE;
    }
    void *ws = nullptr;
    int rc = scratch_pointer(&ws);
    if (rc) return rc;
    return potential_evaluate(op, space, coefficients, n_points, points, rule, k_re, k_im, precision, out, out_dtype,
                              ws, (size_t)BEM_POTENTIAL_SCRATCH_BYTES, (hipStream_t)stream);
}
