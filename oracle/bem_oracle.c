/*
 * bem_oracle.c -- TEST INFRASTRUCTURE ONLY.
 *
 * FP64 CPU restatement of the reference (bempp-cl, Numba path) dense Galerkin
 * assembly and dense potential evaluation. It is the checker the HIP library is
 * compared against and the "port" CPU baseline; no product code links or calls it.
 *
 * It is deliberately independent of the product's host code: it derives element
 * geometry, shared-vertex/edge adjacency, the triangle rule, Gauss-Legendre nodes and
 * the Sauter-Schwab rules itself. Only the dof maps (local2global / multipliers) are
 * inputs.
 *
 * Reference functions restated (paths relative to the bempp-cl source tree):
 *   geometry            bempp_cl/api/grid/grid.py:552-590
 *   adjacency ordering  bempp_cl/api/grid/grid.py:1015-1098
 *   kernels             bempp_cl/core/numba_kernels.py:176-398 (regular), 453-551 (singular)
 *   regular assembly    bempp_cl/core/numba_kernels.py:980-1074 (scalar), 1853-1969 (EFIE),
 *                       1078-1333 (Laplace / Helmholtz hypersingular), 2151-2277 (MFIE)
 *   singular assembly   bempp_cl/core/numba_kernels.py:1471-1526 (scalar), 1973-2061 (EFIE),
 *                       1530-1697 (hypersingular), 2065-2147 (MFIE),
 *                       bempp_cl/core/singular_assembler.py:69-448, dense_assembler.py:78-98
 *   potentials          bempp_cl/core/numba_kernels.py:1788-1849 (scalar), 2281-2359 (EFIE), 2363-2433 (MFIE)
 *   far fields          bempp_cl/core/numba_kernels.py:402-449 (Helmholtz), 2437-2570 (Maxwell)
 *   Piola / edge length bempp_cl/core/numba_kernels.py:96-133
 *   colouring           bempp_cl/api/space/space.py:704-733 (per-colour parallel loop,
 *                       bempp_cl/core/numba_assemblers.py:90-112)
 *   quadrature          bempp_cl/api/integration/triangle_gauss.py:2928-2940,
 *                       gauss.py:1042-1055, duffy_galerkin.py:25-285
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

enum { OP_LAPLACE_SL = 0, OP_LAPLACE_DL, OP_LAPLACE_ADL, OP_HELMHOLTZ_SL, OP_HELMHOLTZ_DL, OP_HELMHOLTZ_ADL, OP_MAXWELL_EFIE,
       OP_LAPLACE_HS, OP_HELMHOLTZ_HS, OP_HELMHOLTZ_FF_SL, OP_HELMHOLTZ_FF_DL, OP_MAXWELL_E_FF, OP_MAXWELL_H_FF,
       OP_MAXWELL_MFIE };
enum { SHAPE_P0 = 0, SHAPE_P1 = 1, SHAPE_RWG = 2 };

#define M_INV_4PI_ 0.079577471545947667884
#define PI_ 3.14159265358979323846
#define MAXQ 80

typedef struct { double re, im; } cplx;

static inline cplx cmul(cplx a, cplx b) { cplx r = {a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re}; return r; }

/* ------------------------------------------------------------------ geometry */
typedef struct {
    double v0[3], e1[3], e2[3]; /* corner 0 and jacobian columns */
    double n[3];                /* unit normal */
    double intel;               /* |e1 x e2| */
    double len[3];              /* |v0-v1|, |v2-v0|, |v1-v2| */
    int vid[3];
} Elem;

static void build_elements(int nf, const double *vtx, const int32_t *elm, Elem *out) {
    for (int e = 0; e < nf; ++e) {
        const double *a = vtx + 3 * elm[3 * e], *b = vtx + 3 * elm[3 * e + 1], *c = vtx + 3 * elm[3 * e + 2];
        Elem *g = out + e;
        for (int d = 0; d < 3; ++d) { g->v0[d] = a[d]; g->e1[d] = b[d] - a[d]; g->e2[d] = c[d] - a[d]; g->vid[d] = elm[3 * e + d]; }
        double cx = g->e1[1] * g->e2[2] - g->e1[2] * g->e2[1];
        double cy = g->e1[2] * g->e2[0] - g->e1[0] * g->e2[2];
        double cz = g->e1[0] * g->e2[1] - g->e1[1] * g->e2[0];
        double nn = sqrt(cx * cx + cy * cy + cz * cz);
        g->n[0] = cx / nn; g->n[1] = cy / nn; g->n[2] = cz / nn;
        g->intel = nn;
        double d01 = 0, d20 = 0, d12 = 0;
        for (int d = 0; d < 3; ++d) {
            d01 += (a[d] - b[d]) * (a[d] - b[d]);
            d20 += (c[d] - a[d]) * (c[d] - a[d]);
            d12 += (b[d] - c[d]) * (b[d] - c[d]);
        }
        g->len[0] = sqrt(d01); g->len[1] = sqrt(d20); g->len[2] = sqrt(d12);
    }
}

static inline void map_point(const Elem *g, double x, double y, double *p) {
    for (int d = 0; d < 3; ++d) p[d] = g->v0[d] + g->e1[d] * x + g->e2[d] * y;
}

/* shapeset values (nshape scalars, or RWG Piola-mapped 3-vectors) at reference point (x, y) */
static inline void scalar_shape(int shape, double x, double y, double *phi) {
    if (shape == SHAPE_P0) { phi[0] = 1.0; }
    else { phi[0] = 1.0 - x - y; phi[1] = x; phi[2] = y; }
}

static inline void rwg_piola(const Elem *g, double x, double y, double out[3][3]) {
    /* reference values (x, y-1), (x-1, y), (x, y); Piola map J ref / |J| */
    double ref[3][2] = {{x, y - 1.0}, {x - 1.0, y}, {x, y}};
    for (int i = 0; i < 3; ++i)
        for (int d = 0; d < 3; ++d) out[i][d] = (g->e1[d] * ref[i][0] + g->e2[d] * ref[i][1]) / g->intel;
}

/* surface curls n x grad(phi_i) of the three P1 shape functions, times the normal multiplier:
 * grad = J (J^T J)^{-1} [[-1, 1, 0], [-1, 0, 1]] (bempp_cl/api/grid/grid.py:585-590,
 * bempp_cl/core/numba_kernels.py:1120-1147) */
static void surface_curls(const Elem *g, double nm, double curl[3][3]) {
    double a = 0, b = 0, c = 0;
    for (int d = 0; d < 3; ++d) { a += g->e1[d] * g->e1[d]; b += g->e1[d] * g->e2[d]; c += g->e2[d] * g->e2[d]; }
    double det = a * c - b * b;
    double i00 = c / det, i01 = -b / det, i11 = a / det;
    double jit[3][2];
    for (int d = 0; d < 3; ++d) {
        jit[d][0] = g->e1[d] * i00 + g->e2[d] * i01;
        jit[d][1] = g->e1[d] * i01 + g->e2[d] * i11;
    }
    static const double rg[3][2] = {{-1, -1}, {1, 0}, {0, 1}};
    for (int i = 0; i < 3; ++i) {
        double gr[3];
        for (int d = 0; d < 3; ++d) gr[d] = jit[d][0] * rg[i][0] + jit[d][1] * rg[i][1];
        curl[i][0] = (g->n[1] * gr[2] - g->n[2] * gr[1]) * nm;
        curl[i][1] = (g->n[2] * gr[0] - g->n[0] * gr[2]) * nm;
        curl[i][2] = (g->n[0] * gr[1] - g->n[1] * gr[0]) * nm;
    }
}

/* hypersingular weight of (i, j) at a point pair: curl_i . curl_j - k^2 phi_i psi_j (n_x . n_y)
 * (bempp_cl/core/numba_kernels.py:1284-1299); Laplace: the curl product only */
static inline cplx hs_weight(int op, double ct[3][3], double cs[3][3], const double *tn, const double *sn,
                             const double *phi, const double *psi, int i, int j, double kr, double ki) {
    cplx w = {ct[i][0] * cs[j][0] + ct[i][1] * cs[j][1] + ct[i][2] * cs[j][2], 0.0};
    if (op == OP_HELMHOLTZ_HS) {
        double nn = tn[0] * sn[0] + tn[1] * sn[1] + tn[2] * sn[2];
        double pp = phi[i] * psi[j] * nn;
        w.re -= (kr * kr - ki * ki) * pp;
        w.im -= 2.0 * kr * ki * pp;
    }
    return w;
}

/* loc[i][j] += g (ik r - 1) / r^2 (x - y) . (tv_i x sv_j), the MFIE integrand at one point pair */
static void mfie_accumulate(const double *x, const double *y, double tv[3][3], double sv[3][3], cplx g, cplx ik,
                            cplx loc[3][3]) {
    double diff[3] = {x[0] - y[0], x[1] - y[1], x[2] - y[2]};
    double r2 = diff[0] * diff[0] + diff[1] * diff[1] + diff[2] * diff[2], r = sqrt(r2);
    cplx f = {(ik.re * r - 1.0) / r2, ik.im * r / r2};
    cplx gf = cmul(g, f);
    for (int i = 0; i < 3; ++i) for (int j = 0; j < 3; ++j) {
        double c0 = tv[i][1] * sv[j][2] - tv[i][2] * sv[j][1];
        double c1 = tv[i][2] * sv[j][0] - tv[i][0] * sv[j][2];
        double c2 = tv[i][0] * sv[j][1] - tv[i][1] * sv[j][0];
        double d = diff[0] * c0 + diff[1] * c1 + diff[2] * c2;
        loc[i][j].re += gf.re * d; loc[i][j].im += gf.im * d;
    }
}

/* ------------------------------------------------------------------ kernels */
/* G(x, y) for the operator. nx / ny are the (multiplied) test / trial normals. */
static inline cplx kernel_eval(int op, const double *x, const double *y, const double *nx, const double *ny, double kr, double ki) {
    double diff[3] = {y[0] - x[0], y[1] - x[1], y[2] - x[2]};
    double r2 = diff[0] * diff[0] + diff[1] * diff[1] + diff[2] * diff[2];
    double r = sqrt(r2);
    cplx out = {0.0, 0.0};
    switch (op) {
    case OP_LAPLACE_SL:
    case OP_LAPLACE_HS: out.re = M_INV_4PI_ / r; break;
    case OP_LAPLACE_DL: out.re = -M_INV_4PI_ * (diff[0] * ny[0] + diff[1] * ny[1] + diff[2] * ny[2]) / (r * r * r); break;
    case OP_LAPLACE_ADL: out.re = M_INV_4PI_ * (diff[0] * nx[0] + diff[1] * nx[1] + diff[2] * nx[2]) / (r * r * r); break;
    case OP_HELMHOLTZ_SL:
    case OP_HELMHOLTZ_HS:
    case OP_MAXWELL_MFIE:
    case OP_MAXWELL_EFIE: {
        double damp = exp(-ki * r);
        out.re = cos(kr * r) * M_INV_4PI_ / r * damp;
        out.im = sin(kr * r) * M_INV_4PI_ / r * damp;
        break;
    }
    case OP_HELMHOLTZ_DL:
    case OP_HELMHOLTZ_ADL: {
        double g;
        if (op == OP_HELMHOLTZ_DL) g = (diff[0] * ny[0] + diff[1] * ny[1] + diff[2] * ny[2]);
        else g = -(diff[0] * nx[0] + diff[1] * nx[1] + diff[2] * nx[2]); /* (x - y) . n_x */
        g *= M_INV_4PI_ / (r * r * r);
        double damp = exp(-ki * r);
        double fr = cos(kr * r) * g * damp, fi = sin(kr * r) * g * damp;
        /* (i k r - 1) e^{ikr} grad, with i k r = -ki r + i kr r */
        out.re = (-1.0 - ki * r) * fr - kr * r * fi;
        out.im = kr * r * fr + (-1.0 - ki * r) * fi;
        break;
    }
    }
    return out;
}

/* ------------------------------------------------------------------ quadrature */
/* Rules of orders 6-20 are data, not closed forms (the reference's HyENA table,
 * bempp_cl/api/integration/triangle_gauss.py:175-2925): the caller passes them through
 * bem_oracle_set_regular_rule and selects them with order 0. */
static int g_user_n = 0;
static double g_user_x[MAXQ], g_user_y[MAXQ], g_user_w[MAXQ];

static int triangle_rule(int order, double *px, double *py, double *w) {
    if (order == 0) {
        if (g_user_n < 1) return -1;
        for (int i = 0; i < g_user_n; ++i) { px[i] = g_user_x[i]; py[i] = g_user_y[i]; w[i] = g_user_w[i]; }
        return g_user_n;
    }
    double b[8][3], wt[8];
    int n = 0;
#define S21(a, ww) do { double a_ = (a); double c_ = 1.0 - 2.0 * a_; \
        b[n][0] = c_; b[n][1] = a_; b[n][2] = a_; wt[n++] = (ww); \
        b[n][0] = a_; b[n][1] = c_; b[n][2] = a_; wt[n++] = (ww); \
        b[n][0] = a_; b[n][1] = a_; b[n][2] = c_; wt[n++] = (ww); } while (0)
    if (order == 1) { b[0][0] = b[0][1] = b[0][2] = 1.0 / 3; wt[0] = 1.0; n = 1; }
    else if (order == 2) { S21(1.0 / 6, 1.0 / 3); }
    else if (order == 3) { b[0][0] = b[0][1] = b[0][2] = 1.0 / 3; wt[0] = -27.0 / 48; n = 1; S21(0.2, 25.0 / 48); }
    else if (order == 4) {
        /* Dunavant degree-4 rule; reference row order: A0, B0, A1, A2, B1, B2 */
        double tb[6][3]; double tw[6];
        n = 0; S21(0.44594849091596488632, 0.22338158967801146570);
        S21(0.091576213509770743460, 0.10995174365532186764);
        int perm[6] = {0, 3, 1, 2, 4, 5};
        for (int i = 0; i < 6; ++i) { memcpy(tb[i], b[perm[i]], sizeof(tb[i])); tw[i] = wt[perm[i]]; }
        memcpy(b, tb, sizeof(tb)); memcpy(wt, tw, sizeof(tw));
    } else if (order == 5) {
        double s15 = sqrt(15.0);
        b[0][0] = b[0][1] = b[0][2] = 1.0 / 3; wt[0] = 9.0 / 40; n = 1;
        S21((6.0 + s15) / 21.0, (155.0 + s15) / 1200.0);
        S21((6.0 - s15) / 21.0, (155.0 - s15) / 1200.0);
        double tb[7][3]; double tw[7]; int perm[7] = {0, 1, 4, 2, 3, 5, 6};
        for (int i = 0; i < 7; ++i) { memcpy(tb[i], b[perm[i]], sizeof(tb[i])); tw[i] = wt[perm[i]]; }
        memcpy(b, tb, sizeof(tb)); memcpy(wt, tw, sizeof(tw));
    } else return -1;
#undef S21
    for (int i = 0; i < n; ++i) { px[i] = b[i][1]; py[i] = b[i][2]; w[i] = 0.5 * wt[i]; }
    return n;
}

/* Gauss-Legendre on [0,1], nodes in descending order (Newton on P_n) */
static void gauss01(int n, double *x, double *w) {
    for (int i = 0; i < n; ++i) {
        double z = cos(PI_ * (i + 0.75) / (n + 0.5)), pp = 0;
        for (int it = 0; it < 100; ++it) {
            double p1 = 1.0, p2 = 0.0;
            for (int j = 1; j <= n; ++j) { double p3 = p2; p2 = p1; p1 = ((2.0 * j - 1.0) * z * p2 - (j - 1.0) * p3) / j; }
            pp = n * (z * p1 - p2) / (z * z - 1.0);
            double dz = p1 / pp;
            z -= dz;
            if (fabs(dz) < 1e-16) break;
        }
        x[i] = 0.5 * (1.0 + z);
        w[i] = 1.0 / ((1.0 - z * z) * pp * pp);
    }
}

typedef struct { int n; double *tx, *ty, *sx, *sy, *w; } SingRule;

static void sing_rule_alloc(SingRule *r, int n) {
    r->n = n;
    r->tx = malloc(sizeof(double) * n); r->ty = malloc(sizeof(double) * n);
    r->sx = malloc(sizeof(double) * n); r->sy = malloc(sizeof(double) * n);
    r->w = malloc(sizeof(double) * n);
}
static void sing_rule_free(SingRule *r) { free(r->tx); free(r->ty); free(r->sx); free(r->sy); free(r->w); }

/* Sauter-Schwab rule; kind 0 coincident (6 regions), 1 edge (5), 2 vertex (2) */
static void duffy_rule(int order, int kind, SingRule *r) {
    double gx[64], gw[64];
    gauss01(order, gx, gw);
    int n = order, nreg = kind == 0 ? 6 : (kind == 1 ? 5 : 2);
    sing_rule_alloc(r, nreg * n * n * n * n);
    int idx = 0;
    for (int ti = 0; ti < n * n; ++ti)
        for (int si = 0; si < n * n; ++si) {
            double xsi = gx[ti % n], e1 = gx[ti / n], e2 = gx[si % n], e3 = gx[si / n];
            double wb = gw[ti % n] * gw[ti / n] * gw[si % n] * gw[si / n];
            double e12 = e1 * e2, e123 = e12 * e3;
            double P[6][5];
            if (kind == 0) {
                double wgt = wb * xsi * xsi * xsi * e1 * e1 * e2;
                double T[6][4] = {
                    {xsi, xsi * (1 - e1 + e12), xsi * (1 - e123), xsi * (1 - e1)},
                    {xsi * (1 - e123), xsi * (1 - e1), xsi, xsi * (1 - e1 + e12)},
                    {xsi, xsi * (e1 - e12 + e123), xsi * (1 - e12), xsi * (e1 - e12)},
                    {xsi * (1 - e12), xsi * (e1 - e12), xsi, xsi * (e1 - e12 + e123)},
                    {xsi * (1 - e123), xsi * (e1 - e123), xsi, xsi * (e1 - e12)},
                    {xsi, xsi * (e1 - e12), xsi * (1 - e123), xsi * (e1 - e123)}};
                for (int k = 0; k < 6; ++k) { memcpy(P[k], T[k], 4 * sizeof(double)); P[k][4] = wgt; }
            } else if (kind == 1) {
                double wgt = wb * xsi * xsi * xsi * e1 * e1;
                double T[5][4] = {
                    {xsi, xsi * e1 * e3, xsi * (1 - e12), xsi * e1 * (1 - e2)},
                    {xsi, xsi * e1, xsi * (1 - e123), xsi * e12 * (1 - e3)},
                    {xsi * (1 - e12), xsi * e1 * (1 - e2), xsi, xsi * e123},
                    {xsi * (1 - e123), xsi * e12 * (1 - e3), xsi, xsi * e1},
                    {xsi * (1 - e123), xsi * e1 * (1 - e2 * e3), xsi, xsi * e12}};
                for (int k = 0; k < 5; ++k) { memcpy(P[k], T[k], 4 * sizeof(double)); P[k][4] = k == 0 ? wgt : wgt * e2; }
            } else {
                double wgt = wb * xsi * xsi * xsi * e2;
                double T[2][4] = {{xsi, xsi * e1, xsi * e2, xsi * e2 * e3}, {xsi * e2, xsi * e2 * e3, xsi, xsi * e1}};
                for (int k = 0; k < 2; ++k) { memcpy(P[k], T[k], 4 * sizeof(double)); P[k][4] = wgt; }
            }
            for (int k = 0; k < nreg; ++k) {
                r->tx[idx] = P[k][0] - P[k][1]; r->ty[idx] = P[k][1];
                r->sx[idx] = P[k][2] - P[k][3]; r->sy[idx] = P[k][3];
                r->w[idx] = P[k][4];
                ++idx;
            }
        }
}

static const double REFV[3][2] = {{0, 0}, {1, 0}, {0, 1}};

static inline void remap_edge(double x, double y, int v0, int v1, double *ox, double *oy) {
    int v2 = 3 - v0 - v1;
    double ax = REFV[v1][0] - REFV[v0][0], ay = REFV[v1][1] - REFV[v0][1];
    double bx = REFV[v2][0] - REFV[v0][0], by = REFV[v2][1] - REFV[v0][1];
    *ox = REFV[v0][0] + ax * x + bx * y;
    *oy = REFV[v0][1] + ay * x + by * y;
}

static inline void remap_vertex(double x, double y, int v, double *ox, double *oy) {
    if (v == 0) { *ox = x; *oy = y; }
    else if (v == 1) { *ox = 1.0 - x - y; *oy = y; }
    else { *ox = x; *oy = 1.0 - x - y; }
}

/* ------------------------------------------------------------------ colouring */
/* Greedy colouring of support elements: elements sharing a dof get different colours. */
static int colour_elements(int nf, int nshape, const int32_t *l2g, const double *mult, const uint8_t *support, int ndofs, int *colour) {
    /* dof -> elements (CSR) */
    int *cnt = calloc(ndofs + 1, sizeof(int));
    for (int e = 0; e < nf; ++e) if (support[e]) for (int i = 0; i < nshape; ++i) if (mult[e * nshape + i] != 0) cnt[l2g[e * nshape + i] + 1]++;
    for (int d = 0; d < ndofs; ++d) cnt[d + 1] += cnt[d];
    int *lst = malloc(sizeof(int) * (cnt[ndofs] + 1)), *fill = calloc(ndofs, sizeof(int));
    for (int e = 0; e < nf; ++e) if (support[e]) for (int i = 0; i < nshape; ++i) if (mult[e * nshape + i] != 0) {
        int d = l2g[e * nshape + i];
        lst[cnt[d] + fill[d]++] = e;
    }
    int ncol = 0;
    char *used = calloc(nf + 1, 1);
    for (int e = 0; e < nf; ++e) colour[e] = -1;
    for (int e = 0; e < nf; ++e) {
        if (!support[e]) continue;
        for (int i = 0; i < nshape; ++i) {
            int d = l2g[e * nshape + i];
            for (int k = cnt[d]; k < cnt[d + 1]; ++k) { int o = lst[k]; if (o != e && colour[o] >= 0) used[colour[o]] = 1; }
        }
        int c = 0;
        while (used[c]) ++c;
        colour[e] = c;
        if (c + 1 > ncol) ncol = c + 1;
        for (int i = 0; i < nshape; ++i) {
            int d = l2g[e * nshape + i];
            for (int k = cnt[d]; k < cnt[d + 1]; ++k) { int o = lst[k]; if (colour[o] >= 0) used[colour[o]] = 0; }
        }
    }
    free(cnt); free(lst); free(fill); free(used);
    return ncol;
}

static inline int share_vertex(const Elem *a, const Elem *b) {
    for (int i = 0; i < 3; ++i) for (int j = 0; j < 3; ++j) if (a->vid[i] == b->vid[j]) return 1;
    return 0;
}

static inline void accumulate(double *out, int cols, int is_complex, int row, int col, cplx v) {
    if (is_complex) { out[2 * ((int64_t)row * cols + col)] += v.re; out[2 * ((int64_t)row * cols + col) + 1] += v.im; }
    else out[(int64_t)row * cols + col] += v.re;
}

/* local 3x3 block of one regular (non-adjacent) pair */
static void regular_pair(int op, const Elem *te, const Elem *se, double tnm, double snm, int tshape, int sshape,
                         int nq, const double *qx, const double *qy, const double *qw, double kr, double ki, cplx loc[3][3]) {
    double tn[3] = {te->n[0] * tnm, te->n[1] * tnm, te->n[2] * tnm};
    double sn[3] = {se->n[0] * snm, se->n[1] * snm, se->n[2] * snm};
    for (int i = 0; i < 3; ++i) for (int j = 0; j < 3; ++j) loc[i][j].re = loc[i][j].im = 0;
    cplx ik = {-ki, kr};
    const int hs = op == OP_LAPLACE_HS || op == OP_HELMHOLTZ_HS;
    double ct[3][3], cs[3][3];
    if (hs) { surface_curls(te, tnm, ct); surface_curls(se, snm, cs); }
    for (int p = 0; p < nq; ++p) {
        double x[3]; map_point(te, qx[p], qy[p], x);
        double phi[3] = {0, 0, 0}, tv[3][3] = {{0}};
        if (op == OP_MAXWELL_EFIE || op == OP_MAXWELL_MFIE) rwg_piola(te, qx[p], qy[p], tv);
        else scalar_shape(tshape, qx[p], qy[p], phi);
        for (int q = 0; q < nq; ++q) {
            double y[3]; map_point(se, qx[q], qy[q], y);
            cplx g = kernel_eval(op, x, y, tn, sn, kr, ki);
            double f = qw[p] * qw[q] * te->intel * se->intel;
            g.re *= f; g.im *= f;
            if (op == OP_MAXWELL_MFIE) {
                /* G (ik r - 1) / r^2 (x - y) . (phi_i x psi_j) (numba_kernels.py:2226-2262) */
                double sv[3][3]; rwg_piola(se, qx[q], qy[q], sv);
                mfie_accumulate(x, y, tv, sv, g, ik, loc);
            } else if (op == OP_MAXWELL_EFIE) {
                double sv[3][3]; rwg_piola(se, qx[q], qy[q], sv);
                double divp = 4.0 / (te->intel * se->intel);
                /* -(ik)(phi.psi) - divp / (ik);  1/(ik) = conj(ik)/|ik|^2 */
                double den = ik.re * ik.re + ik.im * ik.im;
                cplx inv_ik = {ik.re / den, -ik.im / den};
                for (int i = 0; i < 3; ++i) for (int j = 0; j < 3; ++j) {
                    double dot = tv[i][0] * sv[j][0] + tv[i][1] * sv[j][1] + tv[i][2] * sv[j][2];
                    cplx fac = {-ik.re * dot - divp * inv_ik.re, -ik.im * dot - divp * inv_ik.im};
                    cplx c = cmul(g, fac);
                    loc[i][j].re += c.re; loc[i][j].im += c.im;
                }
            } else if (hs) {
                double psi[3] = {0, 0, 0}; scalar_shape(sshape, qx[q], qy[q], psi);
                for (int i = 0; i < 3; ++i) for (int j = 0; j < 3; ++j) {
                    cplx c = cmul(g, hs_weight(op, ct, cs, tn, sn, phi, psi, i, j, kr, ki));
                    loc[i][j].re += c.re; loc[i][j].im += c.im;
                }
            } else {
                double psi[3]; scalar_shape(sshape, qx[q], qy[q], psi);
                int nt = tshape == SHAPE_P0 ? 1 : 3, ns = sshape == SHAPE_P0 ? 1 : 3;
                for (int i = 0; i < nt; ++i) for (int j = 0; j < ns; ++j) {
                    loc[i][j].re += g.re * phi[i] * psi[j];
                    loc[i][j].im += g.im * phi[i] * psi[j];
                }
            }
        }
    }
}

/* local block of one singular pair; (tx,ty),(sx,sy) already remapped, w weights */
static void singular_pair(int op, const Elem *te, const Elem *se, double tnm, double snm, int tshape, int sshape, int np,
                          const double *tx, const double *ty, const double *sx, const double *sy, const double *w,
                          double kr, double ki, cplx loc[3][3]) {
    double tn[3] = {te->n[0] * tnm, te->n[1] * tnm, te->n[2] * tnm};
    double sn[3] = {se->n[0] * snm, se->n[1] * snm, se->n[2] * snm};
    for (int i = 0; i < 3; ++i) for (int j = 0; j < 3; ++j) loc[i][j].re = loc[i][j].im = 0;
    cplx ik = {-ki, kr};
    double den = ik.re * ik.re + ik.im * ik.im;
    cplx inv_ik = {ik.re / den, -ik.im / den};
    const int hs = op == OP_LAPLACE_HS || op == OP_HELMHOLTZ_HS;
    double ct[3][3], cs[3][3];
    if (hs) { surface_curls(te, tnm, ct); surface_curls(se, snm, cs); }
    for (int k = 0; k < np; ++k) {
        double x[3], y[3];
        map_point(te, tx[k], ty[k], x);
        map_point(se, sx[k], sy[k], y);
        cplx g = kernel_eval(op, x, y, tn, sn, kr, ki);
        g.re *= w[k]; g.im *= w[k];
        if (op == OP_MAXWELL_MFIE) {
            double tv[3][3], sv[3][3];
            rwg_piola(te, tx[k], ty[k], tv); rwg_piola(se, sx[k], sy[k], sv);
            mfie_accumulate(x, y, tv, sv, g, ik, loc);
        } else if (op == OP_MAXWELL_EFIE) {
            double tv[3][3], sv[3][3];
            rwg_piola(te, tx[k], ty[k], tv); rwg_piola(se, sx[k], sy[k], sv);
            double divp = 4.0 / (te->intel * se->intel);
            for (int i = 0; i < 3; ++i) for (int j = 0; j < 3; ++j) {
                double dot = tv[i][0] * sv[j][0] + tv[i][1] * sv[j][1] + tv[i][2] * sv[j][2];
                cplx fac = {-ik.re * dot - divp * inv_ik.re, -ik.im * dot - divp * inv_ik.im};
                cplx c = cmul(g, fac);
                loc[i][j].re += c.re; loc[i][j].im += c.im;
            }
        } else if (hs) {
            double phi[3] = {0, 0, 0}, psi[3] = {0, 0, 0};
            scalar_shape(tshape, tx[k], ty[k], phi);
            scalar_shape(sshape, sx[k], sy[k], psi);
            for (int i = 0; i < 3; ++i) for (int j = 0; j < 3; ++j) {
                cplx c = cmul(g, hs_weight(op, ct, cs, tn, sn, phi, psi, i, j, kr, ki));
                loc[i][j].re += c.re; loc[i][j].im += c.im;
            }
        } else {
            double phi[3], psi[3];
            scalar_shape(tshape, tx[k], ty[k], phi);
            scalar_shape(sshape, sx[k], sy[k], psi);
            int nt = tshape == SHAPE_P0 ? 1 : 3, ns = sshape == SHAPE_P0 ? 1 : 3;
            for (int i = 0; i < nt; ++i) for (int j = 0; j < ns; ++j) {
                loc[i][j].re += g.re * phi[i] * psi[j];
                loc[i][j].im += g.im * phi[i] * psi[j];
            }
        }
    }
    double jj = te->intel * se->intel;
    for (int i = 0; i < 3; ++i) for (int j = 0; j < 3; ++j) { loc[i][j].re *= jj; loc[i][j].im *= jj; }
}

/* classify a pair: 0 coincident, 1 edge, 2 vertex, -1 regular; fills the remap indices */
static int classify(const Elem *te, const Elem *se, int *ti, int *si) {
    if (te == se) return 0;
    int nsh = 0, tpos[3], spos[3];
    for (int i = 0; i < 3; ++i) for (int j = 0; j < 3; ++j) if (te->vid[i] == se->vid[j]) { tpos[nsh] = i; spos[nsh] = j; ++nsh; }
    if (nsh == 0) return -1;
    if (nsh == 1) { ti[0] = tpos[0]; si[0] = spos[0]; return 2; }
    if (nsh == 3) return 0;
    /* test positions ascending (scan order), swap if trial positions descend (grid.py:1036-1044) */
    if (spos[1] < spos[0]) { ti[0] = tpos[1]; ti[1] = tpos[0]; si[0] = spos[1]; si[1] = spos[0]; }
    else { ti[0] = tpos[0]; ti[1] = tpos[1]; si[0] = spos[0]; si[1] = spos[1]; }
    return 1;
}

typedef struct {
    int op, tshape, sshape, nt, ns;
    const int32_t *tl2g, *sl2g;
    const double *tmult, *smult;
    const int32_t *tnm, *snm;
    const uint8_t *tsup, *ssup;
} Spaces;

static int nshape_of(int shape) { return shape == SHAPE_P0 ? 1 : 3; }

/* Optional row selection (bem_oracle_set_row_map): global test dof -> row of `out`, or -1 to
 * drop the entry, so that sampled rows of a large operator need only their own storage. */
static const int32_t *g_row_map = NULL;

static void scatter_block(double *out, int cols, int is_complex, const Spaces *S, const Elem *te, const Elem *se, int e, int f, cplx loc[3][3]) {
    for (int i = 0; i < S->nt; ++i) for (int j = 0; j < S->ns; ++j) {
        int row = S->tl2g[e * S->nt + i];
        if (g_row_map) {
            row = g_row_map[row];
            if (row < 0) continue;
        }
        double m = S->tmult[e * S->nt + i] * S->smult[f * S->ns + j];
        if (S->op == OP_MAXWELL_EFIE || S->op == OP_MAXWELL_MFIE) m *= te->len[i] * se->len[j];
        cplx v = {loc[i][j].re * m, loc[i][j].im * m};
        accumulate(out, cols, is_complex, row, S->sl2g[f * S->ns + j], v);
    }
}

void bem_oracle_set_row_map(const int32_t *map) { g_row_map = map; }

/*
 * Dense assembly of one operator on one grid (test and trial spaces on the same grid).
 * out: rows x cols, row-major; interleaved (re, im) doubles when is_complex. It is added to.
 * parts: bit 0 regular pairs, bit 1 singular pairs.
 * test_begin/test_end: restrict to support test elements with index in [begin, end) (row sampling).
 * Returns 0 on success.
 */
int bem_oracle_assemble(int op, int nv, const double *vtx, int nf, const int32_t *elm,
                        int tshape, const int32_t *tl2g, const double *tmult, const int32_t *tnm, const uint8_t *tsup,
                        int sshape, const int32_t *sl2g, const double *smult, const int32_t *snm, const uint8_t *ssup,
                        int regular_order, int singular_order, double kr, double ki,
                        int rows, int cols, int is_complex, double *out, int parts,
                        int test_begin, int test_end, int nthreads) {
    (void)nv;
    double qx[MAXQ], qy[MAXQ], qw[MAXQ];
    int nq = triangle_rule(regular_order, qx, qy, qw);
    if (nq < 0) return -1;
    Elem *G = malloc(sizeof(Elem) * nf);
    build_elements(nf, vtx, elm, G);
    Spaces S = {op, tshape, sshape, nshape_of(tshape), nshape_of(sshape), tl2g, sl2g, tmult, smult, tnm, snm, tsup, ssup};
    if (op == OP_MAXWELL_EFIE || op == OP_MAXWELL_MFIE) { S.nt = 3; S.ns = 3; }
    if ((op == OP_LAPLACE_HS || op == OP_HELMHOLTZ_HS) && (tshape != SHAPE_P1 || sshape != SHAPE_P1)) return -2;
#ifdef _OPENMP
    if (nthreads > 0) omp_set_num_threads(nthreads);
#endif
    if (test_end < 0 || test_end > nf) test_end = nf;

    if (parts & 1) {
        int *colour = malloc(sizeof(int) * nf);
        int ncol = colour_elements(nf, S.nt, tl2g, tmult, tsup, rows, colour);
        int *trial = malloc(sizeof(int) * nf), ntrial = 0;
        for (int f = 0; f < nf; ++f) if (ssup[f]) trial[ntrial++] = f;
        for (int c = 0; c < ncol; ++c) {
#pragma omp parallel for schedule(dynamic, 4)
            for (int e = test_begin; e < test_end; ++e) {
                if (!tsup[e] || colour[e] != c) continue;
                cplx loc[3][3];
                for (int k = 0; k < ntrial; ++k) {
                    int f = trial[k];
                    if (share_vertex(G + e, G + f)) continue;
                    regular_pair(op, G + e, G + f, tnm[e], snm[f], tshape, sshape, nq, qx, qy, qw, kr, ki, loc);
                    scatter_block(out, cols, is_complex, &S, G + e, G + f, e, f, loc);
                }
            }
        }
        free(colour); free(trial);
    }

    if (parts & 2) {
        SingRule R[3];
        for (int k = 0; k < 3; ++k) duffy_rule(singular_order, k, &R[k]);
        int np = R[0].n;
        double *tx = malloc(sizeof(double) * np), *ty = malloc(sizeof(double) * np);
        double *sx = malloc(sizeof(double) * np), *sy = malloc(sizeof(double) * np);
        /* singular pairs are few: loop test elements serially, neighbours by brute force */
        for (int e = test_begin; e < test_end; ++e) {
            if (!tsup[e]) continue;
            for (int f = 0; f < nf; ++f) {
                if (!ssup[f]) continue;
                int ti[2] = {0, 0}, si[2] = {0, 0};
                int kind = e == f ? 0 : classify(G + e, G + f, ti, si);
                if (kind < 0) continue;
                SingRule *r = &R[kind];
                for (int k = 0; k < r->n; ++k) {
                    if (kind == 0) { tx[k] = r->tx[k]; ty[k] = r->ty[k]; sx[k] = r->sx[k]; sy[k] = r->sy[k]; }
                    else if (kind == 1) {
                        remap_edge(r->tx[k], r->ty[k], ti[0], ti[1], tx + k, ty + k);
                        remap_edge(r->sx[k], r->sy[k], si[0], si[1], sx + k, sy + k);
                    } else {
                        remap_vertex(r->tx[k], r->ty[k], ti[0], tx + k, ty + k);
                        remap_vertex(r->sx[k], r->sy[k], si[0], sx + k, sy + k);
                    }
                }
                cplx loc[3][3];
                singular_pair(op, G + e, G + f, tnm[e], snm[f], tshape, sshape, r->n, tx, ty, sx, sy, r->w, kr, ki, loc);
                scatter_block(out, cols, is_complex, &S, G + e, G + f, e, f, loc);
            }
        }
        free(tx); free(ty); free(sx); free(sy);
        for (int k = 0; k < 3; ++k) sing_rule_free(&R[k]);
    }
    free(G);
    return 0;
}

/*
 * Dense potential u(x_m) = sum_e sum_q G(x_m, y_eq) J_e w_q sum_i phi_i(q) c[e*nshape+i]
 * (numba_kernels.py:1788-1849), or the EFIE potential (numba_kernels.py:2281-2359) for
 * op == OP_MAXWELL_EFIE. coeffs: interleaved complex per (element, local shape function),
 * already multiplied by the local multipliers. out: (kdim, M) interleaved complex.
 */
int bem_oracle_potential(int op, int nv, const double *vtx, int nf, const int32_t *elm, int shape,
                         const int32_t *nm, const uint8_t *sup, const double *coeffs, int npts, const double *pts,
                         int order, double kr, double ki, double *out, int nthreads) {
    (void)nv;
    double qx[MAXQ], qy[MAXQ], qw[MAXQ];
    int nq = triangle_rule(order, qx, qy, qw);
    if (nq < 0) return -1;
    Elem *G = malloc(sizeof(Elem) * nf);
    build_elements(nf, vtx, elm, G);
    const int maxwell = op == OP_MAXWELL_EFIE || op == OP_MAXWELL_E_FF || op == OP_MAXWELL_H_FF || op == OP_MAXWELL_MFIE;
    const int far = op == OP_HELMHOLTZ_FF_SL || op == OP_HELMHOLTZ_FF_DL || op == OP_MAXWELL_E_FF || op == OP_MAXWELL_H_FF;
    int nsh = maxwell ? 3 : nshape_of(shape);
    int kdim = maxwell ? 3 : 1;
    int nsrc = 0;
    for (int e = 0; e < nf; ++e) if (sup[e]) ++nsrc;
    nsrc *= nq;
    double *sp = malloc(sizeof(double) * 3 * nsrc), *sn = malloc(sizeof(double) * 3 * nsrc);
    cplx *t2 = malloc(sizeof(cplx) * nsrc), *t1 = malloc(sizeof(cplx) * 3 * nsrc);
    int s = 0;
    for (int e = 0; e < nf; ++e) {
        if (!sup[e]) continue;
        for (int q = 0; q < nq; ++q, ++s) {
            map_point(G + e, qx[q], qy[q], sp + 3 * s);
            for (int d = 0; d < 3; ++d) sn[3 * s + d] = G[e].n[d] * nm[e];
            t2[s].re = t2[s].im = 0;
            for (int d = 0; d < 3; ++d) t1[3 * s + d].re = t1[3 * s + d].im = 0;
            if (maxwell) {
                double bv[3][3]; rwg_piola(G + e, qx[q], qy[q], bv);
                for (int i = 0; i < 3; ++i) {
                    double f = qw[q] * G[e].len[i];
                    cplx c = {coeffs[2 * (e * 3 + i)] * f, coeffs[2 * (e * 3 + i) + 1] * f};
                    for (int d = 0; d < 3; ++d) { t1[3 * s + d].re += c.re * bv[i][d] * G[e].intel; t1[3 * s + d].im += c.im * bv[i][d] * G[e].intel; }
                    t2[s].re += 2 * c.re; t2[s].im += 2 * c.im;
                }
            } else {
                double phi[3]; scalar_shape(shape, qx[q], qy[q], phi);
                for (int i = 0; i < nsh; ++i) {
                    double f = G[e].intel * qw[q] * phi[i];
                    t2[s].re += f * coeffs[2 * (e * nsh + i)];
                    t2[s].im += f * coeffs[2 * (e * nsh + i) + 1];
                }
            }
        }
    }
#ifdef _OPENMP
    if (nthreads > 0) omp_set_num_threads(nthreads);
#endif
    double zero[3] = {0, 0, 0};
    cplx ik = {-ki, kr};
#pragma omp parallel for schedule(static)
    for (int m = 0; m < npts; ++m) {
        const double *x = pts + 3 * m;
        cplx acc[3] = {{0, 0}, {0, 0}, {0, 0}};
        for (int j = 0; j < nsrc; ++j) {
            if (far) {
                /* K = e^{-i Re(k) x.y} / (4 pi), x a unit direction (numba_kernels.py:402-420) */
                const double *y = sp + 3 * j;
                double dot = x[0] * y[0] + x[1] * y[1] + x[2] * y[2];
                cplx K = {M_INV_4PI_ * cos(-kr * dot), M_INV_4PI_ * sin(-kr * dot)};
                if (op == OP_HELMHOLTZ_FF_SL) {
                    cplx c = cmul(K, t2[j]);
                    acc[0].re += c.re; acc[0].im += c.im;
                } else if (op == OP_HELMHOLTZ_FF_DL) {
                    /* factor = -Re(k) x.n_y; value = (-factor sin(-k x.y), factor cos(-k x.y)) / (4 pi) (424-449) */
                    const double *n = sn + 3 * j;
                    double factor = -kr * (x[0] * n[0] + x[1] * n[1] + x[2] * n[2]);
                    cplx v = {-factor * M_INV_4PI_ * sin(-kr * dot), factor * M_INV_4PI_ * cos(-kr * dot)};
                    cplx c = cmul(v, t2[j]);
                    acc[0].re += c.re; acc[0].im += c.im;
                } else if (op == OP_MAXWELL_E_FF) {
                    /* K (ik t1 - x t2) (2437-2500) */
                    for (int d = 0; d < 3; ++d) {
                        cplx a = cmul(ik, t1[3 * j + d]);
                        cplx term = {a.re - x[d] * t2[j].re, a.im - x[d] * t2[j].im};
                        cplx c = cmul(K, term);
                        acc[d].re += c.re; acc[d].im += c.im;
                    }
                } else {
                    /* x cross (K ik t1) (2504-2570) */
                    cplx v[3];
                    for (int d = 0; d < 3; ++d) v[d] = cmul(K, cmul(ik, t1[3 * j + d]));
                    acc[0].re += x[1] * v[2].re - x[2] * v[1].re; acc[0].im += x[1] * v[2].im - x[2] * v[1].im;
                    acc[1].re += x[2] * v[0].re - x[0] * v[2].re; acc[1].im += x[2] * v[0].im - x[0] * v[2].im;
                    acc[2].re += x[0] * v[1].re - x[1] * v[0].re; acc[2].im += x[0] * v[1].im - x[1] * v[0].im;
                }
                continue;
            }
            cplx g = kernel_eval(op, x, sp + 3 * j, zero, sn + 3 * j, kr, ki);
            if (op == OP_MAXWELL_MFIE) {
                /* diff x (G (ik r - 1) t1 / r^2), diff = x - y (numba_kernels.py:2420-2433) */
                double diff[3] = {x[0] - sp[3 * j], x[1] - sp[3 * j + 1], x[2] - sp[3 * j + 2]};
                double r2 = diff[0] * diff[0] + diff[1] * diff[1] + diff[2] * diff[2], r = sqrt(r2);
                cplx f = {(ik.re * r - 1.0) / r2, ik.im * r / r2};
                cplx gf = cmul(g, f), v[3];
                for (int d = 0; d < 3; ++d) v[d] = cmul(gf, t1[3 * j + d]);
                acc[0].re += diff[1] * v[2].re - diff[2] * v[1].re; acc[0].im += diff[1] * v[2].im - diff[2] * v[1].im;
                acc[1].re += diff[2] * v[0].re - diff[0] * v[2].re; acc[1].im += diff[2] * v[0].im - diff[0] * v[2].im;
                acc[2].re += diff[0] * v[1].re - diff[1] * v[0].re; acc[2].im += diff[0] * v[1].im - diff[1] * v[0].im;
                continue;
            }
            if (op == OP_MAXWELL_EFIE) {
                double diff[3] = {x[0] - sp[3 * j], x[1] - sp[3 * j + 1], x[2] - sp[3 * j + 2]};
                double r2 = diff[0] * diff[0] + diff[1] * diff[1] + diff[2] * diff[2], r = sqrt(r2);
                /* (ik r - 1) / (ik r^2) */
                cplx num = {ik.re * r - 1.0, ik.im * r};
                cplx den = {ik.re * r2, ik.im * r2};
                double dd = den.re * den.re + den.im * den.im;
                cplx q = {(num.re * den.re + num.im * den.im) / dd, (num.im * den.re - num.re * den.im) / dd};
                cplx qt2 = cmul(q, t2[j]);
                for (int d = 0; d < 3; ++d) {
                    cplx a = cmul(ik, t1[3 * j + d]);
                    cplx term = {a.re - diff[d] * qt2.re, a.im - diff[d] * qt2.im};
                    cplx c = cmul(g, term);
                    acc[d].re += c.re; acc[d].im += c.im;
                }
            } else {
                cplx c = cmul(g, t2[j]);
                acc[0].re += c.re; acc[0].im += c.im;
            }
        }
        for (int d = 0; d < kdim; ++d) { out[2 * (d * npts + m)] = acc[d].re; out[2 * (d * npts + m) + 1] = acc[d].im; }
    }
    free(G); free(sp); free(sn); free(t1); free(t2);
    return 0;
}

/* 1-D Gauss nodes/weights, for tests of the rule generator */
void bem_oracle_gauss01(int n, double *x, double *w) { gauss01(n, x, w); }

/* Regular rule used by order 0 (points on the reference triangle, weights summing to 1/2). */
int bem_oracle_set_regular_rule(int n, const double *x, const double *y, const double *w) {
    if (n < 1 || n > MAXQ) return -1;
    for (int i = 0; i < n; ++i) { g_user_x[i] = x[i]; g_user_y[i] = y[i]; g_user_w[i] = w[i]; }
    g_user_n = n;
    return 0;
}

/* Duffy rule points for tests: returns number of points; arrays must hold 6 n^4 */
int bem_oracle_duffy(int order, int kind, double *tx, double *ty, double *sx, double *sy, double *w) {
    SingRule r;
    duffy_rule(order, kind, &r);
    memcpy(tx, r.tx, sizeof(double) * r.n); memcpy(ty, r.ty, sizeof(double) * r.n);
    memcpy(sx, r.sx, sizeof(double) * r.n); memcpy(sy, r.sy, sizeof(double) * r.n);
    memcpy(w, r.w, sizeof(double) * r.n);
    int n = r.n;
    sing_rule_free(&r);
    return n;
}
