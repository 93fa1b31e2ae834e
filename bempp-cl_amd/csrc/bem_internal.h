// Host-side helpers shared by the library's translation units: error reporting,
// argument structs passed to kernels by value, dispatch utilities.
#pragma once

#include <hip/hip_runtime.h>
#include <stdarg.h>
#include <stdint.h>
#include <stdio.h>

#include "bempp_hip.h"

namespace bem {

void set_error(const char *fmt, ...);
void clear_error();

// One side of an assembly, copied into kernel arguments.
struct SpaceArgs {
    int n;
    int shape;
    const int32_t *elements;
    const double *geom;
    const int32_t *vids;
    const int32_t *l2g;
    const double *mult;
    const int32_t *nmult;
};

struct RuleArgs {
    int n;
    double x[16], y[16], w[16];
    // a 6-point rule made of two symmetric orbits, reordered to A0 A1 A2 B0 B1 B2 where point k of
    // orbit o has barycentric coordinate a_o at index k and b_o at the other two:
    // orb = (b_A, a_A - b_A, b_B, a_B - b_B) (set by orbit_rule; the kernels that use it check it)
    float orb[4];
};

// Reorders a 6-point rule into the orbit order above when it has that structure (two orbits of
// three points with equal weights, permutations of (a, b, b)); returns false and leaves it unchanged
// otherwise. The order of the points does not change any sum beyond rounding.
bool orbit_rule(RuleArgs &r);

inline SpaceArgs to_args(const bem_space_t *s) {
    SpaceArgs a;
    a.n = s->n_elements;
    a.shape = s->shapeset;
    a.elements = s->elements;
    a.geom = s->geometry;
    a.vids = s->vertex_ids;
    a.l2g = s->local2global;
    a.mult = s->multipliers;
    a.nmult = s->normal_multipliers;
    return a;
}

// Rules longer than the fast kernels' 16 points (regular_wide.hip, potential prep): kernel
// arguments, read with uniform indices (scalar loads).
struct RuleArgsWide {
    int n;
    double x[BEM_QUAD_MAX_POINTS], y[BEM_QUAD_MAX_POINTS], w[BEM_QUAD_MAX_POINTS];
};

inline RuleArgsWide to_args_wide(const bem_quad_rule_t *r) {
    RuleArgsWide a;
    a.n = r->n_points;
    for (int i = 0; i < BEM_QUAD_MAX_POINTS; ++i) {
        a.x[i] = i < r->n_points ? r->x[i] : 0.0;
        a.y[i] = i < r->n_points ? r->y[i] : 0.0;
        a.w[i] = i < r->n_points ? r->w[i] : 0.0;
    }
    return a;
}

inline RuleArgs to_args(const bem_quad_rule_t *r) {
    RuleArgs a;
    a.n = r->n_points;
    a.orb[0] = a.orb[1] = a.orb[2] = a.orb[3] = 0.f;
    for (int i = 0; i < 16; ++i) {
        a.x[i] = i < r->n_points ? r->x[i] : 0.0;
        a.y[i] = i < r->n_points ? r->y[i] : 0.0;
        a.w[i] = i < r->n_points ? r->w[i] : 0.0;
    }
    return a;
}

inline int nshape_of(int shapeset) { return shapeset == BEM_SHAPE_P0 ? 1 : 3; }

// Validate a space description; returns 0 or an error code (message set).
int check_space(const bem_space_t *s, const char *which);

// Map a HIP error to BEM_ERR_DEVICE with message.
int check_launch(const char *what);

// Regular pairs with a rule of more than BEM_QUAD_FAST_POINTS points, FP64 (regular_wide.hip):
// one lane per (test, trial) element pair. atomic: add the local blocks with atomics (spaces with
// shared dofs); otherwise every entry of a pair is stored (accumulate = 0) or added (accumulate = 1).
struct RegularWideArgs {
    SpaceArgs test, trial;
    RuleArgsWide rule;
    double kr, ki;
    void *out;
    int64_t ld;
    int identical;
    int accumulate;
};
int regular_wide_launch(int op, const RegularWideArgs &a, bool atomic, int out_dtype, hipStream_t s);

// Square n x n matrix (leading dimension ld, dtype BEM_F32/F64/C64/C128) on stream s:
// what = 0 sets it to zero, what = 1 replaces A by A + A^T, what = 2 copies the upper triangle
// into the lower one and zeroes the diagonal (matvec.hip).
int square_matrix_op(int what, void *A, int dtype, int64_t n, int64_t ld, hipStream_t s);

// Maxwell magnetic field operator (mfie.hip): regular pairs (FP64, or the far pairs in mixed
// precision), near pairs of the mixed mode and singular pairs (FP64), added to the output.
int mfie_regular_launch(const SpaceArgs &test, const SpaceArgs &trial, const RuleArgs &rule, double kr, double ki,
                        int identical, int precision, double near_factor, void *out, int64_t ld, int out_dtype,
                        hipStream_t s);
int mfie_near_launch(const SpaceArgs &test, const SpaceArgs &trial, const RuleArgs &rule, double kr, double ki,
                     int identical, double near_factor, int n_pairs, const int32_t *pairs, void *out, int64_t ld,
                     int out_dtype, hipStream_t s);
int mfie_singular_launch(const SpaceArgs &test, const SpaceArgs &trial, int n_pairs, const int32_t *pairs,
                         const double *tpts, const double *spts, const double *wts, double kr, double ki, void *out,
                         int64_t ld, int out_dtype, hipStream_t s);

}  // namespace bem
