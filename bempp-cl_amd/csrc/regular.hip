// Entry points of the regular-pair assembly; the kernels live in regular_impl.h and are
// instantiated per operator in regular_op<k>.hip so that they compile in parallel.
#include "regular_impl.h"

namespace bem {

static int regular_entry(int op, const bem_space_t *test, const bem_space_t *trial, const bem_dof_blocks_t *tblocks,
                         const bem_dof_blocks_t *sblocks, int grids_identical, const bem_quad_rule_t *rule,
                         double k_re, double k_im, int precision, double mixed_near_factor, int kahan, void *out,
                         void *out2, int64_t ld, int out_dtype, int accumulate, void *stream,
                         const int32_t *work = nullptr, int nwork = 0, bool upper = false) {
    clear_error();
    if (!test || !trial || !rule || !out) {
        set_error("null argument");
        return BEM_ERR_INVALID_ARGUMENT;
    }
    int rc;
    if ((rc = check_space(test, "test")) || (rc = check_space(trial, "trial"))) return rc;
    if (rule->n_points < 1 || rule->n_points > BEM_QUAD_MAX_POINTS) {
        set_error("quadrature rule must have 1..%d points, got %d", BEM_QUAD_MAX_POINTS, rule->n_points);
        return BEM_ERR_INVALID_ARGUMENT;
    }
    if (precision != BEM_PRECISION_FP64 && precision != BEM_PRECISION_MIXED) {
        set_error("unknown precision mode %d", precision);
        return BEM_ERR_INVALID_ARGUMENT;
    }
    const bool wide = rule->n_points > BEM_QUAD_FAST_POINTS;
    if (wide && (out2 || work || upper || op == BEM_MAXWELL_MAGNETIC_FIELD)) {
        set_error("rules of more than %d points are served by bem_regular_assemble(_blocked) for operators 0-8 only",
                  BEM_QUAD_FAST_POINTS);
        return BEM_ERR_UNSUPPORTED;
    }
    if (wide && precision != BEM_PRECISION_FP64) {
        set_error("rules of more than %d points run in FP64 (BEM_PRECISION_FP64, any output dtype)",
                  BEM_QUAD_FAST_POINTS);
        return BEM_ERR_INVALID_ARGUMENT;
    }
    const bool maxwell = op == BEM_MAXWELL_ELECTRIC_FIELD || op == BEM_MAXWELL_MAGNETIC_FIELD;
    if (maxwell && (test->shapeset != BEM_SHAPE_RWG || trial->shapeset != BEM_SHAPE_RWG)) {
        set_error("the Maxwell operators need RWG/SNC spaces on both sides");
        return BEM_ERR_INVALID_ARGUMENT;
    }
    if (!maxwell && (test->shapeset == BEM_SHAPE_RWG || trial->shapeset == BEM_SHAPE_RWG)) {
        set_error("scalar operators need scalar (P0/P1) spaces");
        return BEM_ERR_INVALID_ARGUMENT;
    }
    if (op == BEM_MAXWELL_MAGNETIC_FIELD && (out2 || work || upper)) {
        set_error("the magnetic field operator has no fused or symmetric pass");
        return BEM_ERR_UNSUPPORTED;
    }
    const bool tile = tblocks != nullptr;
    if (wide) {
        RegularWideArgs w;
        w.test = to_args(test);
        w.trial = to_args(trial);
        w.rule = to_args_wide(rule);
        w.kr = k_re;
        w.ki = k_im;
        w.out = out;
        w.ld = ld;
        w.identical = grids_identical;
        w.accumulate = accumulate;
        return regular_wide_launch(op, w, tile, out_dtype, (hipStream_t)stream);
    }
    if (op == BEM_MAXWELL_MAGNETIC_FIELD && !(tile && precision == BEM_PRECISION_MIXED)) {
        // adds to the output: FP64, or without dof blocks the untiled mixed kernel (far pairs)
        return mfie_regular_launch(to_args(test), to_args(trial), to_args(rule), k_re, k_im, grids_identical,
                                   precision, mixed_near_factor, out, ld, out_dtype, (hipStream_t)stream);
    }
    if ((tblocks == nullptr) != (sblocks == nullptr)) {
        set_error("dof blocks must be given for both sides or for none");
        return BEM_ERR_INVALID_ARGUMENT;
    }
    if (tile && (!tblocks->start || !tblocks->dofs || !tblocks->ndofs || !tblocks->slots || !sblocks->start ||
                 !sblocks->dofs || !sblocks->ndofs || !sblocks->slots || tblocks->n_blocks < 0 || sblocks->n_blocks < 0)) {
        set_error("incomplete dof block description");
        return BEM_ERR_INVALID_ARGUMENT;
    }
    RegularArgs a;
    a.test = to_args(test);
    a.trial = to_args(trial);
    a.rule = to_args(rule);
    a.kr = k_re;
    a.ki = k_im;
    a.near_factor = mixed_near_factor;
    a.out = out;
    a.out2 = out2;
    a.ld = ld;
    a.identical = grids_identical;
    a.accumulate = accumulate;
    a.work = work;
    a.nwork = nwork;
    a.tb = BlockArgs{nullptr, nullptr, nullptr, nullptr, nullptr, 0, 0, 0};
    a.sb = a.tb;
    if (tile) {
        a.tb = BlockArgs{tblocks->start, tblocks->dofs, tblocks->ndofs, tblocks->slots, tblocks->incidence,
                         tblocks->n_blocks, tblocks->max_dofs, tblocks->max_incidence};
        a.sb = BlockArgs{sblocks->start, sblocks->dofs, sblocks->ndofs, sblocks->slots, sblocks->incidence,
                         sblocks->n_blocks, sblocks->max_dofs, sblocks->max_incidence};
    }
    hipStream_t s = (hipStream_t)stream;
    if (op == BEM_MAXWELL_MAGNETIC_FIELD) return regular_launch_mfie(a, out_dtype, s);  // mixed, tiled: far pairs
    if (upper) {
        if (op == BEM_LAPLACE_SINGLE_LAYER) return regular_launch_upper_0(a, precision, kahan, out_dtype, s);
        if (op == BEM_HELMHOLTZ_SINGLE_LAYER) return regular_launch_upper_3(a, precision, kahan, out_dtype, s);
        set_error("the symmetric DP0 pass needs a single-layer operator, got %d", op);
        return BEM_ERR_UNSUPPORTED;
    }
    if (work) {
        if (op == BEM_LAPLACE_SINGLE_LAYER)
            return out2 ? regular_launch_sym_0(a, precision, kahan, out_dtype, s)
                        : regular_launch_symsl_0(a, precision, kahan, out_dtype, s);
        if (op == BEM_HELMHOLTZ_SINGLE_LAYER)
            return out2 ? regular_launch_sym_3(a, precision, kahan, out_dtype, s)
                        : regular_launch_symsl_3(a, precision, kahan, out_dtype, s);
        if (op == BEM_LAPLACE_HYPERSINGULAR && !out2) return regular_launch_symsl_7(a, precision, kahan, out_dtype, s);
        if (op == BEM_HELMHOLTZ_HYPERSINGULAR && !out2) return regular_launch_symsl_8(a, precision, kahan, out_dtype, s);
        if (op == BEM_MAXWELL_ELECTRIC_FIELD && !out2) return regular_launch_symsl_6(a, precision, kahan, out_dtype, s);
        set_error("the symmetric pass needs a single-layer operator (alone also a hypersingular or the electric "
                  "field operator), got %d", op);
        return BEM_ERR_UNSUPPORTED;
    }
    if (out2) {
        if (op == BEM_LAPLACE_SINGLE_LAYER) return regular_launch_dual_0(a, precision, kahan, tile, out_dtype, s);
        if (op == BEM_HELMHOLTZ_SINGLE_LAYER) return regular_launch_dual_3(a, precision, kahan, tile, out_dtype, s);
        set_error("the single/double layer pair needs a single-layer operator, got %d", op);
        return BEM_ERR_UNSUPPORTED;
    }
    switch (op) {
    case 0: return regular_launch_0(a, precision, kahan, tile, out_dtype, s);
    case 1: return regular_launch_1(a, precision, kahan, tile, out_dtype, s);
    case 2: return regular_launch_2(a, precision, kahan, tile, out_dtype, s);
    case 3: return regular_launch_3(a, precision, kahan, tile, out_dtype, s);
    case 4: return regular_launch_4(a, precision, kahan, tile, out_dtype, s);
    case 5: return regular_launch_5(a, precision, kahan, tile, out_dtype, s);
    case 6: return regular_launch_6(a, precision, kahan, tile, out_dtype, s);
    case 7: return regular_launch_7(a, precision, kahan, tile, out_dtype, s);
    case 8: return regular_launch_8(a, precision, kahan, tile, out_dtype, s);
    default: set_error("unsupported operator %d", op); return BEM_ERR_UNSUPPORTED;
    }
}

}  // namespace bem

using namespace bem;

extern "C" int bem_regular_assemble(int op, const bem_space_t *test, const bem_space_t *trial, int grids_identical,
                                    const bem_quad_rule_t *rule, double k_re, double k_im, int precision,
                                    double mixed_near_factor, int kahan, void *out, int64_t ld, int out_dtype,
                                    int accumulate, void *stream) {
    return regular_entry(op, test, trial, nullptr, nullptr, grids_identical, rule, k_re, k_im, precision,
                         mixed_near_factor, kahan, out, nullptr, ld, out_dtype, accumulate, stream);
}

extern "C" int bem_regular_assemble_blocked(int op, const bem_space_t *test, const bem_space_t *trial,
                                            const bem_dof_blocks_t *test_blocks, const bem_dof_blocks_t *trial_blocks,
                                            int grids_identical, const bem_quad_rule_t *rule, double k_re, double k_im,
                                            int precision, double mixed_near_factor, int kahan, void *out, int64_t ld,
                                            int out_dtype, void *stream) {
    if (!test_blocks || !trial_blocks) {
        clear_error();
        set_error("bem_regular_assemble_blocked needs dof blocks for both sides");
        return BEM_ERR_INVALID_ARGUMENT;
    }
    return regular_entry(op, test, trial, test_blocks, trial_blocks, grids_identical, rule, k_re, k_im, precision,
                         mixed_near_factor, kahan, out, nullptr, ld, out_dtype, 0, stream);
}

extern "C" int bem_regular_assemble_sl_dl(int op, const bem_space_t *test, const bem_space_t *trial,
                                          const bem_dof_blocks_t *test_blocks, const bem_dof_blocks_t *trial_blocks,
                                          int grids_identical, const bem_quad_rule_t *rule, double k_re, double k_im,
                                          int precision, double mixed_near_factor, int kahan, void *out_single,
                                          void *out_double, int64_t ld, int out_dtype, int accumulate, void *stream) {
    if (!out_double) {
        clear_error();
        set_error("null argument");
        return BEM_ERR_INVALID_ARGUMENT;
    }
    return regular_entry(op, test, trial, test_blocks, trial_blocks, grids_identical, rule, k_re, k_im, precision,
                         mixed_near_factor, kahan, out_single, out_double, ld, out_dtype,
                         test_blocks ? 0 : accumulate, stream);
}

extern "C" int bem_regular_assemble_sl_dl_symmetric(int op, const bem_space_t *space,
                                                    const bem_dof_blocks_t *test_tiles,
                                                    const bem_dof_blocks_t *trial_blocks, int n_work,
                                                    const int32_t *work, int64_t n_dofs, const bem_quad_rule_t *rule,
                                                    double k_re, double k_im, int precision, double mixed_near_factor,
                                                    int kahan, void *out_single, void *out_double, int64_t ld,
                                                    int out_dtype, void *stream) {
    clear_error();
    if (!space || !test_tiles || !trial_blocks || !out_single || (n_work > 0 && !work)) {
        set_error("null argument");
        return BEM_ERR_INVALID_ARGUMENT;
    }
    if (n_work < 0 || n_dofs < 0 || ld < n_dofs) {
        set_error("invalid work list or matrix shape");
        return BEM_ERR_INVALID_ARGUMENT;
    }
    if (test_tiles->max_dofs > BEM_TILE_SYM_TEST_DOFS) {
        set_error("test tiles of the symmetric pass hold at most %d dofs", BEM_TILE_SYM_TEST_DOFS);
        return BEM_ERR_INVALID_ARGUMENT;
    }
    hipStream_t s = (hipStream_t)stream;
    int rc;
    // both outputs are overwritten: zero, then the pairs (e, f) with e before f in the element
    // order, then the single layer's transpose
    if ((rc = square_matrix_op(0, out_single, out_dtype, n_dofs, ld, s))) return rc;
    if (out_double && (rc = square_matrix_op(0, out_double, out_dtype, n_dofs, ld, s))) return rc;
    if (n_work > 0) {
        rc = regular_entry(op, space, space, test_tiles, trial_blocks, 1, rule, k_re, k_im, precision,
                           mixed_near_factor, kahan, out_single, out_double, ld, out_dtype, 0, stream, work, n_work);
        if (rc) return rc;
    }
    return square_matrix_op(1, out_single, out_dtype, n_dofs, ld, s);
}

extern "C" int bem_regular_assemble_symmetric(int op, const bem_space_t *space, const bem_quad_rule_t *rule,
                                              double k_re, double k_im, int precision, double mixed_near_factor,
                                              int kahan, void *out, int64_t ld, int out_dtype, void *stream) {
    clear_error();
    if (!space || !out) {
        set_error("null argument");
        return BEM_ERR_INVALID_ARGUMENT;
    }
    if (space->shapeset != BEM_SHAPE_P0) {
        set_error("the symmetric DP0 pass needs a piecewise constant space");
        return BEM_ERR_INVALID_ARGUMENT;
    }
    if (ld < space->n_elements) {
        set_error("leading dimension smaller than the number of dofs");
        return BEM_ERR_INVALID_ARGUMENT;
    }
    // the upper triangle and, from the same pairs, its mirror image below a zero diagonal: every
    // entry is overwritten
    return regular_entry(op, space, space, nullptr, nullptr, 1, rule, k_re, k_im, precision, mixed_near_factor,
                         kahan, out, nullptr, ld, out_dtype, 0, stream, nullptr, 0, true);
}
