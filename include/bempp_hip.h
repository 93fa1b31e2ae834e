/*
 * bempp_hip.h -- C ABI of the MI355X (gfx950) dense BEM library libbempp_hip.so.
 *
 * The entry points replace the three backend functions the reference dispatches to
 * (bempp_cl/core/dispatcher.py:4-55):
 *   bem_regular_assemble   <- dense_assembler(...)     bempp_cl/core/numba_assemblers.py:58-112,
 *                                                      bempp_cl/core/opencl_assemblers.py:108-284
 *   bem_singular_assemble  <- singular_assembler(...)  bempp_cl/core/numba_assemblers.py:6-55,
 *                                                      bempp_cl/core/opencl_assemblers.py:12-105
 *                             plus the np.add.at scatter of bempp_cl/core/dense_assembler.py:78-98
 *   bem_potential_evaluate <- potential_assembler(...) bempp_cl/core/numba_assemblers.py:115-164,
 *                                                      bempp_cl/core/opencl_assemblers.py:287-465
 * Extra entry points with no reference counterpart: bem_kbatch_assemble (many Helmholtz
 * wavenumbers in one persistent launch), bem_adjoint_* (gradients), bem_dense_matvec, bem_gmres_cgs2,
 * bem_fill_zero, bem_potential_workspace_size, bem_last_error, bem_version.
 *
 * Conventions
 *  - All array arguments are DEVICE pointers (hipMalloc'ed or torch device memory) unless
 *    the comment says host. The library allocates nothing; callers pass workspaces.
 *  - Work is enqueued on `stream` (a hipStream_t, NULL = default stream) and the call
 *    returns without synchronising. Results are complete once the stream is synchronised.
 *  - Return value 0 = success; otherwise one of the BEM_ERR_* codes, with a message
 *    available from bem_last_error() (thread-local).
 *  - Dense outputs are row-major: entry (row, col) at out[row * ld + col]; rows index
 *    dual_to_range (test) dofs, columns domain (trial) dofs, as the reference's
 *    np.zeros((rows, cols)) result. Complex outputs are interleaved (re, im).
 */
#ifndef BEMPP_HIP_H
#define BEMPP_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* operators (kernel_type / assembly_type of the reference OperatorDescriptor) */
enum bem_operator {
    BEM_LAPLACE_SINGLE_LAYER = 0,
    BEM_LAPLACE_DOUBLE_LAYER = 1,
    BEM_LAPLACE_ADJOINT_DOUBLE_LAYER = 2,
    BEM_HELMHOLTZ_SINGLE_LAYER = 3,
    BEM_HELMHOLTZ_DOUBLE_LAYER = 4,
    BEM_HELMHOLTZ_ADJOINT_DOUBLE_LAYER = 5,
    BEM_MAXWELL_ELECTRIC_FIELD = 6,
    /* P1 x P1 only: the single layer's Green's function with the surface-curl weights of
       bempp_cl/core/numba_kernels.py:1078-1333 (regular) and 1530-1697 (singular) */
    BEM_LAPLACE_HYPERSINGULAR = 7,
    BEM_HELMHOLTZ_HYPERSINGULAR = 8,
    /* far fields, for bem_potential_evaluate only (the points are unit directions):
       bempp_cl/core/numba_kernels.py:402-449 (Helmholtz) and 2437-2570 (Maxwell) */
    BEM_HELMHOLTZ_FAR_FIELD_SINGLE_LAYER = 9,
    BEM_HELMHOLTZ_FAR_FIELD_DOUBLE_LAYER = 10,
    BEM_MAXWELL_ELECTRIC_FAR_FIELD = 11,
    BEM_MAXWELL_MAGNETIC_FAR_FIELD = 12,
    /* RWG domain, SNC dual; bem_regular_assemble (no dof blocks) and bem_singular_assemble add to
       the output, FP64 in both precision modes; bem_potential_evaluate gives the magnetic field
       potential: bempp_cl/core/numba_kernels.py:2065-2277, 2363-2433 */
    BEM_MAXWELL_MAGNETIC_FIELD = 13
};

/* reference shapesets (bempp_cl/api/space/shapesets.py:46-158) */
enum bem_shapeset { BEM_SHAPE_P0 = 0, BEM_SHAPE_P1 = 1, BEM_SHAPE_RWG = 2 };

/* output element types */
enum bem_dtype { BEM_F32 = 0, BEM_F64 = 1, BEM_C64 = 2, BEM_C128 = 3 };

/* arithmetic modes: FP64 everywhere, or FP32 for well-separated regular pairs with FP64
 * for near-singular regular pairs and all singular pairs */
enum bem_precision { BEM_PRECISION_FP64 = 0, BEM_PRECISION_MIXED = 1 };

enum bem_error {
    BEM_OK = 0,
    BEM_ERR_INVALID_ARGUMENT = 1, /* reference raises ValueError */
    BEM_ERR_UNSUPPORTED = 2,      /* reference raises KeyError (unsupported kernel) */
    BEM_ERR_DEVICE = 3,           /* reference raises RuntimeError (device problem) */
    BEM_ERR_WORKSPACE = 4         /* workspace too small */
};

/* Number of doubles per element in the geometry table */
#define BEM_GEOM_STRIDE 16
/* Geometry table row of element e (all float64):
 *  [0..2] v0, [3..5] v1 - v0, [6..8] v2 - v0, [9..11] unit normal, [12] integration element
 *  |(v1-v0) x (v2-v0)|, [13] diameter, [14..15] unused.  (bempp_cl/api/grid/grid.py:552-590) */

/* One side (test or trial) of an assembly: a function space on a grid, device arrays. */
typedef struct bem_space {
    int32_t n_elements;         /* number of entries in `elements` (the support elements)  */
    int32_t shapeset;           /* enum bem_shapeset                                        */
    const int32_t *elements;    /* [n_elements] grid element indices                        */
    const double *geometry;     /* [n_grid_elements * BEM_GEOM_STRIDE]                      */
    const int32_t *vertex_ids;  /* [n_grid_elements * 4] vertex indices (4th unused)       */
    const int32_t *local2global;/* [n_grid_elements * nshape] global dof of shape function */
    const double *multipliers;  /* [n_grid_elements * nshape] local multipliers            */
    const int32_t *normal_multipliers; /* [n_grid_elements] +1/-1, or NULL for all +1      */
} bem_space_t;

/* A triangle quadrature rule, host memory, at most BEM_QUAD_MAX_POINTS points on the reference
 * triangle (the reference's symmetric rules of orders 1-20 have 1..79 points,
 * bempp_cl/api/integration/triangle_gauss.py:143-173). Rules of up to BEM_QUAD_FAST_POINTS points
 * run through every kernel; longer rules are served by bem_regular_assemble,
 * bem_regular_assemble_blocked (FP64 arithmetic only) and bem_potential_evaluate. */
#define BEM_QUAD_MAX_POINTS 79
#define BEM_QUAD_FAST_POINTS 16
typedef struct bem_quad_rule {
    int32_t n_points;
    double x[BEM_QUAD_MAX_POINTS], y[BEM_QUAD_MAX_POINTS], w[BEM_QUAD_MAX_POINTS];
} bem_quad_rule_t;

/* Blocking of one side for bem_regular_assemble_blocked (device arrays). The element list of
 * the bem_space_t is split into n_blocks consecutive blocks; block b holds list positions
 * [start[b], start[b+1]) (at most 256 elements for the trial side, at most 16 for the test side)
 * and the ndofs[b] <= max_dofs distinct global dofs dofs[b * max_dofs + s]; slots[3 * pos + i]
 * is the index s of the dof of shape function i of the element at list position pos. */
typedef struct bem_dof_blocks {
    int32_t n_blocks;
    int32_t max_dofs;
    const int32_t *start;  /* [n_blocks + 1] */
    const int32_t *dofs;   /* [n_blocks * max_dofs] */
    const int32_t *ndofs;  /* [n_blocks] */
    const int32_t *slots;  /* [n_elements * 3] */
    /* trial side, MIXED precision: for block b and slot s, up to max_incidence codes
     * 3 * (pos - start[b]) + i of the (element, shape function) pairs carrying dof s, -1 padded:
     * incidence[(b * max_dofs + s) * max_incidence + k]. May be NULL for the test side. */
    const int16_t *incidence;
    int32_t max_incidence;
} bem_dof_blocks_t;

/* LDS tile limits of bem_regular_assemble_blocked */
#define BEM_TILE_TEST_ELEMENTS 16
#define BEM_TILE_TEST_DOFS 20
#define BEM_TILE_TRIAL_ELEMENTS 256
#define BEM_TILE_TRIAL_DOFS 160
#define BEM_TILE_MAX_INCIDENCE 12
/* test tiles of bem_regular_assemble_sl_dl with blocks (two LDS tiles) */
#define BEM_TILE_PAIR_TEST_ELEMENTS 8
#define BEM_TILE_PAIR_TEST_DOFS 12
/* blocks of the mixed-precision EFIE (RWG/SNC) in bem_regular_assemble_blocked: a 256-element
 * trial block holds up to ~400 edges; test tiles are smaller so that the LDS tile stays small */
#define BEM_TILE_RWG_MIXED_TEST_ELEMENTS 8
#define BEM_TILE_RWG_MIXED_TEST_DOFS 16
#define BEM_TILE_RWG_MIXED_TRIAL_DOFS 416
/* test tiles of bem_regular_assemble_sl_dl_symmetric (three LDS tiles) */
#define BEM_TILE_SYM_TEST_ELEMENTS 8
#define BEM_TILE_SYM_TEST_DOFS 12
/* test elements of one work item of bem_regular_assemble_sl_dl_symmetric (its run of tiles) */
#define BEM_TILE_SYM_RUN_ELEMENTS 48

/*
 * Regular part of a dense Galerkin operator: all (test, trial) element pairs, except,
 * when grids_identical != 0, pairs sharing a vertex (bempp_cl/core/numba_kernels.py:980-1074,
 * 1853-1969; pairs sharing vertices are left to bem_singular_assemble).
 *
 * bem_regular_assemble is for spaces whose (element, shape function) pairs own distinct dofs
 * (DP0, DP1): every entry is written once; it is added to the existing value when
 * accumulate != 0 and stored otherwise (no prior zeroing needed).
 * bem_regular_assemble_blocked is for spaces with dofs shared between elements (P1, RWG/SNC):
 * contributions are summed per (test block, trial block) in LDS and added to `out`, which
 * must be initialised (e.g. bem_fill_zero). Block limits: BEM_TILE_* below.
 * k_re, k_im: wavenumber for Helmholtz/Maxwell (ignored for Laplace).
 * precision: enum bem_precision. mixed_near_factor: in MIXED mode, pairs whose centroid
 *   distance is below factor * (diameter_test + diameter_trial) / 2 use FP64.
 * kahan: compensated FP32 summation in MIXED mode.
 * out_dtype must be BEM_F64/BEM_F32 for Laplace and BEM_C128/BEM_C64 otherwise.
 * Rules of more than BEM_QUAD_FAST_POINTS points (regular orders 7-20): precision must be
 * BEM_PRECISION_FP64 (any output dtype); one lane per element pair evaluates every point pair
 * (blocked entry: atomic adds of the local block; dof blocks are not read). The magnetic field
 * operator, the fused and the symmetric passes take at most BEM_QUAD_FAST_POINTS points.
 * BEM_MAXWELL_MAGNETIC_FIELD (bempp_cl/core/numba_kernels.py:2151-2277) takes the 6-point rule
 * and always adds to `out` through both entry points; in MIXED mode the blocked entry point
 * runs its far pairs through the LDS-tiled kernel.
 */
int bem_regular_assemble(int op, const bem_space_t *test, const bem_space_t *trial, int grids_identical,
                         const bem_quad_rule_t *rule, double k_re, double k_im, int precision,
                         double mixed_near_factor, int kahan, void *out, int64_t ld, int out_dtype,
                         int accumulate, void *stream);
int bem_regular_assemble_blocked(int op, const bem_space_t *test, const bem_space_t *trial,
                                 const bem_dof_blocks_t *test_blocks, const bem_dof_blocks_t *trial_blocks,
                                 int grids_identical, const bem_quad_rule_t *rule, double k_re, double k_im,
                                 int precision, double mixed_near_factor, int kahan, void *out, int64_t ld,
                                 int out_dtype, void *stream);

/*
 * Single layer and double layer of one family in one pass (no single reference counterpart:
 * the reference assembles each operator separately, bempp_cl/core/numba_kernels.py:980-1074
 * with the kernels of numba_kernels.py:176-320). op: BEM_LAPLACE_SINGLE_LAYER or
 * BEM_HELMHOLTZ_SINGLE_LAYER; the matching double layer goes to out_double. Both operators
 * share the spaces, the wavenumber and ld/out_dtype; every point pair's distance, 1/r and
 * e^{ikr} are computed once for both. With test_blocks/trial_blocks NULL this behaves as
 * bem_regular_assemble (localised spaces, `accumulate` as there), otherwise as
 * bem_regular_assemble_blocked (sums added to both outputs) with test blocks of at most
 * BEM_TILE_PAIR_TEST_ELEMENTS elements and BEM_TILE_PAIR_TEST_DOFS dofs.
 */
int bem_regular_assemble_sl_dl(int op, const bem_space_t *test, const bem_space_t *trial,
                               const bem_dof_blocks_t *test_blocks, const bem_dof_blocks_t *trial_blocks,
                               int grids_identical, const bem_quad_rule_t *rule, double k_re, double k_im,
                               int precision, double mixed_near_factor, int kahan, void *out_single, void *out_double,
                               int64_t ld, int out_dtype, int accumulate, void *stream);

/*
 * The same pair of operators for one P1 space on one grid (test space == trial space), with
 * every unordered element pair evaluated once (no reference counterpart; the reference
 * computes each ordered pair, bempp_cl/core/numba_kernels.py:980-1074). The single layer of
 * the pair (f, e) is the transpose of that of (e, f); the double layer of (f, e) shares the
 * distance and e^{ikr} and only needs the other normal.
 *
 * trial_blocks: blocking of the space's element list (as for bem_regular_assemble_blocked,
 * with incidences). test_tiles: a blocking of the same element list whose tiles (at most
 * BEM_TILE_SYM_TEST_ELEMENTS elements, BEM_TILE_SYM_TEST_DOFS dofs) each lie inside one trial
 * block, in list order. work: n_work triples (block b, first tile, end tile); together they must
 * list, for every block b, each tile of blocks 0..b exactly once, and the tiles of one triple may
 * hold at most BEM_TILE_SYM_RUN_ELEMENTS elements (one workgroup per triple keeps the block's
 * trial elements in registers and the run's test elements in LDS). In a tile's own block only pairs
 * with the trial element later in the list than the test element are evaluated. out_single and out_double (n_dofs x n_dofs,
 * leading dimension ld) are OVERWRITTEN with the regular part: pairs sharing a vertex are
 * excluded, and in MIXED precision so are the near pairs (bem_near_assemble_sl_dl adds them).
 * out_double may be NULL: the operator alone, at the cost of its own point pairs only (the near
 * pairs then come from bem_near_assemble). op may then also be BEM_LAPLACE_HYPERSINGULAR or
 * BEM_HELMHOLTZ_HYPERSINGULAR, symmetric like the single layers (W_fe = W_ef^T), or
 * BEM_MAXWELL_ELECTRIC_FIELD on one RWG dof map (space: RWG; the SNC test functions of the same
 * dofs, multipliers and normal orientation give the same matrix). MIXED precision with the
 * electric field operator takes trial blocks of up to BEM_TILE_RWG_MIXED_TRIAL_DOFS dofs.
 */
int bem_regular_assemble_sl_dl_symmetric(int op, const bem_space_t *space, const bem_dof_blocks_t *test_tiles,
                                         const bem_dof_blocks_t *trial_blocks, int n_work, const int32_t *work,
                                         int64_t n_dofs, const bem_quad_rule_t *rule, double k_re, double k_im,
                                         int precision, double mixed_near_factor, int kahan, void *out_single,
                                         void *out_double, int64_t ld, int out_dtype, void *stream);

/*
 * Near-field regular pairs of the MIXED mode, in FP64 (no reference counterpart: the
 * reference computes every regular pair in one precision). The MIXED regular kernels skip
 * every non-adjacent pair with |c_test - c_trial| < factor * (diam_test + diam_trial) / 2;
 * this entry point computes exactly those among the n_pairs candidate pairs
 * pairs[2 p] (test element), pairs[2 p + 1] (trial element) and adds them to `out` with
 * atomics. The candidate list must contain every such pair; other candidates are ignored.
 */
int bem_near_assemble(int op, const bem_space_t *test, const bem_space_t *trial, int n_pairs, const int32_t *pairs,
                      int grids_identical, const bem_quad_rule_t *rule, double k_re, double k_im,
                      double near_factor, void *out, int64_t ld, int out_dtype, void *stream);
/* bem_near_assemble for a single layer (op) and its double layer in one pass (see bem_regular_assemble_sl_dl). */
int bem_near_assemble_sl_dl(int op, const bem_space_t *test, const bem_space_t *trial, int n_pairs,
                            const int32_t *pairs, int grids_identical, const bem_quad_rule_t *rule, double k_re,
                            double k_im, double near_factor, void *out_single, void *out_double, int64_t ld,
                            int out_dtype, void *stream);
/* The near pairs of bem_regular_assemble_sl_dl_symmetric (one space, MIXED precision): each pair
 * (e, f) of the list stands for both orientations and must appear once (e.g. e < f). It adds the
 * single-layer block at (e, f) and, transposed, at (f, e); the double layer at (e, f) and the
 * double layer of the reversed pair at (f, e). No reference counterpart. */
int bem_near_assemble_sl_dl_symmetric(int op, const bem_space_t *space, int n_pairs, const int32_t *pairs,
                                      const bem_quad_rule_t *rule, double k_re, double k_im, double near_factor,
                                      void *out_single, void *out_double, int64_t ld, int out_dtype, void *stream);

/*
 * Singular (Sauter-Schwab) part, always FP64: for each of n_pairs pairs p,
 *   pair_data[6 p + 0..5] = test element, trial element, test point offset, trial point offset,
 *                           weight offset, number of points
 * with reference coordinates test_points[2 * (offset + k) + 0/1], trial_points likewise and
 * weights[weight offset + k] (device arrays, bempp_cl/core/singular_assembler.py:225-448).
 * The local block is multiplied by the local multipliers (and RWG edge lengths) and added
 * atomically to out at (local2global_test, local2global_trial).
 */
int bem_singular_assemble(int op, const bem_space_t *test, const bem_space_t *trial, int n_pairs,
                          const int32_t *pair_data, const double *test_points, const double *trial_points,
                          const double *weights, double k_re, double k_im, void *out, int64_t ld, int out_dtype,
                          void *stream);
/* bem_singular_assemble for a single layer (op) and its double layer in one pass (shared r and e^{ikr}). */
int bem_singular_assemble_sl_dl(int op, const bem_space_t *test, const bem_space_t *trial, int n_pairs,
                                const int32_t *pair_data, const double *test_points, const double *trial_points,
                                const double *weights, double k_re, double k_im, void *out_single, void *out_double,
                                int64_t ld, int out_dtype, void *stream);

/*
 * Dense potential u(x_m) (bempp_cl/core/numba_kernels.py:1788-1849, 2281-2359):
 * coefficients: complex128 [n_grid_elements * nshape] per (element, shape function), already
 *   multiplied by the local multipliers (space.map_to_full_grid @ x).
 * points: float64 [3 * n_points] (x0 y0 z0 x1 ...). out: [kdim * n_points] of out_dtype
 *   (BEM_C128 or BEM_C64), row d holds component d (kdim = 3 for Maxwell, else 1). With
 *   n_points = 0, points and out may be null; a space without elements gives zeros.
 * Only BEM_LAPLACE_SINGLE_LAYER, BEM_LAPLACE_DOUBLE_LAYER, BEM_HELMHOLTZ_SINGLE_LAYER,
 * BEM_HELMHOLTZ_DOUBLE_LAYER and BEM_MAXWELL_ELECTRIC_FIELD are potentials.
 * workspace: device memory of at least bem_potential_workspace_size(...) bytes.
 */
size_t bem_potential_workspace_size(int op, const bem_space_t *space, const bem_quad_rule_t *rule, int n_points,
                                    int precision);
int bem_potential_evaluate(int op, const bem_space_t *space, const void *coefficients, int n_points,
                           const double *points, const bem_quad_rule_t *rule, double k_re, double k_im,
                           int precision, void *out, int out_dtype, void *workspace, size_t workspace_bytes,
                           void *stream);

/*
 * Allocation-free bem_potential_evaluate. No counterpart in the reference: for callers that
 * cannot allocate device memory (code under HIP graph capture, tight serving loops, harnesses
 * that forbid hipMalloc). Same arguments and results as bem_potential_evaluate without the
 * workspace: it uses a library-owned static device buffer of BEM_POTENTIAL_SCRATCH_BYTES per
 * GPU (held until the process ends) and calls no allocator. Returns BEM_ERR_WORKSPACE when
 * bem_potential_workspace_size(...) exceeds that size (the level-5 P1 potential at 16,384
 * points needs 29.5 MB, the level-4 RWG potential at 65,536 points 50.6 MB). All calls on one
 * GPU share the buffer: calls that overlap on one GPU (other streams, other threads) race.
 */
#define BEM_POTENTIAL_SCRATCH_BYTES ((size_t)64 << 20)
int bem_potential_evaluate_noalloc(int op, const bem_space_t *space, const void *coefficients, int n_points,
                                   const double *points, const bem_quad_rule_t *rule, double k_re, double k_im,
                                   int precision, void *out, int out_dtype, void *stream);

/* Launch shape of bem_kbatch_assemble, picked by the autotuner (bempp_amd/core/autotune.py). */
typedef struct bem_kbatch_config {
    int32_t test_tile;          /* test elements staged in LDS per tile: 8, 16 or 32 */
    int32_t workgroups_per_cu;  /* persistent grid = CUs * this (1..16)              */
} bem_kbatch_config_t;

#define BEM_KBATCH_MAX_WAVENUMBERS 64
/* `parts` bits of bem_kbatch_assemble */
#define BEM_KBATCH_REGULAR 1
#define BEM_KBATCH_NEAR 2
#define BEM_KBATCH_SINGULAR 4

/*
 * Frequency sweep of the Helmholtz single layer, DP0 x DP0 (no single reference counterpart:
 * the reference assembles one wavenumber per weak_form() call,
 * bempp_cl/core/numba_kernels.py:980-1074 with helmholtz_single_layer at numba_kernels.py:298-320).
 * out[m * matrix_stride + row * ld + col] receives the matrix of real wavenumber wavenumbers[m]
 * (host array, 1 <= nk <= BEM_KBATCH_MAX_WAVENUMBERS). Regular pairs run in one persistent
 * launch that evaluates the point geometry of a pair once for all wavenumbers; every entry is
 * stored (no prior zeroing). Then, MIXED mode only, the FP64 near-field pairs among near_pairs
 * (as bem_near_assemble), and the FP64 singular pairs (as bem_singular_assemble) are added.
 * precision: BEM_PRECISION_MIXED with out_dtype BEM_C64, or BEM_PRECISION_FP64 with BEM_C128.
 * The regular rule must have 6 points (order 4). `parts` (BEM_KBATCH_* bits) selects which of
 * the three launches run, so that a caller can time them separately; the regular launch must
 * precede the others on the stream.
 */
int bem_kbatch_assemble(const bem_space_t *test, const bem_space_t *trial, int grids_identical,
                        const bem_quad_rule_t *rule, int nk, const double *wavenumbers, int precision, int parts,
                        double mixed_near_factor, const bem_kbatch_config_t *config, void *out, int64_t ld,
                        int64_t matrix_stride, int out_dtype, int n_near, const int32_t *near_pairs, int n_singular,
                        const int32_t *singular_pairs, const double *test_points, const double *trial_points,
                        const double *weights, void *stream);

/*
 * Vertex gradient of the Helmholtz single layer (DP0 x DP0 on one space; no reference
 * counterpart, the reference has no derivatives). Adds to grad_vertices ([n_vertices * 3],
 * float64) the gradient with respect to the vertex coordinates of
 *     Re sum_{e,f} conj(x_e) V_ef y_f,
 * V the Galerkin matrix of bempp_cl/core/numba_kernels.py:298-320 (regular pairs, 980-1074)
 * with the singular pairs of bem_singular_assemble. x, y: complex128 [global dofs]. With x the
 * cotangent of V c and y = c this is the vertex part of the backward of c -> V c.
 * precision: BEM_PRECISION_MIXED (FP32 far pairs, FP64 near pairs among near_pairs and FP64
 * singular pairs, as the forward mixed mode) or BEM_PRECISION_FP64. k: real wavenumber
 * (k = 0 is the Laplace single layer). The regular rule must have 6 points.
 */
int bem_adjoint_single_layer(const bem_space_t *space, const bem_quad_rule_t *rule, double k, int precision,
                             double mixed_near_factor, const void *x, const void *y, int n_near,
                             const int32_t *near_pairs, int n_singular, const int32_t *singular_pairs,
                             const double *test_points, const double *trial_points, const double *weights,
                             double *grad_vertices, void *stream);

/*
 * Backward of the DP0 single-layer potential u = P c at n_points points (FP64;
 * bempp_cl/core/numba_kernels.py:1788-1849 is the forward): for the cotangent g of u
 * (complex128 [n_points]) adds P^H g to grad_coefficients (complex128 [global dofs], may be
 * NULL) and the vertex gradient of Re sum_p conj(g_p) u_p to grad_vertices (may be NULL).
 * coefficients: complex128 [global dofs]; points: float64 [3 * n_points].
 */
int bem_adjoint_potential_single_layer(const bem_space_t *space, const bem_quad_rule_t *rule, double k,
                                       const void *coefficients, int n_points, const double *points,
                                       const void *cotangent, void *grad_coefficients, double *grad_vertices,
                                       void *stream);

/*
 * One classical Gram-Schmidt step with reorthogonalisation (CGS2) for GMRES on device vectors
 * (no reference counterpart: the reference runs scipy.sparse.linalg.gmres on the host,
 * bempp_cl/api/linalg/iterative_solvers.py:137-140). V: k >= 1 complex128 rows of length n with
 * row stride ldv; w: complex128 vector of length n, overwritten with
 *   w2 = w1 - V^T c2,  w1 = w - V^T c1,  c1 = conj(V) w,  c2 = conj(V) w1.
 * out: device buffer of 2k + 2 complex128 values, overwritten with c1 (k values), |w|^2 (real
 * part), c2 (k values), |w2|^2. Sums are FP64 with atomics (summation order not fixed).
 * v_next (optional, length n, may be a row of V past k): receives the next basis vector
 * w2 / |w2|, or w2 itself when |w2| <= eps |w| (breakdown), so that the next step can start
 * before the host has read out.
 */
int bem_gmres_cgs2(const void *V, int64_t ldv, int k, int64_t n, void *w, void *out, void *v_next, double eps,
                   void *stream);

/*
 * y = op(A) x for a dense row-major matrix in HBM (the product of the reference's
 * DenseDiscreteBoundaryOperator, bempp_cl/api/assembly/discrete_boundary_operator.py).
 * trans: 0 = A x, 1 = A^T x, 2 = A^H x. a_dtype: enum bem_dtype. x, y: complex128 device
 * vectors (y is overwritten). Sums are accumulated in FP64.
 */
int bem_dense_matvec(int trans, const void *A, int a_dtype, int64_t rows, int64_t cols, int64_t ld, const void *x,
                     void *y, void *stream);

/*
 * Regular part of a symmetric single layer (BEM_LAPLACE_SINGLE_LAYER, BEM_HELMHOLTZ_SINGLE_LAYER)
 * on one piecewise constant space whose dof of list position p is p (no reference counterpart:
 * the reference computes every ordered pair, bempp_cl/core/numba_kernels.py:980-1074). Only the
 * pairs with the trial element later in the list are evaluated; each is stored in the upper
 * triangle and at its mirror image in the lower one by the same launch. out (n x n,
 * n = space->n_elements, leading dimension ld) is OVERWRITTEN,
 * with zeros on the diagonal and at the pairs that share a vertex (and, in MIXED precision, at
 * the near pairs): add bem_near_assemble and bem_singular_assemble afterwards.
 */
int bem_regular_assemble_symmetric(int op, const bem_space_t *space, const bem_quad_rule_t *rule, double k_re,
                                   double k_im, int precision, double mixed_near_factor, int kahan, void *out,
                                   int64_t ld, int out_dtype, void *stream);

/* A <- A + A^T for a square n x n matrix (leading dimension ld, dtype BEM_F32/F64/C64/C128),
 * in place (completes the single layer of bem_regular_assemble_sl_dl_symmetric). */
int bem_dense_add_transpose(void *A, int dtype, int64_t n, int64_t ld, void *stream);

/* A <- 0 for a square n x n matrix (leading dimension ld), with a fill kernel (no memset node,
 * so the call can be captured in a HIP graph). */
int bem_dense_zero(void *A, int dtype, int64_t n, int64_t ld, void *stream);

/* Copy the upper triangle of a square n x n matrix into its lower triangle and zero the diagonal
 * (no reference counterpart). */
int bem_dense_mirror_upper(void *A, int dtype, int64_t n, int64_t ld, void *stream);

/* Set `bytes` bytes at `ptr` to zero with a fill kernel (safe inside graph capture). */
int bem_fill_zero(void *ptr, size_t bytes, void *stream);

/* Message of the last error on this thread ("" if none). */
const char *bem_last_error(void);

/* Library version string. */
const char *bem_version(void);

#ifdef __cplusplus
}
#endif

#endif /* BEMPP_HIP_H */
