/*
 * vlp2d_mg.h -- C ABI of the gfx950 multigrid Dirichlet solver for the
 * 5-point Poisson/Laplace problem on a rectangle (u_xx + u_yy = -f).
 *
 * Drop-in boundary: this ABI replaces, for one solver, the reference's
 * solver-method interface IScheme<T> = ISchemeInit<T> + ISchemeCalculate<T>
 * (VisualLaplacePoisson2D/Model/CPU/IScheme.cs:16-45) as implemented by
 * MultiGridSchemeCU<T> (VisualLaplacePoisson2D/Model/CUDA/MultiGridSchemeCU.cs)
 * on top of Iterative1DScheme<T> (VisualLaplacePoisson2D/Model/CUDA/
 * Iterative1DScheme.cs:13-70).  A C# P/Invoke class binds these entry points.
 *
 * Conventions
 *   - cdecl, blittable arguments only, no exceptions cross the boundary.
 *   - Every int-returning call returns VLP2D_OK (0) or a negative VLP2D_E_*
 *     code; vlp2d_mg_last_error() then describes the failure.
 *   - Grids: segX x segY segments, nx = segX+1 by ny = segY+1 nodes, host
 *     arrays row-major with index i*ny + j (j contiguous), i along x -- the
 *     layout of the reference GPU schemes (MultiGridSchemeCU.cs:41-45).
 *   - segX and segY must be powers of two >= 4 and <= 32768 (the reference
 *     rounds segments up to powers of two before building the scheme,
 *     VisualLaplacePoisson2D/Model/CPU/StepsRecalculator.cs:168-172).
 *   - Calls on one handle are serialised by the caller (one worker thread,
 *     VisualLaplacePoisson2D/Model/CPU/VLPRectangleModel.cs:494-502), except
 *     vlp2d_mg_cancel(), which may come from any thread.  Handles share no
 *     state: different handles may be driven from different threads at once
 *     (each has its own non-blocking stream; hipGraph capture is
 *     thread-local).
 */
#ifndef VLP2D_MG_H
#define VLP2D_MG_H

#ifdef __cplusplus
extern "C" {
#endif

#define VLP2D_OK 0
#define VLP2D_E_ARG (-1)       /* null pointer, bad option value            */
#define VLP2D_E_SIZE (-2)      /* segments not a power of two in [4, 32768] */
#define VLP2D_E_DEVICE (-3)    /* no such device, or a HIP runtime failure  */
#define VLP2D_E_OOM (-4)       /* device allocation failed                  */
#define VLP2D_E_NONFINITE (-5) /* NaN/Inf in inputs or in the iterate       */
#define VLP2D_E_NOTCONV (-6)   /* max_cycles reached before rtol            */
#define VLP2D_E_CANCELLED (-7) /* vlp2d_mg_cancel() was called              */
#define VLP2D_E_STATE (-8)     /* call out of order                          */

/* Precision of the V-cycles. */
#define VLP2D_PREC_FP64 0  /* plain FP64 multigrid                            */
#define VLP2D_PREC_MIXED 1 /* FP32 V-cycles inside FP64 iterative refinement  */

#define VLP2D_SMOOTH_RBGS 0    /* red-black Gauss-Seidel, red first (the reference's) */
#define VLP2D_SMOOTH_WJACOBI 1 /* weighted Jacobi u += omega (g - u), g the same
                                  point formula; no counterpart in the reference;
                                  fused streaming kernels like Gauss-Seidel's
                                  (VLP2D_FLAG_SIMPLE: one operator per launch) */

#define VLP2D_RESTRICT_HALF 0 /* half weighting 1/8 [0 1 0; 1 4 1; 0 1 0] (the
                                 reference's, MultiGridScheme.cs:145-169)      */
#define VLP2D_RESTRICT_FULL 1 /* full weighting 1/16 [1 2 1; 2 4 2; 1 2 1]; no
                                 counterpart in the reference; in the fused
                                 descent kernels' restriction epilogue       */

#define VLP2D_CYCLE_V 0   /* V-cycle per iteration (the reference's doIteration,
                             MultiGridScheme.cs:65)                          */
#define VLP2D_CYCLE_FMG 1 /* FMG per iteration: the reference's unused FMG()
                             (MultiGridScheme.cs:117-132); FP64, one operator
                             per launch                                      */

/* Option flags (vlp2d_mg_opts.flags). */
#define VLP2D_FLAG_NO_GRAPH 1u   /* launch kernels directly, no hipGraph replay   */
#define VLP2D_FLAG_PROFILE 2u    /* time every kernel class with hipEvents        */
#define VLP2D_FLAG_SIMPLE 4u     /* one operator per launch (no fused kernels)    */
#define VLP2D_FLAG_SEPARATE_REFINE 8u /* mixed: FP64 update as its own pass      */
#define VLP2D_FLAG_COARSE_L2 16u /* small levels' one-workgroup V-cycle in global
                                     memory (L2) instead of LDS             */
#define VLP2D_FLAG_STORE_PRESMOOTH 32u /* zero-initialised levels with one pre-
                                     smoothing sweep: the descent stores the
                                     iterate and the ascent reads it back
                                     (default: the ascent recomputes it)    */
#define VLP2D_FLAG_NO_CYCLE_FUSION 128u /* mixed two-cycle steps: level 0 of
                                     the two cycles in separate launches
                                     (default: one launch between them)     */
#define VLP2D_FLAG_GRAPH_SOLVE 256u /* vlp2d_mg_solve replays hipGraphs of its
                                     steps (default: direct launches, one
                                     step launched ahead)                   */
#define VLP2D_FLAG_STREAM_SMALL 64u /* levels below level 0 with <= 1025^2
                                     nodes ascend with the register-streaming
                                     kernel (default: the LDS-tiled one)    */
#define VLP2D_FLAG_NO_STEP_FOLD 512u /* mixed solves: each refinement step's
                                     first descent in a launch of its own
                                     (default with one pre-sweep: folded into
                                     the previous step's refinement ascent) */
#define VLP2D_FLAG_NO_LT_PAIR 1024u /* FP32 cycles: one launch per LDS-
                                     tiled ascent (default: two such levels
                                     per launch, the coarser recomputed per
                                     tile)                                  */

typedef struct vlp2d_mg_ctx *vlp2d_mg_handle;

typedef struct {
  int nu1;         /* pre-smoothing RB-GS sweeps            (reference: 3) */
  int nu2;         /* post-smoothing RB-GS sweeps           (reference: 3) */
  int nu_coarse;   /* sweeps on the coarsest level          (reference: 3) */
  int precision;   /* VLP2D_PREC_FP64 or VLP2D_PREC_MIXED                  */
  unsigned flags;  /* VLP2D_FLAG_*                                          */
  int coarse_nodes;/* levels with nx*ny <= this run inside one workgroup;
                      0 = library default, -1 = never                       */
  int max_inner;   /* mixed solves: at most this many FP32 V-cycles per
                      FP64 refinement step, in [1,8] (default 1).  Step k
                      runs the fewest cycles that the contraction per cycle
                      measured on step k-1 predicts reach rtol, capped at
                      max_inner (step 0 runs max_inner).                    */
  int smoother;    /* VLP2D_SMOOTH_* (default RBGS)                         */
  double omega;    /* weighted-Jacobi weight in (0, 1] (default 0.8)        */
  int restriction; /* VLP2D_RESTRICT_* (default HALF)                       */
  int cycle;       /* VLP2D_CYCLE_* (default V)                             */
} vlp2d_mg_opts;

typedef struct {
  int cycles;             /* V-cycles (FP64) or refinement steps (mixed) run */
  int levels;             /* multigrid levels                                 */
  double r0;              /* ||fn - Lap_h u0||_2 over interior nodes (FP64)   */
  double rel_residual;    /* final ||r||_2 / r0                               */
  double seconds;         /* device time of the solve, hipEvents              */
  double munknowns_per_s; /* interior unknowns / seconds / 1e6                */
  int vcycles;            /* V-cycles run from level 0 (mixed: over all
                             refinement steps; FMG: two per iteration)       */
} vlp2d_mg_stats;

typedef struct {
  int levels;
  int nx, ny;                    /* level-0 nodes                           */
  long long device_bytes;        /* device memory held by the handle now    */
  long long device_peak_bytes;   /* high-water mark of device_bytes         */
  int checkpoints;               /* forward cycles recorded for the adjoint */
} vlp2d_mg_info;

/* Reference defaults: V(3,3), 3 coarsest sweeps, FP64, graphs on. */
void vlp2d_mg_default_opts(vlp2d_mg_opts *opts);

/* Constructor.  Replaces MultiGridSchemeCU<T>(segX, segY, hx, hy, fKsi, eps,
 * device) (MultiGridSchemeCU.cs:32-134).  f: PDE right-hand side sampled at
 * all nx*ny nodes (boundary entries ignored), or NULL for the Laplace
 * equation; the library stores fn = -f as the reference does
 * (MultiGridSchemeCU.cs:41-45).  eps is kept for the reference stop rule and
 * reported back by vlp2d_mg_cycle's caller contract.  opts may be NULL. */
int vlp2d_mg_create(int segX, int segY, double hx, double hy, const double *f, double eps, int device,
                    const vlp2d_mg_opts *opts, vlp2d_mg_handle *out);

/* Uploads the level-0 iterate: Dirichlet boundary values plus the initial
 * guess, nx*ny doubles.  Replaces initAfterBoundariesAndInitialIterationInited
 * (MultiGridSchemeCU.cs:263-266). */
int vlp2d_mg_upload(vlp2d_mg_handle h, const double *u);

/* One multigrid step.  Replaces doIteration (MultiGridSchemeCU.cs:140-146).
 * FP64: one V(nu1,nu2) cycle; *stop_metric = max over interior nodes of
 * |u - u_before_post_smoothing| on level 0, the quantity the reference
 * compares with eps (MultiGridScheme.cs:69,111).  Mixed: one FP32 V-cycle on
 * the FP64 residual followed by the FP64 update; *stop_metric = max |update|.
 * stop_metric may be NULL. */
int vlp2d_mg_cycle(vlp2d_mg_handle h, double *stop_metric);

/* Downloads the level-0 iterate (nx*ny doubles).  Replaces getArray
 * (MultiGridSchemeCU.cs:252-261). */
int vlp2d_mg_download(vlp2d_mg_handle h, double *u);

/* ||fn - Lap_h u||_2 over interior nodes, FP64 (no counterpart in the
 * reference; the relative-residual stop of vlp2d_mg_solve uses it). */
int vlp2d_mg_residual_norm(vlp2d_mg_handle h, double *norm);

/* Runs cycles from the current iterate until ||r||_2 <= rtol * ||r_0||_2 or
 * max_cycles.  Replaces the caller's loop `for iter < maxIters:
 * doIteration` (VLPRectangleModel.cs:579-614) with a device-side loop.
 * Returns VLP2D_E_NOTCONV (stats filled) if rtol was not reached. */
int vlp2d_mg_solve(vlp2d_mg_handle h, double rtol, int max_cycles, vlp2d_mg_stats *stats);

/* Gradient of a loss L(u) through the exact discrete solve (no counterpart in
 * the reference).  dL_du: nx*ny (interior entries used).  Solves the adjoint
 * system A lambda = dL_du with the same multigrid (A = -Lap_h is symmetric)
 * to rtol, then dL_df = lambda (interior; 0 on the boundary) and dL_dg =
 * B^T lambda (boundary nodes; 0 inside and at the four corners), where g are
 * the Dirichlet values.  The handle's iterate is left unchanged. */
int vlp2d_mg_adjoint_exact(vlp2d_mg_handle h, const double *dL_du, double rtol, int max_cycles, double *dL_df,
                           double *dL_dg);

/* Runs K FP64 V(nu1,nu2) cycles from the current iterate in Poisson form and
 * records K, so that vlp2d_mg_adjoint_cycles can differentiate
 * u_K = Phi^K(u_0; f, g).  One V-cycle is an affine map of (u, f, g), so the
 * transposed sweep needs no forward state: no iterate is stored per cycle
 * and the call allocates nothing beyond the solver's own hierarchy. */
int vlp2d_mg_forward_cycles(vlp2d_mg_handle h, int K);

/* Transposed sweep through the K cycles recorded by vlp2d_mg_forward_cycles.
 * dL_du: adjoint of u_K (nx*ny).  Outputs (nx*ny each, any may be NULL):
 * dL_df (interior), dL_dg (boundary nodes), dL_du0 (interior initial guess). */
int vlp2d_mg_adjoint_cycles(vlp2d_mg_handle h, const double *dL_du, double *dL_df, double *dL_dg,
                            double *dL_du0);

/* Sizes and device-memory accounting of the handle. */
int vlp2d_mg_get_info(vlp2d_mg_handle h, vlp2d_mg_info *info);

/* Per-kernel-class accounting, one entry per (kernel class, level) that was
 * launched: launches, total milliseconds (hipEvents; collected only under
 * VLP2D_FLAG_PROFILE, else 0) and algorithmic bytes (the compulsory HBM
 * traffic of the operators launched).  Returns the number of entries; at
 * most cap are written.  names receives pointers to static strings.  No
 * counterpart in the reference (its CUDA event helper, UtilsCU.cs:215-223,
 * is unused by MultiGrid). */
int vlp2d_mg_kernel_times(vlp2d_mg_handle h, const char **names, int *levels, long long *launches, double *total_ms,
                          double *alg_bytes, int cap);
int vlp2d_mg_reset_kernel_times(vlp2d_mg_handle h);

/* Device-side copy of the level-0 iterate and its restoration (no PCIe
 * transfer), so repeated solves start from the same state.  Restoring also
 * clears a pending cancel request and the recorded forward cycles.  No
 * counterpart in the reference, whose model re-runs the whole setup. */
int vlp2d_mg_save_iterate(vlp2d_mg_handle h);
int vlp2d_mg_restore_iterate(vlp2d_mg_handle h);

/* Requests cancellation; safe from any thread.  The next cycle boundary of a
 * running vlp2d_mg_solve returns VLP2D_E_CANCELLED.  Replaces
 * cancelIterations (MultiGridSchemeCU.cs:270, a no-op in the reference). */
void vlp2d_mg_cancel(vlp2d_mg_handle h);

/* Frees all device memory and the handle.  Replaces cleanup
 * (MultiGridSchemeCU.cs:272-285).  NULL is accepted. */
void vlp2d_mg_destroy(vlp2d_mg_handle h);

/* ---- problem set-up on the device --------------------------------------
 * Expressions of x and y in the syntax of the reference's input files
 * (VLP2D_Data/InputData/<name>.txt, read by ViewModel/InputParser.cs:50-152):
 * + - * / ^ (power binds tightest and is left-associative, as in
 * Substitutor/OperationsRegistry.cs:64-84), unary -, c ? a : b, numbers,
 * x, y, sin cos tan asin acos atan sinh cosh tanh exp ln log log10 sqrt abs,
 * and pi(a) = pi*a (Common/Utils.cs:352-364).  The reference evaluates them
 * with the binary-only ELW library through managed delegates; here they are
 * compiled once and evaluated on the GPU.  Node coordinates are x = hx*i,
 * y = hy*j. */

/* Constructor with the right-hand side as an expression: fn = -f(hx*i, hy*j)
 * is sampled on the GPU (no host array, no PCIe copy).  f_expr NULL or
 * blank = Laplace.  Replaces the constructor's host sampling through fKsi
 * (MultiGridSchemeCU.cs:41-45).  A syntax error returns VLP2D_E_ARG, a
 * non-finite f VLP2D_E_NONFINITE (message in vlp2d_mg_last_error(NULL)). */
int vlp2d_mg_create_expr(int segX, int segY, double hx, double hy, const char *f_expr, double eps, int device,
                         const vlp2d_mg_opts *opts, vlp2d_mg_handle *out);

/* Samples the Dirichlet data into the level-0 iterate on the GPU: bottom
 * (j = 0) and top (j = ny-1) edges from the edge functions of x, or the
 * boundary function at (x, 0) and (x, yMax); then left (i = 0) and right
 * (i = nx-1) from the edge functions of y, or the boundary function at (0, y)
 * and (xMax, y), overwriting the corners.  An edge function may name its
 * variable x or y.  NULL or blank = absent; the boundary function is required
 * unless all four edge functions are given (InputParser.cs:190-203).
 * *vmin / *vmax (may be NULL) receive the min / max of every value written.
 * Replaces initTopBottomBorders + initLeftRightBorders (UtilsBorders.cs:
 * 115-144, 210-239; order VLPRectangleModel.cs:541-542). */
int vlp2d_mg_set_boundary_expr(vlp2d_mg_handle h, const char *bottom, const char *top, const char *left,
                               const char *right, const char *boundary, double *vmin, double *vmax);

#define VLP2D_INIT_MEAN 0            /* interior = value (NaN: (min+max)/2 of the sampled boundary) */
#define VLP2D_INIT_ARITHMETIC_MEAN 1 /* mean of the four edge values of the node's row and column  */
#define VLP2D_INIT_LINEAR 2          /* mean of the linear interpolations along x and along y       */
#define VLP2D_INIT_WEIGHT_LINEAR 3   /* the two interpolations weighted by distance to the edges    */

/* Initial guess of the interior on the GPU from the boundary values in place.
 * Replaces initInitialIteration{Mean, ArithmeticMean, LinearInterpolation,
 * WeightLinearInterpolation} (UtilsII.cs:15-105; weights computed in FP32 as
 * there; Mean value VLPRectangleModel.cs:543-545). */
int vlp2d_mg_init_guess(vlp2d_mg_handle h, int kind, double value);

/* min / max of the level-0 iterate over all nodes (NaN ignored).  Replaces
 * pointsMinMax (UtilsDiff.cs:19-40). */
int vlp2d_mg_points_min_max(vlp2d_mg_handle h, double *vmin, double *vmax);

/* u - analytic(x, y) on nodes i >= 1, j >= 1: its min / max (NaN u skipped)
 * and, when diff is not NULL, the whole array (nx*ny, row 0 and column 0 set
 * to 0).  Replaces calculateDifference (UtilsDiff.cs:42-88). */
int vlp2d_mg_difference(vlp2d_mg_handle h, const char *analytic, double *dmin, double *dmax, double *diff);

/* Host evaluation of an expression at n points (for hosts without the ELW
 * evaluator; the GPU evaluates the same compiled program).  Errors in
 * vlp2d_mg_last_error(NULL). */
int vlp2d_expr_eval(const char *expr, const double *x, const double *y, long long n, double *out);

/* ---- allocation-free solve ----------------------------------------------
 * No counterpart in the reference.  For callers that cannot allocate device
 * memory (code run between hipGraph captures, tight serving loops, harnesses
 * that forbid hipMalloc).  The solve of vlp2d_mg_create + vlp2d_mg_upload +
 * vlp2d_mg_solve + vlp2d_mg_download on the same inputs and options (the
 * same iterate, bit for bit, and the same cycle count), on caller-owned
 * device arrays.  Every other buffer (the level hierarchy, partial sums,
 * result scalars) is carved from one static device array of
 * VLP2D_NOALLOC_SCRATCH_BYTES per GPU, which the library holds from its first
 * use on that GPU until the process ends; the call itself calls no hipMalloc,
 * hipFree or hipHostMalloc and captures no graph.
 *   f_dev  device, PDE right-hand side f (as vlp2d_mg_create), NULL = Laplace
 *   u_dev  device, in: Dirichlet boundary values + initial guess; out: the
 *          iterate at the end of the solve
 *   pitch  elements per row of both arrays: ny rounded up to a multiple of 32
 *          (the library's row layout); rows i = 0..nx-1, 16-byte aligned
 *   stream a hipStream_t, NULL = the null stream; the call returns after the
 *          solve has finished on it
 * Calls on one GPU share its scratch array, which keeps the last call's
 * contents (every buffer is cleared before use); the library serialises
 * them, so overlapping calls on one GPU wait for each other rather than
 * race.  A grid whose hierarchy does not fit returns VLP2D_E_SIZE before any
 * kernel runs (64 MiB holds 1025^2 in either precision and 2049 x 1025 in
 * FP64).  opts as for vlp2d_mg_create (NULL = defaults); VLP2D_FLAG_PROFILE
 * is rejected.  Returns as vlp2d_mg_solve (stats may be NULL); the message of
 * a failure in vlp2d_mg_last_error(NULL). */
#define VLP2D_NOALLOC_SCRATCH_BYTES (64u << 20)
int vlp2d_mg_solve_noalloc(int segX, int segY, double hx, double hy, const double *f_dev, double *u_dev, int pitch,
                           double rtol, int max_cycles, const vlp2d_mg_opts *opts, int device, void *stream,
                           vlp2d_mg_stats *stats);

/* Message of the last failure on h; with h == NULL, of the last failed
 * vlp2d_mg_create / vlp2d_mg_create_expr / vlp2d_expr_eval on the calling
 * thread.  Never NULL. */
const char *vlp2d_mg_last_error(vlp2d_mg_handle h);

/* ABI version, (major << 16) | minor. */
int vlp2d_mg_version(void);

#ifdef __cplusplus
}
#endif

#endif /* VLP2D_MG_H */
