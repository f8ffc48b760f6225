/* Host layer above the solver's C ABI (libvlp2d_host.so): the reference's
 * input files, its model loop and its IScheme-shaped scheme object, so a run
 * of VLP2D's MultiGrid method can be driven end to end without the WPF
 * application.
 *
 * C++ users include vlp2d_host.hpp (MultiGridSchemeHIP, parse_input_file,
 * run_model); the functions below are the same entry points for FFI.  All
 * types are blittable. */
#ifndef VLP2D_HOST_H
#define VLP2D_HOST_H

#include "vlp2d_mg.h"

#ifdef __cplusplus
extern "C" {
#endif

#define VLP2D_EXPR_LEN 1024

/* One input file (VLP2D_Data/InputData/<name>.txt), as InputParser.parseFile
 * (ViewModel/InputParser.cs:50-152) reads it: "key=value" lines, unknown keys
 * ignored, empty values absent.  Function texts have xMax / yMax replaced by
 * the texts of those keys (InputParser.cs:107-114). */
typedef struct {
  char name[256];                     /* file name without extension */
  char x_max[VLP2D_EXPR_LEN];         /* expressions, e.g. "exp(1)" */
  char y_max[VLP2D_EXPR_LEN];
  int seg_x, seg_y;                   /* segmentsX, segmentsY as written */
  char f_rhs[VLP2D_EXPR_LEN];         /* funcRHS (f of u_xx + u_yy = -f); "" = Laplace */
  char f_left[VLP2D_EXPR_LEN], f_right[VLP2D_EXPR_LEN], f_top[VLP2D_EXPR_LEN], f_bottom[VLP2D_EXPR_LEN];
  char f_boundary[VLP2D_EXPR_LEN];
  char f_analytic[VLP2D_EXPR_LEN];    /* "" = none */
  double eps_double;                  /* epsilonDouble, else epsilon, else 1e-13
                                         (VLPRectangleModel.cs:171-182) */
  int mg_iters[4];                    /* MultiGridDouble (else MultiGrid) per initial guess
                                         (Mean, ArithmeticMean, Linear, WeightLinear);
                                         5000 when absent (InputParser.cs:268-286) */
} vlp2d_input;

/* Parses path into *out.  Returns VLP2D_OK, or VLP2D_E_ARG with the messages
 * (missing required keys xMax yMax segmentsX segmentsY, missing boundary
 * functions, unparsable numbers) in err (errlen bytes, may be NULL). */
int vlp2d_input_parse(const char *path, vlp2d_input *out, char *err, int errlen);

typedef struct {
  int seg_x, seg_y;      /* after power-of-two sizing (StepsRecalculator.cs:26-30) */
  double hx, hy;         /* xMax / seg_x, yMax / seg_y (StepsRecalculator.cs:226-233) */
  double eps;            /* stop threshold of doIteration */
  int cycles;            /* doIteration calls made */
  int elapsed_iters;     /* the reference's count: index of the cycle that met eps (min 1),
                            VLPRectangleModel.cs:579-634 */
  int converged;         /* the last update norm was below eps */
  double last_metric;    /* update norm of the last cycle */
  double init_seconds;   /* create + boundary + initial guess */
  double solve_seconds;  /* the cycles */
  double vmin, vmax;     /* pointsMinMax of the result */
  double dmin, dmax;     /* calculateDifference against funcAnalytic (NaN without one) */
} vlp2d_run_result;

/* Runs the MultiGrid method on one input: power-of-two sizing, steps, device
 * sampling of f and of the boundary, initial guess init_kind
 * (VLP2D_INIT_*; VLP2D_INIT_MEAN takes (min + max) / 2 of the boundary),
 * doIteration until the update norm is below eps or max_cycles (<= 0: the
 * file's budget for that initial guess), then pointsMinMax and
 * calculateDifference.  Mirrors VLPRectangleModel.allIterations_DoWork
 * (VLPRectangleModel.cs:560-648).  opts may be NULL (reference parity mode). */
int vlp2d_run_input(const vlp2d_input *in, int init_kind, int device, const vlp2d_mg_opts *opts, int max_cycles,
                    vlp2d_run_result *out, char *err, int errlen);

#ifdef __cplusplus
}
#endif

#endif /* VLP2D_HOST_H */
