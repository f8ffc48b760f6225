// Problem set-up and post-processing on the GPU (gfx950): the steps the
// reference runs on the host before and after the solve.
//
//   k_expr_rhs      fn = -f(hx i, hy j) on interior nodes
//                   (Model/CUDA/MultiGridSchemeCU.cs:41-45)
//   k_border_rows   bottom (j = 0) and top (j = ny-1) edges, then
//   k_border_cols   left (i = 0) and right (i = nx-1) edges, which overwrite
//                   the corners (Common/UtilsBorders.cs:115-144, 210-239,
//                   called in that order at Model/CPU/VLPRectangleModel.cs:541-542),
//                   with the min / max of every value written
//   k_init_*        the four initial guesses of the interior
//                   (Common/UtilsII.cs:15-105)
//   k_minmax        min / max over all nodes (Common/UtilsDiff.cs:19-40)
//   k_difference    u - analytic(x, y) on nodes i >= 1, j >= 1 and its min / max
//                   (Common/UtilsDiff.cs:42-88)
//
// Expressions run through a small interpreter of the postfix programs that
// expr.hpp compiles; its evaluation stack lives in LDS (kExprMaxDepth doubles
// per thread, strided by the block size).  Node coordinates are products
// x = hx * i, y = hy * j.  The reference accumulates x += hx along a row in
// each of its host threads, so its coordinates (and sampled values) depend on
// the thread count; the product is the thread-independent value it
// approximates.  Min / max ignore NaN, as the reference's comparisons do.
#pragma once
#include "expr.hpp"
#include "mg_kernels.hpp"

namespace vlp2d {

constexpr int SETUP_THREADS = 256;
constexpr int SETUP_MAX_BLOCKS = 1024;  // partial min / max per launch

__device__ inline double dexpr_fn(int fn, double a) {
  switch (fn) {
    case FN_SIN: return sin(a);
    case FN_COS: return cos(a);
    case FN_TAN: return tan(a);
    case FN_ASIN: return asin(a);
    case FN_ACOS: return acos(a);
    case FN_ATAN: return atan(a);
    case FN_SINH: return sinh(a);
    case FN_COSH: return cosh(a);
    case FN_TANH: return tanh(a);
    case FN_EXP: return exp(a);
    case FN_LN: return log(a);
    case FN_LOG10: return log10(a);
    case FN_SQRT: return sqrt(a);
    case FN_ABS: return fabs(a);
    case FN_PI: return 3.141592653589793 * a;
    default: return __longlong_as_double(0x7ff8000000000000LL);
  }
}

// Evaluates program ops[0..n) at (x, y); stk is this thread's LDS stack
// (element d at stk[d * SETUP_THREADS]).  The host compiler guarantees a
// well-formed program no deeper than kExprMaxDepth.
__device__ inline double dexpr_eval(const ExprOp *__restrict__ ops, int n, double x, double y, double *stk) {
  constexpr int S = SETUP_THREADS;
  int sp = 0;  // number of values on the stack; element d at stk[d * S]
  for (int k = 0; k < n; ++k) {
    const ExprOp o = ops[k];
    switch (o.code) {
      case EX_CONST: stk[sp++ * S] = o.val; break;
      case EX_X: stk[sp++ * S] = x; break;
      case EX_Y: stk[sp++ * S] = y; break;
      case EX_ADD: --sp; stk[(sp - 1) * S] = stk[(sp - 1) * S] + stk[sp * S]; break;
      case EX_SUB: --sp; stk[(sp - 1) * S] = stk[(sp - 1) * S] - stk[sp * S]; break;
      case EX_MUL: --sp; stk[(sp - 1) * S] = stk[(sp - 1) * S] * stk[sp * S]; break;
      case EX_DIV: --sp; stk[(sp - 1) * S] = stk[(sp - 1) * S] / stk[sp * S]; break;
      case EX_POW: --sp; stk[(sp - 1) * S] = pow(stk[(sp - 1) * S], stk[sp * S]); break;
      case EX_POWI: stk[(sp - 1) * S] = expr_powi(stk[(sp - 1) * S], o.arg); break;
      case EX_NEG: stk[(sp - 1) * S] = -stk[(sp - 1) * S]; break;
      case EX_FN: stk[(sp - 1) * S] = dexpr_fn(o.arg, stk[(sp - 1) * S]); break;
      case EX_COND:
        sp -= 2;
        stk[(sp - 1) * S] = stk[(sp - 1) * S] != 0.0 ? stk[sp * S] : stk[(sp + 1) * S];
        break;
      default: break;
    }
  }
  return stk[0];
}

struct DevProg {
  const ExprOp *ops;  // device copy of the program (null when absent)
  int n;
};

__device__ inline void mm_update(double v, double &mn, double &mx) {
  if (v < mn) mn = v;  // NaN compares false and is skipped
  if (v > mx) mx = v;
}

// Block min / max -> part_negmin[blk] = -min, part_max[blk] = max.
__device__ inline void mm_store(double mn, double mx, double *part_negmin, double *part_max, int blk) {
  __shared__ double sm[SETUP_THREADS / 64];
  double a = -mn;
  for (int off = 32; off > 0; off >>= 1) a = fmax(a, __shfl_down(a, off, 64));
  double b = mx;
  for (int off = 32; off > 0; off >>= 1) b = fmax(b, __shfl_down(b, off, 64));
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  __syncthreads();
  if (lane == 0) sm[w] = a;
  __syncthreads();
  if (threadIdx.x == 0) {
    double r = sm[0];
    for (int k = 1; k < SETUP_THREADS / 64; ++k) r = fmax(r, sm[k]);
    part_negmin[blk] = r;
  }
  __syncthreads();
  if (lane == 0) sm[w] = b;
  __syncthreads();
  if (threadIdx.x == 0) {
    double r = sm[0];
    for (int k = 1; k < SETUP_THREADS / 64; ++k) r = fmax(r, sm[k]);
    part_max[blk] = r;
  }
}

// out[0] = min, out[1] = max from n partials (fmax ignores NaN; the empty
// min / max are +inf / -inf)
__global__ void __launch_bounds__(SETUP_THREADS) k_mm_final(const double *__restrict__ part_negmin,
                                                            const double *__restrict__ part_max, int n,
                                                            double *__restrict__ out) {
  double a = -INFINITY, b = -INFINITY;
  for (int k = threadIdx.x; k < n; k += SETUP_THREADS) {
    a = fmax(a, part_negmin[k]);
    b = fmax(b, part_max[k]);
  }
  __shared__ double sa[SETUP_THREADS], sb[SETUP_THREADS];
  sa[threadIdx.x] = a;
  sb[threadIdx.x] = b;
  __syncthreads();
  for (int s = SETUP_THREADS / 2; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) {
      sa[threadIdx.x] = fmax(sa[threadIdx.x], sa[threadIdx.x + s]);
      sb[threadIdx.x] = fmax(sb[threadIdx.x], sb[threadIdx.x + s]);
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    out[0] = -sa[0];
    out[1] = sb[0];
  }
}

// fn = -f(hx i, hy j) on interior nodes; grid-stride over rows
__global__ void __launch_bounds__(SETUP_THREADS) k_expr_rhs(DevProg f, double *__restrict__ rhs, int nx, int ny,
                                                            int pitch, double hx, double hy) {
  __shared__ double stk[kExprMaxDepth * SETUP_THREADS];
  double *my = stk + threadIdx.x;
  for (int i = 1 + blockIdx.y; i <= nx - 2; i += gridDim.y) {
    const double x = hx * (double)i;
    for (int j = 1 + blockIdx.x * SETUP_THREADS + threadIdx.x; j <= ny - 2; j += gridDim.x * SETUP_THREADS)
      rhs[(size_t)i * pitch + j] = -dexpr_eval(f.ops, f.n, x, hy * (double)j, my);
  }
}

// Bottom and top edges for i = 0..nx-1: an edge function of x when given,
// else the boundary function at (x, 0) / (x, yMax), yMax = hy (ny - 1).
__global__ void __launch_bounds__(SETUP_THREADS) k_border_rows(DevProg bottom, DevProg top, DevProg border,
                                                               double *__restrict__ u, int nx, int ny, int pitch,
                                                               double hx, double hy, double *part_negmin,
                                                               double *part_max) {
  __shared__ double stk[kExprMaxDepth * SETUP_THREADS];
  double *my = stk + threadIdx.x;
  const double ymax = hy * (double)(ny - 1);
  double mn = INFINITY, mx = -INFINITY;
  for (int i = blockIdx.x * SETUP_THREADS + threadIdx.x; i < nx; i += gridDim.x * SETUP_THREADS) {
    const double x = hx * (double)i;
    const double b = bottom.n ? dexpr_eval(bottom.ops, bottom.n, x, x, my) : dexpr_eval(border.ops, border.n, x, 0.0, my);
    const double t = top.n ? dexpr_eval(top.ops, top.n, x, x, my) : dexpr_eval(border.ops, border.n, x, ymax, my);
    u[(size_t)i * pitch] = b;
    u[(size_t)i * pitch + ny - 1] = t;
    mm_update(b, mn, mx);
    mm_update(t, mn, mx);
  }
  mm_store(mn, mx, part_negmin, part_max, blockIdx.x);
}

// Left and right edges for j = 0..ny-1 (corners overwritten): an edge
// function of y when given, else the boundary function at (0, y) / (xMax, y),
// xMax = hx (nx - 1).
__global__ void __launch_bounds__(SETUP_THREADS) k_border_cols(DevProg left, DevProg right, DevProg border,
                                                               double *__restrict__ u, int nx, int ny, int pitch,
                                                               double hx, double hy, double *part_negmin,
                                                               double *part_max) {
  __shared__ double stk[kExprMaxDepth * SETUP_THREADS];
  double *my = stk + threadIdx.x;
  const double xmax = hx * (double)(nx - 1);
  double mn = INFINITY, mx = -INFINITY;
  for (int j = blockIdx.x * SETUP_THREADS + threadIdx.x; j < ny; j += gridDim.x * SETUP_THREADS) {
    const double y = hy * (double)j;
    const double l = left.n ? dexpr_eval(left.ops, left.n, y, y, my) : dexpr_eval(border.ops, border.n, 0.0, y, my);
    const double r = right.n ? dexpr_eval(right.ops, right.n, y, y, my) : dexpr_eval(border.ops, border.n, xmax, y, my);
    u[j] = l;
    u[(size_t)(nx - 1) * pitch + j] = r;
    mm_update(l, mn, mx);
    mm_update(r, mn, mx);
  }
  mm_store(mn, mx, part_negmin, part_max, blockIdx.x);
}

// Initial guesses of the interior (Common/UtilsII.cs): 0 mean value, 1
// arithmetic mean of the four edge values of the node's row and column, 2
// linear interpolation, 3 weighted linear interpolation.  The interpolation
// weights are computed in FP32 as in the reference ((float)i / upperX).
__global__ void __launch_bounds__(SETUP_THREADS) k_init_guess(double *__restrict__ u, int nx, int ny, int pitch,
                                                              int kind, double value) {
  const int ux = nx - 1, uy = ny - 1;
  for (int i = 1 + blockIdx.y; i <= nx - 2; i += gridDim.y) {
    const double *r0 = u;                      // row 0
    const double *r1 = u + (size_t)ux * pitch;  // row upperX
    double *ri = u + (size_t)i * pitch;
    for (int j = 1 + blockIdx.x * SETUP_THREADS + threadIdx.x; j <= ny - 2; j += gridDim.x * SETUP_THREADS) {
      double v;
      if (kind == 0) {
        v = value;
      } else if (kind == 1) {
        v = (r0[j] + r1[j] + ri[0] + ri[uy]) / 4.0;
      } else {
        const double sx = (double)((float)i / (float)ux), sy = (double)((float)j / (float)uy);
        const double lx = r0[j] * (1.0 - sx) + r1[j] * sx;
        const double ly = ri[0] * (1.0 - sy) + ri[uy] * sy;
        if (kind == 2) {
          v = (lx + ly) / 2.0;
        } else {
          const int lngX = ux > i + i ? i : ux - i;
          const int lngY = uy > j + j ? j : uy - j;
          const double w = (double)((float)lngX / (float)(lngX + lngY));
          v = lx * (1.0 - w) + ly * w;
        }
      }
      ri[j] = v;
    }
  }
}

// min / max over all nodes of u
__global__ void __launch_bounds__(SETUP_THREADS) k_minmax(const double *__restrict__ u, int nx, int ny, int pitch,
                                                          double *part_negmin, double *part_max) {
  double mn = INFINITY, mx = -INFINITY;
  const int blk = blockIdx.y * gridDim.x + blockIdx.x;
  for (int i = blockIdx.y; i < nx; i += gridDim.y)
    for (int j = blockIdx.x * SETUP_THREADS + threadIdx.x; j < ny; j += gridDim.x * SETUP_THREADS)
      mm_update(u[(size_t)i * pitch + j], mn, mx);
  mm_store(mn, mx, part_negmin, part_max, blk);
}

// diff = u - analytic(hx i, hy j) for i in 1..nx-1, j in 1..ny-1 (NaN u is
// copied and skipped by min / max); optional diff array (row 0 and column 0
// hold 0).
__global__ void __launch_bounds__(SETUP_THREADS) k_difference(const double *__restrict__ u, DevProg an, int nx, int ny,
                                                              int pitch, double hx, double hy, double *diff,
                                                              double *part_negmin, double *part_max) {
  __shared__ double stk[kExprMaxDepth * SETUP_THREADS];
  double *my = stk + threadIdx.x;
  double mn = INFINITY, mx = -INFINITY;
  const int blk = blockIdx.y * gridDim.x + blockIdx.x;
  for (int i = blockIdx.y; i < nx; i += gridDim.y) {
    const double x = hx * (double)i;
    for (int j = blockIdx.x * SETUP_THREADS + threadIdx.x; j < ny; j += gridDim.x * SETUP_THREADS) {
      double d = 0.0;
      if (i >= 1 && j >= 1) {
        const double v = u[(size_t)i * pitch + j];
        const double a = dexpr_eval(an.ops, an.n, x, hy * (double)j, my);
        d = v != v ? v : v - a;
        mm_update(d, mn, mx);
      }
      if (diff) diff[(size_t)i * pitch + j] = d;
    }
  }
  mm_store(mn, mx, part_negmin, part_max, blk);
}

}  // namespace vlp2d
