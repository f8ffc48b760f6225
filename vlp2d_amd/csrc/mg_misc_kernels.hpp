// Non-template gfx950 kernels of the one-operator path: the separate mixed-
// precision update and the operators of the transposed cycle and of the exact
// adjoint.  Included by mg_lib.hip only (each TU that includes a non-template
// kernel compiles and registers its own copy).
#pragma once
#include "mg_kernels.hpp"

namespace vlp2d {

// u64 += (double) e32 on interior nodes; partial max |e|.
__global__ void __launch_bounds__(BX *BY) k_update_mixed(double *__restrict__ u, const float *__restrict__ e, int nx,
                                                         int ny, int pitch, double *__restrict__ part) {
  __shared__ double sm[BX * BY / 64];
  const int j = blockIdx.x * BX + threadIdx.x + 1;
  double m = 0.0;
  if (j <= ny - 2)
    for (int i = blockIdx.y * BY + threadIdx.y + 1; i <= nx - 2; i += gridDim.y * BY) {
      const int c = i * pitch + j;
      double d = (double)e[c];
      u[c] += d;
      m = nanmax(m, fabs(d));
    }
  const int tid = threadIdx.y * BX + threadIdx.x;
  m = block_reduce<true>(m, sm, tid, BX * BY);
  if (tid == 0) part[blockIdx.y * gridDim.x + blockIdx.x] = m;
}

// fn = -src on interior nodes (the adjoint solve's right-hand side).
__global__ void __launch_bounds__(BX *BY) k_neg_interior(double *__restrict__ dst, const double *__restrict__ src,
                                                         int nx, int ny, int pitch) {
  const int j = blockIdx.x * BX + threadIdx.x + 1;
  const int i = blockIdx.y * BY + threadIdx.y + 1;
  if (i > nx - 2 || j > ny - 2) return;
  dst[i * pitch + j] = -src[i * pitch + j];
}

// ---------------------------------------------------------------------------
// Adjoint (transposed) operators, FP64.  See the oracle's Adj for the
// derivation; the operation order matches it so results are bitwise equal.
// ---------------------------------------------------------------------------
// Transposed half-sweep, step 1: every node Q with (i+j)%2 != color (boundary
// nodes included) gathers ax/ay * ubar from its interior colour-c neighbours.
__global__ void __launch_bounds__(BX *BY) k_adj_gather(double *ub, int nx, int ny, int pitch, double ax, double ay,
                                                       int color) {
  const int j = blockIdx.x * BX + threadIdx.x;
  const int i = blockIdx.y * BY + threadIdx.y;
  if (i > nx - 1 || j > ny - 1 || ((i + j) & 1) == color) return;
  const int c = i * pitch + j;
  double s = 0.0;
  // neighbours of Q have colour c automatically; they must be interior
  if (i - 1 >= 1 && j >= 1 && j <= ny - 2) s += ax * ub[c - pitch];
  if (i + 1 <= nx - 2 && j >= 1 && j <= ny - 2) s += ax * ub[c + pitch];
  if (j - 1 >= 1 && i >= 1 && i <= nx - 2) s += ay * ub[c - 1];
  if (j + 1 <= ny - 2 && i >= 1 && i <= nx - 2) s += ay * ub[c + 1];
  ub[c] += s;
}

// Transposed half-sweep, step 2: interior colour-c nodes: rb += -b*ub; ub = 0.
__global__ void __launch_bounds__(BX *BY) k_adj_color(double *__restrict__ ub, double *__restrict__ rb, int nx, int ny,
                                                      int pitch, double b, int color) {
  const int j = blockIdx.x * BX + threadIdx.x + 1;
  const int i = blockIdx.y * BY + threadIdx.y + 1;
  if (i > nx - 2 || j > ny - 2 || ((i + j) & 1) != color) return;
  const int c = i * pitch + j;
  rb[c] += -b * ub[c];
  ub[c] = 0.0;
}

// ubar_c = P^T ubar_f on the coarse interior; coarse boundary = 0.
__global__ void __launch_bounds__(BX *BY) k_adj_prolongT(double *__restrict__ ubc, const double *__restrict__ ubf,
                                                         int ncx, int ncy, int pc, int pf) {
  const int J = blockIdx.x * BX + threadIdx.x;
  const int I = blockIdx.y * BY + threadIdx.y;
  if (I > ncx - 1 || J > ncy - 1) return;
  double s = 0.0;
  if (I >= 1 && I <= ncx - 2 && J >= 1 && J <= ncy - 2) {
    const int k = 2 * I * pf + 2 * J;
    s = ubf[k] + 0.5 * (ubf[k - pf] + ubf[k + pf] + ubf[k - 1] + ubf[k + 1]) +
        0.25 * (ubf[k - pf - 1] + ubf[k - pf + 1] + ubf[k + pf - 1] + ubf[k + pf + 1]);
  }
  ubc[I * pc + J] = s;
}

// resbar_f = R^T rhsbar_c (fine interior; zero elsewhere), gather form.
__global__ void __launch_bounds__(BX *BY) k_adj_restrictT(double *__restrict__ resb, const double *__restrict__ rbc,
                                                          int nx, int ny, int pf, int ncx, int ncy, int pc) {
  const int j = blockIdx.x * BX + threadIdx.x;
  const int i = blockIdx.y * BY + threadIdx.y;
  if (i > nx - 1 || j > ny - 1) return;
  double v = 0.0;
  if (i >= 1 && i <= nx - 2 && j >= 1 && j <= ny - 2) {
    const bool ie = (i & 1) == 0, je = (j & 1) == 0;
    auto g = [&](int I, int J) -> double {
      return (I >= 1 && I <= ncx - 2 && J >= 1 && J <= ncy - 2) ? rbc[I * pc + J] * 0.125 : 0.0;
    };
    if (ie && je) {
      if (i / 2 >= 1 && i / 2 <= ncx - 2 && j / 2 >= 1 && j / 2 <= ncy - 2) v = 0.0 + 4.0 * g(i / 2, j / 2);
    } else if (!ie && je) {
      int I0 = (i - 1) / 2, I1 = (i + 1) / 2, J = j / 2;
      bool a = I0 >= 1 && J >= 1 && J <= ncy - 2, c = I1 <= ncx - 2 && J >= 1 && J <= ncy - 2;
      if (a) v = v + g(I0, J);
      if (c) v = v + g(I1, J);
    } else if (ie && !je) {
      int J0 = (j - 1) / 2, J1 = (j + 1) / 2, I = i / 2;
      bool a = J0 >= 1 && I >= 1 && I <= ncx - 2, c = J1 <= ncy - 2 && I >= 1 && I <= ncx - 2;
      if (a) v = v + g(I, J0);
      if (c) v = v + g(I, J1);
    }
  }
  resb[i * pf + j] = v;
}

// Transpose of res = rhs - Lap(u): rhsbar += resbar (interior) and
// ubar -= Lap(resbar) on all nodes (resbar is zero outside the interior).
__global__ void __launch_bounds__(BX *BY) k_adj_residualT(double *__restrict__ ub, double *__restrict__ rb,
                                                          const double *__restrict__ resb, int nx, int ny, int pitch,
                                                          double ix2, double iy2) {
  const int j = blockIdx.x * BX + threadIdx.x;
  const int i = blockIdx.y * BY + threadIdx.y;
  if (i > nx - 1 || j > ny - 1) return;
  const int c = i * pitch + j;
  const bool inner = i > 0 && i < nx - 1 && j > 0 && j < ny - 1;
  double s = 0.0;
  if (i > 0) s += resb[c - pitch] * ix2;
  if (i < nx - 1) s += resb[c + pitch] * ix2;
  if (j > 0) s += resb[c - 1] * iy2;
  if (j < ny - 1) s += resb[c + 1] * iy2;
  if (inner) s -= (2.0 * ix2 + 2.0 * iy2) * resb[c];
  ub[c] -= s;
  if (inner) rb[c] += resb[c];
}

// gbar = B^T lambda: boundary node Q gets lambda of its interior neighbour
// divided by h^2 (corners 0); interior nodes 0.
__global__ void __launch_bounds__(BX *BY) k_adj_bT(double *__restrict__ gb, const double *__restrict__ lam, int nx,
                                                   int ny, int pitch, double ix2, double iy2) {
  const int j = blockIdx.x * BX + threadIdx.x;
  const int i = blockIdx.y * BY + threadIdx.y;
  if (i > nx - 1 || j > ny - 1) return;
  const int c = i * pitch + j;
  double v = 0.0;
  const bool jin = j >= 1 && j <= ny - 2, iin = i >= 1 && i <= nx - 2;
  if (i == 0 && jin) v = lam[c + pitch] * ix2;
  else if (i == nx - 1 && jin) v = lam[c - pitch] * ix2;
  else if (j == 0 && iin) v = lam[c + 1] * iy2;
  else if (j == ny - 1 && iin) v = lam[c - 1] * iy2;
  gb[c] = v;
}

// Outputs of the transposed sweep, packed nx x ny (row pitch ny): kind 0 the
// right-hand-side gradient -rbar (interior), 1 the Dirichlet-data gradient
// ubar (boundary), 2 the initial-iterate gradient ubar (interior); 0 elsewhere.
__global__ void __launch_bounds__(BX *BY) k_adj_output(double *__restrict__ out, const double *__restrict__ ub,
                                                       const double *__restrict__ rb, int nx, int ny, int pitch,
                                                       int kind) {
  const int j = blockIdx.x * BX + threadIdx.x;
  const int i = blockIdx.y * BY + threadIdx.y;
  if (i > nx - 1 || j > ny - 1) return;
  const int c = i * pitch + j;
  const bool inner = i > 0 && i < nx - 1 && j > 0 && j < ny - 1;
  double v = 0.0;
  if (kind == 0) v = inner ? -rb[c] : 0.0;
  else if (kind == 1) v = inner ? 0.0 : ub[c];
  else v = inner ? ub[c] : 0.0;
  out[(size_t)i * ny + j] = v;
}

}  // namespace vlp2d

namespace vlp2d {
// Out-of-place refinement update: uout = uin + (double) e on interior nodes
// (boundary nodes of uout already hold the Dirichlet values); partial max |e|.
__global__ void __launch_bounds__(BX *BY) k_update_oop(const double *__restrict__ uin, double *__restrict__ uout,
                                                       const float *__restrict__ e, int nx, int ny, int pitch,
                                                       double *__restrict__ part) {
  __shared__ double sm[BX * BY / 64];
  const int j = blockIdx.x * BX + threadIdx.x + 1;
  double m = 0.0;
  if (j <= ny - 2)
    for (int i = blockIdx.y * BY + threadIdx.y + 1; i <= nx - 2; i += gridDim.y * BY) {
      const int c = i * pitch + j;
      const double d = (double)e[c];
      uout[c] = uin[c] + d;
      m = nanmax(m, fabs(d));
    }
  const int tid = threadIdx.y * BX + threadIdx.x;
  m = block_reduce<true>(m, sm, tid, BX * BY);
  if (tid == 0) part[blockIdx.y * gridDim.x + blockIdx.x] = m;
}
}  // namespace vlp2d
