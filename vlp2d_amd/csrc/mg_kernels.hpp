// gfx950 kernels of the multigrid path: one operator per launch.
//
// Layout: every level is a row-major (nx, pitch) array, pitch >= ny a multiple
// of 32 elements, node (i, j) at i*pitch + j, j contiguous.  Thread blocks are
// 64 x 4: the 64 lanes of a wavefront walk consecutive j (or consecutive
// same-colour j for the smoother) of one row, so every load is a coalesced
// row segment.
//
// Arithmetic follows the reference point formulas operation by operation
// (the library is built with -ffp-contract=off), so results are bitwise equal
// to a CPU evaluation of the same formulas:
//   smoother   VisualLaplacePoisson2D/Model/CPU/MultiGridSlidingIteration.cs:53-57
//   residual   VisualLaplacePoisson2D/Common/UtilsOpLap.cs:15-20
//   restrict   VisualLaplacePoisson2D/Model/CPU/MultiGridScheme.cs:145-169
//   prolongate VisualLaplacePoisson2D/Model/CPU/MultiGridScheme.cs:171-186
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace vlp2d {

constexpr int BX = 64;  // lanes along j
constexpr int BY = 4;   // rows per block
constexpr int RED_THREADS = 1024;

// ---------------------------------------------------------------------------
// Reductions (fixed shapes, fixed order: deterministic)
// ---------------------------------------------------------------------------
__device__ inline double wave_sum(double v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
  return v;
}
__device__ inline double nanmax(double a, double b) { return (b > a || b != b) ? b : a; }
__device__ inline double wave_max(double v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v = nanmax(v, __shfl_down(v, off, 64));
  return v;
}

// Block-wide reduction; result valid in thread 0.  nthreads multiple of 64.
template <bool MAX>
__device__ inline double block_reduce(double v, double *sm, int tid, int nthreads) {
  v = MAX ? wave_max(v) : wave_sum(v);
  int lane = tid & 63, w = tid >> 6, nw = nthreads >> 6;
  __syncthreads();
  if (lane == 0) sm[w] = v;
  __syncthreads();
  if (w == 0) {
    v = lane < nw ? sm[lane] : 0.0;
    v = MAX ? wave_max(v) : wave_sum(v);
  }
  return v;
}

// out[0] = sum (or max) of part[0..n), one block of RED_THREADS threads.
template <bool MAX>
__global__ void __launch_bounds__(RED_THREADS) k_reduce(const double *__restrict__ part, int n,
                                                        double *__restrict__ out) {
  __shared__ double sm[RED_THREADS / 64];
  double v = 0.0;
  for (int k = threadIdx.x; k < n; k += RED_THREADS) v = MAX ? nanmax(v, part[k]) : v + part[k];
  v = block_reduce<MAX>(v, sm, threadIdx.x, RED_THREADS);
  if (threadIdx.x == 0) out[0] = v;
}

// The two result scalars of a cycle, published to mapped, coherent pinned
// host memory (k_finish): values first, then, after a system-scope fence, the
// sequence number the host polls for.  The host thus learns a cycle's result
// without a device-to-host copy and a stream synchronisation.
struct HostPub {
  double v[3];  // dscal[0..2]
  unsigned long long seq;
};
constexpr int kPubRing = 4;  // result s goes to slot s % kPubRing

// The reductions that end a cycle or a refinement step, in one launch:
// dscal[sum_slot] = sum of part_sum[0..nsum); with MAX also dscal[0] = max of
// part_max[0..nmax) (NaN-propagating); then, with pub, dscal[0..2] are
// published as result s = dseq + 1 into ring slot s % kPubRing.  Same
// summation order as k_reduce.
template <bool MAX>
__global__ void __launch_bounds__(RED_THREADS) k_finish(const double *__restrict__ part_max, int nmax,
                                                        const double *__restrict__ part_sum, int nsum,
                                                        double *__restrict__ dscal, int sum_slot,
                                                        unsigned long long *dseq, HostPub *pub) {
  __shared__ double sm[RED_THREADS / 64];
  double v = 0.0;
  for (int k = threadIdx.x; k < nsum; k += RED_THREADS) v = v + part_sum[k];
  v = block_reduce<false>(v, sm, threadIdx.x, RED_THREADS);
  if (threadIdx.x == 0) dscal[sum_slot] = v;
  if constexpr (MAX) {
    double w = 0.0;
    for (int k = threadIdx.x; k < nmax; k += RED_THREADS) w = nanmax(w, part_max[k]);
    w = block_reduce<true>(w, sm, threadIdx.x, RED_THREADS);
    if (threadIdx.x == 0) dscal[0] = w;
  }
  if (pub && threadIdx.x == 0) {
    const unsigned long long s = dseq[0] + 1;
    dseq[0] = s;
    HostPub *slot = pub + (s % kPubRing);
    slot->v[0] = dscal[0];
    slot->v[1] = dscal[1];
    slot->v[2] = dscal[2];
    __threadfence_system();
    __hip_atomic_store(&slot->seq, s, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_SYSTEM);
  }
}

// ---------------------------------------------------------------------------
// Red-black Gauss-Seidel half-sweep (one colour).  Red = (i+j) even, updated
// first (VisualLaplacePoisson2D/Common/GridIterator.cs:184-188).  Thread
// (m, i) updates j = 2m + 1 + ((i + color + 1) & 1).
// ---------------------------------------------------------------------------
template <typename T, bool EQ, bool LAP>
__device__ inline T rb_point(T uim, T uip, T ujm, T ujp, T f, T hx2, T hy2, T coef) {
  if (EQ) {
    if (LAP) return (uim + uip + ujm + ujp) * T(0.25);
    return (uim + uip + ujm + ujp - f * hx2) * T(0.25);
  } else {
    if (LAP) return coef * ((uim + uip) / hx2 + (ujm + ujp) / hy2);
    return coef * ((uim + uip) / hx2 + (ujm + ujp) / hy2 - f);
  }
}

template <typename T, bool EQ, bool LAP>
__global__ void __launch_bounds__(BX *BY)
    k_rb_half(T *u, const T *__restrict__ f, int nx, int ny, int pitch, T hx2, T hy2, T coef, int color) {
  const int m = blockIdx.x * BX + threadIdx.x;
  const int i = blockIdx.y * BY + threadIdx.y + 1;
  if (i > nx - 2) return;
  const int j = 2 * m + 1 + ((i + color + 1) & 1);
  if (j > ny - 2) return;
  const int c = i * pitch + j;
  u[c] = rb_point<T, EQ, LAP>(u[c - pitch], u[c + pitch], u[c - 1], u[c + 1], LAP ? T(0) : f[c], hx2, hy2, coef);
}

// Weighted-Jacobi sweep (an addition: the reference smooths with red-black
// Gauss-Seidel only): every interior node from the previous iterate,
// out = u + w (g - u) with g the Gauss-Seidel point formula.
template <typename T, bool EQ, bool LAP>
__global__ void __launch_bounds__(BX *BY) k_wjacobi(T *__restrict__ out, const T *__restrict__ u,
                                                    const T *__restrict__ f, int nx, int ny, int pitch, T hx2, T hy2,
                                                    T coef, T w) {
  const int j = blockIdx.x * BX + threadIdx.x + 1;
  const int i = blockIdx.y * BY + threadIdx.y + 1;
  if (i > nx - 2 || j > ny - 2) return;
  const int c = i * pitch + j;
  const T o = u[c];
  const T g = rb_point<T, EQ, LAP>(u[c - pitch], u[c + pitch], u[c - 1], u[c + 1], LAP ? T(0) : f[c], hx2, hy2, coef);
  out[c] = o + w * (g - o);
}

// ---------------------------------------------------------------------------
// Residual res = rhs - ((uW + uE - 2u)/hx2 + (uS + uN - 2u)/hy2), interior.
// ---------------------------------------------------------------------------
template <typename T>
__device__ inline T resid_point(const T *u, const T *rhs, int c, int pitch, T hx2, T hy2) {
  T u2 = u[c] * T(2);
  return rhs[c] - ((u[c - pitch] + u[c + pitch] - u2) / hx2 + (u[c - 1] + u[c + 1] - u2) / hy2);
}

template <typename T>
__global__ void __launch_bounds__(BX *BY) k_residual(T *__restrict__ res, const T *__restrict__ rhs,
                                                     const T *__restrict__ u, int nx, int ny, int pitch, T hx2,
                                                     T hy2) {
  const int j = blockIdx.x * BX + threadIdx.x + 1;
  const int i = blockIdx.y * BY + threadIdx.y + 1;
  if (i > nx - 2 || j > ny - 2) return;
  const int c = i * pitch + j;
  res[c] = resid_point(u, rhs, c, pitch, hx2, hy2);
}

// ---------------------------------------------------------------------------
// Half-weighting restriction to the coarse interior:
// rhs_c = 1/8 (4 r[2i,2j] + (r[2i+1,2j] + r[2i-1,2j] + r[2i,2j+1] + r[2i,2j-1]))
// ---------------------------------------------------------------------------
template <typename T>
__global__ void __launch_bounds__(BX *BY) k_restrict(T *__restrict__ rhsc, const T *__restrict__ res, int ncx,
                                                     int ncy, int pc, int pf) {
  const int j = blockIdx.x * BX + threadIdx.x + 1;
  const int i = blockIdx.y * BY + threadIdx.y + 1;
  if (i > ncx - 2 || j > ncy - 2) return;
  const int k = 2 * i * pf + 2 * j;
  rhsc[i * pc + j] = T(0.125) * (res[k] * T(4) + (res[k + pf] + res[k - pf] + res[k + 1] + res[k - 1]));
}

// Full-weighting restriction (no counterpart in the reference):
// rhs_c = 1/16 (4 r + 2 (N + S + E + W) + (NE + NW + SE + SW)) around r[2i,2j]
template <typename T>
__global__ void __launch_bounds__(BX *BY) k_restrict_fw(T *__restrict__ rhsc, const T *__restrict__ res, int ncx,
                                                        int ncy, int pc, int pf) {
  const int j = blockIdx.x * BX + threadIdx.x + 1;
  const int i = blockIdx.y * BY + threadIdx.y + 1;
  if (i > ncx - 2 || j > ncy - 2) return;
  const int k = 2 * i * pf + 2 * j;
  rhsc[i * pc + j] = T(0.0625) * (res[k] * T(4) + (res[k + pf] + res[k - pf] + res[k + 1] + res[k - 1]) * T(2) +
                                  (res[k + pf + 1] + res[k + pf - 1] + res[k - pf + 1] + res[k - pf - 1]));
}

// p[0..n) = v
template <typename T>
__global__ void k_fill(T *__restrict__ p, long long n, T v) {
  long long k = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  long long stride = (long long)gridDim.x * blockDim.x;
  for (; k < n; k += stride) p[k] = v;
}

// ---------------------------------------------------------------------------
// Additive bilinear prolongation over coarse cells (i, j) in
// [0, ncx-1) x [0, ncy-1); each fine interior node is written exactly once.
// ---------------------------------------------------------------------------
template <typename T>
__global__ void __launch_bounds__(BX *BY) k_prolong(T *__restrict__ uf, const T *__restrict__ uc, int ncx, int ncy,
                                                    int pf, int pc) {
  const int j = blockIdx.x * BX + threadIdx.x;
  const int i = blockIdx.y * BY + threadIdx.y;
  if (i > ncx - 2 || j > ncy - 2) return;
  const T a = uc[i * pc + j], b = uc[i * pc + j + 1];
  const T d = uc[(i + 1) * pc + j], e = uc[(i + 1) * pc + j + 1];
  const int k = 2 * i * pf + 2 * j;
  if (i > 0 && j > 0) uf[k] += a;
  if (j > 0) uf[k + pf] += (a + d) * T(0.5);
  if (i > 0) uf[k + 1] += (a + b) * T(0.5);
  uf[k + pf + 1] += (a + b + d + e) * T(0.25);
}

// ---------------------------------------------------------------------------
// Level-0 helpers
// ---------------------------------------------------------------------------
template <typename T>
__global__ void __launch_bounds__(BX *BY) k_copy_interior(T *__restrict__ dst, const T *__restrict__ src, int nx,
                                                          int ny, int pitch) {
  const int j = blockIdx.x * BX + threadIdx.x + 1;
  const int i = blockIdx.y * BY + threadIdx.y + 1;
  if (i > nx - 2 || j > ny - 2) return;
  dst[i * pitch + j] = src[i * pitch + j];
}

// Partial max over interior of |a - b| (NaN propagates).  Grid-stride in i.
template <typename T>
__global__ void __launch_bounds__(BX *BY) k_absdiff_max(const T *__restrict__ a, const T *__restrict__ b, int nx,
                                                        int ny, int pitch, double *__restrict__ part) {
  __shared__ double sm[BX * BY / 64];
  const int j = blockIdx.x * BX + threadIdx.x + 1;
  double m = 0.0;
  if (j <= ny - 2)
    for (int i = blockIdx.y * BY + threadIdx.y + 1; i <= nx - 2; i += gridDim.y * BY) {
      const int c = i * pitch + j;
      m = nanmax(m, fabs((double)(a[c] - b[c])));
    }
  const int tid = threadIdx.y * BX + threadIdx.x;
  m = block_reduce<true>(m, sm, tid, BX * BY);
  if (tid == 0) part[blockIdx.y * gridDim.x + blockIdx.x] = m;
}

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
