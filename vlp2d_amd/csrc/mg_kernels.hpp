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

}  // namespace vlp2d
