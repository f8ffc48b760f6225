// One-workgroup V-cycle over the small levels of a hierarchy (gfx950).
#pragma once
#include "mg_tiles.hpp"

namespace vlp2d {

// ---------------------------------------------------------------------------
// Small levels: one workgroup runs the V-cycle from level l0 down to the
// coarsest and back, in global memory (the working set sits in L2), with a
// workgroup barrier between dependent steps.  Level l0 is zero-initialised
// unless it is level 0 of the problem (zero_first = 0).
// ---------------------------------------------------------------------------
#ifndef VLP2D_COARSE_THREADS
#define VLP2D_COARSE_THREADS 1024
#endif
constexpr int CNT = VLP2D_COARSE_THREADS;  // threads of the one workgroup
constexpr int kMaxLevels = 16;

template <typename T> struct CoarseLevel {
  T *u;
  const T *f;
  int nx, ny, pitch;
  T hx2, hy2, coef;
  T ihx2, ihy2;
  int eq, lap, pw2;
  T *u2;  // second iterate buffer (k_coarse's weighted-Jacobi sweeps), or null
};
template <typename T> struct CoarseArgs {
  CoarseLevel<T> lv[kMaxLevels];
  int l0, L;
  int nu1, nu2, nuc;
  int zero_first;
  double *snap;  // level-0 snapshot before post-smoothing (l0 == 0), or null
  int jac;       // weighted-Jacobi smoother (weight w) instead of red-black Gauss-Seidel
  T w;
  int fw;  // full-weighting restriction instead of half weighting
};

// Gauss-Seidel point formula of node c (MultiGridSlidingIteration.cs:53-57),
// row pitch P; the weighted-Jacobi update uses the same g.
template <typename T> __device__ inline T c_point(const T *u, const T *f, int c, int P, const CoarseLevel<T> &g) {
  if (g.eq) {
    if (g.lap) return (u[c - P] + u[c + P] + u[c - 1] + u[c + 1]) * T(0.25);
    return (u[c - P] + u[c + P] + u[c - 1] + u[c + 1] - f[c] * g.hx2) * T(0.25);
  }
  const T sx = divh(u[c - P] + u[c + P], g.hx2, g.ihx2, g.pw2);
  const T sy = divh(u[c - 1] + u[c + 1], g.hy2, g.ihy2, g.pw2);
  if (g.lap) return g.coef * (sx + sy);
  return g.coef * (sx + sy - f[c]);
}

// Restriction of the residual r(c) around fine node k (row pitch P):
// half weighting (MultiGridScheme.cs:158-167) or full weighting (the oracle's
// restrict_fw), terms summed in the oracle's order.
template <typename T, typename R> __device__ inline T c_restrict(R r, int k, int P, int fw) {
  const T r0 = r(k), rip = r(k + P), rim = r(k - P), rjp = r(k + 1), rjm = r(k - 1);
  if (!fw) return T(0.125) * (r0 * T(4) + (rip + rim + rjp + rjm));
  return T(0.0625) * (r0 * T(4) + (rip + rim + rjp + rjm) * T(2) + (r(k + P + 1) + r(k + P - 1) + r(k - P + 1) + r(k - P - 1)));
}

template <typename T>
__device__ inline void cg_zero(const CoarseLevel<T> &v) {
  const int n = (v.nx - 2) * (v.ny - 2);
  for (int p = threadIdx.x; p < n; p += CNT) {
    const int i = 1 + p / (v.ny - 2), j = 1 + p - (p / (v.ny - 2)) * (v.ny - 2);
    v.u[i * v.pitch + j] = T(0);
  }
  __syncthreads();
}

template <typename T>
__device__ inline void cg_smooth(const CoarseLevel<T> &v, int n, const CoarseArgs<T> &a) {
  if (a.jac) {
    // weighted Jacobi: new values into u2, then back into u
    const int ni = (v.nx - 2) * (v.ny - 2);
    for (int s = 0; s < n; ++s) {
      for (int p = threadIdx.x; p < ni; p += CNT) {
        const int i = 1 + p / (v.ny - 2), j = 1 + p - (p / (v.ny - 2)) * (v.ny - 2);
        const int c = i * v.pitch + j;
        const T o = v.u[c];
        v.u2[c] = o + a.w * (c_point(v.u, v.f, c, v.pitch, v) - o);
      }
      __syncthreads();
      for (int p = threadIdx.x; p < ni; p += CNT) {
        const int i = 1 + p / (v.ny - 2), j = 1 + p - (p / (v.ny - 2)) * (v.ny - 2);
        v.u[i * v.pitch + j] = v.u2[i * v.pitch + j];
      }
      __syncthreads();
    }
    return;
  }
  const int wc = (v.ny - 1) / 2;  // items per row and colour (ny odd)
  const int items = (v.nx - 2) * wc;
  for (int s = 0; s < n; ++s)
    for (int color = 0; color < 2; ++color) {
      for (int p = threadIdx.x; p < items; p += CNT) {
        const int i = 1 + p / wc, m = p - (p / wc) * wc;
        const int j = 2 * m + 1 + ((i + color + 1) & 1);
        if (j > v.ny - 2) continue;
        const int c = i * v.pitch + j;
        T *u = v.u;
        T r;
        if (v.eq) {
          if (v.lap) r = (u[c - v.pitch] + u[c + v.pitch] + u[c - 1] + u[c + 1]) * T(0.25);
          else r = (u[c - v.pitch] + u[c + v.pitch] + u[c - 1] + u[c + 1] - v.f[c] * v.hx2) * T(0.25);
        } else {
          const T sx = divh(u[c - v.pitch] + u[c + v.pitch], v.hx2, v.ihx2, v.pw2);
          const T sy = divh(u[c - 1] + u[c + 1], v.hy2, v.ihy2, v.pw2);
          if (v.lap) r = v.coef * (sx + sy);
          else r = v.coef * (sx + sy - v.f[c]);
        }
        u[c] = r;
      }
      __syncthreads();
    }
}

template <typename T>
__device__ inline T cg_resid(const CoarseLevel<T> &v, int c) {
  const T *u = v.u;
  const T u2 = u[c] * T(2);
  return v.f[c] - (divh(u[c - v.pitch] + u[c + v.pitch] - u2, v.hx2, v.ihx2, v.pw2) +
                   divh(u[c - 1] + u[c + 1] - u2, v.hy2, v.ihy2, v.pw2));
}

template <typename T>
__global__ void __launch_bounds__(CNT) k_coarse(CoarseArgs<T> a) {
  for (int l = a.l0; l <= a.L - 2; ++l) {
    const CoarseLevel<T> &f = a.lv[l];
    const CoarseLevel<T> &g = a.lv[l + 1];
    if (l > a.l0 || a.zero_first) cg_zero(f);
    cg_smooth(f, a.nu1, a);
    const int n = (g.nx - 2) * (g.ny - 2);
    for (int p = threadIdx.x; p < n; p += CNT) {
      const int I = 1 + p / (g.ny - 2), J = 1 + p - (p / (g.ny - 2)) * (g.ny - 2);
      const int c = 2 * I * f.pitch + 2 * J;
      ((T *)g.f)[I * g.pitch + J] = c_restrict<T>([&](int k) { return cg_resid(f, k); }, c, f.pitch, a.fw);
    }
    __syncthreads();
  }
  {
    const CoarseLevel<T> &v = a.lv[a.L - 1];
    if (a.L - 1 > a.l0 || a.zero_first) cg_zero(v);
    cg_smooth(v, a.nuc, a);
  }
  for (int l = a.L - 2; l >= a.l0; --l) {
    const CoarseLevel<T> &f = a.lv[l];
    const CoarseLevel<T> &g = a.lv[l + 1];
    const int n = (f.nx - 2) * (f.ny - 2);
    for (int p = threadIdx.x; p < n; p += CNT) {
      const int i = 1 + p / (f.ny - 2), j = 1 + p - (p / (f.ny - 2)) * (f.ny - 2);
      const int c = i * f.pitch + j;
      const T v = f.u[c] + prolong_corr(g.u, g.pitch, i, j);
      f.u[c] = v;
      if (l == 0 && a.snap) a.snap[c] = (double)v;
    }
    __syncthreads();
    cg_smooth(f, a.nu2, a);
  }
}

// ---------------------------------------------------------------------------
// The same V-cycle with every level resident in LDS: per level the iterate
// and the right-hand side, unpadded (row pitch ny), in level order.  Level l0
// is loaded from global memory (its iterate unless zero-initialised, its
// right-hand side, restricted by the launch above it), all coarser levels are
// built in LDS, and only level l0's interior iterate is written back (the
// next ascent prolongates from it).  Same formulas and order as k_coarse, so
// the same bits; each phase costs an LDS round trip instead of an L2 one.
// Launched with the LDS footprint sum_l 2 nx_l ny_l sizeof(T) as dynamic
// shared memory (at most 160 KB on gfx950).
// ---------------------------------------------------------------------------
template <typename T> struct LdsLevel {
  T *u, *f;
  int nx, ny;
  int ly;  // ny = 2^ly + 1 (every level of a power-of-two grid)
};

// LDS bytes k_coarse_lds needs for levels [l0, L) of the given sizes (jac:
// plus one level-l0 array for the new values of a Jacobi sweep).
inline size_t coarse_lds_bytes(const int *nx, const int *ny, int l0, int L, size_t elem, bool jac = false) {
  size_t n = 0;
  for (int l = l0; l < L; ++l) n += 2 * (size_t)nx[l] * (size_t)ny[l];
  if (jac) n += (size_t)nx[l0] * (size_t)ny[l0];
  return n * elem;
}

__device__ __forceinline__ int ilog2(int x) { return 31 - __builtin_clz(x); }

// The table the kernel builds in LDS at its start: each level's parameters
// (one thread per level copies them from the kernel arguments, all in
// flight together) and the offset of its arrays in LDS, so that no phase
// reads the kernel arguments again.  Level l's arrays follow those of levels
// l0 .. l-1.
template <typename T> struct ClTab {
  int off[kMaxLevels], nx[kMaxLevels], ny[kMaxLevels];
  CoarseLevel<T> lv[kMaxLevels];
};
template <typename T> __device__ inline LdsLevel<T> cl_level(const ClTab<T> &tb, T *sm, int l) {
  const int o = tb.off[l], nx = tb.nx[l], ny = tb.ny[l];
  return LdsLevel<T>{sm + o, sm + o + nx * ny, nx, ny, ilog2(ny - 1)};
}

// Work items are mapped to nodes with shifts and masks only (row lengths are
// powers of two plus one): one item per node of one colour in a half-sweep,
// one per interior node (column 0 of each row idle) elsewhere.
// The LDS kernel's phases take the level's parameters and the smoother by
// value: copied once per phase, they stay in scalar registers instead of
// being reloaded from the kernel arguments inside the item loops.
template <typename T>
__device__ inline void cl_smooth(const LdsLevel<T> &v, const CoarseLevel<T> g, int n, const int jac, const T w, T *tmp) {
  if (jac) {
    // weighted Jacobi: new values into tmp (LDS after the levels), then back
    const int items = (v.nx - 2) << v.ly;
    for (int s = 0; s < n; ++s) {
      for (int p = threadIdx.x; p < items; p += CNT) {
        const int i = 1 + (p >> v.ly), j = p & ((1 << v.ly) - 1);
        if (j == 0) continue;
        const int c = i * v.ny + j;
        const T o = v.u[c];
        tmp[c] = o + w * (c_point(v.u, v.f, c, v.ny, g) - o);
      }
      __syncthreads();
      for (int p = threadIdx.x; p < items; p += CNT) {
        const int i = 1 + (p >> v.ly), j = p & ((1 << v.ly) - 1);
        if (j == 0) continue;
        v.u[i * v.ny + j] = tmp[i * v.ny + j];
      }
      __syncthreads();
    }
    return;
  }
  const int lw = v.ly - 1;  // 2^lw items per row and colour
  const int items = (v.nx - 2) << lw;
  const int P = v.ny;
  for (int s = 0; s < n; ++s)
    for (int color = 0; color < 2; ++color) {
      for (int p = threadIdx.x; p < items; p += CNT) {
        const int i = 1 + (p >> lw), m = p & ((1 << lw) - 1);
        const int j = 2 * m + 1 + ((i + color + 1) & 1);
        if (j > v.ny - 2) continue;
        const int c = i * P + j;
        T *u = v.u;
        T r;
        if (g.eq) {
          if (g.lap) r = (u[c - P] + u[c + P] + u[c - 1] + u[c + 1]) * T(0.25);
          else r = (u[c - P] + u[c + P] + u[c - 1] + u[c + 1] - v.f[c] * g.hx2) * T(0.25);
        } else {
          const T sx = divh(u[c - P] + u[c + P], g.hx2, g.ihx2, g.pw2);
          const T sy = divh(u[c - 1] + u[c + 1], g.hy2, g.ihy2, g.pw2);
          if (g.lap) r = g.coef * (sx + sy);
          else r = g.coef * (sx + sy - v.f[c]);
        }
        u[c] = r;
      }
      __syncthreads();
    }
}

template <typename T> __device__ inline T cl_resid(const LdsLevel<T> &v, const CoarseLevel<T> g, int c) {
  const T *u = v.u;
  const int P = v.ny;
  const T u2 = u[c] * T(2);
  return v.f[c] - (divh(u[c - P] + u[c + P] - u2, g.hx2, g.ihx2, g.pw2) +
                   divh(u[c - 1] + u[c + 1] - u2, g.hy2, g.ihy2, g.pw2));
}

template <typename T>
__global__ void __launch_bounds__(CNT) k_coarse_lds(CoarseArgs<T> a) {
  extern __shared__ __align__(16) unsigned char smem_raw[];
  T *sm = reinterpret_cast<T *>(smem_raw);
  // the level table in LDS, built once (the kernel arguments are read here
  // only, not inside every phase)
  __shared__ ClTab<T> tb;
  if ((int)threadIdx.x < a.L - a.l0) tb.lv[a.l0 + threadIdx.x] = a.lv[a.l0 + threadIdx.x];
  __syncthreads();
  if (threadIdx.x == 0) {
    int o = 0;
    for (int k = a.l0; k < a.L; ++k) {
      tb.off[k] = o;
      tb.nx[k] = tb.lv[k].nx;
      tb.ny[k] = tb.lv[k].ny;
      o += 2 * tb.nx[k] * tb.ny[k];
    }
  }
  __syncthreads();
  // Jacobi's new values: after the last level
  T *const tmp = cl_level(tb, sm, a.L - 1).f + tb.nx[a.L - 1] * tb.ny[a.L - 1];
  const int jac = a.jac, fw = a.fw;
  const T w = a.w;
  {
    // level l0: the whole iterate (its boundary feeds the stencils; interior
    // zeroed when zero_first) and the right-hand side (0 on Laplace levels).
    // Its parameters are copied once; each thread then issues the global
    // loads of kLdBatch nodes before it writes any of them to LDS, so the
    // memory latency is paid once per batch instead of once per node (one
    // batch covers 65 x 65 nodes).
    constexpr int kLdBatch = 8;
    const CoarseLevel<T> g = tb.lv[a.l0];
    const LdsLevel<T> v = cl_level(tb, sm, a.l0);
    const int n = v.nx * v.ny;
    const bool zf = a.zero_first != 0;
    for (int base = 0; base < n; base += kLdBatch * CNT) {
      T uu[kLdBatch], ff[kLdBatch];
#pragma unroll
      for (int k = 0; k < kLdBatch; ++k) {
        const int p = base + k * CNT + (int)threadIdx.x;
        const int i = p / v.ny, j = p - i * v.ny;
        const bool in = i > 0 && i < v.nx - 1 && j > 0 && j < v.ny - 1;
        uu[k] = T(0);
        ff[k] = T(0);
        if (p < n) {
          if (!(zf && in)) uu[k] = g.u[i * g.pitch + j];
          if (!g.lap) ff[k] = g.f[i * g.pitch + j];
        }
      }
#pragma unroll
      for (int k = 0; k < kLdBatch; ++k) {
        const int p = base + k * CNT + (int)threadIdx.x;
        if (p < n) {
          v.u[p] = uu[k];
          v.f[p] = ff[k];
        }
      }
    }
    __syncthreads();
  }
  for (int l = a.l0; l <= a.L - 2; ++l) {
    const LdsLevel<T> f = cl_level(tb, sm, l), g = cl_level(tb, sm, l + 1);
    const CoarseLevel<T> fg = tb.lv[l];
    cl_smooth(f, fg, a.nu1, jac, w, tmp);
    for (int p = threadIdx.x; p < (g.nx - 2) << g.ly; p += CNT) {
      const int I = 1 + (p >> g.ly), J = p & ((1 << g.ly) - 1);
      if (J == 0) continue;
      const int c = 2 * I * f.ny + 2 * J;
      g.f[I * g.ny + J] = c_restrict<T>([&](int k) { return cl_resid(f, fg, k); }, c, f.ny, fw);
    }
    // the coarse iterate, boundary included, starts at 0 (the coarse
    // right-hand side's boundary is never read)
    for (int p = threadIdx.x; p < g.nx * g.ny; p += CNT) g.u[p] = T(0);
    __syncthreads();
  }
  cl_smooth(cl_level(tb, sm, a.L - 1), tb.lv[a.L - 1], a.nuc, jac, w, tmp);
  for (int l = a.L - 2; l >= a.l0; --l) {
    const LdsLevel<T> f = cl_level(tb, sm, l), g = cl_level(tb, sm, l + 1);
    double *const snap = l == 0 ? a.snap : nullptr;
    const int sp = a.lv[0].pitch;
    for (int p = threadIdx.x; p < (f.nx - 2) << f.ly; p += CNT) {
      const int i = 1 + (p >> f.ly), j = p & ((1 << f.ly) - 1);
      if (j == 0) continue;
      const int c = i * f.ny + j;
      const T v = f.u[c] + prolong_corr(g.u, g.ny, i, j);
      f.u[c] = v;
      if (snap) snap[i * sp + j] = (double)v;
    }
    __syncthreads();
    cl_smooth(f, tb.lv[l], a.nu2, jac, w, tmp);
  }
  {
    // level l0's interior iterate back to global memory; the coarser levels
    // live and die in LDS
    const CoarseLevel<T> g = tb.lv[a.l0];
    const LdsLevel<T> v = cl_level(tb, sm, a.l0);
    for (int p = threadIdx.x; p < (v.nx - 2) << v.ly; p += CNT) {
      const int i = 1 + (p >> v.ly), j = p & ((1 << v.ly) - 1);
      if (j == 0) continue;
      g.u[i * g.pitch + j] = v.u[i * v.ny + j];
    }
  }
}

}  // namespace vlp2d
