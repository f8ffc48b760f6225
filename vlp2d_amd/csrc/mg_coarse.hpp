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
constexpr int CNT = 1024;
constexpr int kMaxLevels = 16;

template <typename T> struct CoarseLevel {
  T *u;
  const T *f;
  int nx, ny, pitch;
  T hx2, hy2, coef;
  T ihx2, ihy2;
  int eq, lap, pw2;
};
template <typename T> struct CoarseArgs {
  CoarseLevel<T> lv[kMaxLevels];
  int l0, L;
  int nu1, nu2, nuc;
  int zero_first;
  double *snap;  // level-0 snapshot before post-smoothing (l0 == 0), or null
};

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
__device__ inline void cg_smooth(const CoarseLevel<T> &v, int n) {
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
    cg_smooth(f, a.nu1);
    const int n = (g.nx - 2) * (g.ny - 2);
    for (int p = threadIdx.x; p < n; p += CNT) {
      const int I = 1 + p / (g.ny - 2), J = 1 + p - (p / (g.ny - 2)) * (g.ny - 2);
      const int c = 2 * I * f.pitch + 2 * J;
      const T r0 = cg_resid(f, c), rip = cg_resid(f, c + f.pitch), rim = cg_resid(f, c - f.pitch);
      const T rjp = cg_resid(f, c + 1), rjm = cg_resid(f, c - 1);
      ((T *)g.f)[I * g.pitch + J] = T(0.125) * (r0 * T(4) + (rip + rim + rjp + rjm));
    }
    __syncthreads();
  }
  {
    const CoarseLevel<T> &v = a.lv[a.L - 1];
    if (a.L - 1 > a.l0 || a.zero_first) cg_zero(v);
    cg_smooth(v, a.nuc);
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
    cg_smooth(f, a.nu2);
  }
}

}  // namespace vlp2d
