// Fused, LDS-tiled multigrid kernels for gfx950.
//
// One workgroup (256 threads = 4 wavefronts) owns a 64 x 64 output tile of a
// level and stages the tile plus a halo of H nodes into LDS.  Inside LDS it
// runs several red-black Gauss-Seidel half-sweeps back to back (temporal
// blocking): half-sweep k updates only nodes at distance >= k from the staged
// edge, so after S half-sweeps every node at distance >= S holds exactly the
// value a global sweep-by-sweep schedule produces.  Within one colour the
// Gauss-Seidel update is order-independent, so the result is bitwise the
// reference's (VisualLaplacePoisson2D/Model/CPU/MultiGridSlidingIteration.cs:
// 53-57, GridIterator.cs:184-188: red = (i+j) even first).
//
// LDS layout splits the colours: U[c][li*LJC + (lj>>1)] with c = (li+lj)&1.
// The tile origin (gi0, gj0) always has gi0 + gj0 even, so local and global
// colours agree.  A red node's four neighbours are black nodes at
// (li+-1, m), (li, m-1+d), (li, m+d) with m = lj>>1 and d = lj&1, so the
// 64 lanes of a wavefront read consecutive LDS words.
//
// Tiles read a halo that neighbouring tiles write, so every fused kernel is
// out of place (uin -> uout); each level keeps two iterate buffers.
//
// Kernels:
//   k_down  : [zero or staged iterate] -> S = 2*NS half-sweeps -> residual
//             and half-weighting restriction of the coarse right-hand side
//             owned by the tile (MultiGridScheme.cs:134-169) -> write tile.
//   k_up    : staged fine iterate + additive bilinear prolongation of the
//             coarse correction (MultiGridScheme.cs:171-186) -> S half-sweeps
//             -> epilogue (none | residual-norm partials | update-norm
//             partials | FP64 refinement update of a mixed-precision step).
//   k_coarse: one workgroup runs the V-cycle on the small levels entirely.
#pragma once
#include "mg_kernels.hpp"

namespace vlp2d {

constexpr int TT = 64;        // output tile edge
constexpr int TNT = 256;      // threads per tile workgroup

template <typename T> struct LevelView {
  const T *uin;   // staged iterate (ignored when zero-initialised)
  T *uout;        // written iterate (tile interior nodes of the domain)
  const T *f;     // right-hand side (fn on level 0, restricted residual below)
  int nx, ny, pitch;
  T hx2, hy2, coef;
  int eq, lap;
  int tiles_i, tiles_j, chunk;  // tile grid and per-XCD chunk of tiles
};

// Mixed-precision epilogue operands (level 0 of the FP32 correction).
struct MixView {
  const double *u64in;
  double *u64out;
  const double *fn64;
  float *r32out;
  double hx2, hy2;
  double *part_max;  // partial max |e|
};

// Workgroup -> tile: consecutive workgroup ids land on different XCDs
// (round-robin), so each XCD is given one contiguous band of tiles; tiles
// sharing a halo then share that XCD's L2.
__device__ inline int tile_of_block(int tiles, int chunk) {
  const int b = blockIdx.x;
  const int t = (b & 7) * chunk + (b >> 3);
  return t < tiles ? t : -1;
}

template <typename T>
__device__ inline T tile_point(const T *o, int li, int lj, int LJC, T f, const LevelView<T> &a) {
  const int m = lj >> 1;
  const T uim = o[(li - 1) * LJC + m];
  const T uip = o[(li + 1) * LJC + m];
  const T ujm = o[li * LJC + ((lj - 1) >> 1)];
  const T ujp = o[li * LJC + ((lj + 1) >> 1)];
  if (a.eq) {
    if (a.lap) return (uim + uip + ujm + ujp) * T(0.25);
    return (uim + uip + ujm + ujp - f * a.hx2) * T(0.25);
  }
  if (a.lap) return a.coef * ((uim + uip) / a.hx2 + (ujm + ujp) / a.hy2);
  return a.coef * ((uim + uip) / a.hx2 + (ujm + ujp) / a.hy2 - f);
}

// One half-sweep of colour C over nodes at distance >= lo from the staged
// edge.  fr[k] holds f for this thread's k-th colour-C item.
template <typename T, int LI, int LJ, int KPT, int C>
__device__ inline void tile_half(T (*U)[LI * (LJ / 2)], const T (&fr)[KPT], int lo, int gi0, int gj0,
                                 const LevelView<T> &a) {
  constexpr int LJC = LJ / 2, NC = LI * LJC;
  const T *o = U[1 - C];
  T *w = U[C];
#pragma unroll
  for (int k = 0; k < KPT; ++k) {
    const int p = threadIdx.x + TNT * k;
    if (p < NC) {
      const int li = p / LJC, m = p - (p / LJC) * LJC;
      const int lj = 2 * m + ((li + C) & 1);
      const int gi = gi0 + li, gj = gj0 + lj;
      if (li >= lo && li < LI - lo && lj >= lo && lj < LJ - lo && gi >= 1 && gi <= a.nx - 2 && gj >= 1 &&
          gj <= a.ny - 2)
        w[li * LJC + m] = tile_point(o, li, lj, LJC, fr[k], a);
    }
  }
  __syncthreads();
}

// S (even) red-black half-sweeps over a staged (LI x LJ) tile with origin
// (gi0, gj0), red first.
template <typename T, int LI, int LJ, int S, int KPT>
__device__ inline void tile_sweeps(T (*U)[LI * (LJ / 2)], const T (&fr)[2][KPT], int gi0, int gj0,
                                   const LevelView<T> &a) {
#pragma unroll 1
  for (int s = 0; s < S; s += 2) {
    tile_half<T, LI, LJ, KPT, 0>(U, fr[0], s + 1, gi0, gj0, a);
    tile_half<T, LI, LJ, KPT, 1>(U, fr[1], s + 2, gi0, gj0, a);
  }
}

// residual fn - Lap(u) at local node (li, lj) from the staged tile
template <typename T, int LI, int LJ>
__device__ inline T tile_resid(T (*U)[LI * (LJ / 2)], int li, int lj, T rhs, T hx2, T hy2) {
  constexpr int LJC = LJ / 2;
  const int c = (li + lj) & 1;
  const T *o = U[1 - c];
  const int m = lj >> 1;
  const T u = U[c][li * LJC + m];
  const T u2 = u * T(2);
  const T uim = o[(li - 1) * LJC + m], uip = o[(li + 1) * LJC + m];
  const T ujm = o[li * LJC + ((lj - 1) >> 1)], ujp = o[li * LJC + ((lj + 1) >> 1)];
  return rhs - ((uim + uip - u2) / hx2 + (ujm + ujp - u2) / hy2);
}

// Stage the (LI x LJ) region of uin (0 outside the domain or when ZI).
template <typename T, int LI, int LJ, bool ZI>
__device__ inline void tile_load(T (*U)[LI * (LJ / 2)], const LevelView<T> &a, int gi0, int gj0) {
  constexpr int LJC = LJ / 2;
  for (int p = threadIdx.x; p < LI * LJ; p += TNT) {
    const int li = p / LJ, lj = p - (p / LJ) * LJ;
    const int gi = gi0 + li, gj = gj0 + lj;
    T v = T(0);
    if (!ZI && gi >= 0 && gi < a.nx && gj >= 0 && gj < a.ny) v = a.uin[gi * a.pitch + gj];
    U[(li + lj) & 1][li * LJC + (lj >> 1)] = v;
  }
}

template <typename T, int LI, int LJ, int KPT>
__device__ inline void tile_load_f(T (&fr)[2][KPT], const LevelView<T> &a, int gi0, int gj0) {
  constexpr int LJC = LJ / 2, NC = LI * LJC;
#pragma unroll
  for (int c = 0; c < 2; ++c)
#pragma unroll
    for (int k = 0; k < KPT; ++k) {
      const int p = threadIdx.x + TNT * k;
      T v = T(0);
      if (p < NC && !a.lap) {
        const int li = p / LJC, m = p - (p / LJC) * LJC;
        const int lj = 2 * m + ((li + c) & 1);
        const int gi = gi0 + li, gj = gj0 + lj;
        if (gi >= 1 && gi <= a.nx - 2 && gj >= 1 && gj <= a.ny - 2) v = a.f[gi * a.pitch + gj];
      }
      fr[c][k] = v;
    }
}

template <typename T, int LI, int LJ, int H>
__device__ inline void tile_store(T (*U)[LI * (LJ / 2)], const LevelView<T> &a, int gi0, int gj0) {
  constexpr int LJC = LJ / 2;
  for (int p = threadIdx.x; p < TT * TT; p += TNT) {
    const int li = H + p / TT, lj = H + (p & (TT - 1));
    const int gi = gi0 + li, gj = gj0 + lj;
    if (gi >= 1 && gi <= a.nx - 2 && gj >= 1 && gj <= a.ny - 2)
      a.uout[gi * a.pitch + gj] = U[(li + lj) & 1][li * LJC + (lj >> 1)];
  }
}

// ---------------------------------------------------------------------------
// Descent: smoothing + residual + restriction.  Coarse interior node (I, J)
// belongs to the tile containing its fine centre (2I, 2J).
// ---------------------------------------------------------------------------
template <typename T, int NS, bool ZI>
__global__ void __launch_bounds__(TNT) k_down(LevelView<T> a, T *__restrict__ rhsc, int ncx, int ncy, int pc) {
  constexpr int S = 2 * NS, H = S + 2, LI = TT + 2 * H, LJ = TT + 2 * H, LJC = LJ / 2, NC = LI * LJC;
  constexpr int KPT = (NC + TNT - 1) / TNT;
  __shared__ T U[2][NC];
  const int t = tile_of_block(a.tiles_i * a.tiles_j, a.chunk);
  if (t < 0) return;
  const int ti = t / a.tiles_j, tj = t - (t / a.tiles_j) * a.tiles_j;
  const int gi0 = 1 + ti * TT - H, gj0 = 1 + tj * TT - H;
  tile_load<T, LI, LJ, ZI>(U, a, gi0, gj0);
  T fr[2][KPT];
  tile_load_f<T, LI, LJ, KPT>(fr, a, gi0, gj0);
  __syncthreads();
  tile_sweeps<T, LI, LJ, S, KPT>(U, fr, gi0, gj0, a);
  // restriction of the residual to the coarse nodes owned by this tile
  for (int q = threadIdx.x; q < (TT / 2) * (TT / 2); q += TNT) {
    const int I = 1 + ti * (TT / 2) + q / (TT / 2);
    const int J = 1 + tj * (TT / 2) + (q & (TT / 2 - 1));
    if (I > ncx - 2 || J > ncy - 2) continue;
    const int gi = 2 * I, gj = 2 * J;
    const int li = gi - gi0, lj = gj - gj0;
    const T *f = a.f;
    const int c = gi * a.pitch + gj;
    const T r0 = tile_resid<T, LI, LJ>(U, li, lj, f[c], a.hx2, a.hy2);
    const T rip = tile_resid<T, LI, LJ>(U, li + 1, lj, f[c + a.pitch], a.hx2, a.hy2);
    const T rim = tile_resid<T, LI, LJ>(U, li - 1, lj, f[c - a.pitch], a.hx2, a.hy2);
    const T rjp = tile_resid<T, LI, LJ>(U, li, lj + 1, f[c + 1], a.hx2, a.hy2);
    const T rjm = tile_resid<T, LI, LJ>(U, li, lj - 1, f[c - 1], a.hx2, a.hy2);
    rhsc[I * pc + J] = T(0.125) * (r0 * T(4) + (rip + rim + rjp + rjm));
  }
  tile_store<T, LI, LJ, H>(U, a, gi0, gj0);
}

// ---------------------------------------------------------------------------
// Ascent: prolongation-add + smoothing (+ epilogue).
// EPI 0: write the tile.  EPI 1: also partial sums of the squared residual of
// the result (FP64 stop test).  EPI 2: also partial max of |u - u_after_
// prolongation| (the reference update norm, MultiGridScheme.cs:69,111).
// EPI 3 (FP32, level 0 of a mixed step): u64out = u64in + e, r32out =
// (float)(fn64 - Lap u64out), partial sums of r^2 and partial max |e|.
// ---------------------------------------------------------------------------
template <typename T>
__device__ inline T prolong_corr(const T *uc, int pc, int gi, int gj) {
  const int I = gi >> 1, J = gj >> 1;
  const T *r0 = uc + I * pc + J;
  if (!(gi & 1)) {
    if (!(gj & 1)) return r0[0];
    return (r0[0] + r0[1]) * T(0.5);
  }
  if (!(gj & 1)) return (r0[0] + r0[pc]) * T(0.5);
  return (r0[0] + r0[1] + r0[pc] + r0[pc + 1]) * T(0.25);
}

template <typename T, int NS, int EPI>
__global__ void __launch_bounds__(TNT) k_up(LevelView<T> a, const T *__restrict__ uc, int pc, double *__restrict__ part,
                                            MixView mx) {
  constexpr int S = 2 * NS, H = S + ((EPI == 1 || EPI == 3) ? 1 : 0);
  constexpr int LI = TT + 2 * H, LJ = TT + 2 * H, LJC = LJ / 2, NC = LI * LJC;
  constexpr int KPT = (NC + TNT - 1) / TNT;
  constexpr int OPT = TT * TT / TNT;  // output nodes per thread
  __shared__ T U[2][NC];
  __shared__ double red[TNT / 64];
  const int t = tile_of_block(a.tiles_i * a.tiles_j, a.chunk);
  if (t < 0) {
    // idle workgroup of the XCD-strided grid: neutral partials
    if (threadIdx.x == 0 && EPI != 0) {
      part[blockIdx.x] = 0.0;
      if (EPI == 3) mx.part_max[blockIdx.x] = 0.0;
    }
    return;
  }
  const int ti = t / a.tiles_j, tj = t - (t / a.tiles_j) * a.tiles_j;
  const int gi0 = 1 + ti * TT - H, gj0 = 1 + tj * TT - H;
  // stage and prolongate
  for (int p = threadIdx.x; p < LI * LJ; p += TNT) {
    const int li = p / LJ, lj = p - (p / LJ) * LJ;
    const int gi = gi0 + li, gj = gj0 + lj;
    T v = T(0);
    if (gi >= 0 && gi < a.nx && gj >= 0 && gj < a.ny) {
      v = a.uin[gi * a.pitch + gj];
      if (gi >= 1 && gi <= a.nx - 2 && gj >= 1 && gj <= a.ny - 2) v = v + prolong_corr(uc, pc, gi, gj);
    }
    U[(li + lj) & 1][li * LJC + (lj >> 1)] = v;
  }
  T fr[2][KPT];
  tile_load_f<T, LI, LJ, KPT>(fr, a, gi0, gj0);
  __syncthreads();
  T snap[EPI == 2 ? OPT : 1];
  if (EPI == 2) {
#pragma unroll
    for (int k = 0; k < OPT; ++k) {
      const int p = threadIdx.x + TNT * k;
      const int li = H + p / TT, lj = H + (p & (TT - 1));
      snap[k] = U[(li + lj) & 1][li * LJC + (lj >> 1)];
    }
  }
  tile_sweeps<T, LI, LJ, S, KPT>(U, fr, gi0, gj0, a);
  const int tid = threadIdx.x;
  if (EPI == 0 || EPI == 1 || EPI == 2) {
    double acc = 0.0;
#pragma unroll
    for (int k = 0; k < OPT; ++k) {
      const int p = tid + TNT * k;
      const int li = H + p / TT, lj = H + (p & (TT - 1));
      const int gi = gi0 + li, gj = gj0 + lj;
      if (gi >= 1 && gi <= a.nx - 2 && gj >= 1 && gj <= a.ny - 2) {
        const T u = U[(li + lj) & 1][li * LJC + (lj >> 1)];
        a.uout[gi * a.pitch + gj] = u;
        if (EPI == 1) {
          const double r = (double)tile_resid<T, LI, LJ>(U, li, lj, a.f[gi * a.pitch + gj], a.hx2, a.hy2);
          acc += r * r;
        } else if (EPI == 2) {
          acc = nanmax(acc, fabs((double)(u - snap[k])));
        }
      }
    }
    if (EPI == 1 || EPI == 2) {
      acc = block_reduce<EPI == 2>(acc, red, tid, TNT);
      if (tid == 0) part[blockIdx.x] = acc;
    }
    return;
  }
  // EPI == 3: FP64 refinement update and FP64 residual of the new iterate.
  constexpr int W = TT + 2;
  __shared__ double Wn[W * W];
  for (int p = tid; p < W * W; p += TNT) {
    const int li = H - 1 + p / W, lj = H - 1 + (p - (p / W) * W);
    const int gi = gi0 + li, gj = gj0 + lj;
    double v = 0.0;
    if (gi >= 0 && gi < a.nx && gj >= 0 && gj < a.ny) {
      v = mx.u64in[gi * a.pitch + gj];
      if (gi >= 1 && gi <= a.nx - 2 && gj >= 1 && gj <= a.ny - 2)
        v += (double)U[(li + lj) & 1][li * LJC + (lj >> 1)];
    }
    Wn[p] = v;
  }
  __syncthreads();
  double ss = 0.0, em = 0.0;
#pragma unroll
  for (int k = 0; k < OPT; ++k) {
    const int p = tid + TNT * k;
    const int oi = p / TT, oj = p & (TT - 1);
    const int li = H + oi, lj = H + oj;
    const int gi = gi0 + li, gj = gj0 + lj;
    if (gi >= 1 && gi <= a.nx - 2 && gj >= 1 && gj <= a.ny - 2) {
      const int w = (oi + 1) * W + (oj + 1);
      const double u = Wn[w];
      const double u2 = u * 2.0;
      const int g = gi * a.pitch + gj;
      const double r = mx.fn64[g] - ((Wn[w - W] + Wn[w + W] - u2) / mx.hx2 + (Wn[w - 1] + Wn[w + 1] - u2) / mx.hy2);
      mx.u64out[g] = u;
      mx.r32out[g] = (float)r;
      ss += r * r;
      em = nanmax(em, fabs((double)U[(li + lj) & 1][li * LJC + (lj >> 1)]));
    }
  }
  ss = block_reduce<false>(ss, red, tid, TNT);
  if (tid == 0) part[blockIdx.x] = ss;
  em = block_reduce<true>(em, red, tid, TNT);
  if (tid == 0) mx.part_max[blockIdx.x] = em;
}

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
  int eq, lap;
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
          if (v.lap) r = v.coef * ((u[c - v.pitch] + u[c + v.pitch]) / v.hx2 + (u[c - 1] + u[c + 1]) / v.hy2);
          else r = v.coef * ((u[c - v.pitch] + u[c + v.pitch]) / v.hx2 + (u[c - 1] + u[c + 1]) / v.hy2 - v.f[c]);
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
  return v.f[c] - ((u[c - v.pitch] + u[c + v.pitch] - u2) / v.hx2 + (u[c - 1] + u[c + 1] - u2) / v.hy2);
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
