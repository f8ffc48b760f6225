// Fused, LDS-tiled multigrid kernels for gfx950.
//
// A workgroup of 256 threads (4 wavefronts) owns an output tile of TI rows by
// TJ columns of one level and stages the tile plus a halo of H nodes into LDS
// as LI = TI + 2H rows of 128 columns (TJ = 128 - 2H).  Inside LDS it runs
// several red-black Gauss-Seidel half-sweeps back to back (temporal
// blocking): half-sweep k updates only nodes at distance >= k from the staged
// edge, so after S half-sweeps every node at distance >= S holds exactly the
// value a global sweep-by-sweep schedule produces.  Within one colour the
// Gauss-Seidel update is order-independent, so the result is bitwise the
// reference's (VisualLaplacePoisson2D/Model/CPU/MultiGridSlidingIteration.cs:
// 53-57, GridIterator.cs:184-188: red = (i+j) even first).
//
// LDS layout splits the colours: colour c = (li+lj)&1 of staged node (li, lj)
// lives at U[c*NC + li*64 + (lj>>1)], NC = LI*64.  Lane m of a wavefront owns
// the column pair (2m, 2m+1); wavefront w handles the rows li = w + 4k.  All
// rows of a wavefront have the parity of w, so row bounds and the column
// parity of a colour are wavefront-uniform, and the four neighbours of a node
// are the LDS words b -+ 64, b - 1 + d, b + d of the other colour (b = li*64 +
// m, d the column parity): consecutive lanes read consecutive words.  The
// tile origin (gi0, gj0) always has gi0 + gj0 even, so local and global
// colours agree.
//
// Tiles read a halo that neighbouring tiles write, so every fused kernel is
// out of place (uin -> uout); each level keeps two iterate buffers.
//
// Kernels:
//   k_down : [zero or staged iterate] -> S = 2*NS half-sweeps -> residual and
//            half-weighting restriction of the coarse right-hand side owned by
//            the tile (MultiGridScheme.cs:134-169) -> write the tile.
//   k_up   : staged fine iterate + additive bilinear prolongation of the
//            coarse correction (MultiGridScheme.cs:171-186) -> S half-sweeps
//            -> epilogue (none | residual-norm partials | update-norm partials
//            | FP64 refinement update of a mixed-precision step).
#pragma once
#include "mg_kernels.hpp"

namespace vlp2d {

constexpr int TNT = 256;  // threads per tile workgroup (4 wavefronts)
constexpr int SJ = 128;   // staged columns per tile
constexpr int SJC = 64;   // staged columns per colour = lanes per wavefront

// Output tile rows, 32 for both types (measured on MI355X: 64-row tiles
// lower occupancy more than they save in halo traffic).  A staged FP64 tile
// stays near 45 KB of LDS, three workgroups per CU.
#ifndef VLP2D_ROWS64
#define VLP2D_ROWS64 32
#endif
#ifndef VLP2D_ROWS32
#define VLP2D_ROWS32 32
#endif
#ifndef VLP2D_TILE_COLMAJOR
#define VLP2D_TILE_COLMAJOR 0
#endif
template <typename T> __host__ __device__ constexpr int tile_rows() { return sizeof(T) == 8 ? VLP2D_ROWS64 : VLP2D_ROWS32; }
__host__ __device__ constexpr int tile_cols(int H) { return SJ - 2 * H; }
__host__ __device__ constexpr int down_halo(int ns) { return 2 * ns + 2; }
__host__ __device__ constexpr int up_halo(int ns, int epi) { return 2 * ns + ((epi == 1 || epi == 3) ? 1 : 0); }

template <typename T> struct LevelView {
  const T *uin;  // staged iterate (ignored when zero-initialised)
  T *uout;       // written iterate (tile interior nodes of the domain)
  const T *f;    // right-hand side (fn on level 0, restricted residual below)
  int nx, ny, pitch;
  T hx2, hy2, coef;
  T ihx2, ihy2;  // 1/hx2, 1/hy2: exact when pw2 (both squares are powers of two)
  int eq, lap, pw2;
  int tiles_i, tiles_j, chunk;  // tile grid of this launch and per-XCD chunk
};

// v / h2.  When h2 is a power of two, v * (1/h2) is the same correctly
// rounded value, and a multiply is far cheaper than an IEEE division.
template <typename T> __device__ inline T divh(T v, T h2, T ih2, int pw2) { return pw2 ? v * ih2 : v / h2; }

// Mixed-precision epilogue operands (level 0 of the FP32 correction).
struct MixView {
  const double *u64in;
  double *u64out;
  const double *fn64;
  float *r32out;
  double hx2, hy2, ihx2, ihy2;
  int pw2;
  double *part_max;  // partial max |e|
};

// Workgroup -> tile: consecutive workgroup ids land on different XCDs
// (round-robin), so each XCD is given one contiguous run of tiles, and the
// workgroups an XCD runs together are neighbours that share halos in its L2.
__device__ inline int tile_of_block(int tiles, int chunk) {
  const int b = blockIdx.x;
  const int t = (b & 7) * chunk + (b >> 3);
  return t < tiles ? t : -1;
}

// tile number -> (tile row, tile column): row-major (measured faster on
// MI355X than column-major for these tile shapes), or column-major.
__device__ inline void tile_coords(int t, int tiles_i, int tiles_j, int &ti, int &tj) {
  if (VLP2D_TILE_COLMAJOR) {
    tj = t / tiles_i;
    ti = t - tj * tiles_i;
  } else {
    ti = t / tiles_j;
    tj = t - ti * tiles_j;
  }
}

template <typename T, int H> struct Geo {
  static constexpr int TI = tile_rows<T>();
  static constexpr int TJ = tile_cols(H);
  static constexpr int LI = TI + 2 * H;
  static constexpr int NC = LI * SJC;  // LDS words per colour
  static constexpr int RPW = (LI + 3) / 4;  // staged rows per wavefront
};

template <typename T>
__device__ inline T gs_point(T uim, T uip, T ujm, T ujp, T f, const LevelView<T> &a) {
  if (a.eq) {
    if (a.lap) return (uim + uip + ujm + ujp) * T(0.25);
    return (uim + uip + ujm + ujp - f * a.hx2) * T(0.25);
  }
  if (a.lap) return a.coef * (divh(uim + uip, a.hx2, a.ihx2, a.pw2) + divh(ujm + ujp, a.hy2, a.ihy2, a.pw2));
  return a.coef * (divh(uim + uip, a.hx2, a.ihx2, a.pw2) + divh(ujm + ujp, a.hy2, a.ihy2, a.pw2) - f);
}

__device__ inline bool inside(int gi, int gj, int nx, int ny) {
  return gi >= 1 && gi <= nx - 2 && gj >= 1 && gj <= ny - 2;
}

// correction of fine interior node (gi, gj) from the coarse grid
// (MultiGridScheme.cs:180-183; sums in the reference's order)
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

template <typename T, int H> __device__ inline T lds_at(const T *U, int li, int lj) {
  return U[((li + lj) & 1) * Geo<T, H>::NC + li * SJC + (lj >> 1)];
}

// Stage rows li = w + 4k (columns 2m, 2m+1 per lane) of uin into U.  With
// uc != nullptr the additive bilinear prolongation is applied to interior
// nodes on the way in.  All loads are issued before the LDS stores.
template <typename T, int H, bool ZI>
__device__ inline void stage(T *U, const LevelView<T> &a, int gi0, int gj0, const T *uc, int pc) {
  using G = Geo<T, H>;
  const int w = threadIdx.x >> 6, m = threadIdx.x & 63;
  const int gj = gj0 + 2 * m;
  T v0[G::RPW], v1[G::RPW];
#pragma unroll
  for (int k = 0; k < G::RPW; ++k) {
    const int li = w + 4 * k, gi = gi0 + li;
    T x0 = T(0), x1 = T(0);
    if (!ZI && li < G::LI && gi >= 0 && gi < a.nx) {
      const T *row = a.uin + gi * a.pitch;
      if (gj >= 0 && gj < a.ny) x0 = row[gj];
      if (gj + 1 >= 0 && gj + 1 < a.ny) x1 = row[gj + 1];
    }
    v0[k] = x0;
    v1[k] = x1;
  }
  if (uc) {
#pragma unroll
    for (int k = 0; k < G::RPW; ++k) {
      const int li = w + 4 * k, gi = gi0 + li;
      if (li < G::LI && gi >= 1 && gi <= a.nx - 2) {
        if (gj >= 1 && gj <= a.ny - 2) v0[k] = v0[k] + prolong_corr(uc, pc, gi, gj);
        if (gj + 1 >= 1 && gj + 1 <= a.ny - 2) v1[k] = v1[k] + prolong_corr(uc, pc, gi, gj + 1);
      }
    }
  }
#pragma unroll
  for (int k = 0; k < G::RPW; ++k) {
    const int li = w + 4 * k;
    if (li < G::LI) {
      U[(li & 1) * G::NC + li * SJC + m] = v0[k];
      U[((li + 1) & 1) * G::NC + li * SJC + m] = v1[k];
    }
  }
}

// f at this thread's colour-c nodes: fr[c][k] at (w + 4k, 2m + ((w + c) & 1)).
template <typename T, int H, int RPW>
__device__ inline void load_f(T (&fr)[2][RPW], const LevelView<T> &a, int gi0, int gj0) {
  using G = Geo<T, H>;
  const int w = threadIdx.x >> 6, m = threadIdx.x & 63;
#pragma unroll
  for (int c = 0; c < 2; ++c)
#pragma unroll
    for (int k = 0; k < RPW; ++k) {
      const int li = w + 4 * k;
      const int gi = gi0 + li, gj = gj0 + 2 * m + ((li + c) & 1);
      T v = T(0);
      if (!a.lap && li < G::LI && inside(gi, gj, a.nx, a.ny)) v = a.f[gi * a.pitch + gj];
      fr[c][k] = v;
    }
}

// One half-sweep of colour C over nodes at distance >= lo from the staged edge.
template <typename T, int H, int C, int RPW>
__device__ inline void half_sweep_tile(T *U, const T (&fr)[RPW], int lo, int gi0, int gj0, const LevelView<T> &a) {
  using G = Geo<T, H>;
  const int w = threadIdx.x >> 6, m = threadIdx.x & 63;
  const int d = (w + C) & 1;  // column parity of colour C on this wavefront's rows
  const int lj = 2 * m + d, gj = gj0 + lj;
  const bool col_ok = lj >= lo && lj < SJ - lo && gj >= 1 && gj <= a.ny - 2;
  const T *o = U + (1 - C) * G::NC;
  T *wr = U + C * G::NC;
#pragma unroll
  for (int k = 0; k < RPW; ++k) {
    const int li = w + 4 * k, gi = gi0 + li;
    if (li >= lo && li < G::LI - lo && gi >= 1 && gi <= a.nx - 2 && col_ok) {
      const int b = li * SJC + m;
      wr[b] = gs_point(o[b - SJC], o[b + SJC], o[b - 1 + d], o[b + d], fr[k], a);
    }
  }
  __syncthreads();
}

template <typename T, int H, int S, int RPW>
__device__ inline void sweeps_tile(T *U, const T (&fr)[2][RPW], int gi0, int gj0, const LevelView<T> &a) {
#pragma unroll 1
  for (int s = 0; s < S; s += 2) {
    half_sweep_tile<T, H, 0, RPW>(U, fr[0], s + 1, gi0, gj0, a);
    half_sweep_tile<T, H, 1, RPW>(U, fr[1], s + 2, gi0, gj0, a);
  }
}

// residual rhs - Lap(u) at staged node (li, lj) (UtilsOpLap.cs:15-20 order)
template <typename T, int H>
__device__ inline T resid_tile(const T *U, int li, int lj, T rhs, const LevelView<T> &a) {
  const T u2 = lds_at<T, H>(U, li, lj) * T(2);
  const T uim = lds_at<T, H>(U, li - 1, lj), uip = lds_at<T, H>(U, li + 1, lj);
  const T ujm = lds_at<T, H>(U, li, lj - 1), ujp = lds_at<T, H>(U, li, lj + 1);
  return rhs - (divh(uim + uip - u2, a.hx2, a.ihx2, a.pw2) + divh(ujm + ujp - u2, a.hy2, a.ihy2, a.pw2));
}

// write the output part of the staged tile
template <typename T, int H>
__device__ inline void store_tile(const T *U, const LevelView<T> &a, int gi0, int gj0) {
  using G = Geo<T, H>;
  const int w = threadIdx.x >> 6, m = threadIdx.x & 63;
#pragma unroll
  for (int k = 0; k < G::RPW; ++k) {
    const int li = w + 4 * k, gi = gi0 + li;
    if (li >= H && li < H + G::TI && gi >= 1 && gi <= a.nx - 2) {
#pragma unroll
      for (int e = 0; e < 2; ++e) {
        const int lj = 2 * m + e, gj = gj0 + lj;
        if (lj >= H && lj < H + G::TJ && gj >= 1 && gj <= a.ny - 2)
          a.uout[gi * a.pitch + gj] = U[((li + e) & 1) * G::NC + li * SJC + m];
      }
    }
  }
}

// ---------------------------------------------------------------------------
// Descent: smoothing + residual + restriction.  Coarse interior node (I, J)
// belongs to the tile whose output part contains its fine centre (2I, 2J).
// ---------------------------------------------------------------------------
template <typename T, int NS, bool ZI>
__global__ void __launch_bounds__(TNT) k_down(LevelView<T> a, T *__restrict__ rhsc, int ncx, int ncy, int pc) {
  constexpr int S = 2 * NS, H = down_halo(NS);
  using G = Geo<T, H>;
  __shared__ T U[2 * G::NC];
  const int t = tile_of_block(a.tiles_i * a.tiles_j, a.chunk);
  if (t < 0) return;
  int ti, tj;
  tile_coords(t, a.tiles_i, a.tiles_j, ti, tj);
  const int gi0 = 1 + ti * G::TI - H, gj0 = 1 + tj * G::TJ - H;
  stage<T, H, ZI>(U, a, gi0, gj0, nullptr, 0);
  T fr[2][G::RPW];
  load_f<T, H, G::RPW>(fr, a, gi0, gj0);
  __syncthreads();
  sweeps_tile<T, H, S, G::RPW>(U, fr, gi0, gj0, a);
  const int w = threadIdx.x >> 6, m = threadIdx.x & 63;
  const int J = 1 + tj * (G::TJ / 2) + m;
#pragma unroll
  for (int k = 0; k < G::TI / 8; ++k) {
    const int I = 1 + ti * (G::TI / 2) + w + 4 * k;
    if (m < G::TJ / 2 && I <= ncx - 2 && J <= ncy - 2) {
      const int gi = 2 * I, gj = 2 * J;
      const int li = gi - gi0, lj = gj - gj0;
      const T *f = a.f + gi * a.pitch + gj;
      const T r0 = resid_tile<T, H>(U, li, lj, f[0], a);
      const T rip = resid_tile<T, H>(U, li + 1, lj, f[a.pitch], a);
      const T rim = resid_tile<T, H>(U, li - 1, lj, f[-a.pitch], a);
      const T rjp = resid_tile<T, H>(U, li, lj + 1, f[1], a);
      const T rjm = resid_tile<T, H>(U, li, lj - 1, f[-1], a);
      rhsc[I * pc + J] = T(0.125) * (r0 * T(4) + (rip + rim + rjp + rjm));
    }
  }
  store_tile<T, H>(U, a, gi0, gj0);
}

// ---------------------------------------------------------------------------
// Ascent: prolongation-add + smoothing (+ epilogue).
// EPI 0: write the tile.  EPI 1: also partial sums of the squared residual of
// the result (FP64 stop test).  EPI 2: also partial max of |u - u_after_
// prolongation| (the reference update norm, MultiGridScheme.cs:69,111).
// EPI 3 (FP32, level 0 of a mixed step): u64out = u64in + e, r32out =
// (float)(fn64 - Lap u64out), partial sums of r^2 and partial max |e|; the
// correction e itself is not written.
// ---------------------------------------------------------------------------
template <typename T, int NS, int EPI>
__global__ void __launch_bounds__(TNT) k_up(LevelView<T> a, const T *__restrict__ uc, int pc, double *__restrict__ part,
                                            MixView mx) {
  constexpr int S = 2 * NS, H = up_halo(NS, EPI);
  using G = Geo<T, H>;
  __shared__ T U[2 * G::NC];
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
  int ti, tj;
  tile_coords(t, a.tiles_i, a.tiles_j, ti, tj);
  const int gi0 = 1 + ti * G::TI - H, gj0 = 1 + tj * G::TJ - H;
  const int tid = threadIdx.x, w = tid >> 6, m = tid & 63;
  stage<T, H, false>(U, a, gi0, gj0, uc, pc);
  T fr[2][G::RPW];
  load_f<T, H, G::RPW>(fr, a, gi0, gj0);
  __syncthreads();
  constexpr int NSNAP = EPI == 2 ? G::RPW : 1;
  T snap[NSNAP][2];
  if (EPI == 2) {
#pragma unroll
    for (int k = 0; k < NSNAP; ++k) {
      const int li = w + 4 * k;
#pragma unroll
      for (int e = 0; e < 2; ++e)
        snap[k][e] = (li >= H && li < H + G::TI) ? U[((li + e) & 1) * G::NC + li * SJC + m] : T(0);
    }
  }
  sweeps_tile<T, H, S, G::RPW>(U, fr, gi0, gj0, a);
  double acc = 0.0, em = 0.0;
#pragma unroll
  for (int k = 0; k < G::RPW; ++k) {
    const int li = w + 4 * k, gi = gi0 + li;
    if (!(li >= H && li < H + G::TI && gi >= 1 && gi <= a.nx - 2)) continue;
#pragma unroll
    for (int e = 0; e < 2; ++e) {
      const int lj = 2 * m + e, gj = gj0 + lj;
      if (!(lj >= H && lj < H + G::TJ && gj >= 1 && gj <= a.ny - 2)) continue;
      const int g = gi * a.pitch + gj;
      const T u = U[((li + e) & 1) * G::NC + li * SJC + m];
      if (EPI == 0 || EPI == 1 || EPI == 2) a.uout[g] = u;
      if (EPI == 1) {
        const double r = (double)resid_tile<T, H>(U, li, lj, a.f[g], a);
        acc += r * r;
      } else if (EPI == 2) {
        acc = nanmax(acc, fabs((double)(u - snap[EPI == 2 ? k : 0][e])));
      } else if (EPI == 3) {
        // u_next = u64 + e at the node and its four neighbours (boundary
        // neighbours keep their Dirichlet value), then the FP64 residual.
        const double uc0 = mx.u64in[g] + (double)u;
        double uim = mx.u64in[g - a.pitch], uip = mx.u64in[g + a.pitch];
        double ujm = mx.u64in[g - 1], ujp = mx.u64in[g + 1];
        if (gi - 1 >= 1) uim = uim + (double)lds_at<T, H>(U, li - 1, lj);
        if (gi + 1 <= a.nx - 2) uip = uip + (double)lds_at<T, H>(U, li + 1, lj);
        if (gj - 1 >= 1) ujm = ujm + (double)lds_at<T, H>(U, li, lj - 1);
        if (gj + 1 <= a.ny - 2) ujp = ujp + (double)lds_at<T, H>(U, li, lj + 1);
        const double u2 = uc0 * 2.0;
        const double r =
            mx.fn64[g] - (divh(uim + uip - u2, mx.hx2, mx.ihx2, mx.pw2) + divh(ujm + ujp - u2, mx.hy2, mx.ihy2, mx.pw2));
        mx.u64out[g] = uc0;
        mx.r32out[g] = (float)r;
        acc += r * r;
        em = nanmax(em, fabs((double)u));
      }
    }
  }
  if (EPI == 1 || EPI == 3) {
    acc = block_reduce<false>(acc, red, tid, TNT);
    if (tid == 0) part[blockIdx.x] = acc;
  }
  if (EPI == 2) {
    acc = block_reduce<true>(acc, red, tid, TNT);
    if (tid == 0) part[blockIdx.x] = acc;
  }
  if (EPI == 3) {
    em = block_reduce<true>(em, red, tid, TNT);
    if (tid == 0) mx.part_max[blockIdx.x] = em;
  }
}

}  // namespace vlp2d

namespace vlp2d {

// ---------------------------------------------------------------------------
// Refinement update of a mixed-precision step, after the FP32 correction
// cycle: uout = uin + (double) e on interior nodes (boundary nodes of uout
// already hold the Dirichlet values), r32out = (float)(fn - Lap uout) in FP64,
// partial sums of r^2 and partial max |e| per workgroup.  A workgroup stages
// its 32 x 64 output tile plus a one-node ring of uout in LDS, so every input
// is read from HBM about once.
// ---------------------------------------------------------------------------
constexpr int RF_TI = 32, RF_TJ = 64, RF_W = RF_TJ + 2, RF_N = (RF_TI + 2) * RF_W;

__global__ void __launch_bounds__(TNT) k_refine(const double *__restrict__ uin, double *__restrict__ uout,
                                                const float *__restrict__ e, const double *__restrict__ fn,
                                                float *__restrict__ r32out, int nx, int ny, int pitch, double hx2,
                                                double hy2, double ihx2, double ihy2, int pw2, int tiles_i,
                                                double *__restrict__ part_sum, double *__restrict__ part_max) {
  __shared__ double W[RF_N];
  __shared__ double red[TNT / 64];
  const int t = blockIdx.x;
  const int tj = t / tiles_i, ti = t - (t / tiles_i) * tiles_i;
  const int gi0 = ti * RF_TI, gj0 = tj * RF_TJ;  // staged origin: output origin minus the ring
  const int tid = threadIdx.x;
  constexpr int KS = (RF_N + TNT - 1) / TNT;
  double v[KS];
#pragma unroll
  for (int k = 0; k < KS; ++k) {
    const int p = tid + TNT * k;
    const int li = p / RF_W, lj = p - (p / RF_W) * RF_W;
    const int gi = gi0 + li, gj = gj0 + lj;
    double x = 0.0;
    if (p < RF_N && gi < nx && gj < ny) {
      const int g = gi * pitch + gj;
      x = uin[g];
      if (inside(gi, gj, nx, ny)) x = x + (double)e[g];
    }
    v[k] = x;
  }
#pragma unroll
  for (int k = 0; k < KS; ++k) {
    const int p = tid + TNT * k;
    if (p < RF_N) W[p] = v[k];
  }
  __syncthreads();
  const int w = tid >> 6, m = tid & 63;
  double ss = 0.0, em = 0.0;
#pragma unroll
  for (int k = 0; k < RF_TI / 4; ++k) {
    const int oi = w + 4 * k;
    const int gi = gi0 + 1 + oi, gj = gj0 + 1 + m;
    if (inside(gi, gj, nx, ny)) {
      const int c = (oi + 1) * RF_W + (m + 1);
      const int g = gi * pitch + gj;
      const double u = W[c];
      const double u2 = u * 2.0;
      const double r = fn[g] - (divh(W[c - RF_W] + W[c + RF_W] - u2, hx2, ihx2, pw2) +
                                divh(W[c - 1] + W[c + 1] - u2, hy2, ihy2, pw2));
      uout[g] = u;
      r32out[g] = (float)r;
      ss += r * r;
      em = nanmax(em, fabs((double)e[g]));
    }
  }
  ss = block_reduce<false>(ss, red, tid, TNT);
  if (tid == 0) part_sum[blockIdx.x] = ss;
  em = block_reduce<true>(em, red, tid, TNT);
  if (tid == 0) part_max[blockIdx.x] = em;
}

}  // namespace vlp2d
