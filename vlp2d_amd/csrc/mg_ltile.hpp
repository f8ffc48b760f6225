// LDS-tiled ascent for the small levels of a hierarchy (gfx950).
//
// On levels of about 1025^2 nodes and fewer the register-streaming ascent
// (k_sup) is bound by latency, not bandwidth: its wavefront walks R + 2H rows
// one after another, and each step waits on a load and on a chain of 2 NS
// dependent half-sweeps.  Here a workgroup of kLtWaves wavefronts instead
// stages its whole tile at once (every load of the tile in flight together),
// adds the prolongation, runs the 2 NS half-sweeps in LDS with one barrier
// each, and writes the output nodes.
//
// Tile: kLtRows = kLtWaves * kLtRpw staged rows by SCOLS = 128 staged
// columns; output rows [ti TI, ti TI + TI), output columns [tj TJ, tj TJ + TJ)
// with TI = kLtRows - 2H, TJ = 128 - 2H and H = 2 NS (the last half-sweep's
// result is exact at distance >= 2 NS from the staged edge).  The staged
// origin (ti TI - H, tj TJ - H) has both coordinates even, so node
// (gi0 + p, cb + 2m + e) has colour (p + e) & 1 as in the streaming kernels.
// Wavefront w owns staged rows w kLtRpw .. w kLtRpw + kLtRpw - 1 and lane m
// the column pair cb + 2m, cb + 2m + 1 of each; it keeps its pairs and their
// right-hand side in registers and publishes its updated nodes in LDS, where
// the neighbours read them.  A Gauss-Seidel half-sweep writes only nodes of
// one colour and reads only nodes of the other, so the in-place update in LDS
// has no race within a half-sweep.
//
// Arithmetic: the prolongation (MultiGridScheme.cs:171-186) and the point
// update (MultiGridSlidingIteration.cs:53-57) in the reference's operation
// order, with the same helpers as the streaming kernels; within one colour the
// update is order independent, so the result is bitwise that of k_sup with
// N1 = 0 and no epilogue.
#pragma once
#include "mg_stream.hpp"

namespace vlp2d {

constexpr int kLtWaves = 4;                  // wavefronts per workgroup
constexpr int kLtRpw = 8;                    // staged rows per wavefront
constexpr int kLtRows = kLtWaves * kLtRpw;   // staged rows per tile

__host__ __device__ constexpr int lt_halo(int ns) { return 2 * ns; }

struct LtGrid {
  int ti, tj, tiles;  // output rows / columns per tile, tile count
  int tiles_j;
};
inline LtGrid lt_grid(int nx, int ny, int H) {
  LtGrid g;
  g.ti = kLtRows - 2 * H;
  g.tj = SCOLS - 2 * H;
  const int tiles_i = (nx - 2) / g.ti + 1;  // output rows cover 1..nx-2
  g.tiles_j = (ny - 2) / g.tj + 1;
  g.tiles = tiles_i * g.tiles_j;
  return g;
}

// Every load and LDS read below is unconditional (clamped addresses, the
// value selected afterwards), so a wavefront issues all of them back to back
// and waits once; only stores are predicated.
// lt_pass: one pass over the staged tile with origin (gi0, cb) (both even) of
// level a: stage a.uin plus the prolongation of the coarse correction
// cval(I, J) (coarse grid ncx x ncy), NS sweeps in su; OUT: write the output
// nodes to a.uout, else the result stays in su only.
template <typename T, int NS, bool EQ, bool PW2, bool OUT, typename CV>
__device__ __forceinline__ void lt_pass(const LevelView<T> &a, int gi0, int cb, int ncx, int ncy, CV &&cval, T *su) {
  constexpr int H = lt_halo(NS), TI = kLtRows - 2 * H, TJ = SCOLS - 2 * H;
  const int w = (int)(threadIdx.x >> 6), m = (int)(threadIdx.x & 63);
  const int gj = cb + 2 * m;
  const bool ld = gj >= 0 && gj <= a.pitch - 2;
  const bool in0 = gj >= 1 && gj <= a.ny - 2, in1 = gj + 1 >= 1 && gj + 1 <= a.ny - 2;
  const unsigned boff = ld ? (unsigned)gj * (unsigned)sizeof(T) : 0u;
  const int p0 = w * kLtRpw;  // first owned staged row (even, as gi0)
  // coarse columns J = gj >> 1 and J + 1, clamped into the row (exact for
  // interior fine nodes, the only ones that take the prolongation)
  const int J0 = std::min(std::max(gj, 0) >> 1, ncy - 1), J1 = std::min(J0 + 1, ncy - 1);

  // stage: iterate + prolongation of the coarse correction, right-hand side
  T U[kLtRpw][2], F[kLtRpw][2];
#pragma unroll
  for (int k = 0; k < kLtRpw; ++k) {
    const int gi = gi0 + p0 + k;
    const bool rok = ld && gi >= 0 && gi <= a.nx - 1;
    const int gic = std::min(std::max(gi, 0), a.nx - 1);
    Pair<T> u = ld_pair(row_of(a.uin, gic, a.pitch), boff);
    const Pair<T> f = ld_pair(row_of(a.f, gic, a.fpitch), boff);
    const int I0 = std::min(gic >> 1, ncx - 1), I1 = std::min(I0 + 1, ncx - 1);
    const T c00 = cval(I0, J0), c01 = cval(I0, J1);
    T q0, q1;
    if ((k & 1) == 0) {  // even row
      q0 = c00;
      q1 = (c00 + c01) * T(0.5);
    } else {
      const T c10 = cval(I1, J0), c11 = cval(I1, J1);
      q0 = (c00 + c10) * T(0.5);
      q1 = (c00 + c01 + c10 + c11) * T(0.25);
    }
    const bool rin = gi >= 1 && gi <= a.nx - 2;
    T v0 = rok ? u.x0 : T(0), v1 = rok ? u.x1 : T(0);
    if (rin && in0) v0 = v0 + q0;
    if (rin && in1) v1 = v1 + q1;
    U[k][0] = v0;
    U[k][1] = v1;
    F[k][0] = rok ? f.x0 : T(0);
    F[k][1] = rok ? f.x1 : T(0);
  }
#pragma unroll
  for (int k = 0; k < kLtRpw; ++k)
    st_pair(su + (p0 + k) * SCOLS, (unsigned)(2 * m) * (unsigned)sizeof(T), U[k][0], U[k][1]);
  __syncthreads();

  // NS sweeps of two half-sweeps, red (colour 0) first: in row p the colour
  // cs node of the lane's pair is element (p + cs) & 1 = (k + cs) & 1
  auto half = [&](auto csc) {
    constexpr int cs = decltype(csc)::value;
    T G[kLtRpw];
#pragma unroll
    for (int k = 0; k < kLtRpw; ++k) {
      const int e = (k + cs) & 1;
      const int p = std::min(std::max(p0 + k, 1), kLtRows - 2);  // clamped: rows 0, kLtRows-1 stay fixed
      const int j = std::min(std::max(2 * m + e, 1), SCOLS - 2);
      const T *row = su + p * SCOLS + j;
      G[k] = s_gs<T, EQ, PW2>(row[-SCOLS], row[SCOLS], row[-1], row[1], F[k][e], a);
    }
#pragma unroll
    for (int k = 0; k < kLtRpw; ++k) {
      const int e = (k + cs) & 1;
      const int p = p0 + k, gi = gi0 + p, j = 2 * m + e;
      const bool go = p >= 1 && p <= kLtRows - 2 && j >= 1 && j <= SCOLS - 2 && gi >= 1 && gi <= a.nx - 2 &&
                      (e == 0 ? in0 : in1);
      if (go) {
        su[p * SCOLS + j] = G[k];
        U[k][e] = G[k];
      }
    }
    __syncthreads();
  };
#pragma unroll 1
  for (int s = 0; s < NS; ++s) {
    half(std::integral_constant<int, 0>{});
    half(std::integral_constant<int, 1>{});
  }

  if constexpr (!OUT) return;
  // output nodes: rows [H, H + TI), lanes [H/2, H/2 + TJ/2), interior nodes
  const bool oc = m >= H / 2 && m < H / 2 + TJ / 2;
#pragma unroll
  for (int k = 0; k < kLtRpw; ++k) {
    const int p = p0 + k, gi = gi0 + p;
    if (!oc || p < H || p >= H + TI || gi < 1 || gi > a.nx - 2) continue;
    T *row = row_of(a.uout, gi, a.pitch);
    if (in0 && in1) st_pair(row, boff, U[k][0], U[k][1]);
    else if (in0) row[gj] = U[k][0];
    else if (in1) row[gj + 1] = U[k][1];
  }
}

template <typename T, int NS, bool EQ, bool PW2>
__device__ __forceinline__ void lt_up_body(const LevelView<T> &a, const T *__restrict__ uc, int ncx, int ncy, int pc,
                                           int tiles_j, T *su) {
  constexpr int H = lt_halo(NS), TI = kLtRows - 2 * H, TJ = SCOLS - 2 * H;
  const int t = blockIdx.x;
  const int ti = t / tiles_j, tj = t - ti * tiles_j;
  lt_pass<T, NS, EQ, PW2, true>(a, ti * TI - H, tj * TJ - H, ncx, ncy, [&](int I, int J) { return uc[I * pc + J]; },
                                su);
}

// Ascent of a zero-initialised level whose pre-smoothed iterate the descent
// stored (a.uin): prolongation + NS post-smoothing sweeps -> a.uout.
template <typename T, int NS, bool EQ>
__global__ void __launch_bounds__(64 * kLtWaves) k_lt_up(LevelView<T> a, const T *__restrict__ uc, int ncx, int ncy,
                                                        int pc, int tiles_j) {
  __shared__ __align__(16) T su[kLtRows * SCOLS];
  if (a.pw2) lt_up_body<T, NS, EQ, true>(a, uc, ncx, ncy, pc, tiles_j, su);
  else lt_up_body<T, NS, EQ, false>(a, uc, ncx, ncy, pc, tiles_j, su);
}

// Two levels in one launch: the ascent of level b (the coarser, l + 1; its
// stored pre-smoothed iterate b.uin, right-hand side b.f, correction uc2 of
// level l + 2) and then of level a (l), whose tile is the launch's.  Each
// workgroup first runs b's pass over the coarse tile that holds every coarse
// node its fine tile's prolongation reads, exact after the NS sweeps (origin
// even, H coarse nodes further out on every side; neighbouring workgroups
// compute overlapping coarse tiles, the same values), keeps it in LDS (sc),
// and prolongates from there.  b's result is never written: the cycle reads
// level l + 1 only through this prolongation.  Same operations per node as
// k_lt_up on each level, so the same bits, with one launch instead of two.
template <typename T, int NS, bool EQ>
__global__ void __launch_bounds__(64 * kLtWaves) k_lt_up2(LevelView<T> a, LevelView<T> b, const T *__restrict__ uc2,
                                                         int ncx2, int ncy2, int pc2, int tiles_j) {
  __shared__ __align__(16) T su[kLtRows * SCOLS];
  __shared__ __align__(16) T sc[kLtRows * SCOLS];
  constexpr int H = lt_halo(NS), TI = kLtRows - 2 * H, TJ = SCOLS - 2 * H;
  static_assert(2 * H < 15, "the coarse tile holds the fine tile's footprint");
  const int t = blockIdx.x;
  const int ti = t / tiles_j, tj = t - ti * tiles_j;
  const int gi0 = ti * TI - H, cb = tj * TJ - H;
  // coarse tile origin: even, at most H above / left of the first coarse
  // node the fine tile reads (>> 1 rounds down for negative origins)
  const int ci0 = ((gi0 >> 1) - H) & ~1, cj0 = ((cb >> 1) - H) & ~1;
  auto pass = [&](auto pw2) {
    constexpr bool PW2 = decltype(pw2)::value;
    lt_pass<T, NS, EQ, PW2, false>(b, ci0, cj0, ncx2, ncy2, [&](int I, int J) { return uc2[I * pc2 + J]; }, sc);
    lt_pass<T, NS, EQ, PW2, true>(a, gi0, cb, b.nx, b.ny,
                                  [&](int I, int J) { return sc[(I - ci0) * SCOLS + (J - cj0)]; }, su);
  };
  if (a.pw2) pass(std::true_type{});
  else pass(std::false_type{});
}

}  // namespace vlp2d
