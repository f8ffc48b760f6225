// Operands shared by the fused kernels (gfx950): the view of one level, the
// mixed-precision refinement operands, and the workgroup -> tile mapping.
#pragma once
#include "mg_kernels.hpp"

namespace vlp2d {

template <typename T> struct LevelView {
  const T *uin;  // staged iterate (ignored when zero-initialised)
  T *uout;       // written iterate (tile interior nodes of the domain)
  const T *f;    // right-hand side (fn on level 0, restricted residual below)
  T *rc;         // coarse right-hand side an ascent continuing into the next descent restricts into
  int nx, ny, pitch;
  T hx2, hy2, coef;
  T ihx2, ihy2;  // 1/hx2, 1/hy2: exact when pw2 (both squares are powers of two)
  T w;           // weighted-Jacobi weight (SM_JAC kernels)
  int eq, pw2;
  // row stride of f: pitch, or 0 on Laplace levels, whose right-hand side
  // array is all zeros (every row then reads row 0, which stays in L2), so
  // the right-hand-side loads need no branch
  int fpitch;
  int tiles_i, tiles_j, chunk;  // tile grid of this launch and per-XCD chunk
  int rows;                     // output rows per tile
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

// tile number -> (tile row, tile column), row-major: tiles that share a
// column halo are consecutive, and vertically adjacent tiles are one tile
// row apart, both inside the run one XCD executes at a time.
__device__ inline void tile_coords(int t, int tiles_i, int tiles_j, int &ti, int &tj) {
  (void)tiles_i;
  ti = t / tiles_j;
  tj = t - ti * tiles_j;
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

}  // namespace vlp2d
