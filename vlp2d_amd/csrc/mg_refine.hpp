// Streaming refinement update of a mixed-precision step (gfx950).
//
// uout = uin + (double) e on interior nodes, r32out = (float)(fn - Lap uout)
// in FP64 with the reference's association (VisualLaplacePoisson2D/Common/
// UtilsOpLap.cs:15-20), partial sums of r^2 and partial max |e|.
//
// Layout of the work: a workgroup owns a tile of RS_COLS = 512 columns
// (four wavefronts side by side, 128 columns each, lane m holding the column
// pair 2m, 2m+1) and RS_ROWS output rows.  Each lane slides a three-row
// window of the updated iterate down its column pair, so every row of uin,
// e and fn is loaded once per tile (plus the two halo rows); the horizontal
// neighbours come from lanes m-1 and m+1 through cross-lane shuffles (lanes 0
// and 63 load theirs).  Tile columns start on multiples of 512, so with the
// pitch a multiple of 32 every row segment is aligned to whole 128-byte lines
// and the pair loads and stores are 16-byte (FP64) and 8-byte (FP32) vectors.
// Tiles are numbered column-strip-major and handed to XCDs in contiguous
// runs, so vertically adjacent tiles (which share halo rows) run on the same
// XCD and meet in its L2.  Compulsory traffic: 8 (uin) + 4 (e) + 8 (fn) +
// 8 (uout) + 4 (r32) bytes per node.
#pragma once
#include "mg_tiles.hpp"

#ifndef VLP2D_REFINE_ROWS
#define VLP2D_REFINE_ROWS 128  // output rows per tile (32 and 64 measured slower on 8193^2)
#endif
// bit 0: non-temporal stores of uout / r32out; bit 1: non-temporal loads of
// uin / e / fn (all are streamed once per launch)
#ifndef VLP2D_REFINE_NT
#define VLP2D_REFINE_NT 1
#endif

namespace vlp2d {

typedef double dvec2 __attribute__((ext_vector_type(2)));
typedef float fvec2 __attribute__((ext_vector_type(2)));

template <typename V> __device__ inline V ld_stream(const V *p) {
  if (VLP2D_REFINE_NT & 2) return __builtin_nontemporal_load(p);
  return *p;
}
template <typename V> __device__ inline void st_stream(V *p, V v) {
  if (VLP2D_REFINE_NT & 1) __builtin_nontemporal_store(v, p);
  else *p = v;
}

constexpr int RS_ROWS = VLP2D_REFINE_ROWS;  // output rows per tile
constexpr int RS_COLS = 512;                // columns per tile (4 wavefronts x 64 lanes x 2)

// Number of column strips of a level with ny columns: output columns are
// 1..ny-2, strip s covers columns [512 s, 512 s + 512).
inline __host__ __device__ int refine_strips(int ny) { return (ny - 2) / RS_COLS + 1; }
inline __host__ __device__ int refine_row_blocks(int nx) { return (nx - 2 + RS_ROWS - 1) / RS_ROWS; }

// Updated iterate at (i, j0) and (i, j0 + 1): x = uin + (double) e on
// interior nodes, uin on boundary nodes (UPDATE = false: x = uin).  j0 is
// even and j0 <= ny - 1, so both columns lie inside the row's pitch (pitch
// >= ny rounded up to 32).
template <bool UPDATE>
__device__ inline void refined_pair(const double *__restrict__ uin, const float *__restrict__ e, int g, int i, int j0,
                                    int nx, int ny, double &x0, double &x1, float &e0, float &e1) {
  const dvec2 u = ld_stream(reinterpret_cast<const dvec2 *>(uin + g));
  x0 = u.x;
  x1 = u.y;
  e0 = 0.0f;
  e1 = 0.0f;
  if (UPDATE && i >= 1 && i <= nx - 2) {
    const fvec2 ee = ld_stream(reinterpret_cast<const fvec2 *>(e + g));
    if (j0 >= 1 && j0 <= ny - 2) {
      e0 = ee.x;
      x0 = x0 + (double)e0;
    }
    if (j0 + 1 <= ny - 2) {
      e1 = ee.y;
      x1 = x1 + (double)e1;
    }
  }
}

template <bool UPDATE>
__device__ inline double refined_one(const double *__restrict__ uin, const float *__restrict__ e, int i, int j, int nx,
                                     int ny, int pitch) {
  const int g = i * pitch + j;
  double x = uin[g];
  if (UPDATE && i >= 1 && i <= nx - 2 && j >= 1 && j <= ny - 2) x = x + (double)e[g];
  return x;
}

// grid: 8 * chunk workgroups of 256 threads (see tile_of_block); tiles =
// strips * row_blocks; partials are written per tile.  UPDATE = false is the
// residual of uin alone (e, uout and part_max unused; r32out may be null):
// the initial residual of a solve, the same per-node formula and order.
template <bool UPDATE>
__global__ void __launch_bounds__(256) k_refine_stream(const double *__restrict__ uin, double *__restrict__ uout,
                                                       const float *__restrict__ e, const double *__restrict__ fn,
                                                       float *__restrict__ r32out, int nx, int ny, int pitch,
                                                       double hx2, double hy2, double ihx2, double ihy2, int pw2,
                                                       int row_blocks, int tiles, int chunk,
                                                       double *__restrict__ part_sum, double *__restrict__ part_max) {
  __shared__ double red[4];
  const int t = tile_of_block(tiles, chunk);
  if (t < 0) return;  // whole workgroup
  const int strip = t / row_blocks, rb = t - strip * row_blocks;
  const int tid = threadIdx.x, w = tid >> 6, m = tid & 63;
  const int j0 = strip * RS_COLS + w * 128 + 2 * m;  // this lane's column pair j0, j0 + 1
  const int i0 = 1 + rb * RS_ROWS;                   // first output row (<= nx - 2)
  const bool lane_in = j0 <= ny - 1;                 // holds a node (its values feed the neighbours' stencils)
  const bool out0 = j0 >= 1 && j0 <= ny - 2, out1 = j0 + 1 <= ny - 2;
  double ss = 0.0, em = 0.0;
  double a0 = 0.0, a1 = 0.0, b0 = 0.0, b1 = 0.0;
  float eb0 = 0.0f, eb1 = 0.0f, ed0, ed1;
  if (lane_in) {
    refined_pair<UPDATE>(uin, e, (i0 - 1) * pitch + j0, i0 - 1, j0, nx, ny, a0, a1, ed0, ed1);
    refined_pair<UPDATE>(uin, e, i0 * pitch + j0, i0, j0, nx, ny, b0, b1, eb0, eb1);
  }
  // No early exit from the row loop: rows past the grid are predicated off
  // (wavefront-uniform), so the loads of later rows can be issued ahead.
#pragma unroll
  for (int r = 0; r < RS_ROWS; ++r) {
    const int i = i0 + r;
    const bool row_in = i <= nx - 2;
    const int g = i * pitch + j0;
    double c0 = 0.0, c1 = 0.0;
    float ec0 = 0.0f, ec1 = 0.0f;
    dvec2 f = {0.0, 0.0};
    if (lane_in && row_in) {
      refined_pair<UPDATE>(uin, e, g + pitch, i + 1, j0, nx, ny, c0, c1, ec0, ec1);
      f = ld_stream(reinterpret_cast<const dvec2 *>(fn + g));
    }
    double l = __shfl(b1, (m + 63) & 63, 64);
    double rt = __shfl(b0, (m + 1) & 63, 64);
    if (m == 0 && row_in && j0 >= 1) l = refined_one<UPDATE>(uin, e, i, j0 - 1, nx, ny, pitch);
    if (m == 63 && row_in && j0 + 2 <= ny - 1) rt = refined_one<UPDATE>(uin, e, i, j0 + 2, nx, ny, pitch);
    if (row_in) {
      const double u20 = b0 * 2.0, u21 = b1 * 2.0;
      const double res0 = f.x - (divh(a0 + c0 - u20, hx2, ihx2, pw2) + divh(l + b1 - u20, hy2, ihy2, pw2));
      const double res1 = f.y - (divh(a1 + c1 - u21, hx2, ihx2, pw2) + divh(b0 + rt - u21, hy2, ihy2, pw2));
      if (UPDATE) {
        if (out0 && out1) st_stream(reinterpret_cast<dvec2 *>(uout + g), dvec2{b0, b1});
        else if (out0) st_stream(uout + g, b0);
        else if (out1) st_stream(uout + g + 1, b1);
      }
      if (UPDATE || r32out) {
        if (out0 && out1) st_stream(reinterpret_cast<fvec2 *>(r32out + g), fvec2{(float)res0, (float)res1});
        else if (out0) st_stream(r32out + g, (float)res0);
        else if (out1) st_stream(r32out + g + 1, (float)res1);
      }
      if (out0) {
        ss += res0 * res0;
        em = nanmax(em, fabs((double)eb0));
      }
      if (out1) {
        ss += res1 * res1;
        em = nanmax(em, fabs((double)eb1));
      }
    }
    a0 = b0;
    a1 = b1;
    b0 = c0;
    b1 = c1;
    eb0 = ec0;
    eb1 = ec1;
  }
  ss = block_reduce<false>(ss, red, tid, 256);
  if (tid == 0) part_sum[t] = ss;
  if (UPDATE) {
    em = block_reduce<true>(em, red, tid, 256);
    if (tid == 0) part_max[t] = em;
  }
}

}  // namespace vlp2d
