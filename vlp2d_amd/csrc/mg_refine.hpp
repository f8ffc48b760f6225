// Streaming refinement update of a mixed-precision step (gfx950).
//
// uout = uin + (double) e on interior nodes, r32out = (float)(fn - Lap uout)
// in FP64 with the reference's association (VisualLaplacePoisson2D/Common/
// UtilsOpLap.cs:15-20), partial sums of r^2 and partial max |e|.
//
// Each wavefront owns 64 consecutive columns and RR consecutive rows.  Lane m
// slides a three-row window (rows i-1, i, i+1 of column j) of the updated
// iterate down its column, so each input value is loaded once; the two
// horizontal neighbours come from lanes m-1 and m+1 through cross-lane
// shuffles (lanes 0 and 63 load theirs).  Inputs are read from HBM once,
// outputs written once: 8 (uin) + 4 (e) + 8 (fn) + 8 (uout) + 4 (r32) bytes
// per node.
#pragma once
#include "mg_tiles.hpp"

namespace vlp2d {

constexpr int RR = 16;  // rows per wavefront

__device__ inline double refined(const double *uin, const float *e, int i, int j, int nx, int ny, int pitch,
                                 float &ev) {
  const int g = i * pitch + j;
  double x = uin[g];
  ev = 0.0f;
  if (i >= 1 && i <= nx - 2 && j >= 1 && j <= ny - 2) {
    ev = e[g];
    x = x + (double)ev;
  }
  return x;
}

// grid: x = column blocks of 64, y = row blocks of 4*RR; block = 256 threads
__global__ void __launch_bounds__(256) k_refine_stream(const double *__restrict__ uin, double *__restrict__ uout,
                                                       const float *__restrict__ e, const double *__restrict__ fn,
                                                       float *__restrict__ r32out, int nx, int ny, int pitch,
                                                       double hx2, double hy2, double ihx2, double ihy2, int pw2,
                                                       double *__restrict__ part_sum, double *__restrict__ part_max) {
  __shared__ double red[4];
  const int tid = threadIdx.x, w = tid >> 6, m = tid & 63;
  const int j = 1 + blockIdx.x * 64 + m;             // this lane's column
  const int i0 = 1 + (blockIdx.y * 4 + w) * RR;      // first output row of this wavefront
  const bool lane_in = j <= ny - 1;                  // holds a node (interior or right boundary)
  double ss = 0.0, em = 0.0;
  float ev;
  double a = 0.0, b = 0.0;
  if (lane_in && i0 - 1 <= nx - 1) a = refined(uin, e, i0 - 1, j, nx, ny, pitch, ev);
  if (lane_in && i0 <= nx - 1) b = refined(uin, e, i0, j, nx, ny, pitch, ev);
  float eb = ev;
  // No early exit from the row loop: rows past the grid are predicated off,
  // so the unrolled loads of later rows can be issued ahead.
#pragma unroll
  for (int r = 0; r < RR; ++r) {
    const int i = i0 + r;
    const bool row_in = i <= nx - 2;  // wavefront-uniform
    double c = 0.0;
    float ec = 0.0f;
    if (lane_in && row_in) c = refined(uin, e, i + 1, j, nx, ny, pitch, ec);
    double l = __shfl(b, (m + 63) & 63, 64);
    double rt = __shfl(b, (m + 1) & 63, 64);
    float dummy;
    if (m == 0 && row_in) l = refined(uin, e, i, j - 1, nx, ny, pitch, dummy);
    if (m == 63 && row_in && j + 1 <= ny - 1) rt = refined(uin, e, i, j + 1, nx, ny, pitch, dummy);
    if (row_in && j <= ny - 2) {
      const int g = i * pitch + j;
      const double u2 = b * 2.0;
      const double res = fn[g] - (divh(a + c - u2, hx2, ihx2, pw2) + divh(l + rt - u2, hy2, ihy2, pw2));
      uout[g] = b;
      r32out[g] = (float)res;
      ss += res * res;
      em = nanmax(em, fabs((double)eb));
    }
    a = b;
    b = c;
    eb = ec;
  }
  ss = block_reduce<false>(ss, red, tid, 256);
  const int blk = blockIdx.y * gridDim.x + blockIdx.x;
  if (tid == 0) part_sum[blk] = ss;
  em = block_reduce<true>(em, red, tid, 256);
  if (tid == 0) part_max[blk] = em;
}

}  // namespace vlp2d
