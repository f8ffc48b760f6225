// Launchers of the transposed-cycle streaming kernels (k_tdown, k_tup of
// mg_tstream.hpp) on FP64 levels, 1..3 sweeps per pass.
#include "mg_ctx.hpp"
#include "mg_tstream.hpp"

namespace vlp2d {

namespace {

// Tiles of R output rows by 128 - 2H columns covering every node of the
// level, R chosen so that the tiles fill the resident wavefront slots about
// once (as the forward streaming kernels).
TLevel tview(Ctx *c, const void *fn, const Lvl<double> &v, const double *ubin, double *ubout, double *rb, int H,
             int P) {
  TLevel a;
  a.ubin = ubin;
  a.ubout = ubout;
  a.rb = rb;
  a.nx = v.nx;
  a.ny = v.ny;
  a.pitch = v.pitch;
  a.ax = v.ax;
  a.ay = v.ay;
  a.b = v.b;
  a.ix2 = v.ix2;
  a.iy2 = v.iy2;
  const int tj = (v.ny - 1) / (SCOLS - 2 * H) + 1;
  const int ti = std::max(1, resident_waves(c, fn) / tj);
  const int R = s_rows_at_least((v.nx - 1) / ti + 1, H, P);
  const STileGrid g = t_grid(v.nx, v.ny, R, H);
  a.rows = R;
  a.tiles_i = g.tiles_i;
  a.tiles_j = g.tiles_j;
  a.chunk = g.chunk;
  return a;
}

template <int NS>
void tdown_ns(Ctx *c, const Lvl<double> &f, const double *ubin, double *ubout, double *rb, const Lvl<double> &g,
              double *ubc, double *rbc) {
  constexpr int H = t_halo(NS), P = t_period(NS);
  const void *fn = (const void *)k_tdown<NS>;
  const TLevel a = tview(c, fn, f, ubin, ubout, rb, H, P);
  const TCoarse cg{ubc, rbc, g.nx, g.ny, g.pitch};
  hipLaunchKernelGGL(k_tdown<NS>, dim3(8 * a.chunk), dim3(64 * kWaves), 0, c->st, a, cg);
}

template <int NS>
void tup_ns(Ctx *c, const Lvl<double> &f, const double *ubin, double *ubout, double *rb, const Lvl<double> &g,
            double *rbc) {
  constexpr int H = t_halo(NS), P = t_period(NS);
  const void *fn = (const void *)k_tup<NS>;
  const TLevel a = tview(c, fn, f, ubin, ubout, rb, H, P);
  const TCoarse cg{nullptr, rbc, g.nx, g.ny, g.pitch};
  hipLaunchKernelGGL(k_tup<NS>, dim3(8 * a.chunk), dim3(64 * kWaves), 0, c->st, a, cg);
}

}  // namespace

void launch_tdown(Ctx *c, int ns, const Lvl<double> &f, const double *ubin, double *ubout, double *rb,
                  const Lvl<double> &g, double *ubc, double *rbc) {
  switch (ns) {
    case 1: tdown_ns<1>(c, f, ubin, ubout, rb, g, ubc, rbc); break;
    case 2: tdown_ns<2>(c, f, ubin, ubout, rb, g, ubc, rbc); break;
    default: tdown_ns<3>(c, f, ubin, ubout, rb, g, ubc, rbc); break;
  }
}

void launch_tup(Ctx *c, int ns, const Lvl<double> &f, const double *ubin, double *ubout, double *rb,
                const Lvl<double> &g, double *rbc) {
  switch (ns) {
    case 1: tup_ns<1>(c, f, ubin, ubout, rb, g, rbc); break;
    case 2: tup_ns<2>(c, f, ubin, ubout, rb, g, rbc); break;
    default: tup_ns<3>(c, f, ubin, ubout, rb, g, rbc); break;
  }
}

}  // namespace vlp2d
