// Descent launchers (k_sdown instances) for float levels, 1 pre-sweeps.
// One translation unit per precision and sweep count, compiled in parallel.
#include "mg_launch.hpp"

namespace vlp2d {

template void launch_down_ns<float, 1>(Ctx *, Lvl<float> &, Lvl<float> &, bool, bool);

int launch_down_r64(Ctx *c, Lvl<float> &f, Lvl<float> &g, const MixView &mx) {
  constexpr int H = s_down_halo(1), P = s_period(1);
  using K = void (*)(LevelView<float>, float *, int, int, int, MixView, double *);
  const K k = f.eq ? (K)k_sdown_r64<1, true> : (K)k_sdown_r64<1, false>;
  const LevelView<float> a = sview(c, (const void *)k, f, f.u, f.u2, H, P);
  hipLaunchKernelGGL(k, dim3(8 * a.chunk), dim3(64 * kWaves), 0, c->st, a, g.rhs, g.nx, g.ny, g.pitch, mx, c->part);
  return a.tiles_i * a.tiles_j;
}

}  // namespace vlp2d
