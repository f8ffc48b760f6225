// Ascent launchers (k_sup instances) for float levels, 2 post-sweeps.
// One translation unit per precision and sweep count, compiled in parallel.
#include "mg_launch.hpp"

namespace vlp2d {

// Level 0 between the two FP32 cycles of a mixed step, one launch: the first
// cycle's ascent (pre-smoothed iterate recomputed from zero, prolongation,
// nu2 sweeps) continues into the second cycle's descent (one sweep on e, its
// result -> z.u2 for the second cycle's ascent, residual, restriction into
// level 1's right-hand side).
template <int NS> void launch_up_down0_ns(Ctx *c, Lvl<float> &f, Lvl<float> &g) {
  constexpr int NB = 1, H = s_up_halo(1 + NS + NB, 4), P = s_period(1 + NS + NB);
  using K = void (*)(LevelView<float>, const float *, int, int, int, double *, MixView);
  const K k = f.eq ? (K)k_sup<float, NS, 4, true, 1, true, NB> : (K)k_sup<float, NS, 4, false, 1, true, NB>;
  LevelView<float> a = sview(c, (const void *)k, f, f.u, f.u2, H, P);  // zero start: f.u is not read
  a.rc = g.rhs;
  hipLaunchKernelGGL(k, dim3(8 * a.chunk), dim3(64 * kWaves), 0, c->st, a, g.u, g.nx, g.ny, g.pitch, c->part, MixView{});
}

template void launch_up_down0_ns<2>(Ctx *, Lvl<float> &, Lvl<float> &);
template int launch_up_ns<float, 2>(Ctx *, Lvl<float> &, const Lvl<float> &, int, bool, bool, const MixView &);

}  // namespace vlp2d
