// Ascent launchers (k_sup instances) for float levels, 2 post-sweeps.
// One translation unit per precision and sweep count, compiled in parallel.
#include "mg_ctx.hpp"

namespace vlp2d {

template <typename T, int NS, int EPI, int N1, bool ZS = true>
int launch_sup(Ctx *c, Lvl<T> &f, const Lvl<T> &g, const MixView &mx) {
  constexpr int H = s_up_halo(N1 + NS, EPI), P = s_period(N1 + NS);
  using K = void (*)(LevelView<T>, const T *, int, int, int, double *, MixView);
  const K k = f.eq ? (K)k_sup<T, NS, EPI, true, N1, ZS> : (K)k_sup<T, NS, EPI, false, N1, ZS>;
  // a recomputed non-zero start is staged from the descent's input f.u
  const LevelView<T> a = sview(c, (const void *)k, f, ZS ? f.u2 : f.u, f.u, H, P);
  hipLaunchKernelGGL(k, dim3(8 * a.chunk), dim3(64 * kWaves), 0, c->st, a, g.u, g.nx, g.ny, g.pitch, c->part, mx);
  return a.tiles_i * a.tiles_j;
}

// FP64 levels run epilogues none / residual norm / update norm; FP32 levels
// (the correction cycle of a mixed step) none / fused refinement.  rz: the
// level's pre-smoothed iterate is recomputed (one pre-smoothing sweep).
template <typename T, int NS>
int launch_up_ns(Ctx *c, Lvl<T> &f, const Lvl<T> &g, int epi, bool rz, bool zi, const MixView &mx) {
  if constexpr (sizeof(T) == 8) {
    switch (epi) {
      case EPI_RESNORM:
        return rz ? launch_sup<T, NS, EPI_RESNORM, 1>(c, f, g, mx) : launch_sup<T, NS, EPI_RESNORM, 0>(c, f, g, mx);
      case EPI_METRIC: return launch_sup<T, NS, EPI_METRIC, 0>(c, f, g, mx);
      default: return rz ? launch_sup<T, NS, EPI_NONE, 1>(c, f, g, mx) : launch_sup<T, NS, EPI_NONE, 0>(c, f, g, mx);
    }
  } else {
    if (epi == EPI_MIXED) {
      if (rz && !zi) return launch_sup<T, NS, EPI_MIXED, 1, false>(c, f, g, mx);
      return rz ? launch_sup<T, NS, EPI_MIXED, 1>(c, f, g, mx) : launch_sup<T, NS, EPI_MIXED, 0>(c, f, g, mx);
    }
    return rz ? launch_sup<T, NS, EPI_NONE, 1>(c, f, g, mx) : launch_sup<T, NS, EPI_NONE, 0>(c, f, g, mx);
  }
}

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
