// Launchers of the streaming kernels (k_sdown, k_sup), instantiated per
// precision and sweep count in mg_k_down_*.hip / mg_k_up_*.hip, which are
// compiled in parallel.  The smoother and the restriction of the handle's
// options select the kernel instance.
#pragma once
#include "mg_ctx.hpp"

namespace vlp2d {

// The Gauss-Seidel FP32 ascents without an epilogue on levels of 4097^2 up to
// (not including) 8193^2 nodes prefetch three rows ahead instead of two:
// measured on MI355X, the 4097^2 level 49 -> 43.5 us per launch, while 2049^2
// went 28 -> 29 us (larger levels are bandwidth-bound and keep two).  The FP32
// Gauss-Seidel descents below 8193^2 do too (launch_down_ns).
constexpr size_t kDeepPfMin = (size_t)4097 * 4097, kDeepPfEnd = (size_t)8193 * 8193;

// Descent of level f into g's right-hand side (zi: f starts from zero;
// store: the pre-smoothed iterate is written to f.u2); PF: prefetch distance.
template <typename T, int NS, int SM, bool FW, int PF = kPf>
void launch_down_k(Ctx *c, Lvl<T> &f, Lvl<T> &g, bool zi, bool store) {
  constexpr int H = s_down_halo(NS, SM), P = s_period(NS, SM, PF);
  using K = void (*)(LevelView<T>, T *, int, int, int);
  K k;
  if constexpr (SM == SM_JAC) {
    // a Jacobi-smoothed level is always stored (the host never asks otherwise)
    (void)store;
    if (!zi)
      k = f.eq ? (K)k_sdown<T, NS, false, true, true, SM, FW, PF> : (K)k_sdown<T, NS, false, false, true, SM, FW, PF>;
    else k = f.eq ? (K)k_sdown<T, NS, true, true, true, SM, FW, PF> : (K)k_sdown<T, NS, true, false, true, SM, FW, PF>;
  } else {
    if (!zi && store)
      k = f.eq ? (K)k_sdown<T, NS, false, true, true, SM, FW, PF> : (K)k_sdown<T, NS, false, false, true, SM, FW, PF>;
    else if (!zi)
      k = f.eq ? (K)k_sdown<T, NS, false, true, false, SM, FW, PF>
               : (K)k_sdown<T, NS, false, false, false, SM, FW, PF>;
    else if (store)
      k = f.eq ? (K)k_sdown<T, NS, true, true, true, SM, FW, PF> : (K)k_sdown<T, NS, true, false, true, SM, FW, PF>;
    else
      k = f.eq ? (K)k_sdown<T, NS, true, true, false, SM, FW, PF> : (K)k_sdown<T, NS, true, false, false, SM, FW, PF>;
  }
  const LevelView<T> a = sview(c, (const void *)k, f, f.u, f.u2, H, P);
  hipLaunchKernelGGL(k, dim3(8 * a.chunk), dim3(64 * kWaves), 0, c->st, a, g.rhs, g.nx, g.ny, g.pitch);
}

template <typename T, int NS>
void launch_down_ns(Ctx *c, Lvl<T> &f, Lvl<T> &g, bool zi, bool store) {
  const bool fw = c->o.restriction == VLP2D_RESTRICT_FULL;
  if (c->o.smoother == VLP2D_SMOOTH_WJACOBI) {
    if (fw) launch_down_k<T, NS, SM_JAC, true>(c, f, g, zi, store);
    else launch_down_k<T, NS, SM_JAC, false>(c, f, g, zi, store);
  } else if (fw) {
    launch_down_k<T, NS, SM_GS, true>(c, f, g, zi, store);
  } else {
    if constexpr (sizeof(T) == 4) {
      // FP32 Gauss-Seidel descents below 8193^2 nodes prefetch three rows
      // ahead (measured on MI355X: 4097^2 25.4 -> 25.0 us, 2049^2 12.9 -> 11.9
      // us per launch; the bandwidth-bound 8193^2 descent 81 -> 89 us keeps two)
      if ((size_t)f.nx * (size_t)f.ny < kDeepPfEnd) {
        launch_down_k<T, NS, SM_GS, false, 3>(c, f, g, zi, store);
        return;
      }
    }
    launch_down_k<T, NS, SM_GS, false>(c, f, g, zi, store);
  }
}

// ND > 0: the next refinement step's first descent folded into the
// refinement ascent, restricting into g's right-hand side.
template <typename T, int NS, int EPI, int N1, bool ZS = true, int SM = SM_GS, int PF = kPf, int ND = 0>
int launch_sup(Ctx *c, Lvl<T> &f, const Lvl<T> &g, const MixView &mx) {
  constexpr int H = s_up_halo(N1 + NS, EPI, SM, ND), P = s_period(N1 + NS, SM, PF);
  using K = void (*)(LevelView<T>, const T *, int, int, int, double *, MixView);
  const K k = f.eq ? (K)k_sup<T, NS, EPI, true, N1, ZS, 0, SM, PF, ND>
                   : (K)k_sup<T, NS, EPI, false, N1, ZS, 0, SM, PF, ND>;
  // a recomputed non-zero start is staged from the descent's input f.u
  LevelView<T> a = sview(c, (const void *)k, f, ZS ? f.u2 : f.u, f.u, H, P);
  if constexpr (ND > 0) a.rc = g.rhs;
  hipLaunchKernelGGL(k, dim3(8 * a.chunk), dim3(64 * kWaves), 0, c->st, a, g.u, g.nx, g.ny, g.pitch, c->part, mx);
  return a.tiles_i * a.tiles_j;
}

// FP64 levels run epilogues none / residual norm / update norm; FP32 levels
// (the correction cycle of a mixed step) none / fused refinement.  rz: the
// level's pre-smoothed iterate is recomputed (one Gauss-Seidel pre-sweep; a
// Jacobi-smoothed level is always stored, rz false).  nd = 1 (FP32 refinement
// ascents that start from zero or stage the stored pre-smoothed iterate, one
// pre-sweep): the next step's first descent is folded in.
template <typename T, int NS>
int launch_up_ns(Ctx *c, Lvl<T> &f, const Lvl<T> &g, int epi, bool rz, bool zi, const MixView &mx, int nd) {
  if (c->o.smoother == VLP2D_SMOOTH_WJACOBI) {
    if constexpr (sizeof(T) == 8) {
      switch (epi) {
        case EPI_RESNORM: return launch_sup<T, NS, EPI_RESNORM, 0, true, SM_JAC>(c, f, g, mx);
        case EPI_METRIC: return launch_sup<T, NS, EPI_METRIC, 0, true, SM_JAC>(c, f, g, mx);
        default: return launch_sup<T, NS, EPI_NONE, 0, true, SM_JAC>(c, f, g, mx);
      }
    } else {
      if (epi == EPI_MIXED) return launch_sup<T, NS, EPI_MIXED, 0, true, SM_JAC>(c, f, g, mx);
      return launch_sup<T, NS, EPI_NONE, 0, true, SM_JAC>(c, f, g, mx);
    }
  }
  if constexpr (sizeof(T) == 8) {
    switch (epi) {
      case EPI_RESNORM:
        return rz ? launch_sup<T, NS, EPI_RESNORM, 1>(c, f, g, mx) : launch_sup<T, NS, EPI_RESNORM, 0>(c, f, g, mx);
      case EPI_METRIC: return launch_sup<T, NS, EPI_METRIC, 0>(c, f, g, mx);
      default: return rz ? launch_sup<T, NS, EPI_NONE, 1>(c, f, g, mx) : launch_sup<T, NS, EPI_NONE, 0>(c, f, g, mx);
    }
  } else {
    if (epi == EPI_MIXED) {
      if (nd == 1 && rz && zi) return launch_sup<T, NS, EPI_MIXED, 1, true, SM_GS, kPf, 1>(c, f, g, mx);
      if (nd == 1 && !rz) return launch_sup<T, NS, EPI_MIXED, 0, true, SM_GS, kPf, 1>(c, f, g, mx);
      if (rz && !zi) return launch_sup<T, NS, EPI_MIXED, 1, false>(c, f, g, mx);
      return rz ? launch_sup<T, NS, EPI_MIXED, 1>(c, f, g, mx) : launch_sup<T, NS, EPI_MIXED, 0>(c, f, g, mx);
    }
    const size_t n = (size_t)f.nx * (size_t)f.ny;
    if (n >= kDeepPfMin && n < kDeepPfEnd)
      return rz ? launch_sup<T, NS, EPI_NONE, 1, true, SM_GS, 3>(c, f, g, mx)
                : launch_sup<T, NS, EPI_NONE, 0, true, SM_GS, 3>(c, f, g, mx);
    return rz ? launch_sup<T, NS, EPI_NONE, 1>(c, f, g, mx) : launch_sup<T, NS, EPI_NONE, 0>(c, f, g, mx);
  }
}

// Level 0 between the two FP32 cycles of a mixed step, one launch: the first
// cycle's ascent (N1 pre-sweeps recomputed from zero, prolongation, NS
// post-sweeps) continues into the second cycle's descent (N1 sweeps on e,
// their result -> z.u2 for the second cycle's ascent, residual, restriction
// into level 1's right-hand side).
template <int N1, int NS> void launch_up_down0_k(Ctx *c, Lvl<float> &f, Lvl<float> &g) {
  constexpr int NB = N1, H = s_up_halo(N1 + NS + NB, 4), P = s_period(N1 + NS + NB);
  using K = void (*)(LevelView<float>, const float *, int, int, int, double *, MixView);
  const K k = f.eq ? (K)k_sup<float, NS, 4, true, N1, true, NB> : (K)k_sup<float, NS, 4, false, N1, true, NB>;
  LevelView<float> a = sview(c, (const void *)k, f, f.u, f.u2, H, P);  // zero start: f.u is not read
  a.rc = g.rhs;
  hipLaunchKernelGGL(k, dim3(8 * a.chunk), dim3(64 * kWaves), 0, c->st, a, g.u, g.nx, g.ny, g.pitch, c->part, MixView{});
}

// (the host runs it only for 2 nu1 + NS <= 7)
template <int NS> void launch_up_down0_ns(Ctx *c, Lvl<float> &f, Lvl<float> &g, int nu1) {
  if (nu1 == 1) launch_up_down0_k<1, NS>(c, f, g);
  if constexpr (4 + NS <= 7) {
    if (nu1 == 2) launch_up_down0_k<2, NS>(c, f, g);
  }
  if constexpr (6 + NS <= 7) {
    if (nu1 == 3) launch_up_down0_k<3, NS>(c, f, g);
  }
}

// (prefetch distance 3: measured on MI355X at 8193^2, 333 -> 324 us)
template <int NS> int launch_down_r64_ns(Ctx *c, Lvl<float> &f, Lvl<float> &g, const MixView &mx) {
  constexpr int PF = 3, H = s_down_halo(NS), P = s_period(NS, SM_GS, PF);
  using K = void (*)(LevelView<float>, float *, int, int, int, MixView, double *);
  const K k = f.eq ? (K)k_sdown_r64<NS, true, PF> : (K)k_sdown_r64<NS, false, PF>;
  const LevelView<float> a = sview(c, (const void *)k, f, f.u, f.u2, H, P);
  hipLaunchKernelGGL(k, dim3(8 * a.chunk), dim3(64 * kWaves), 0, c->st, a, g.rhs, g.nx, g.ny, g.pitch, mx, c->part);
  return a.tiles_i * a.tiles_j;
}

}  // namespace vlp2d
