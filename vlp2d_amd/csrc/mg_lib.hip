// libvlp2d_mg: gfx950 multigrid Dirichlet solver behind the C ABI declared in
// include/vlp2d_mg.h.
//
// One handle = one device, one non-blocking HIP stream, one level hierarchy
// resident in HBM.  Large levels run two fused LDS-tiled kernels per V-cycle
// (descent: smoothing + residual + restriction; ascent: prolongation +
// smoothing + epilogue); the small levels run inside one workgroup
// (k_coarse).  The fixed launch sequence of a cycle (or of one mixed-
// precision refinement step) is captured once into a hipGraph and replayed;
// the host reads one 16-byte result per cycle to decide whether to stop.
//
// Reference behaviour mirrored (paths relative to VisualLaplacePoisson2D/):
//   Model/CPU/MultiGridScheme.cs:23-61    hierarchy: nLevels = min(log2 segX,
//                                         log2 segY), h_l = (h*seg)/seg_l
//   Model/CPU/MultiGridScheme.cs:93-115   V-cycle order, snapshot before the
//                                         level-0 post-smoothing
//   Model/CUDA/MultiGridSchemeCU.cs:140-146  doIteration -> update-norm test
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/vlp2d_mg.h"
#include "mg_ctx.hpp"
#include "mg_coarse.hpp"
#include "mg_refine.hpp"
#include "mg_ltile.hpp"
#include "mg_misc_kernels.hpp"
#include "mg_setup.hpp"


using namespace vlp2d;

namespace {

thread_local std::string g_create_error;

const char *kKClassName[KC_COUNT] = {"rb_half",   "residual",     "restrict", "fill",    "prolong",
                                     "copy_interior", "absdiff_max", "resnorm64", "reduce", "update_mixed",
                                     "adjoint",   "down_fused",   "up_fused", "coarse_vcycle", "setup",
                                     "up_lds",    "up_fused_refine", "up_down_fused", "down_fused_resnorm64",
                                     "adjoint_down_fused", "adjoint_up_fused"};

#define HIPCHK(ctx, call)                                                \
  do {                                                                   \
    hipError_t e_ = (call);                                              \
    if (e_ != hipSuccess) {                                              \
      (ctx)->err = std::string(#call) + ": " + hipGetErrorString(e_);    \
      return e_ == hipErrorOutOfMemory ? VLP2D_E_OOM : VLP2D_E_DEVICE;   \
    }                                                                    \
  } while (0)

#define RCHK(call)                 \
  do {                             \
    int r_ = (call);               \
    if (r_ != VLP2D_OK) return r_; \
  } while (0)

int fail(Ctx *c, int code, const std::string &msg) {
  c->err = msg;
  return code;
}

template <typename T> int dalloc(Ctx *c, T **p, size_t n) {
  if (c->arena) {
    // allocation-free calls: carve 256-byte aligned pieces of the static array
    const size_t off = (c->arena_used + 255) / 256 * 256, need = n * sizeof(T);
    if (off + need > c->arena_cap) {
      *p = nullptr;
      c->err = "the grid needs more than the " + std::to_string(c->arena_cap >> 20) +
               " MiB of device scratch the allocation-free entry point holds";
      return VLP2D_E_SIZE;
    }
    *p = reinterpret_cast<T *>(c->arena + off);
    c->arena_used = off + need;
    c->bytes += (long long)need;
    c->peak = std::max(c->peak, c->bytes);
    return VLP2D_OK;
  }
  hipError_t e = hipMalloc((void **)p, n * sizeof(T));
  if (e != hipSuccess) {
    *p = nullptr;
    (void)hipGetLastError();
    c->err = "hipMalloc of " + std::to_string(n * sizeof(T)) + " bytes failed: " + hipGetErrorString(e);
    return e == hipErrorOutOfMemory ? VLP2D_E_OOM : VLP2D_E_DEVICE;
  }
  c->bytes += (long long)(n * sizeof(T));
  c->peak = std::max(c->peak, c->bytes);
  return VLP2D_OK;
}
template <typename T> void dfree(Ctx *c, T *&p, size_t n) {
  if (p) {
    if (!c->arena) (void)hipFree(p);
    c->bytes -= (long long)(n * sizeof(T));
    p = nullptr;
  }
}

inline int round_pitch(int ny) { return (ny + 31) / 32 * 32; }

// ---- launches -----------------------------------------------------------
// Every launch is tagged with its kernel class, the level it works on and its
// algorithmic bytes (compulsory HBM traffic of the operators it performs:
// each array it must read or write streamed once).  Under VLP2D_FLAG_PROFILE
// each launch is bracketed by hipEvents on the handle's stream.
template <typename F> void launch(Ctx *c, int kc, int lvl, double bytes, F &&f) {
  lvl = std::max(0, std::min(lvl, kMaxLv - 1));
  if (c->profile) {
    PendingEv pe{kc, lvl, nullptr, nullptr};
    (void)hipEventCreate(&pe.a);
    (void)hipEventCreate(&pe.b);
    (void)hipEventRecord(pe.a, c->st);
    f();
    (void)hipEventRecord(pe.b, c->st);
    c->pend.push_back(pe);
  } else {
    f();
  }
  c->kc_launch[kc][lvl] += 1;
  c->kc_bytes[kc][lvl] += bytes;
}

void collect_profile(Ctx *c) {
  for (auto &pe : c->pend) {
    float ms = 0.f;
    if (hipEventElapsedTime(&ms, pe.a, pe.b) == hipSuccess) c->kc_ms[pe.kc][pe.lvl] += ms;
    (void)hipEventDestroy(pe.a);
    (void)hipEventDestroy(pe.b);
  }
  c->pend.clear();
}

inline dim3 grid2(int nj, int ni) {
  return dim3((unsigned)std::max(1, (nj + BX - 1) / BX), (unsigned)std::max(1, (ni + BY - 1) / BY));
}
const dim3 kBlock(BX, BY);

// partial-reduction grid: x over j chunks, y capped rows of blocks
inline dim3 grid_red(int ny, int nx) {
  int gx = std::max(1, (ny - 2 + BX - 1) / BX);
  int gy = std::max(1, std::min((nx - 2 + BY - 1) / BY, 64));
  return dim3(gx, gy);
}

template <typename T> double nodes(const Lvl<T> &v) { return (double)v.nx * (double)v.ny; }

template <typename T> void fill(Ctx *c, T *p, size_t n, T v, int lvl = 0) {
  launch(c, KC_FILL, lvl, (double)n * sizeof(T), [&] {
    int blocks = (int)std::min<size_t>((n + 255) / 256, 4096);
    hipLaunchKernelGGL(k_fill<T>, dim3(std::max(blocks, 1)), dim3(256), 0, c->st, p, (long long)n, v);
  });
}

// ---- one operator per launch ------------------------------------------------
// one colour: reads the other colour's u and this colour's f, writes this
// colour's u -> 1.5 arrays
template <typename T> void half_sweep(Ctx *c, Lvl<T> &v, int lvl, int color) {
  int mcount = (v.ny - 2 + 1) / 2;
  dim3 g = grid2(mcount, v.nx - 2);
  launch(c, KC_SMOOTH, lvl, 1.5 * sizeof(T) * nodes(v), [&] {
    if (v.eq) {
      if (v.lap)
        hipLaunchKernelGGL((k_rb_half<T, true, true>), g, kBlock, 0, c->st, v.u, v.rhs, v.nx, v.ny, v.pitch, v.hx2, v.hy2, v.coef, color);
      else
        hipLaunchKernelGGL((k_rb_half<T, true, false>), g, kBlock, 0, c->st, v.u, v.rhs, v.nx, v.ny, v.pitch, v.hx2, v.hy2, v.coef, color);
    } else {
      if (v.lap)
        hipLaunchKernelGGL((k_rb_half<T, false, true>), g, kBlock, 0, c->st, v.u, v.rhs, v.nx, v.ny, v.pitch, v.hx2, v.hy2, v.coef, color);
      else
        hipLaunchKernelGGL((k_rb_half<T, false, false>), g, kBlock, 0, c->st, v.u, v.rhs, v.nx, v.ny, v.pitch, v.hx2, v.hy2, v.coef, color);
    }
  });
}

// Weighted-Jacobi sweep: u -> res (the level's scratch, free while it
// smooths), then the interior back into u.
template <typename T> void jacobi_sweep(Ctx *c, Lvl<T> &v, int lvl) {
  const dim3 g = grid2(v.ny - 2, v.nx - 2);
  const T w = (T)c->o.omega;
  launch(c, KC_SMOOTH, lvl, 3.0 * sizeof(T) * nodes(v), [&] {
    if (v.eq) {
      if (v.lap)
        hipLaunchKernelGGL((k_wjacobi<T, true, true>), g, kBlock, 0, c->st, v.res, v.u, v.rhs, v.nx, v.ny, v.pitch, v.hx2,
                           v.hy2, v.coef, w);
      else
        hipLaunchKernelGGL((k_wjacobi<T, true, false>), g, kBlock, 0, c->st, v.res, v.u, v.rhs, v.nx, v.ny, v.pitch,
                           v.hx2, v.hy2, v.coef, w);
    } else {
      if (v.lap)
        hipLaunchKernelGGL((k_wjacobi<T, false, true>), g, kBlock, 0, c->st, v.res, v.u, v.rhs, v.nx, v.ny, v.pitch,
                           v.hx2, v.hy2, v.coef, w);
      else
        hipLaunchKernelGGL((k_wjacobi<T, false, false>), g, kBlock, 0, c->st, v.res, v.u, v.rhs, v.nx, v.ny, v.pitch,
                           v.hx2, v.hy2, v.coef, w);
    }
  });
  launch(c, KC_COPY, lvl, 2.0 * sizeof(T) * nodes(v), [&] {
    hipLaunchKernelGGL(k_copy_interior<T>, g, kBlock, 0, c->st, v.u, v.res, v.nx, v.ny, v.pitch);
  });
}

template <typename T> void smooth(Ctx *c, Lvl<T> &v, int lvl, int n) {
  for (int s = 0; s < n; ++s) {
    if (c->o.smoother == VLP2D_SMOOTH_WJACOBI) {
      jacobi_sweep(c, v, lvl);
      continue;
    }
    half_sweep(c, v, lvl, 0);
    half_sweep(c, v, lvl, 1);
  }
}

// Residual of level l, restricted into level l + 1's right-hand side; the
// coarse iterate set to zero.
template <typename T> void resid_restrict(Ctx *c, Lvl<T> &f, Lvl<T> &g, int l) {
  const double s = sizeof(T);
  launch(c, KC_RESID, l, 3 * s * nodes(f), [&] {
    hipLaunchKernelGGL(k_residual<T>, grid2(f.ny - 2, f.nx - 2), kBlock, 0, c->st, f.res, f.rhs, f.u, f.nx, f.ny, f.pitch, f.hx2, f.hy2);
  });
  launch(c, KC_RESTRICT, l, s * (nodes(f) + nodes(g)), [&] {
    if (c->o.restriction == VLP2D_RESTRICT_FULL)
      hipLaunchKernelGGL(k_restrict_fw<T>, grid2(g.ny - 2, g.nx - 2), kBlock, 0, c->st, g.rhs, f.res, g.nx, g.ny, g.pitch, f.pitch);
    else
      hipLaunchKernelGGL(k_restrict<T>, grid2(g.ny - 2, g.nx - 2), kBlock, 0, c->st, g.rhs, f.res, g.nx, g.ny, g.pitch, f.pitch);
  });
  fill(c, g.u, g.elems(), T(0), l + 1);
}

// Additive prolongation of level l + 1's iterate into level l.
template <typename T> void prolong(Ctx *c, Lvl<T> &f, Lvl<T> &g, int l) {
  launch(c, KC_PROLONG, l, (double)sizeof(T) * (2 * nodes(f) + nodes(g)), [&] {
    hipLaunchKernelGGL(k_prolong<T>, grid2(g.ny - 1, g.nx - 1), kBlock, 0, c->st, f.u, g.u, g.nx, g.ny, f.pitch, g.pitch);
  });
}

// V-cycle over levels from..L-1, one operator per launch
// (MultiGridScheme.cs:93-115, startLevel = from).  If snap is given (from ==
// 0) the level-0 iterate is copied there before post-smoothing.  zero0: level
// 0 starts from u = 0 (correction cycle of a mixed step).
template <typename T> void vcycle_simple(Ctx *c, Hier<T> &H, bool zero0, double *snap, int from = 0) {
  const int L = (int)H.lv.size();
  const double s = sizeof(T);
  if (zero0) fill(c, H.lv[0].u, H.lv[0].elems(), T(0));
  for (int l = from; l <= L - 2; ++l) {
    smooth(c, H.lv[l], l, c->o.nu1);
    resid_restrict(c, H.lv[l], H.lv[l + 1], l);
  }
  smooth(c, H.lv[L - 1], L - 1, c->o.nu_coarse);
  for (int l = L - 2; l >= from; --l) {
    Lvl<T> &f = H.lv[l];
    prolong(c, f, H.lv[l + 1], l);
    if (l == 0 && snap) {
      launch(c, KC_COPY, l, 2 * s * nodes(f), [&] {
        hipLaunchKernelGGL(k_copy_interior<T>, grid2(f.ny - 2, f.nx - 2), kBlock, 0, c->st, (T *)(void *)snap, f.u, f.nx, f.ny, f.pitch);
      });
    }
    smooth(c, f, l, c->o.nu2);
  }
}

// FMG iteration (MultiGridScheme.cs:117-132), one operator per launch: the
// level-0 residual restricted down every level without smoothing (each coarse
// iterate is zero, so its residual is its right-hand side), one sweep on the
// coarsest level, then from coarse to fine a prolongation and a V-cycle
// started at that level, and a last V-cycle from level 0 (snapshot taken
// there).
void fmg_simple(Ctx *c, Hier<double> &H, double *snap) {
  const int L = (int)H.lv.size();
  for (int l = 0; l <= L - 2; ++l) resid_restrict(c, H.lv[l], H.lv[l + 1], l);
  smooth(c, H.lv[L - 1], L - 1, 1);
  for (int l = L - 2; l >= 0; --l) {
    prolong(c, H.lv[l], H.lv[l + 1], l);
    vcycle_simple(c, H, false, nullptr, l);
  }
  vcycle_simple(c, H, false, snap, 0);
}


bool is_pow2(int v) { return v > 0 && (v & (v - 1)) == 0; }



// A zero-initialised level pre-smoothed with one sweep is not stored by the
// descent; the ascent recomputes it from the right-hand side (k_sup N1 = 1).
// Not with the update metric, which needs the iterate before post-smoothing.
bool recompute_presmooth(const Ctx *c, bool zi, int epi) {
  // a non-zero start is recomputed only by the refinement ascent, which
  // writes no iterate and so may stage the descent's input in place
  // (Gauss-Seidel only: a Jacobi-smoothed level is stored)
  return (zi || epi == EPI_MIXED) && c->o.nu1 == 1 && epi != EPI_METRIC && c->o.smoother == VLP2D_SMOOTH_RBGS &&
         !(c->o.flags & VLP2D_FLAG_STORE_PRESMOOTH);
}

// Levels below level 0 with at most kLtNodes nodes (1025^2 and smaller) run
// the LDS-tiled ascent k_lt_up: there the streaming ascent is bound by the
// latency of its row pipeline.  Their descent stores the pre-smoothed
// iterate, which k_lt_up stages.
#ifndef VLP2D_LT_NODES
#define VLP2D_LT_NODES 1100000
#endif
constexpr long long kLtNodes = VLP2D_LT_NODES;
template <typename T> bool lds_ascent(const Ctx *c, const Hier<T> &H, int l) {
  const Lvl<T> &v = H.lv[l];
  return l >= 1 && l < H.lc && (long long)v.nx * v.ny <= kLtNodes && c->o.nu2 >= 1 && c->o.nu2 <= 3 &&
         c->o.smoother == VLP2D_SMOOTH_RBGS && !(c->o.flags & VLP2D_FLAG_STREAM_SMALL);
}


template <typename T> void down(Ctx *c, Lvl<T> &f, Lvl<T> &g, int lvl, bool zi, bool store) {
  const double s = sizeof(T);
  const double rw = (zi ? 1.0 : 2.0) + (store ? 1.0 : 0.0);  // [u] + f + [u out]
  launch(c, KC_DOWN, lvl, s * (rw * nodes(f) + nodes(g)), [&] {
    switch (c->o.nu1) {
      case 1: launch_down_ns<T, 1>(c, f, g, zi, store); break;
      case 2: launch_down_ns<T, 2>(c, f, g, zi, store); break;
      default: launch_down_ns<T, 3>(c, f, g, zi, store); break;
    }
  });
}


// returns the number of partials the epilogue writes; nd = 1: the next
// refinement step's first descent folded in (writes g.rhs)
template <typename T>
int up(Ctx *c, Lvl<T> &f, Lvl<T> &g, int lvl, int epi, bool rz, bool zi, const MixView &mx, int nd = 0) {
  const double s = sizeof(T);
  const double rin = rz && zi ? 0.0 : 1.0;  // the staged iterate, unless recomputed from zero
  double bytes = s * ((rin + 2.0) * nodes(f) + nodes(g));
  if (epi == EPI_MIXED) bytes = s * ((rin + 1.0) * nodes(f) + nodes(g)) + 28.0 * nodes(f);
  if (nd) bytes += s * nodes(g);
  int np = 0;
  launch(c, epi == EPI_MIXED ? KC_UP_REFINE : KC_UP, lvl, bytes, [&] {
    switch (c->o.nu2) {
      case 1: np = launch_up_ns<T, 1>(c, f, g, epi, rz, zi, mx, nd); break;
      case 2: np = launch_up_ns<T, 2>(c, f, g, epi, rz, zi, mx, nd); break;
      default: np = launch_up_ns<T, 3>(c, f, g, epi, rz, zi, mx, nd); break;
    }
  });
  return np;
}

template <typename T, int NS> void launch_lt_up_ns(Ctx *c, Lvl<T> &f, const Lvl<T> &g) {
  constexpr int H = lt_halo(NS);
  const LtGrid lg = lt_grid(f.nx, f.ny, H);
  const LevelView<T> a = view(f, f.u2, f.u, H);
  using K = void (*)(LevelView<T>, const T *, int, int, int, int);
  const K k = f.eq ? (K)k_lt_up<T, NS, true> : (K)k_lt_up<T, NS, false>;
  hipLaunchKernelGGL(k, dim3(lg.tiles), dim3(64 * kLtWaves), 0, c->st, a, g.u, g.nx, g.ny, g.pitch, lg.tiles_j);
}

// LDS-tiled ascent (prolongation + nu2 sweeps) of a zero-initialised level
// whose descent stored the pre-smoothed iterate in u2: u2 -> u.
template <typename T> void up_lds(Ctx *c, Lvl<T> &f, Lvl<T> &g, int lvl) {
  launch(c, KC_UP_LDS, lvl, sizeof(T) * (3.0 * nodes(f) + nodes(g)), [&] {
    switch (c->o.nu2) {
      case 1: launch_lt_up_ns<T, 1>(c, f, g); break;
      case 2: launch_lt_up_ns<T, 2>(c, f, g); break;
      default: launch_lt_up_ns<T, 3>(c, f, g); break;
    }
  });
}

// Levels f (l) and g (l + 1), both LDS-tiled ascents, in one launch
// (k_lt_up2): g's ascent recomputed per tile of f in LDS, h = level l + 2.
template <typename T, int NS> void launch_lt_up2_ns(Ctx *c, Lvl<T> &f, const Lvl<T> &g, const Lvl<T> &h) {
  constexpr int H = lt_halo(NS);
  const LtGrid lg = lt_grid(f.nx, f.ny, H);
  const LevelView<T> a = view(f, f.u2, f.u, H);
  const LevelView<T> b = view(g, g.u2, g.u, H);
  using K = void (*)(LevelView<T>, LevelView<T>, const T *, int, int, int, int);
  const K k = f.eq ? (K)k_lt_up2<T, NS, true> : (K)k_lt_up2<T, NS, false>;
  hipLaunchKernelGGL(k, dim3(lg.tiles), dim3(64 * kLtWaves), 0, c->st, a, b, h.u, h.nx, h.ny, h.pitch, lg.tiles_j);
}

// Whether the LDS-tiled ascents of levels l and l + 1 run as one launch: FP32
// levels (the FP32 cycles of a mixed solve), both LDS-tiled, the same step
// flags, and not VLP2D_FLAG_NO_LT_PAIR.  Measured on MI355X at 8193^2 (levels
// 3-6 as two pairs): the mixed solve 3.652 -> 3.638 ms; in FP64 the pairs are
// slower (the FP64 solve 6.216 -> 6.256 ms: the recomputed coarse tile and
// 64 KB of LDS per workgroup cost more than the launch they save).
template <typename T> bool lt_pair(const Ctx *c, const Hier<T> &H, int l) {
  static_assert(sizeof(T) == 4, "FP32 levels only");
  if (c->o.flags & VLP2D_FLAG_NO_LT_PAIR) return false;
  if (!lds_ascent(c, H, l) || !lds_ascent(c, H, l + 1)) return false;
  const Lvl<T> &f = H.lv[l], &g = H.lv[l + 1];
  return f.eq == g.eq && f.pw2 == g.pw2;
}

template <typename T> void up_lds2(Ctx *c, Lvl<T> &f, Lvl<T> &g, const Lvl<T> &h, int lvl) {
  launch(c, KC_UP_LDS, lvl, sizeof(T) * (3.0 * nodes(f) + 3.0 * nodes(g) + nodes(h)), [&] {
    switch (c->o.nu2) {
      case 1: launch_lt_up2_ns<T, 1>(c, f, g, h); break;
      case 2: launch_lt_up2_ns<T, 2>(c, f, g, h); break;
      default: launch_lt_up2_ns<T, 3>(c, f, g, h); break;
    }
  });
}

template <typename T> void coarse(Ctx *c, Hier<T> &H, bool zero_first, double *snap) {
  CoarseArgs<T> a;
  std::memset(&a, 0, sizeof(a));
  const int L = (int)H.lv.size();
  double bytes = 0;
  for (int l = H.lc; l < L; ++l) {
    const Lvl<T> &v = H.lv[l];
    a.lv[l] = CoarseLevel<T>{v.u,    v.rhs,  v.nx,         v.ny,          v.pitch,       v.hx2, v.hy2,
                             v.coef, v.ihx2, v.ihy2, v.eq ? 1 : 0, v.lap ? 1 : 0, v.pw2 ? 1 : 0, v.u2};
    bytes += 3.0 * sizeof(T) * nodes(v);
  }
  a.l0 = H.lc;
  a.L = L;
  a.nu1 = c->o.nu1;
  a.nu2 = c->o.nu2;
  a.nuc = c->o.nu_coarse;
  a.zero_first = zero_first ? 1 : 0;
  a.snap = snap;
  a.jac = c->o.smoother == VLP2D_SMOOTH_WJACOBI ? 1 : 0;
  a.w = (T)c->o.omega;
  a.fw = c->o.restriction == VLP2D_RESTRICT_FULL ? 1 : 0;
  // every level in LDS when it fits in one workgroup's share, else in global
  // memory (the working set then sits in L2)
  // (the LDS kernel maps work items with shifts: rows of 2^k + 1 nodes)
  int nx[kMaxLevels], ny[kMaxLevels];
  bool pow2_rows = true;
  for (int l = H.lc; l < L; ++l) {
    nx[l] = H.lv[l].nx, ny[l] = H.lv[l].ny;
    pow2_rows = pow2_rows && ny[l] >= 3 && is_pow2(ny[l] - 1);
  }
  const size_t lds = coarse_lds_bytes(nx, ny, H.lc, L, sizeof(T), a.jac != 0);
  // (the kernel's static level table shares the workgroup's LDS)
  if (!(c->o.flags & VLP2D_FLAG_COARSE_L2) && pow2_rows && lds + sizeof(ClTab<T>) <= (size_t)c->max_lds) {
    if (lds > 65536) (void)hipFuncSetAttribute(reinterpret_cast<const void *>(k_coarse_lds<T>),
                                               hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    launch(c, KC_COARSE, H.lc, bytes,
           [&] { hipLaunchKernelGGL(k_coarse_lds<T>, dim3(1), dim3(CNT), lds, c->st, a); });
    return;
  }
  launch(c, KC_COARSE, H.lc, bytes, [&] { hipLaunchKernelGGL(k_coarse<T>, dim3(1), dim3(CNT), 0, c->st, a); });
}

// V-cycle with fused kernels.  Levels [0, lc) run k_down / k_up (out of place:
// u -> u2 on the way down, u2 -> u on the way up); [lc, L) run in k_coarse on
// u.  The level-0 ascent performs epilogue `epi` (partials into c->part);
// returns the number of partials it writes (0 when lc == 0).  from / to:
// first descent / last ascent level (1: level 0's kernel is fused with the
// neighbouring cycle's, up_down0).
void down_r64(Ctx *c, Lvl<float> &f, Lvl<float> &g, const MixView &mx);

// r64 (FP32 hierarchies, zero0, from == 0): the level-0 descent takes the FP64
// residual of r64->u64in as its right-hand side (down_r64).  stored0 (from ==
// 1): level 0's pre-smoothed iterate is in its u2 (written by up_down0), and
// the level-0 ascent stages it.  recomputed0 (to == 1): level 0's ascent
// (up_down0) recomputes its pre-smoothing, so its descent does not store it.
// nd0 = 1 (epi EPI_MIXED): the level-0 ascent also runs the next refinement
// step's first descent into level 1's right-hand side (step_fold).
template <typename T>
int vcycle_fused(Ctx *c, Hier<T> &H, bool zero0, int epi, const MixView &mx, double *snap, int from = 0, int to = 0,
                 const MixView *r64 = nullptr, bool stored0 = false, bool recomputed0 = false, int nd0 = 0) {
  const int lc = H.lc;
  auto zi = [&](int l) { return l > 0 || zero0; };
  auto ep = [&](int l) { return l == 0 ? epi : (int)EPI_NONE; };
  auto rz = [&](int l) {
    if (l == 0 && (stored0 || recomputed0)) return recomputed0;
    return recompute_presmooth(c, zi(l), ep(l)) && !lds_ascent(c, H, l);
  };
  for (int l = from; l < lc; ++l) {
    if constexpr (sizeof(T) == 4) {
      if (l == 0 && r64) {
        down_r64(c, H.lv[0], H.lv[1], *r64);
        continue;
      }
    }
    down(c, H.lv[l], H.lv[l + 1], l, zi(l), !rz(l));
  }
  coarse(c, H, lc > 0 || zero0, lc == 0 ? snap : nullptr);
  int np = 0;
  for (int l = lc - 1; l >= to; --l) {
    if constexpr (sizeof(T) == 4) {
      if (l - 1 >= to && lt_pair(c, H, l - 1)) {
        // levels l and l - 1 in one launch (level l's result stays in LDS)
        up_lds2(c, H.lv[l - 1], H.lv[l], H.lv[l + 1], l - 1);
        --l;
        continue;
      }
    }
    if (lds_ascent(c, H, l)) {
      up_lds(c, H.lv[l], H.lv[l + 1], l);
      continue;
    }
    const int n = up(c, H.lv[l], H.lv[l + 1], l, ep(l), rz(l), zi(l), mx, l == 0 ? nd0 : 0);
    if (l == 0) np = n;
  }
  return np;
}

// dscal[1] = sum of psum[0..nsum); with pmax also dscal[0] = max of
// pmax[0..nmax); pub: then publish both for the solve loop (publish_wait).
void finish(Ctx *c, const double *pmax, int nmax, const double *psum, int nsum, bool pub, int sum_slot = 1) {
  HostPub *hp = pub && !c->profile && !c->nomap ? c->hpub_dev : nullptr;
  launch(c, KC_REDUCE, 0, 8.0 * (nmax + nsum), [&] {
    if (pmax)
      hipLaunchKernelGGL(k_finish<true>, dim3(1), dim3(RED_THREADS), 0, c->st, pmax, nmax, psum, nsum, c->dscal,
                         sum_slot, c->dseq, hp);
    else
      hipLaunchKernelGGL(k_finish<false>, dim3(1), dim3(RED_THREADS), 0, c->st, pmax, nmax, psum, nsum, c->dscal,
                         sum_slot, c->dseq, hp);
  });
}

// residual norm^2 of level 0 (FP64) -> dscal[slot] (1 unless given);
// optional FP32 copy; pub: published for the solve loop
void resnorm(Ctx *c, Lvl<double> &v, const double *u, float *r32, bool pub = false, int slot = 1) {
  // column-streaming tiles of the refinement kernel without the update
  const int rows = refine_row_blocks(v.nx);
  const int np = refine_strips(v.ny) * rows;
  const int chunk = (np + 7) / 8;
  launch(c, KC_RESNORM, 0, (16.0 + (r32 ? 4.0 : 0.0)) * nodes(v), [&] {
    hipLaunchKernelGGL(k_refine_stream<false>, dim3(8 * chunk), dim3(256), 0, c->st, u, nullptr, nullptr, v.rhs, r32,
                       v.nx, v.ny, v.pitch, (double)v.hx2, (double)v.hy2, (double)v.ihx2, (double)v.ihy2,
                       v.pw2 ? 1 : 0, rows, np, chunk, c->part, nullptr);
  });
  finish(c, nullptr, 0, c->part, np, pub, slot);
}

void reduce_part(Ctx *c, const double *part, int np, int slot, bool mx) {
  launch(c, KC_REDUCE, 0, 8.0 * np, [&] {
    if (mx)
      hipLaunchKernelGGL(k_reduce<true>, dim3(1), dim3(RED_THREADS), 0, c->st, part, np, c->dscal + slot);
    else
      hipLaunchKernelGGL(k_reduce<false>, dim3(1), dim3(RED_THREADS), 0, c->st, part, np, c->dscal + slot);
  });
}

// First FP32 descent of a mixed solve with the FP64 residual computed inside
// (u64 and fn read, r32 written: 20 B per fine node, plus the coarse
// right-hand side); ||r0||^2 -> dscal[2].
void down_r64(Ctx *c, Lvl<float> &f, Lvl<float> &g, const MixView &mx) {
  int np = 0;
  launch(c, KC_DOWN_R64, 0, 20.0 * nodes(f) + 4.0 * nodes(g), [&] {
    switch (c->o.nu1) {
      case 1: np = launch_down_r64_ns<1>(c, f, g, mx); break;
      case 2: np = launch_down_r64_ns<2>(c, f, g, mx); break;
      default: np = launch_down_r64_ns<3>(c, f, g, mx); break;
    }
  });
  reduce_part(c, c->part, np, 2, false);
}

// Whether a mixed step of m FP32 cycles runs level 0 of its two cycles as one
// launch (up_down0, k_sup epilogue 4): the refinement fused into the level-0
// ascent, Gauss-Seidel and half weighting, and the boundary kernel's 2(2 nu1 +
// nu2) half-sweeps within the longest pipeline (2 nu1 + nu2 <= 7).
bool fused_two_cycle(const Ctx *c, const Lvl<double> &P, int m) {
  const Hier<float> &Z = c->h32;
  if (m != 2 || c->simple || Z.lv.empty() || Z.lc == 0) return false;
  const Lvl<float> &z = Z.lv[0];
  return !(c->o.flags & (VLP2D_FLAG_SEPARATE_REFINE | VLP2D_FLAG_STORE_PRESMOOTH | VLP2D_FLAG_NO_CYCLE_FUSION)) &&
         z.pw2 == P.pw2 && z.pitch == P.pitch && c->o.smoother == VLP2D_SMOOTH_RBGS &&
         c->o.restriction == VLP2D_RESTRICT_HALF && c->o.nu2 <= 3 && 2 * c->o.nu1 + c->o.nu2 <= 7;
}

// Whether the first refinement step of a mixed solve computes the initial
// FP64 residual inside its first descent (the descent stores no pre-smoothed
// iterate: one pre-sweep, which the ascent recomputes, or the first cycle of a
// two-cycle step, whose up_down0 recomputes it; the FP32 and FP64 level 0
// sharing pitch and power-of-two steps) instead of a separate residual pass
// before it.
bool fused_first_residual(const Ctx *c, const Lvl<double> &P) {
  const Hier<float> &Z = c->h32;
  if (c->simple || Z.lv.empty() || Z.lc == 0) return false;
  const Lvl<float> &z = Z.lv[0];
  return z.pw2 == P.pw2 && z.pitch == P.pitch && c->o.restriction == VLP2D_RESTRICT_HALF &&
         (recompute_presmooth(c, true, EPI_NONE) || fused_two_cycle(c, P, c->o.max_inner));
}

// Whether a mixed step's refinement ascent also runs the next step's first
// FP32 descent (zero start, one pre-sweep not stored, restriction into level
// 1's right-hand side), so that the next step starts at level 1: fused
// refinement, Gauss-Seidel with one pre-sweep, half weighting, and a next
// step whose level-0 ascent recomputes that pre-sweep (up_down0, or the
// refinement ascent with one recomputed sweep).
bool step_fold(const Ctx *c, const Lvl<double> &P) {
  const Hier<float> &Z = c->h32;
  if (c->simple || Z.lv.empty() || Z.lc == 0) return false;
  const Lvl<float> &z = Z.lv[0];
  return !(c->o.flags & (VLP2D_FLAG_SEPARATE_REFINE | VLP2D_FLAG_NO_STEP_FOLD)) && z.pw2 == P.pw2 &&
         z.pitch == P.pitch && c->o.restriction == VLP2D_RESTRICT_HALF && recompute_presmooth(c, true, EPI_MIXED);
}

// Whether a step of m FP32 cycles folds the next step's first descent into
// its refinement ascent: that ascent starts from zero (m = 1) or stages the
// stored pre-smoothed iterate (the two-cycle step).
bool step_folds_out(const Ctx *c, const Lvl<double> &P, int m) {
  return step_fold(c, P) && (m == 1 || fused_two_cycle(c, P, m));
}

// max |u - snap| over interior nodes -> dscal[0]; pub: published for the
// host (publish_wait)
void maxdiff(Ctx *c, Lvl<double> &v, const double *snap, bool pub) {
  dim3 g = grid_red(v.ny, v.nx);
  int np = (int)(g.x * g.y);
  launch(c, KC_MAXDIFF, 0, 16.0 * nodes(v), [&] {
    hipLaunchKernelGGL(k_absdiff_max<double>, g, kBlock, 0, c->st, v.u, snap, v.nx, v.ny, v.pitch, c->part);
  });
  if (pub) finish(c, c->part, np, nullptr, 0, true);
  else reduce_part(c, c->part, np, 0, true);
}

// FP64 cycle on H; epi: EPI_NONE, EPI_RESNORM (||r||^2 -> dscal[1]) or
// EPI_METRIC (update norm -> dscal[0]); pub: the result is published for the
// host (publish_wait).
void cycle64(Ctx *c, Hier<double> &H, int epi, bool pub = false) {
  Lvl<double> &P = H.lv[0];
  MixView mx{};
  if (c->simple || H.lc == 0) {
    double *snap = epi == EPI_METRIC ? c->snap : nullptr;
    if (c->simple && c->o.cycle == VLP2D_CYCLE_FMG) fmg_simple(c, H, snap);
    else if (c->simple) vcycle_simple(c, H, false, snap);
    else vcycle_fused(c, H, false, EPI_NONE, mx, snap);
    if (epi == EPI_RESNORM) resnorm(c, P, P.u, nullptr, pub);
    if (epi == EPI_METRIC) maxdiff(c, P, c->snap, pub);
    return;
  }
  const int np = vcycle_fused(c, H, false, epi, mx, nullptr);
  if (epi == EPI_RESNORM) finish(c, nullptr, 0, c->part, np, pub);
  if (epi == EPI_METRIC) {
    if (pub) finish(c, c->part, np, nullptr, 0, true);
    else reduce_part(c, c->part, np, 0, true);
  }
}

void up_down0(Ctx *c, Lvl<float> &f, Lvl<float> &g) {
  launch(c, KC_UP_DOWN, 0, 4.0 * (2.0 * nodes(f) + 2.0 * nodes(g)), [&] {
    switch (c->o.nu2) {
      case 1: launch_up_down0_ns<1>(c, f, g, c->o.nu1); break;
      case 2: launch_up_down0_ns<2>(c, f, g, c->o.nu1); break;
      default: launch_up_down0_ns<3>(c, f, g, c->o.nu1); break;
    }
  });
}

// One mixed-precision refinement step on the FP64 problem level P, using h32:
// e = V-cycle(rhs32) from e = 0; u_next = u + e (max|e| -> dscal[0]);
// r32_next = fn - Lap u_next, ||r||^2 -> dscal[1].  u/u2 of P and rhs/rhs2 of
// h32 level 0 are swapped by the caller afterwards.
// m >= 2: the correction equation gets m FP32 V-cycles, the first from
// e = 0, the refinement update follows the last.
// first: the first step of a solve; when fused_first_residual holds, its first
// descent also computes the initial FP64 residual (r32 into z.rhs, ||r0||^2
// into dscal[2]), which the caller then does not compute beforehand.
// fold_in: the previous step's refinement ascent left this step's first
// restriction in level 1's right-hand side (step_folds_out), so the step
// starts at level 1.  fold_out: this step's refinement ascent does the same
// for the next step when step_folds_out(c, P, m) holds.
void mixed_step(Ctx *c, Lvl<double> &P, int m = 1, bool pub = false, bool first = false, bool fold_in = false,
                bool fold_out = false) {
  Hier<float> &Z = c->h32;
  Lvl<float> &z = Z.lv[0];
  MixView none{};
  MixView mx;
  mx.u64in = P.u;
  mx.u64out = P.u2;
  mx.fn64 = P.rhs;
  mx.r32out = z.rhs2;
  mx.hx2 = P.hx2;
  mx.hy2 = P.hy2;
  mx.ihx2 = P.ihx2;
  mx.ihy2 = P.ihy2;
  mx.pw2 = P.pw2 ? 1 : 0;
  mx.part_max = c->part2;
  MixView rv = mx;
  rv.r32out = z.rhs;
  const MixView *r64 = first && !fold_in && fused_first_residual(c, P) ? &rv : nullptr;
  const bool fused_refine = !c->simple && Z.lc > 0 && !(c->o.flags & VLP2D_FLAG_SEPARATE_REFINE) &&
                            z.pw2 == P.pw2 && z.pitch == P.pitch;
  const int from0 = fold_in ? 1 : 0;  // the first cycle's first descent level
  const int nd = fold_out && step_folds_out(c, P, m) ? 1 : 0;
  if (fused_two_cycle(c, P, m)) {
    // first cycle down to level 1 (its level-0 pre-smoothing recomputed by
    // up_down0, not stored), level 0 of both cycles in between in one launch
    // (it stores the second cycle's pre-smoothed iterate in z.u2), second
    // cycle from level 1 with the refinement ascent staging it
    vcycle_fused(c, Z, true, EPI_NONE, none, nullptr, from0, 1, r64, false, true);
    up_down0(c, z, Z.lv[1]);
    const int np = vcycle_fused(c, Z, false, EPI_MIXED, mx, nullptr, 1, 0, nullptr, true, false, nd);
    finish(c, c->part2, np, c->part, np, pub);
    return;
  }
  // all cycles but the last (correction e -> z.u)
  for (int k = 0; k + 1 < m; ++k) {
    if (c->simple) vcycle_simple(c, Z, k == 0, nullptr);
    else vcycle_fused(c, Z, k == 0, EPI_NONE, none, nullptr, k == 0 ? from0 : 0, 0, k == 0 ? r64 : nullptr);
  }
  const bool zero_last = m == 1;
  if (c->simple || Z.lc == 0) {
    if (c->simple) vcycle_simple(c, Z, zero_last, nullptr);
    else vcycle_fused(c, Z, zero_last, EPI_NONE, none, nullptr);
    dim3 g = grid_red(P.ny, P.nx);
    int np = (int)(g.x * g.y);
    launch(c, KC_UPDATE, 0, 20.0 * nodes(P), [&] {
      hipLaunchKernelGGL(k_update_oop, g, kBlock, 0, c->st, P.u, P.u2, z.u, P.nx, P.ny, P.pitch, c->part);
    });
    reduce_part(c, c->part, np, 0, true);
    resnorm(c, P, P.u2, z.rhs2, pub);
    return;
  }
  if (fused_refine) {
    // refinement update inside the level-0 ascent of the FP32 cycle (the
    // kernel evaluates both precisions' divisions the same way, as
    // multiplications when the steps are powers of two)
    const int np = vcycle_fused(c, Z, zero_last, EPI_MIXED, mx, nullptr, m == 1 ? from0 : 0, 0, m == 1 ? r64 : nullptr,
                                false, false, nd);
    finish(c, c->part2, np, c->part, np, pub);
    return;
  }
  vcycle_fused(c, Z, zero_last, EPI_NONE, mx, nullptr, 0, 0, m == 1 ? r64 : nullptr);  // correction e -> z.u
  const int rows = refine_row_blocks(P.nx);
  const int np = refine_strips(P.ny) * rows;
  const int chunk = (np + 7) / 8;
  launch(c, KC_UPDATE, 0, 32.0 * nodes(P), [&] {
    hipLaunchKernelGGL(k_refine_stream<true>, dim3(8 * chunk), dim3(256), 0, c->st, P.u, P.u2, z.u, P.rhs, z.rhs2, P.nx,
                       P.ny, P.pitch, (double)P.hx2, (double)P.hy2, (double)P.ihx2, (double)P.ihy2, P.pw2 ? 1 : 0, rows,
                       np, chunk, c->part, c->part2);
  });
  finish(c, c->part2, np, c->part, np, pub);
}

void swap_mixed(Ctx *c, Lvl<double> &P) {
  Lvl<float> &z = c->h32.lv[0];
  std::swap(P.u, P.u2);
  P.par ^= 1;
  std::swap(z.rhs, z.rhs2);
  z.rpar ^= 1;
}

// ---- graph cache ---------------------------------------------------------
constexpr int kPubKey = 1 << 20;  // added to the keys of the solve loop's programs
// Programs of the solve loop (key >= kPubKey) launch their kernels directly
// unless VLP2D_FLAG_GRAPH_SOLVE: the loop keeps one step launched ahead, and
// graph replay measured slower there (a graph's launch stalls the queue in
// the middle of a step); single cycles replay their captured graph.
template <typename F> int run_program(Ctx *c, int key, F &&body) {
  bool use_graph = !(c->o.flags & VLP2D_FLAG_NO_GRAPH) && !c->profile && !c->nomap;
  if (key >= kPubKey && !(c->o.flags & VLP2D_FLAG_GRAPH_SOLVE)) use_graph = false;
  if (!use_graph) {
    body();
    HIPCHK(c, hipGetLastError());
    return VLP2D_OK;
  }
  auto it = c->graphs.find(key);
  if (it == c->graphs.end()) {
    HIPCHK(c, hipStreamBeginCapture(c->st, hipStreamCaptureModeThreadLocal));
    body();
    hipGraph_t g = nullptr;
    hipError_t e = hipStreamEndCapture(c->st, &g);
    if (e != hipSuccess) return fail(c, VLP2D_E_DEVICE, std::string("hipStreamEndCapture: ") + hipGetErrorString(e));
    hipGraphExec_t x = nullptr;
    HIPCHK(c, hipGraphInstantiate(&x, g, nullptr, nullptr, 0));
    it = c->graphs.emplace(key, std::make_pair(g, x)).first;
  }
  HIPCHK(c, hipGraphLaunch(it->second.second, c->st));
  return VLP2D_OK;
}

// graph keys: program id * 4 + swap parities of the buffers it captured
enum Prog { PG_CYCLE_METRIC = 0, PG_CYCLE_NORM = 1, PG_MIXED = 2, PG_ADJ_CYCLE_NORM = 3, PG_ADJ_MIXED = 4 };
constexpr int kProgFirst = 8;  // or-ed into a mixed program: the first step of a solve
constexpr int kProgFold = 16;     // or-ed into a mixed program: the step starts at level 1 (fold_in)
constexpr int kProgFoldOut = 32;  // or-ed into a mixed program: the step folds the next one's first descent
int gkey(int prog, int par, int rpar, int m = 1) { return ((prog * 16 + m) * 2 + par) * 2 + rpar; }

int sync_read(Ctx *c, int nslots) {
  HIPCHK(c, hipMemcpyAsync(c->hscal, c->dscal, sizeof(double) * nslots, hipMemcpyDeviceToHost, c->st));
  HIPCHK(c, hipStreamSynchronize(c->st));
  if (c->profile) collect_profile(c);
  return VLP2D_OK;
}

// dscal[0..1] -> hscal[0..1] as the last kernel of a solve-loop program
// (k_finish with pub) published them: the host spins on the sequence number
// instead of synchronising the stream (profiling runs, which publish
// nothing, keep sync_read).
// Wait for published result `want` and copy its dscal[0..2] into hscal.  At
// most two results are in flight, so ring slot want % kPubRing holds result
// want or an older one.
int wait_result(Ctx *c, unsigned long long want) {
  HostPub *slot = c->hpub + (want % kPubRing);
  for (unsigned long long spin = 1;; ++spin) {
    if (__atomic_load_n(&slot->seq, __ATOMIC_ACQUIRE) >= want) break;
    if ((spin & 4095) == 0) {
      // the stream drained without the result: report, do not spin forever
      const hipError_t q = hipStreamQuery(c->st);
      if (q == hipSuccess) {
        if (__atomic_load_n(&slot->seq, __ATOMIC_ACQUIRE) >= want) break;
        return fail(c, VLP2D_E_DEVICE, "published cycle result did not arrive");
      }
      if (q != hipErrorNotReady) HIPCHK(c, q);
    }
    __builtin_ia32_pause();
  }
  for (int k = 0; k < 3; ++k) c->hscal[k] = slot->v[k];
  return VLP2D_OK;
}

// The result of the program just launched (profiling runs, which publish
// nothing, synchronise instead).
int publish_wait(Ctx *c) {
  if (c->profile || c->nomap) return sync_read(c, 3);
  return wait_result(c, ++c->pub_seq);
}

// ---- hierarchy construction ------------------------------------------------
// Steps follow MultiGridScheme.cs:44-59: lng = h*seg, h_l = lng/seg_l in the
// level's arithmetic type.
template <typename T> void level_params(Lvl<T> &v, T hx0, T hy0, int segX, int segY, int sx, int sy) {
  T lngX = hx0 * (T)segX, lngY = hy0 * (T)segY;
  T hx = lngX / (T)sx, hy = lngY / (T)sy;
  v.hx2 = hx * hx;
  v.hy2 = hy * hy;
  v.coef = T(1) / (T(2) / v.hx2 + T(2) / v.hy2);
  T d = hx - hy;
  v.eq = (d < T(0) ? -d : d) < (T)1e-15;
  int ex = 0, ey = 0;
  v.pw2 = std::frexp((double)v.hx2, &ex) == 0.5 && std::frexp((double)v.hy2, &ey) == 0.5;
  v.ihx2 = T(1) / v.hx2;
  v.ihy2 = T(1) / v.hy2;
  v.ix2 = 1.0 / (double)v.hx2;
  v.iy2 = 1.0 / (double)v.hy2;
  if (v.eq) {
    v.ax = 0.25;
    v.ay = 0.25;
    v.b = 0.25 * (double)v.hx2;
  } else {
    v.ax = (double)v.coef / (double)v.hx2;
    v.ay = (double)v.coef / (double)v.hy2;
    v.b = (double)v.coef;
  }
}

// Levels at or below this size run in the one-workgroup coarse kernel; larger
// ones in the streaming kernels (measured on MI355X at 8193^2: with 65^2 the
// coarse cycle takes 41-44 us against 46-48 us for it at 33^2 plus the two
// streaming launches of the 65^2 level).
int coarse_threshold(const Ctx *c) {
  if (c->o.coarse_nodes > 0) return c->o.coarse_nodes;
  return 65 * 65;
}

template <typename T> void shape_hier(Ctx *c, Hier<T> &H, T hx0, T hy0) {
  H.lv.resize(c->L);
  int sx = c->segX, sy = c->segY;
  const int thr = coarse_threshold(c);
  H.lc = c->L - 1;
  bool found = false;
  for (int l = 0; l < c->L; ++l) {
    Lvl<T> &v = H.lv[l];
    v.nx = sx + 1;
    v.ny = sy + 1;
    v.pitch = round_pitch(v.ny);
    level_params(v, hx0, hy0, c->segX, c->segY, sx, sy);
    if (!found && c->o.coarse_nodes >= 0 && (long long)v.nx * v.ny <= thr) {
      H.lc = l;
      found = true;
    }
    sx /= 2;
    sy /= 2;
  }
}

// Allocate every buffer the cycles of H need; level-0 u and rhs already exist.
// All buffers are allocated before any is cleared, so a failed allocation
// returns before anything runs on the device.
template <typename T> int complete_hier(Ctx *c, Hier<T> &H) {
  if (H.full) return VLP2D_OK;
  struct New {
    T *p;
    size_t n;
    bool copy0;  // level-0 second iterate buffer: a copy of u (boundary values)
  };
  std::vector<New> made;
  auto need = [&](T *&p, size_t n, bool copy0 = false) -> int {
    if (p) return VLP2D_OK;
    RCHK(dalloc(c, &p, n));
    made.push_back(New{p, n, copy0});
    return VLP2D_OK;
  };
  for (int l = 0; l < c->L; ++l) {
    Lvl<T> &v = H.lv[l];
    RCHK(need(v.u, v.elems()));
    RCHK(need(v.rhs, v.elems()));
    // (weighted Jacobi on the coarse levels: k_coarse's second iterate buffer)
    if (l == 0 || l < H.lc || c->o.smoother == VLP2D_SMOOTH_WJACOBI) RCHK(need(v.u2, v.elems(), l == 0));
    if (c->simple && (l < c->L - 1 || c->o.smoother == VLP2D_SMOOTH_WJACOBI)) RCHK(need(v.res, v.elems()));
  }
  for (const New &m : made) {
    if (m.copy0) HIPCHK(c, hipMemcpyAsync(m.p, H.lv[0].u, m.n * sizeof(T), hipMemcpyDeviceToDevice, c->st));
    else fill(c, m.p, m.n, T(0));
  }
  HIPCHK(c, hipStreamSynchronize(c->st));
  H.full = true;
  return VLP2D_OK;
}

template <typename T> void free_hier(Ctx *c, Hier<T> &H) {
  for (auto &v : H.lv) {
    dfree(c, v.u, v.elems());
    dfree(c, v.u2, v.elems());
    dfree(c, v.rhs, v.elems());
    dfree(c, v.rhs2, v.elems());
    dfree(c, v.res, v.elems());
  }
  H.lv.clear();
  H.full = false;
}

int ensure_h32(Ctx *c) {
  if (!c->h32.lv.empty() && c->h32.full) return VLP2D_OK;
  shape_hier(c, c->h32, (float)c->hx, (float)c->hy);
  Lvl<float> &z = c->h32.lv[0];
  z.lap = false;
  RCHK(dalloc(c, &z.rhs2, z.elems()));
  RCHK(complete_hier(c, c->h32));
  fill(c, z.rhs2, z.elems(), 0.0f);
  return VLP2D_OK;
}

int ensure_aux(Ctx *c) {
  if (c->snap) return VLP2D_OK;
  Lvl<double> &v = c->h64.lv[0];
  RCHK(dalloc(c, &c->snap, v.elems()));
  fill(c, c->snap, v.elems(), 0.0);
  return VLP2D_OK;
}

void invalidate_graphs(Ctx *c) {
  for (auto &kv : c->graphs) {
    if (kv.second.second) (void)hipGraphExecDestroy(kv.second.second);
    if (kv.second.first) (void)hipGraphDestroy(kv.second.first);
  }
  c->graphs.clear();
}

int check_finite(const double *a, size_t n) {
  for (size_t k = 0; k < n; ++k)
    if (!std::isfinite(a[k])) return 0;
  return 1;
}

int upload_grid(Ctx *c, double *dst, int pitch, const double *src) {
  HIPCHK(c, hipMemcpy2DAsync(dst, (size_t)pitch * sizeof(double), src, (size_t)c->ny * sizeof(double),
                             (size_t)c->ny * sizeof(double), (size_t)c->nx, hipMemcpyHostToDevice, c->st));
  HIPCHK(c, hipStreamSynchronize(c->st));
  return VLP2D_OK;
}
int download_grid(Ctx *c, double *dst, const double *src, int pitch) {
  HIPCHK(c, hipMemcpy2DAsync(dst, (size_t)c->ny * sizeof(double), src, (size_t)pitch * sizeof(double),
                             (size_t)c->ny * sizeof(double), (size_t)c->nx, hipMemcpyDeviceToHost, c->st));
  HIPCHK(c, hipStreamSynchronize(c->st));
  return VLP2D_OK;
}

// Both iterate buffers of level 0 must carry the boundary values.
int mirror_level0(Ctx *c, Lvl<double> &P) {
  if (P.u2) HIPCHK(c, hipMemcpyAsync(P.u2, P.u, P.elems() * sizeof(double), hipMemcpyDeviceToDevice, c->st));
  HIPCHK(c, hipStreamSynchronize(c->st));
  return VLP2D_OK;
}

// FP32 V-cycles of the next refinement step: the fewest that the contraction
// per cycle of the last step, rho = (rel / rel_prev)^(1/m_prev), predicts
// bring rel to rtol, at most max_inner (all max_inner when rho >= 1).
int inner_cycles(double rel, double rel_prev, int m_prev, int max_inner, double rtol) {
  if (max_inner <= 1) return 1;
  const double rho = std::pow(rel / rel_prev, 1.0 / m_prev);
  if (!(rho < 1.0)) return max_inner;
  int m = 1;
  for (double pr = rel * rho; m < max_inner && pr > rtol; pr *= rho) ++m;
  return m;
}

// Whether the refinement step after the one in flight is certain to be
// needed and to run max_inner cycles, so that it may be launched before the
// result of the one in flight is known: the predicted residual after the one
// in flight, times rho^(max_inner - 1), stays 100 x above rtol.  rho is the
// contraction per cycle of the last accepted step (0.2 before any).
bool speculate_next(double rel, double rel_prev, int m_prev, bool have_rate, int m_flight, int max_inner,
                    double rtol) {
  const double rho = have_rate ? std::pow(rel / rel_prev, 1.0 / m_prev) : 0.2;
  if (!(rho < 1.0)) return true;
  return rel * std::pow(rho, m_flight + max_inner - 1) > 100.0 * rtol;
}

// Whether a step of m cycles launched after m_before cycles still in flight is
// predicted to reach rtol (the last step of the solve), from the contraction
// per cycle of the last accepted step: such a step does not fold the next
// step's first descent into its refinement ascent.  Only a prediction: when
// another step turns out to be needed, it starts at level 0.
bool predicted_last(double rel, double rel_prev, int m_prev, bool have_rate, int m_before, int m, double rtol) {
  if (!have_rate) return false;
  const double rho = std::pow(rel / rel_prev, 1.0 / m_prev);
  return rho < 1.0 && rel * std::pow(rho, m_before + m) <= rtol;
}

// Mixed solve: refinement steps are launched one ahead of the host while the
// next one is certain to be needed (speculate_next), so the host's wait for a
// step's result overlaps the next step.  A step writes the iterate and the
// FP32 residual into the other buffer of each pair, so a launched step that
// turns out not to be needed (rtol reached) or to run the wrong number of
// cycles (the schedule of inner_cycles decides from the actual residual) is
// dropped by restoring the buffer parity; the accepted steps are exactly the
// ones the sequential loop runs.
int solve_mixed(Ctx *c, Lvl<double> &P, int prog_mixed, double rtol, int max_cycles, int &k, int &vcycles,
                double &r0, double &rel) {
  const int M = c->o.max_inner;
  // ||r0||^2 -> dscal[2], published with step 0: computed by step 0's first
  // descent, or by a residual pass before it
  const bool fuse_r0 = fused_first_residual(c, P);
  if (!fuse_r0) resnorm(c, P, P.u, c->h32.lv[0].rhs, false, 2);
  HIPCHK(c, hipGetLastError());
  struct Step {
    int m;
    unsigned long long seq;
  };
  Step fl[2] = {};
  int nfl = 0;
  // level 1's right-hand side holds the next step's first restriction (the
  // last launched step folded it in); a dropped step may have overwritten
  // it, so its relaunch starts at level 0
  bool fold_ready = false;
  // (rel, rel_prev, m_prev: the loop's state below, read when a step is launched)
  double rel_prev = 1.0;
  int m_prev = 1;
  auto launch_step = [&](int m, bool first) -> int {
    first = first && fuse_r0;
    const bool fin = fold_ready;
    const bool fout = step_folds_out(c, P, m) &&
                      !predicted_last(rel, rel_prev, m_prev, k > 0, nfl ? fl[0].m : 0, m, rtol);
    const int prog = (first ? prog_mixed | kProgFirst : prog_mixed) | (fin ? kProgFold : 0) | (fout ? kProgFoldOut : 0);
    RCHK(run_program(c, gkey(prog, P.par, c->h32.lv[0].rpar, m) + kPubKey,
                     [&] { mixed_step(c, P, m, true, first, fin, fout); }));
    fold_ready = fout;
    swap_mixed(c, P);
    fl[nfl++] = Step{m, ++c->pub_seq};
    return VLP2D_OK;
  };
  auto drop = [&](int n) {
    for (int q = 0; q < n; ++q) swap_mixed(c, P);
    fold_ready = false;
  };
  k = 0;
  vcycles = 0;
  rel = 1.0;
  r0 = NAN;
  for (;;) {
    if (c->cancel.load()) {
      drop(nfl);
      return fail(c, VLP2D_E_CANCELLED, "cancelled");
    }
    if (nfl == 0) {
      if (k >= max_cycles) return fail(c, VLP2D_E_NOTCONV, "max_cycles reached before rtol");
      RCHK(launch_step(k == 0 ? M : inner_cycles(rel, rel_prev, m_prev, M, rtol), k == 0));
    }
    if (nfl == 1 && k + 1 < max_cycles && speculate_next(rel, rel_prev, m_prev, k > 0, fl[0].m, M, rtol))
      RCHK(launch_step(M, false));
    RCHK(wait_result(c, fl[0].seq));
    const Step s0 = fl[0];
    fl[0] = fl[1];
    --nfl;
    if (k == 0) {
      r0 = std::sqrt(c->hscal[2]);
      if (!std::isfinite(r0)) {
        drop(1 + nfl);
        return fail(c, VLP2D_E_NONFINITE, "initial residual is not finite");
      }
      if (!(r0 > 0)) {  // already the discrete solution: no step
        drop(1 + nfl);
        rel = 0.0;
        return VLP2D_OK;
      }
    }
    if (s0.m != (k == 0 ? M : inner_cycles(rel, rel_prev, m_prev, M, rtol))) {
      drop(1 + nfl);  // launched with the wrong cycle count: relaunch
      nfl = 0;
      continue;
    }
    ++k;
    vcycles += s0.m;
    m_prev = s0.m;
    rel_prev = rel;
    rel = std::sqrt(c->hscal[1]) / r0;
    if (!std::isfinite(rel)) {
      drop(nfl);
      return fail(c, VLP2D_E_NONFINITE, "residual became non-finite");
    }
    if (rel <= rtol) {
      drop(nfl);
      return VLP2D_OK;
    }
  }
}

// Solve from the current level-0 iterate of H until ||r|| <= rtol ||r0||.
int solve_impl(Ctx *c, Hier<double> &H, int prog_fp64, int prog_mixed, double rtol, int max_cycles,
               vlp2d_mg_stats *st) {
  Lvl<double> &P = H.lv[0];
  const bool mixed = c->o.precision == VLP2D_PREC_MIXED;
  if (mixed) RCHK(ensure_h32(c));
  else RCHK(complete_hier(c, H));  // mixed steps use only level 0 of H
  // the handle's two timing events, created on its first solve
  for (hipEvent_t &e : c->ev)
    if (!e) HIPCHK(c, hipEventCreate(&e));
  hipEvent_t e0 = c->ev[0], e1 = c->ev[1];
  HIPCHK(c, hipEventRecord(e0, c->st));
  double r0 = 0.0, rel = 0.0;
  int k = 0, vcycles = 0;
  int rc = VLP2D_OK;
  if (mixed && !c->profile && !c->nomap && max_cycles >= 1) {
    c->r32_valid = &H == &c->h64;
    rc = solve_mixed(c, P, prog_mixed, rtol, max_cycles, k, vcycles, r0, rel);
  } else {
    // a mixed solve's first step may compute ||r0|| itself (as solve_mixed)
    const bool fuse_r0 = mixed && max_cycles >= 1 && fused_first_residual(c, P);
    c->r32_valid = mixed && (&H == &c->h64);
    if (fuse_r0) {
      r0 = 1.0;  // set from the first step's result
    } else {
      resnorm(c, P, P.u, mixed ? c->h32.lv[0].rhs : nullptr, true);
      HIPCHK(c, hipGetLastError());
      RCHK(publish_wait(c));
      r0 = std::sqrt(c->hscal[1]);
      if (!std::isfinite(r0)) rc = fail(c, VLP2D_E_NONFINITE, "initial residual is not finite");
    }
    rel = r0 > 0 ? 1.0 : 0.0;
    double rel_prev = 1.0;
    int m_prev = 1;
    bool fold_ready = false;  // as in solve_mixed
    while (rc == VLP2D_OK && r0 > 0 && rel > rtol) {
      if (k >= max_cycles) {
        rc = fail(c, VLP2D_E_NOTCONV, "max_cycles reached before rtol");
        break;
      }
      if (c->cancel.load()) {
        rc = fail(c, VLP2D_E_CANCELLED, "cancelled");
        break;
      }
      const bool first = fuse_r0 && k == 0;
      if (mixed) {
        const int m = k == 0 ? c->o.max_inner : inner_cycles(rel, rel_prev, m_prev, c->o.max_inner, rtol);
        const bool fin = fold_ready;
        const bool fout = step_folds_out(c, P, m) && !predicted_last(rel, rel_prev, m_prev, k > 0, 0, m, rtol);
        const int prog =
            (first ? prog_mixed | kProgFirst : prog_mixed) | (fin ? kProgFold : 0) | (fout ? kProgFoldOut : 0);
        RCHK(run_program(c, gkey(prog, P.par, c->h32.lv[0].rpar, m) + kPubKey,
                         [&] { mixed_step(c, P, m, true, first, fin, fout); }));
        fold_ready = fout;
        swap_mixed(c, P);
        vcycles += m;
        m_prev = m;
      } else {
        RCHK(run_program(c, gkey(prog_fp64, P.par, 0) + kPubKey, [&] { cycle64(c, H, EPI_RESNORM, true); }));
        vcycles += c->o.cycle == VLP2D_CYCLE_FMG ? 2 : 1;
      }
      RCHK(publish_wait(c));
      if (first) {
        r0 = std::sqrt(c->hscal[2]);
        if (!std::isfinite(r0)) {
          rc = fail(c, VLP2D_E_NONFINITE, "initial residual is not finite");
          break;
        }
        if (!(r0 > 0)) {  // already the discrete solution: the step is undone
          swap_mixed(c, P);
          vcycles = 0;
          rel = 0.0;
          break;
        }
      }
      ++k;
      rel_prev = rel;
      rel = std::sqrt(c->hscal[1]) / r0;
      if (!std::isfinite(rel)) rc = fail(c, VLP2D_E_NONFINITE, "residual became non-finite");
    }
  }
  if (rc != VLP2D_OK && rc != VLP2D_E_NOTCONV && rc != VLP2D_E_CANCELLED && rc != VLP2D_E_NONFINITE) return rc;
  HIPCHK(c, hipEventRecord(e1, c->st));
  HIPCHK(c, hipEventSynchronize(e1));
  float ms = 0.f;
  HIPCHK(c, hipEventElapsedTime(&ms, e0, e1));
  if (st) {
    st->cycles = k;
    st->vcycles = vcycles;
    st->levels = c->L;
    st->r0 = r0;
    st->rel_residual = rel;
    st->seconds = ms * 1e-3;
    double unknowns = (double)(c->nx - 2) * (double)(c->ny - 2);
    st->munknowns_per_s = st->seconds > 0 ? unknowns / st->seconds / 1e6 : 0.0;
  }
  return rc;
}

// ---- transposed cycle -------------------------------------------------------
// Levels above the coarsest run the transposed cycle as two streaming
// kernels (k_tdown, k_tup), ubar alternating between aub and aub2, rbar in
// place in arb; VLP2D_FLAG_SIMPLE and sweep counts outside 1..3 keep one
// operator per launch in place, with resbar in aresb.
bool t_streamed(const Ctx *c, int l) { return !c->simple && l < c->L - 1; }

int ensure_adj_buffers(Ctx *c) {
  if (!c->aub.empty()) return VLP2D_OK;
  c->aub.assign(c->L, nullptr);
  c->arb.assign(c->L, nullptr);
  c->aresb.assign(c->L, nullptr);
  c->aub2.assign(c->L, nullptr);
  for (int l = 0; l < c->L; ++l) {
    size_t n = c->h64.lv[l].elems();
    RCHK(dalloc(c, &c->aub[l], n));
    RCHK(dalloc(c, &c->arb[l], n));
    fill(c, c->aub[l], n, 0.0);
    fill(c, c->arb[l], n, 0.0);
    if (t_streamed(c, l)) {
      RCHK(dalloc(c, &c->aub2[l], n));
      fill(c, c->aub2[l], n, 0.0);
    } else if (l < c->L - 1 || l == 0) {
      // resbar of the one-operator levels; level 0's also stages the outputs
      RCHK(dalloc(c, &c->aresb[l], n));
      fill(c, c->aresb[l], n, 0.0);
    }
  }
  return VLP2D_OK;
}

void half_sweep_T(Ctx *c, int l, int color) {
  Lvl<double> &v = c->h64.lv[l];
  launch(c, KC_ADJ, l, 16.0 * nodes(v), [&] {
    hipLaunchKernelGGL(k_adj_gather, grid2(v.ny, v.nx), kBlock, 0, c->st, c->aub[l], v.nx, v.ny, v.pitch, v.ax, v.ay, color);
  });
  launch(c, KC_ADJ, l, 16.0 * nodes(v), [&] {
    hipLaunchKernelGGL(k_adj_color, grid2(v.ny - 2, v.nx - 2), kBlock, 0, c->st, c->aub[l], c->arb[l], v.nx, v.ny, v.pitch, v.b, color);
  });
}
void smooth_T(Ctx *c, int l, int n) {
  for (int s = 0; s < n; ++s) {
    half_sweep_T(c, l, 1);
    half_sweep_T(c, l, 0);
  }
}
void cycle_T(Ctx *c, int l) {
  if (l == c->L - 1) {
    smooth_T(c, l, c->o.nu_coarse);
    return;
  }
  Lvl<double> &f = c->h64.lv[l];
  Lvl<double> &g = c->h64.lv[l + 1];
  if (t_streamed(c, l)) {
    // aub -> aub2, coarse ubar = P^T ubar, coarse rbar = 0
    launch(c, KC_ADJ_DOWN, l, 32.0 * nodes(f) + 16.0 * nodes(g), [&] {
      launch_tdown(c, c->o.nu2, f, c->aub[l], c->aub2[l], c->arb[l], g, c->aub[l + 1], c->arb[l + 1]);
    });
    cycle_T(c, l + 1);
    // aub2 + R^T coarse rbar -> aub
    launch(c, KC_ADJ_UP, l, 32.0 * nodes(f) + 8.0 * nodes(g), [&] {
      launch_tup(c, c->o.nu1, f, c->aub2[l], c->aub[l], c->arb[l], g, c->arb[l + 1]);
    });
    return;
  }
  smooth_T(c, l, c->o.nu2);
  launch(c, KC_ADJ, l, 8.0 * (nodes(f) + nodes(g)), [&] {
    hipLaunchKernelGGL(k_adj_prolongT, grid2(g.ny, g.nx), kBlock, 0, c->st, c->aub[l + 1], c->aub[l], g.nx, g.ny, g.pitch, f.pitch);
  });
  fill(c, c->arb[l + 1], g.elems(), 0.0, l + 1);
  cycle_T(c, l + 1);
  launch(c, KC_ADJ, l, 8.0 * (nodes(f) + nodes(g)), [&] {
    hipLaunchKernelGGL(k_adj_restrictT, grid2(f.ny, f.nx), kBlock, 0, c->st, c->aresb[l], c->arb[l + 1], f.nx, f.ny, f.pitch, g.nx, g.ny, g.pitch);
  });
  launch(c, KC_ADJ, l, 40.0 * nodes(f), [&] {
    hipLaunchKernelGGL(k_adj_residualT, grid2(f.ny, f.nx), kBlock, 0, c->st, c->aub[l], c->arb[l], c->aresb[l], f.nx, f.ny, f.pitch, f.ix2, f.iy2);
  });
  smooth_T(c, l, c->o.nu1);
}

}  // namespace

// =============================================================================
// C ABI
// =============================================================================
// ---- problem set-up on the GPU (expressions, boundary, initial guess) ----
constexpr int kProgSlots = 6;

int compile_expr(Ctx *c, const char *text, const char *what, Expr &e) {
  std::string err;
  if (!expr_compile(text ? text : "", e, err)) return fail(c, VLP2D_E_ARG, std::string(what) + ": " + err);
  return VLP2D_OK;
}

// copies program e into device slot k
int upload_prog(Ctx *c, int k, const Expr &e, DevProg &d) {
  if (!c->dprog) RCHK(dalloc(c, &c->dprog, (size_t)kProgSlots * kExprMaxOps));
  d.ops = c->dprog + (size_t)k * kExprMaxOps;
  d.n = (int)e.ops.size();
  if (d.n) HIPCHK(c, hipMemcpyAsync((void *)d.ops, e.ops.data(), e.ops.size() * sizeof(ExprOp), hipMemcpyHostToDevice,
                                    c->st));
  return VLP2D_OK;
}

// 2-D set-up grid: x over columns, y over rows, at most SETUP_MAX_BLOCKS blocks
inline dim3 setup_grid(int nx, int ny) {
  const int gx = std::max(1, std::min((ny + SETUP_THREADS - 1) / SETUP_THREADS, 32));
  const int gy = std::max(1, std::min(nx, SETUP_MAX_BLOCKS / gx));
  return dim3(gx, gy);
}

// min / max of np partials -> *mn, *mx (synchronises)
int read_minmax(Ctx *c, int np, double *mn, double *mx) {
  launch(c, KC_SETUP, 0, 16.0 * np, [&] {
    hipLaunchKernelGGL(k_mm_final, dim3(1), dim3(SETUP_THREADS), 0, c->st, c->part, c->part2, np, c->dscal + 4);
  });
  HIPCHK(c, hipGetLastError());
  HIPCHK(c, hipMemcpyAsync(c->hscal + 4, c->dscal + 4, 2 * sizeof(double), hipMemcpyDeviceToHost, c->st));
  HIPCHK(c, hipStreamSynchronize(c->st));
  if (c->profile) collect_profile(c);
  if (mn) *mn = c->hscal[4];
  if (mx) *mx = c->hscal[5];
  return VLP2D_OK;
}

// fn = -f(hx i, hy j) on level 0 from a compiled program; fails on a
// non-finite value
int sample_rhs(Ctx *c, const Expr &fe) {
  Lvl<double> &P = c->h64.lv[0];
  DevProg d;
  RCHK(upload_prog(c, 0, fe, d));
  launch(c, KC_SETUP, 0, 8.0 * nodes(P), [&] {
    hipLaunchKernelGGL(k_expr_rhs, setup_grid(P.nx, P.ny), dim3(SETUP_THREADS), 0, c->st, d, P.rhs, P.nx, P.ny,
                       P.pitch, c->hx, c->hy);
  });
  HIPCHK(c, hipGetLastError());
  const dim3 g = setup_grid(P.nx, P.ny);
  launch(c, KC_SETUP, 0, 8.0 * nodes(P), [&] {
    hipLaunchKernelGGL(k_minmax, g, dim3(SETUP_THREADS), 0, c->st, P.rhs, P.nx, P.ny, P.pitch, c->part, c->part2);
  });
  double mn, mx;
  RCHK(read_minmax(c, (int)(g.x * g.y), &mn, &mx));
  if (!std::isfinite(mn) || !std::isfinite(mx)) return fail(c, VLP2D_E_NONFINITE, "f evaluates to NaN or Inf");
  return VLP2D_OK;
}

// state after the level-0 iterate changed on the device (as after upload)
int iterate_replaced(Ctx *c) {
  c->r32_valid = false;
  c->K = 0;
  c->cancel.store(0);
  return mirror_level0(c, c->h64.lv[0]);
}

// ---- allocation-free solve ----------------------------------------------------
// The device array that every allocation-free call on a GPU carves its
// buffers from (each GPU holds its own copy, at its own address).  In host
// code the name is a host-side stand-in, not a device address: host code uses
// only the address hipGetSymbolAddress returns for the current device.
static __device__ __align__(256) unsigned char scratch[VLP2D_NOALLOC_SCRATCH_BYTES];

namespace {

constexpr int kNoAllocDevices = 64;
struct NoAllocDevice {
  std::mutex mu;        // serialises the calls on this device
  Ctx *c = nullptr;     // host state reused by every call on the device
  char *ptr = nullptr;  // device address of scratch on this device
};
NoAllocDevice g_noalloc[kNoAllocDevices];

// Once per device, with that device current: the device address of scratch,
// and the host context with its timing events (created here, never per call).
int noalloc_setup(NoAllocDevice &d, int device) {
  if (d.c) return VLP2D_OK;
  void *p = nullptr;
  const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
  if (e != hipSuccess || !p) {
    (void)hipGetLastError();
    g_create_error = std::string("hipGetSymbolAddress(scratch) failed: ") + hipGetErrorString(e);
    return VLP2D_E_DEVICE;
  }
  Ctx *c = new Ctx();
  c->dev = device;
  for (hipEvent_t &ev : c->ev)
    if (hipEventCreate(&ev) != hipSuccess) {
      (void)hipGetLastError();
      for (hipEvent_t &x : c->ev)
        if (x) (void)hipEventDestroy(x);
      delete c;
      g_create_error = "hipEventCreate failed";
      return VLP2D_E_DEVICE;
    }
  if (hipDeviceGetAttribute(&c->max_lds, hipDeviceAttributeMaxSharedMemoryPerBlock, device) != hipSuccess) {
    (void)hipGetLastError();
    c->max_lds = 65536;
  }
  if (hipDeviceGetAttribute(&c->num_cu, hipDeviceAttributeMultiprocessorCount, device) != hipSuccess) {
    (void)hipGetLastError();
    c->num_cu = 0;
  }
  c->h64.lv.reserve(kMaxLv);
  c->h32.lv.reserve(kMaxLv);
  c->arena = static_cast<char *>(p);
  c->arena_cap = VLP2D_NOALLOC_SCRATCH_BYTES;
  c->nomap = true;
  c->hscal = c->hloc;
  d.ptr = static_cast<char *>(p);
  d.c = c;
  return VLP2D_OK;
}

// One solve on the device's context; caller holds d.mu with the device current.
int noalloc_solve(NoAllocDevice &d, int segX, int segY, double hx, double hy, const double *f_dev, double *u_dev,
                  double rtol, int max_cycles, const vlp2d_mg_opts &o, hipStream_t st, vlp2d_mg_stats *stats);

}  // namespace

extern "C" {

void vlp2d_mg_default_opts(vlp2d_mg_opts *o) {
  if (!o) return;
  o->nu1 = 3;
  o->nu2 = 3;
  o->nu_coarse = 3;
  o->precision = VLP2D_PREC_FP64;
  o->flags = 0;
  o->coarse_nodes = 0;
  o->max_inner = 1;
  o->smoother = VLP2D_SMOOTH_RBGS;
  o->omega = 0.8;
  o->restriction = VLP2D_RESTRICT_HALF;
  o->cycle = VLP2D_CYCLE_V;
}

int vlp2d_mg_version(void) { return (1 << 16) | 4; }

// Options with defaults for NULL; false (message in g_create_error) when a
// value is out of range.
static bool checked_opts(const vlp2d_mg_opts *opts, vlp2d_mg_opts &o) {
  vlp2d_mg_default_opts(&o);
  if (opts) o = *opts;
  if (o.nu1 < 0 || o.nu2 < 0 || o.nu_coarse < 1 || o.nu1 > 32 || o.nu2 > 32 || o.nu_coarse > 1024 ||
      (o.precision != VLP2D_PREC_FP64 && o.precision != VLP2D_PREC_MIXED) || o.coarse_nodes < -1 ||
      o.max_inner < 1 || o.max_inner > 8 || (o.smoother != VLP2D_SMOOTH_RBGS && o.smoother != VLP2D_SMOOTH_WJACOBI) ||
      !(o.omega > 0.0 && o.omega <= 1.0) ||
      (o.restriction != VLP2D_RESTRICT_HALF && o.restriction != VLP2D_RESTRICT_FULL) ||
      (o.cycle != VLP2D_CYCLE_V && o.cycle != VLP2D_CYCLE_FMG) ||
      (o.cycle == VLP2D_CYCLE_FMG && o.precision != VLP2D_PREC_FP64)) {
    g_create_error = "invalid options (nu1, nu2 in [0,32], nu_coarse in [1,1024], precision 0 or 1, coarse_nodes >= -1, "
                     "max_inner in [1,8], smoother 0 or 1, omega in (0,1], restriction 0 or 1, cycle 0 or 1, "
                     "cycle FMG with precision FP64 only)";
    return false;
  }
  return true;
}

// The fused kernels are instantiated for 1..3 sweeps per pass, red-black
// Gauss-Seidel and half weighting; everything else runs one operator per launch.
static bool one_operator_path(const vlp2d_mg_opts &o) {
  return (o.flags & VLP2D_FLAG_SIMPLE) != 0 || o.nu1 < 1 || o.nu2 < 1 || o.nu1 > 3 || o.nu2 > 3 ||
         o.cycle == VLP2D_CYCLE_FMG;
}

// nLevels = min(log2 segX, log2 segY) (MultiGridScheme.cs:23-61)
static int level_count(int segX, int segY) {
  return std::min((int)std::log2((double)segX + 0.5), (int)std::log2((double)segY + 0.5));
}

static bool valid_segments(int segX, int segY) {
  return is_pow2(segX) && is_pow2(segY) && segX >= 4 && segY >= 4 && segX <= 32768 && segY <= 32768;
}

// Partials of the tile reductions: one per tile of the largest launch.
// Streaming tiles number at most max(resident wavefronts, column strips), the
// strips being narrowest for the widest halo (ascent with a recomputed
// pre-sweep, 3 post-sweeps and a residual epilogue); refinement tiles are
// 512 x RS_ROWS.
static int partial_capacity(int nx, int ny, int num_cu) {
  constexpr int kMinStripCols = SCOLS - 2 * s_up_halo(4, 3);
  const long long strips = (ny - 2) / kMinStripCols + 1;
  const long long stiles = std::max<long long>(strips, 64LL * std::max(1, num_cu));
  const long long rtiles = (long long)refine_strips(ny) * refine_row_blocks(nx);
  return (int)std::max<long long>(1 << 16, std::max(stiles, rtiles) + 64);
}

static int create_impl(int segX, int segY, double hx, double hy, const double *f, const Expr *fe, double eps,
                       int device, const vlp2d_mg_opts *opts, vlp2d_mg_handle *out) {
  if (!out) {
    g_create_error = "out handle pointer is NULL";
    return VLP2D_E_ARG;
  }
  *out = nullptr;
  if (!valid_segments(segX, segY)) {
    g_create_error = "segX and segY must be powers of two in [4, 32768]";
    return VLP2D_E_SIZE;
  }
  if (!(hx > 0) || !(hy > 0) || !std::isfinite(hx) || !std::isfinite(hy)) {
    g_create_error = "hx and hy must be finite and positive";
    return VLP2D_E_ARG;
  }
  vlp2d_mg_opts o;
  if (!checked_opts(opts, o)) return VLP2D_E_ARG;
  const int nx = segX + 1, ny = segY + 1;
  if (f && !check_finite(f, (size_t)nx * ny)) {
    // only interior values are used, but a NaN anywhere is a caller error
    g_create_error = "f contains NaN or Inf";
    return VLP2D_E_NONFINITE;
  }
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || device < 0 || device >= ndev) {
    (void)hipGetLastError();
    g_create_error = "device " + std::to_string(device) + " not available (" + std::to_string(ndev) + " visible)";
    return VLP2D_E_DEVICE;
  }
  Ctx *c = new Ctx();
  c->dev = device;
  c->segX = segX;
  c->segY = segY;
  c->nx = nx;
  c->ny = ny;
  c->hx = hx;
  c->hy = hy;
  c->eps = eps;
  c->laplace = fe ? fe->empty() : (f == nullptr);
  c->o = o;
  c->profile = (o.flags & VLP2D_FLAG_PROFILE) != 0;
  c->simple = one_operator_path(o);
  c->L = level_count(segX, segY);
  auto bail = [&](int rc) {
    g_create_error = c->err;
    vlp2d_mg_destroy(c);
    return rc;
  };
  if (hipSetDevice(device) != hipSuccess) {
    c->err = "hipSetDevice failed";
    return bail(VLP2D_E_DEVICE);
  }
  if (hipStreamCreateWithFlags(&c->st, hipStreamNonBlocking) != hipSuccess) {
    c->err = "hipStreamCreate failed";
    return bail(VLP2D_E_DEVICE);
  }
  if (hipDeviceGetAttribute(&c->max_lds, hipDeviceAttributeMaxSharedMemoryPerBlock, device) != hipSuccess) {
    (void)hipGetLastError();
    c->max_lds = 65536;
  }
  if (hipDeviceGetAttribute(&c->num_cu, hipDeviceAttributeMultiprocessorCount, device) != hipSuccess) {
    (void)hipGetLastError();
    c->num_cu = 0;
  }
  shape_hier(c, c->h64, hx, hy);
  Lvl<double> &v = c->h64.lv[0];
  v.lap = c->laplace;
  int rc;
  if ((rc = dalloc(c, &v.u, v.elems())) != VLP2D_OK) return bail(rc);
  if ((rc = dalloc(c, &v.u2, v.elems())) != VLP2D_OK) return bail(rc);
  if ((rc = dalloc(c, &v.rhs, v.elems())) != VLP2D_OK) return bail(rc);
  fill(c, v.u, v.elems(), 0.0);
  fill(c, v.u2, v.elems(), 0.0);
  fill(c, v.rhs, v.elems(), 0.0);
  c->part_cap = partial_capacity(nx, ny, c->num_cu);
  if ((rc = dalloc(c, &c->part, (size_t)c->part_cap)) != VLP2D_OK) return bail(rc);
  if ((rc = dalloc(c, &c->part2, (size_t)c->part_cap)) != VLP2D_OK) return bail(rc);
  if ((rc = dalloc(c, &c->dscal, 8)) != VLP2D_OK) return bail(rc);
  fill(c, c->dscal, 8, 0.0);
  if (hipHostMalloc((void **)&c->hscal, 8 * sizeof(double), hipHostMallocDefault) != hipSuccess) {
    c->err = "hipHostMalloc failed";
    return bail(VLP2D_E_OOM);
  }
  if (hipHostMalloc((void **)&c->hpub, kPubRing * sizeof(HostPub), hipHostMallocMapped | hipHostMallocCoherent) !=
          hipSuccess ||
      hipHostGetDevicePointer((void **)&c->hpub_dev, c->hpub, 0) != hipSuccess) {
    c->err = "hipHostMalloc (mapped result buffer) failed";
    return bail(VLP2D_E_OOM);
  }
  std::memset(c->hpub, 0, kPubRing * sizeof(HostPub));
  if ((rc = dalloc(c, &c->dseq, 1)) != VLP2D_OK) return bail(rc);
  if (hipMemsetAsync(c->dseq, 0, sizeof(unsigned long long), c->st) != hipSuccess) {
    c->err = "hipMemsetAsync failed";
    return bail(VLP2D_E_DEVICE);
  }
  if (f) {
    // fn = -f on interior nodes, 0 on the boundary (MultiGridSchemeCU.cs:41-45)
    std::vector<double> fn((size_t)nx * ny, 0.0);
    for (int i = 1; i < nx - 1; ++i)
      for (int j = 1; j < ny - 1; ++j) fn[(size_t)i * ny + j] = -f[(size_t)i * ny + j];
    if ((rc = upload_grid(c, v.rhs, v.pitch, fn.data())) != VLP2D_OK) return bail(rc);
  }
  if (fe && !fe->empty()) {
    if ((rc = sample_rhs(c, *fe)) != VLP2D_OK) return bail(rc);
  }
  if (hipStreamSynchronize(c->st) != hipSuccess) {
    c->err = "stream synchronisation failed";
    return bail(VLP2D_E_DEVICE);
  }
  *out = c;
  return VLP2D_OK;
}

int vlp2d_mg_upload(vlp2d_mg_handle c, const double *u) {
  if (!c) return VLP2D_E_ARG;
  if (!u) return fail(c, VLP2D_E_ARG, "u is NULL");
  if (!check_finite(u, (size_t)c->nx * c->ny)) return fail(c, VLP2D_E_NONFINITE, "u contains NaN or Inf");
  HIPCHK(c, hipSetDevice(c->dev));
  c->r32_valid = false;
  c->K = 0;
  c->cancel.store(0);  // a cancel request holds until the next upload
  Lvl<double> &P = c->h64.lv[0];
  RCHK(upload_grid(c, P.u, P.pitch, u));
  return mirror_level0(c, P);
}

int vlp2d_mg_download(vlp2d_mg_handle c, double *u) {
  if (!c) return VLP2D_E_ARG;
  if (!u) return fail(c, VLP2D_E_ARG, "u is NULL");
  HIPCHK(c, hipSetDevice(c->dev));
  return download_grid(c, u, c->h64.lv[0].u, c->h64.lv[0].pitch);
}

int vlp2d_mg_cycle(vlp2d_mg_handle c, double *stop_metric) {
  if (!c) return VLP2D_E_ARG;
  HIPCHK(c, hipSetDevice(c->dev));
  Lvl<double> &P = c->h64.lv[0];
  if (c->o.precision == VLP2D_PREC_MIXED) {
    RCHK(ensure_h32(c));
    if (!c->r32_valid) {
      resnorm(c, P, P.u, c->h32.lv[0].rhs);
      HIPCHK(c, hipGetLastError());
      c->r32_valid = true;
    }
    RCHK(run_program(c, gkey(PG_MIXED, P.par, c->h32.lv[0].rpar), [&] { mixed_step(c, P, 1, true); }));
    swap_mixed(c, P);
  } else {
    RCHK(complete_hier(c, c->h64));
    RCHK(ensure_aux(c));
    RCHK(run_program(c, gkey(PG_CYCLE_METRIC, P.par, 0), [&] { cycle64(c, c->h64, EPI_METRIC, true); }));
  }
  // the program's last kernel publishes the metric in mapped host memory:
  // no copy and no stream synchronisation per call
  RCHK(publish_wait(c));
  if (stop_metric) *stop_metric = c->hscal[0];
  // the metric is a max over the iterate's update: NaN or +Inf in the
  // iterate reaches it (the max reduction propagates NaN)
  if (!std::isfinite(c->hscal[0])) return fail(c, VLP2D_E_NONFINITE, "iterate became non-finite");
  return VLP2D_OK;
}

int vlp2d_mg_residual_norm(vlp2d_mg_handle c, double *norm) {
  if (!c || !norm) return VLP2D_E_ARG;
  HIPCHK(c, hipSetDevice(c->dev));
  Lvl<double> &P = c->h64.lv[0];
  resnorm(c, P, P.u, nullptr);
  HIPCHK(c, hipGetLastError());
  RCHK(sync_read(c, 2));
  *norm = std::sqrt(c->hscal[1]);
  return VLP2D_OK;
}

int vlp2d_mg_solve(vlp2d_mg_handle c, double rtol, int max_cycles, vlp2d_mg_stats *st) {
  if (!c) return VLP2D_E_ARG;
  if (!(rtol >= 0) || max_cycles < 0) return fail(c, VLP2D_E_ARG, "rtol must be >= 0 and max_cycles >= 0");
  HIPCHK(c, hipSetDevice(c->dev));
  return solve_impl(c, c->h64, PG_CYCLE_NORM, PG_MIXED, rtol, max_cycles, st);
}

int vlp2d_mg_adjoint_exact(vlp2d_mg_handle c, const double *dL_du, double rtol, int max_cycles, double *dL_df,
                           double *dL_dg) {
  if (!c) return VLP2D_E_ARG;
  if (!dL_du) return fail(c, VLP2D_E_ARG, "dL_du is NULL");
  if (!(rtol >= 0) || max_cycles < 0) return fail(c, VLP2D_E_ARG, "rtol must be >= 0 and max_cycles >= 0");
  if (!check_finite(dL_du, (size_t)c->nx * c->ny)) return fail(c, VLP2D_E_NONFINITE, "dL_du contains NaN or Inf");
  HIPCHK(c, hipSetDevice(c->dev));
  // Adjoint system A lambda = dL_du, A = -Lap_h on the interior with zero
  // boundary: in the scheme's form Lap_h lambda = fn with fn = -dL_du.
  if (c->ha.lv.empty()) {
    shape_hier(c, c->ha, c->hx, c->hy);
    Lvl<double> &a = c->ha.lv[0];
    a.lap = false;
    RCHK(dalloc(c, &a.u, a.elems()));
    RCHK(dalloc(c, &a.u2, a.elems()));
    RCHK(dalloc(c, &a.rhs, a.elems()));
  }
  Lvl<double> &a = c->ha.lv[0];
  fill(c, a.u2, a.elems(), 0.0);
  fill(c, a.rhs, a.elems(), 0.0);
  HIPCHK(c, hipGetLastError());
  // stage dL_du in a.u, negate into a.rhs, then clear both iterate buffers
  RCHK(upload_grid(c, a.u, a.pitch, dL_du));
  launch(c, KC_COPY, 0, 16.0 * nodes(a), [&] {
    hipLaunchKernelGGL(k_neg_interior, grid2(a.ny - 2, a.nx - 2), kBlock, 0, c->st, a.rhs, a.u, a.nx, a.ny, a.pitch);
  });
  fill(c, a.u, a.elems(), 0.0);
  HIPCHK(c, hipGetLastError());
  c->r32_valid = false;
  vlp2d_mg_stats st{};
  int rc = solve_impl(c, c->ha, PG_ADJ_CYCLE_NORM, PG_ADJ_MIXED, rtol, max_cycles, &st);
  c->r32_valid = false;
  if (rc != VLP2D_OK && rc != VLP2D_E_NOTCONV) return rc;
  if (dL_df) {
    RCHK(download_grid(c, dL_df, a.u, a.pitch));
    for (int i = 0; i < c->nx; ++i)
      for (int j = 0; j < c->ny; ++j)
        if (i == 0 || j == 0 || i == c->nx - 1 || j == c->ny - 1) dL_df[(size_t)i * c->ny + j] = 0.0;
  }
  if (dL_dg) {
    // B^T lambda into the second iterate buffer (scratch here)
    Lvl<double> &v = c->h64.lv[0];
    launch(c, KC_ADJ, 0, 16.0 * nodes(a), [&] {
      hipLaunchKernelGGL(k_adj_bT, grid2(a.ny, a.nx), kBlock, 0, c->st, a.u2, a.u, a.nx, a.ny, a.pitch, v.ix2, v.iy2);
    });
    HIPCHK(c, hipGetLastError());
    RCHK(download_grid(c, dL_dg, a.u2, a.pitch));
    // the scratch use broke the boundary mirror of a.u2; it is re-zeroed on the
    // next call before use
  }
  if (c->profile) collect_profile(c);
  return rc;
}

int vlp2d_mg_forward_cycles(vlp2d_mg_handle c, int K) {
  if (!c) return VLP2D_E_ARG;
  if (K < 1 || K > 4096) return fail(c, VLP2D_E_ARG, "K must be in [1, 4096]");
  if (c->o.smoother != VLP2D_SMOOTH_RBGS || c->o.restriction != VLP2D_RESTRICT_HALF || c->o.cycle != VLP2D_CYCLE_V)
    return fail(c, VLP2D_E_STATE,
                "the truncated-cycle adjoint transposes the red-black Gauss-Seidel, half-weighting V-cycle only");
  HIPCHK(c, hipSetDevice(c->dev));
  RCHK(complete_hier(c, c->h64));
  // The V-cycle is an affine map of (u, fn, g), so its transpose reads no
  // forward state: nothing is stored per cycle, only the cycle count K that
  // the reverse sweep transposes.
  for (int k = 0; k < K; ++k) {
    if (c->cancel.load()) return fail(c, VLP2D_E_CANCELLED, "cancelled");
    cycle64(c, c->h64, EPI_NONE);
    HIPCHK(c, hipGetLastError());
  }
  HIPCHK(c, hipStreamSynchronize(c->st));
  if (c->profile) collect_profile(c);
  c->K = K;
  c->r32_valid = false;
  return VLP2D_OK;
}

int vlp2d_mg_adjoint_cycles(vlp2d_mg_handle c, const double *dL_du, double *dL_df, double *dL_dg, double *dL_du0) {
  if (!c) return VLP2D_E_ARG;
  if (!dL_du) return fail(c, VLP2D_E_ARG, "dL_du is NULL");
  if (c->K < 1) return fail(c, VLP2D_E_STATE, "vlp2d_mg_forward_cycles must run first");
  if (!check_finite(dL_du, (size_t)c->nx * c->ny)) return fail(c, VLP2D_E_NONFINITE, "dL_du contains NaN or Inf");
  HIPCHK(c, hipSetDevice(c->dev));
  RCHK(ensure_adj_buffers(c));
  Lvl<double> &P = c->h64.lv[0];
  RCHK(upload_grid(c, c->aub[0], P.pitch, dL_du));
  fill(c, c->arb[0], P.elems(), 0.0);
  // The V-cycle is affine in (u, fn, g): its transpose needs none of the
  // forward states, so the reverse sweep applies the transposed cycle K times
  // without reading the checkpoints back.
  for (int k = c->K - 1; k >= 0; --k) {
    if (c->cancel.load()) return fail(c, VLP2D_E_CANCELLED, "cancelled");
    cycle_T(c, 0);
    HIPCHK(c, hipGetLastError());
  }
  // outputs formed on the device, one packed nx x ny array at a time in a
  // staging buffer, then copied to the caller
  double *stage = c->aub2[0] ? c->aub2[0] : c->aresb[0];
  double *outs[3] = {dL_df, dL_dg, dL_du0};
  for (int k = 0; k < 3; ++k) {
    if (!outs[k]) continue;
    hipLaunchKernelGGL(k_adj_output, grid2(P.ny, P.nx), kBlock, 0, c->st, stage, c->aub[0], c->arb[0], P.nx, P.ny,
                       P.pitch, k);
    HIPCHK(c, hipGetLastError());
    HIPCHK(c, hipMemcpyAsync(outs[k], stage, (size_t)P.nx * P.ny * sizeof(double), hipMemcpyDeviceToHost, c->st));
  }
  HIPCHK(c, hipStreamSynchronize(c->st));
  if (c->profile) collect_profile(c);
  return VLP2D_OK;
}

int vlp2d_mg_get_info(vlp2d_mg_handle c, vlp2d_mg_info *info) {
  if (!c || !info) return VLP2D_E_ARG;
  info->levels = c->L;
  info->nx = c->nx;
  info->ny = c->ny;
  info->device_bytes = c->bytes;
  info->device_peak_bytes = c->peak;
  info->checkpoints = c->K;
  return VLP2D_OK;
}

int vlp2d_mg_kernel_times(vlp2d_mg_handle c, const char **names, int *levels, long long *launches, double *total_ms,
                          double *alg_bytes, int cap) {
  if (!c) return VLP2D_E_ARG;
  int n = 0;
  for (int k = 0; k < KC_COUNT; ++k)
    for (int l = 0; l < kMaxLv; ++l) {
      if (c->kc_launch[k][l] == 0) continue;
      if (n < cap) {
        if (names) names[n] = kKClassName[k];
        if (levels) levels[n] = l;
        if (launches) launches[n] = c->kc_launch[k][l];
        if (total_ms) total_ms[n] = c->kc_ms[k][l];
        if (alg_bytes) alg_bytes[n] = c->kc_bytes[k][l];
      }
      ++n;
    }
  return n;
}

int vlp2d_mg_reset_kernel_times(vlp2d_mg_handle c) {
  if (!c) return VLP2D_E_ARG;
  for (int k = 0; k < KC_COUNT; ++k)
    for (int l = 0; l < kMaxLv; ++l) {
      c->kc_launch[k][l] = 0;
      c->kc_ms[k][l] = 0;
      c->kc_bytes[k][l] = 0;
    }
  return VLP2D_OK;
}

int vlp2d_mg_save_iterate(vlp2d_mg_handle c) {
  if (!c) return VLP2D_E_ARG;
  HIPCHK(c, hipSetDevice(c->dev));
  Lvl<double> &P = c->h64.lv[0];
  if (!c->saved) RCHK(dalloc(c, &c->saved, P.elems()));
  HIPCHK(c, hipMemcpyAsync(c->saved, P.u, P.elems() * sizeof(double), hipMemcpyDeviceToDevice, c->st));
  HIPCHK(c, hipStreamSynchronize(c->st));
  return VLP2D_OK;
}

int vlp2d_mg_restore_iterate(vlp2d_mg_handle c) {
  if (!c) return VLP2D_E_ARG;
  if (!c->saved) return fail(c, VLP2D_E_STATE, "vlp2d_mg_save_iterate must run first");
  HIPCHK(c, hipSetDevice(c->dev));
  Lvl<double> &P = c->h64.lv[0];
  HIPCHK(c, hipMemcpyAsync(P.u, c->saved, P.elems() * sizeof(double), hipMemcpyDeviceToDevice, c->st));
  RCHK(mirror_level0(c, P));
  c->r32_valid = false;
  c->K = 0;
  c->cancel.store(0);
  return VLP2D_OK;
}

void vlp2d_mg_cancel(vlp2d_mg_handle c) {
  if (c) c->cancel.store(1);
}

void vlp2d_mg_destroy(vlp2d_mg_handle c) {
  if (!c) return;
  (void)hipSetDevice(c->dev);
  if (c->st) (void)hipStreamSynchronize(c->st);
  collect_profile(c);
  invalidate_graphs(c);
  size_t n0 = (size_t)c->nx * round_pitch(c->ny);
  free_hier(c, c->h64);
  free_hier(c, c->h32);
  free_hier(c, c->ha);
  dfree(c, c->snap, n0);
  dfree(c, c->saved, n0);
  for (size_t l = 0; l < c->aub.size(); ++l) {
    size_t n = (size_t)((c->segX >> l) + 1) * round_pitch((c->segY >> l) + 1);
    dfree(c, c->aub[l], n);
    dfree(c, c->arb[l], n);
    dfree(c, c->aresb[l], n);
    dfree(c, c->aub2[l], n);
  }
  dfree(c, c->part, (size_t)c->part_cap);
  dfree(c, c->part2, (size_t)c->part_cap);
  dfree(c, c->dscal, 8);
  dfree(c, c->dprog, (size_t)kProgSlots * kExprMaxOps);
  if (c->hscal) (void)hipHostFree(c->hscal);
  if (c->hpub) (void)hipHostFree(c->hpub);
  dfree(c, c->dseq, 1);
  for (hipEvent_t &e : c->ev)
    if (e) (void)hipEventDestroy(e);
  if (c->st) (void)hipStreamDestroy(c->st);
  delete c;
}

int vlp2d_mg_create(int segX, int segY, double hx, double hy, const double *f, double eps, int device,
                    const vlp2d_mg_opts *opts, vlp2d_mg_handle *out) {
  g_create_error.clear();
  return create_impl(segX, segY, hx, hy, f, nullptr, eps, device, opts, out);
}

int vlp2d_mg_create_expr(int segX, int segY, double hx, double hy, const char *f_expr, double eps, int device,
                         const vlp2d_mg_opts *opts, vlp2d_mg_handle *out) {
  g_create_error.clear();
  Expr fe;
  std::string err;
  if (!expr_compile(f_expr ? f_expr : "", fe, err)) {
    g_create_error = "f: " + err;
    if (out) *out = nullptr;
    return VLP2D_E_ARG;
  }
  return create_impl(segX, segY, hx, hy, nullptr, &fe, eps, device, opts, out);
}

int vlp2d_expr_eval(const char *expr, const double *x, const double *y, long long n, double *out) {
  g_create_error.clear();
  if (!expr || (n > 0 && (!x || !y || !out)) || n < 0) {
    g_create_error = "expr, x, y and out must be non-NULL and n >= 0";
    return VLP2D_E_ARG;
  }
  Expr e;
  std::string err;
  if (!expr_compile(expr, e, err)) {
    g_create_error = err;
    return VLP2D_E_ARG;
  }
  if (e.empty()) {
    g_create_error = "empty expression";
    return VLP2D_E_ARG;
  }
  for (long long k = 0; k < n; ++k) out[k] = e.eval(x[k], y[k]);
  return VLP2D_OK;
}

int vlp2d_mg_set_boundary_expr(vlp2d_mg_handle c, const char *bottom, const char *top, const char *left,
                               const char *right, const char *boundary, double *vmin, double *vmax) {
  if (!c) return VLP2D_E_ARG;
  Expr eb, et, el, er, eg;
  RCHK(compile_expr(c, bottom, "bottom", eb));
  RCHK(compile_expr(c, top, "top", et));
  RCHK(compile_expr(c, left, "left", el));
  RCHK(compile_expr(c, right, "right", er));
  RCHK(compile_expr(c, boundary, "boundary", eg));
  if (eg.empty() && (eb.empty() || et.empty() || el.empty() || er.empty()))
    return fail(c, VLP2D_E_ARG, "the boundary needs the boundary function or all four edge functions");
  HIPCHK(c, hipSetDevice(c->dev));
  Lvl<double> &P = c->h64.lv[0];
  DevProg db, dt, dl, dr, dg;
  RCHK(upload_prog(c, 1, eb, db));
  RCHK(upload_prog(c, 2, et, dt));
  RCHK(upload_prog(c, 3, el, dl));
  RCHK(upload_prog(c, 4, er, dr));
  RCHK(upload_prog(c, 5, eg, dg));
  const int nb1 = std::max(1, std::min((P.nx + SETUP_THREADS - 1) / SETUP_THREADS, SETUP_MAX_BLOCKS / 2));
  const int nb2 = std::max(1, std::min((P.ny + SETUP_THREADS - 1) / SETUP_THREADS, SETUP_MAX_BLOCKS / 2));
  launch(c, KC_SETUP, 0, 16.0 * (P.nx + P.ny), [&] {
    hipLaunchKernelGGL(k_border_rows, dim3(nb1), dim3(SETUP_THREADS), 0, c->st, db, dt, dg, P.u, P.nx, P.ny, P.pitch,
                       c->hx, c->hy, c->part, c->part2);
    hipLaunchKernelGGL(k_border_cols, dim3(nb2), dim3(SETUP_THREADS), 0, c->st, dl, dr, dg, P.u, P.nx, P.ny, P.pitch,
                       c->hx, c->hy, c->part + nb1, c->part2 + nb1);
  });
  HIPCHK(c, hipGetLastError());
  double mn, mx;
  RCHK(read_minmax(c, nb1 + nb2, &mn, &mx));
  c->bmin = mn;
  c->bmax = mx;
  if (vmin) *vmin = mn;
  if (vmax) *vmax = mx;
  return iterate_replaced(c);
}

int vlp2d_mg_init_guess(vlp2d_mg_handle c, int kind, double value) {
  if (!c) return VLP2D_E_ARG;
  if (kind < VLP2D_INIT_MEAN || kind > VLP2D_INIT_WEIGHT_LINEAR) return fail(c, VLP2D_E_ARG, "unknown initial guess");
  if (kind == VLP2D_INIT_MEAN && std::isnan(value)) {
    if (std::isnan(c->bmin)) return fail(c, VLP2D_E_STATE, "no boundary sampled on the device: pass the mean value");
    value = (c->bmin + c->bmax) / 2.0;  // VLPRectangleModel.cs:545
  }
  HIPCHK(c, hipSetDevice(c->dev));
  Lvl<double> &P = c->h64.lv[0];
  launch(c, KC_SETUP, 0, 8.0 * nodes(P), [&] {
    hipLaunchKernelGGL(k_init_guess, setup_grid(P.nx, P.ny), dim3(SETUP_THREADS), 0, c->st, P.u, P.nx, P.ny, P.pitch,
                       kind, value);
  });
  HIPCHK(c, hipGetLastError());
  return iterate_replaced(c);
}

int vlp2d_mg_points_min_max(vlp2d_mg_handle c, double *vmin, double *vmax) {
  if (!c) return VLP2D_E_ARG;
  HIPCHK(c, hipSetDevice(c->dev));
  Lvl<double> &P = c->h64.lv[0];
  const dim3 g = setup_grid(P.nx, P.ny);
  launch(c, KC_SETUP, 0, 8.0 * nodes(P), [&] {
    hipLaunchKernelGGL(k_minmax, g, dim3(SETUP_THREADS), 0, c->st, P.u, P.nx, P.ny, P.pitch, c->part, c->part2);
  });
  HIPCHK(c, hipGetLastError());
  return read_minmax(c, (int)(g.x * g.y), vmin, vmax);
}

int vlp2d_mg_difference(vlp2d_mg_handle c, const char *analytic, double *dmin, double *dmax, double *diff) {
  if (!c) return VLP2D_E_ARG;
  Expr ea;
  RCHK(compile_expr(c, analytic, "analytic", ea));
  if (ea.empty()) return fail(c, VLP2D_E_ARG, "analytic: empty expression");
  HIPCHK(c, hipSetDevice(c->dev));
  Lvl<double> &P = c->h64.lv[0];
  DevProg d;
  RCHK(upload_prog(c, 0, ea, d));
  double *dd = nullptr;
  if (diff) RCHK(dalloc(c, &dd, P.elems()));
  const dim3 g = setup_grid(P.nx, P.ny);
  launch(c, KC_SETUP, 0, (diff ? 16.0 : 8.0) * nodes(P), [&] {
    hipLaunchKernelGGL(k_difference, g, dim3(SETUP_THREADS), 0, c->st, P.u, d, P.nx, P.ny, P.pitch, c->hx, c->hy, dd,
                       c->part, c->part2);
  });
  int rc = (hipGetLastError() == hipSuccess) ? VLP2D_OK : fail(c, VLP2D_E_DEVICE, "difference kernel failed");
  if (rc == VLP2D_OK) rc = read_minmax(c, (int)(g.x * g.y), dmin, dmax);
  if (rc == VLP2D_OK && diff) rc = download_grid(c, diff, dd, P.pitch);
  if (dd) dfree(c, dd, P.elems());
  return rc;
}

const char *vlp2d_mg_last_error(vlp2d_mg_handle c) {
  if (!c) return g_create_error.c_str();
  return c->err.c_str();
}

int vlp2d_mg_solve_noalloc(int segX, int segY, double hx, double hy, const double *f_dev, double *u_dev, int pitch,
                           double rtol, int max_cycles, const vlp2d_mg_opts *opts, int device, void *stream,
                           vlp2d_mg_stats *stats) {
  if (!valid_segments(segX, segY)) {
    g_create_error = "segX and segY must be powers of two in [4, 32768]";
    return VLP2D_E_SIZE;
  }
  if (!(hx > 0) || !(hy > 0) || !std::isfinite(hx) || !std::isfinite(hy)) {
    g_create_error = "hx and hy must be finite and positive";
    return VLP2D_E_ARG;
  }
  if (!u_dev) {
    g_create_error = "u_dev is NULL";
    return VLP2D_E_ARG;
  }
  if (pitch != round_pitch(segY + 1)) {
    g_create_error = "pitch must be ny = segY + 1 rounded up to a multiple of 32 (" +
                     std::to_string(round_pitch(segY + 1)) + ")";
    return VLP2D_E_ARG;
  }
  if ((reinterpret_cast<uintptr_t>(u_dev) & 15) || (reinterpret_cast<uintptr_t>(f_dev) & 15)) {
    g_create_error = "u_dev and f_dev must be 16-byte aligned";
    return VLP2D_E_ARG;
  }
  if (!(rtol >= 0) || max_cycles < 0) {
    g_create_error = "rtol must be >= 0 and max_cycles >= 0";
    return VLP2D_E_ARG;
  }
  vlp2d_mg_opts o;
  if (!checked_opts(opts, o)) return VLP2D_E_ARG;
  if (o.flags & VLP2D_FLAG_PROFILE) {
    g_create_error = "VLP2D_FLAG_PROFILE is not available without allocation";
    return VLP2D_E_ARG;
  }
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || device < 0 || device >= ndev || device >= kNoAllocDevices) {
    (void)hipGetLastError();
    g_create_error = "device " + std::to_string(device) + " not available (" + std::to_string(ndev) + " visible)";
    return VLP2D_E_DEVICE;
  }
  int prev = -1;
  if (hipGetDevice(&prev) != hipSuccess) prev = -1;
  if (hipSetDevice(device) != hipSuccess) {
    (void)hipGetLastError();
    g_create_error = "hipSetDevice failed";
    return VLP2D_E_DEVICE;
  }
  NoAllocDevice &d = g_noalloc[device];
  int rc;
  {
    std::lock_guard<std::mutex> lock(d.mu);
    rc = noalloc_setup(d, device);
    if (rc == VLP2D_OK)
      rc = noalloc_solve(d, segX, segY, hx, hy, f_dev, u_dev, rtol, max_cycles, o, static_cast<hipStream_t>(stream),
                         stats);
  }
  if (prev >= 0 && prev != device) (void)hipSetDevice(prev);
  return rc;
}

}  // extern "C"

namespace {

int noalloc_solve(NoAllocDevice &d, int segX, int segY, double hx, double hy, const double *f_dev, double *u_dev,
                  double rtol, int max_cycles, const vlp2d_mg_opts &o, hipStream_t st, vlp2d_mg_stats *stats) {
  Ctx *c = d.c;
  // the state of a fresh handle over the caller's arrays (no graphs, no
  // mapped memory, every buffer from scratch, cleared before use)
  c->segX = segX;
  c->segY = segY;
  c->nx = segX + 1;
  c->ny = segY + 1;
  c->hx = hx;
  c->hy = hy;
  c->eps = 0.0;
  c->laplace = f_dev == nullptr;
  c->o = o;
  c->o.flags |= VLP2D_FLAG_NO_GRAPH;
  c->profile = false;
  c->simple = one_operator_path(o);
  c->L = level_count(segX, segY);
  c->st = st;
  c->arena_used = 0;
  c->bytes = c->peak = 0;
  c->r32_valid = false;
  c->K = 0;
  c->cancel.store(0);
  c->err.clear();
  c->part = c->part2 = c->dscal = nullptr;
  c->dseq = nullptr;
  for (Hier<double> *H : {&c->h64, &c->ha}) {
    H->lv.clear();
    H->full = false;
  }
  c->h32.lv.clear();
  c->h32.full = false;
  shape_hier(c, c->h64, hx, hy);
  Lvl<double> &v = c->h64.lv[0];
  v.lap = c->laplace;
  v.u = u_dev;
  auto bail = [&](int rc) {
    g_create_error = c->err;
    return rc;
  };
  int rc;
  c->part_cap = partial_capacity(c->nx, c->ny, c->num_cu);
  if ((rc = dalloc(c, &v.u2, v.elems())) != VLP2D_OK) return bail(rc);
  if ((rc = dalloc(c, &v.rhs, v.elems())) != VLP2D_OK) return bail(rc);
  if ((rc = dalloc(c, &c->part, (size_t)c->part_cap)) != VLP2D_OK) return bail(rc);
  if ((rc = dalloc(c, &c->part2, (size_t)c->part_cap)) != VLP2D_OK) return bail(rc);
  if ((rc = dalloc(c, &c->dscal, 8)) != VLP2D_OK) return bail(rc);
  // the rest of the hierarchy: allocated (from scratch) before anything is
  // cleared or runs, so a grid that does not fit returns here
  if (o.precision == VLP2D_PREC_MIXED) rc = ensure_h32(c);
  else rc = complete_hier(c, c->h64);
  if (rc != VLP2D_OK) return bail(rc);
  fill(c, c->dscal, 8, 0.0);
  fill(c, v.rhs, v.elems(), 0.0);
  if (f_dev) {
    // fn = -f on interior nodes, 0 on the boundary (MultiGridSchemeCU.cs:41-45)
    launch(c, KC_COPY, 0, 16.0 * nodes(v), [&] {
      hipLaunchKernelGGL(k_neg_interior, grid2(v.ny - 2, v.nx - 2), kBlock, 0, c->st, v.rhs, f_dev, v.nx, v.ny,
                         v.pitch);
    });
  }
  if (hipGetLastError() != hipSuccess) return bail(fail(c, VLP2D_E_DEVICE, "set-up kernels failed to launch"));
  if ((rc = mirror_level0(c, v)) != VLP2D_OK) return bail(rc);
  rc = solve_impl(c, c->h64, PG_CYCLE_NORM, PG_MIXED, rtol, max_cycles, stats);
  if (rc == VLP2D_OK || rc == VLP2D_E_NOTCONV || rc == VLP2D_E_NONFINITE) {
    // a mixed solve alternates the iterate between u_dev and its second buffer
    if (v.u != u_dev &&
        hipMemcpyAsync(u_dev, v.u, v.elems() * sizeof(double), hipMemcpyDeviceToDevice, c->st) != hipSuccess) {
      (void)hipGetLastError();
      rc = fail(c, VLP2D_E_DEVICE, "copy of the final iterate failed");
    }
  }
  if (hipStreamSynchronize(c->st) != hipSuccess) {
    (void)hipGetLastError();
    if (rc == VLP2D_OK) rc = fail(c, VLP2D_E_DEVICE, "stream synchronisation failed");
  }
  return rc == VLP2D_OK ? rc : bail(rc);
}

}  // namespace
