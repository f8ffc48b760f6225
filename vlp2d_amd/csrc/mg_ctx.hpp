// Host-side state of one solver handle, shared by the library's translation
// units: the level hierarchy, the handle (struct vlp2d_mg_ctx, opaque behind
// the C ABI), and the tile geometry of the streaming kernels.  The streaming
// kernels are instantiated in their own translation units (mg_k_down.hip,
// mg_k_up.hip), one per precision and sweep count, which define the launchers
// declared at the end of this header; mg_lib.hip drives them.
#pragma once
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <cmath>
#include <map>
#include <string>
#include <vector>

#include "../../include/vlp2d_mg.h"
#include "expr.hpp"
#include "mg_kernels.hpp"
#include "mg_stream.hpp"
#include "mg_tiles.hpp"

namespace vlp2d {

enum KClass {
  KC_SMOOTH,
  KC_RESID,
  KC_RESTRICT,
  KC_FILL,
  KC_PROLONG,
  KC_COPY,
  KC_MAXDIFF,
  KC_RESNORM,
  KC_REDUCE,
  KC_UPDATE,
  KC_ADJ,
  KC_DOWN,
  KC_UP,
  KC_COARSE,
  KC_SETUP,
  KC_UP_LDS,
  KC_UP_REFINE,
  KC_UP_DOWN,
  KC_DOWN_R64,
  KC_ADJ_DOWN,
  KC_ADJ_UP,
  KC_COUNT
};

// Epilogues of the level-0 ascent.
enum Epi { EPI_NONE = 0, EPI_RESNORM = 1, EPI_METRIC = 2, EPI_MIXED = 3 };

template <typename T> struct Lvl {
  T *u = nullptr, *u2 = nullptr;      // iterate and its second buffer
  T *rhs = nullptr, *rhs2 = nullptr;  // right-hand side (rhs2: mixed level 0)
  T *res = nullptr;                   // residual (one-operator-per-launch path)
  int nx = 0, ny = 0, pitch = 0;
  T hx2 = 0, hy2 = 0, coef = 0;
  bool eq = true, lap = false;
  T ihx2 = 0, ihy2 = 0;          // 1/hx^2, 1/hy^2 in T (exact when pw2)
  bool pw2 = false;              // hx^2 and hy^2 are powers of two
  double ix2 = 0, iy2 = 0;       // 1/hx^2, 1/hy^2 in double (adjoint)
  double ax = 0, ay = 0, b = 0;  // smoother coefficients in double (adjoint)
  int par = 0, rpar = 0;         // swap parity of u/u2 and rhs/rhs2
  int tiles_i = 0, tiles_j = 0, chunk = 0;
  size_t elems() const { return (size_t)nx * pitch; }
};

template <typename T> struct Hier {
  std::vector<Lvl<T>> lv;
  bool full = false;  // every buffer of every level allocated
  int lc = 0;         // first level handled by the one-workgroup kernel
};

struct PendingEv {
  int kc, lvl;
  hipEvent_t a, b;
};
constexpr int kMaxLv = 16;

}  // namespace vlp2d

struct vlp2d_mg_ctx {
  int dev = 0;
  hipStream_t st = nullptr;
  int segX = 0, segY = 0, nx = 0, ny = 0, L = 0;
  double hx = 0, hy = 0, eps = 0;
  bool laplace = false;
  bool simple = false;  // one operator per launch
  vlp2d_mg_opts o{};
  vlp2d::Hier<double> h64;  // the problem: level 0 u (with boundary), rhs = fn
  vlp2d::Hier<float> h32;   // FP32 correction hierarchy (mixed precision)
  vlp2d::Hier<double> ha;   // adjoint-solve hierarchy (exact adjoint)
  double *snap = nullptr;
  double *saved = nullptr;
  double *part = nullptr, *part2 = nullptr;
  int part_cap = 0;
  double *dscal = nullptr;  // device scalars
  double *hscal = nullptr;  // pinned host scalars
  vlp2d::HostPub *hpub = nullptr;  // mapped pinned ring of kPubRing results the solve loop polls
  vlp2d::HostPub *hpub_dev = nullptr;
  unsigned long long *dseq = nullptr;  // device count of published results
  unsigned long long pub_seq = 0;      // results published so far
  bool r32_valid = false;   // h32 level-0 rhs holds (float)(fn - Lap u64)
  // transposed-cycle buffers
  std::vector<double *> aub, arb, aresb;
  std::vector<double *> aub2;  // second ubar buffers of the streaming transposed levels
  int K = 0;  // forward cycles recorded for the transposed sweep
  std::map<int, std::pair<hipGraph_t, hipGraphExec_t>> graphs;
  int num_cu = 0;                    // compute units of the device
  int max_lds = 65536;               // dynamic LDS bytes one workgroup may use
  vlp2d::ExprOp *dprog = nullptr;           // device copies of compiled expressions (kProgSlots programs)
  double bmin = NAN, bmax = NAN;     // min / max of the last boundary sampled on the device
  std::map<const void *, int> occ;   // resident 64-thread workgroups per GPU, per kernel
  // profiling
  bool profile = false;
  std::vector<vlp2d::PendingEv> pend;
  long long kc_launch[vlp2d::KC_COUNT][vlp2d::kMaxLv] = {};
  double kc_ms[vlp2d::KC_COUNT][vlp2d::kMaxLv] = {};
  double kc_bytes[vlp2d::KC_COUNT][vlp2d::kMaxLv] = {};
  hipEvent_t ev[2] = {nullptr, nullptr};  // timing events of solves
  // Allocation-free calls (vlp2d_mg_solve_noalloc): device buffers are carved
  // from the library's static device array (arena) instead of hipMalloc, and
  // results are copied back with a stream synchronisation (nomap: no mapped
  // host memory; hscal points at hloc).
  char *arena = nullptr;
  size_t arena_cap = 0, arena_used = 0;
  bool nomap = false;
  double hloc[8] = {};
  std::atomic<int> cancel{0};
  std::string err;
  long long bytes = 0, peak = 0;
};

namespace vlp2d {

using Ctx = vlp2d_mg_ctx;

// ---- fused kernels -------------------------------------------------------------
// Operands of one fused launch (the tile grid is set by sview).
template <typename T> LevelView<T> view(const Lvl<T> &v, const T *uin, T *uout, int H) {
  LevelView<T> a;
  a.uin = uin;
  a.uout = uout;
  a.f = v.rhs;
  a.nx = v.nx;
  a.ny = v.ny;
  a.pitch = v.pitch;
  a.hx2 = v.hx2;
  a.hy2 = v.hy2;
  a.coef = v.coef;
  a.eq = v.eq ? 1 : 0;
  a.fpitch = v.lap ? 0 : v.pitch;
  a.ihx2 = v.ihx2;
  a.ihy2 = v.ihy2;
  a.pw2 = v.pw2 ? 1 : 0;
  a.rc = nullptr;
  a.w = T(0);
  a.tiles_i = a.tiles_j = a.chunk = a.rows = 0;
  return a;
}
// Resident 64-thread workgroups (wavefronts) of a kernel on the whole GPU.
inline int resident_waves(Ctx *c, const void *fn) {
  auto it = c->occ.find(fn);
  if (it != c->occ.end()) return it->second;
  int nb = 0;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, fn, 64 * kWaves, 0) != hipSuccess || nb < 1) {
    (void)hipGetLastError();
    nb = 4;
  }
  const int q = nb * kWaves * std::max(1, c->num_cu);
  c->occ[fn] = q;
  return q;
}

// Streaming-kernel view: tiles of R output rows by 128 - 2H columns, R
// chosen so that the tiles fill the resident wavefront slots about once (one
// wave of tiles: each staged halo row is re-read once per tile boundary, and
// no tail of late tiles).  The staged rows R + 2H are a multiple of P.
template <typename T>
LevelView<T> sview(Ctx *c, const void *fn, const Lvl<T> &v, const T *uin, T *uout, int H, int P) {
  LevelView<T> a = view(v, uin, uout, H);
  a.w = (T)c->o.omega;
  const int tj = (v.ny - 2) / (SCOLS - 2 * H) + 1;
  const int ti = std::max(1, resident_waves(c, fn) / tj);
  const int R = s_rows_at_least((v.nx - 2) / ti + 1, H, P);
  const STileGrid g = s_grid(v.nx, v.ny, R, H);
  a.rows = R;
  a.tiles_i = g.tiles_i;
  a.tiles_j = g.tiles_j;
  a.chunk = g.chunk;
  return a;
}

// ---- streaming-kernel launchers (defined in mg_k_down.hip / mg_k_up.hip) ----
// Descent of level f into g's right-hand side: zi, f starts from zero; store,
// the pre-smoothed iterate is written (else the ascent recomputes it).
template <typename T, int NS> void launch_down_ns(Ctx *c, Lvl<T> &f, Lvl<T> &g, bool zi, bool store);
// Ascent of level f with epilogue epi (Epi); rz: the pre-smoothed iterate is
// recomputed; zi: f started from zero; nd = 1: the next refinement step's
// first descent folded into the refinement ascent (its restriction into
// g.rhs).  Returns the partials written.
template <typename T, int NS>
int launch_up_ns(Ctx *c, Lvl<T> &f, const Lvl<T> &g, int epi, bool rz, bool zi, const MixView &mx, int nd);
// Level 0 between the two FP32 cycles of a mixed step (k_sup epilogue 4;
// nu1 pre-sweeps recomputed and run on the next cycle, 2 nu1 + NS <= 7).
template <int NS> void launch_up_down0_ns(Ctx *c, Lvl<float> &f, Lvl<float> &g, int nu1);
// First FP32 descent of a mixed solve (NS pre-sweeps, not stored), its
// right-hand side the FP64 residual of mx.u64in (k_sdown_r64).  Returns the
// partials written.
template <int NS> int launch_down_r64_ns(Ctx *c, Lvl<float> &f, Lvl<float> &g, const MixView &mx);
// Transposed cycle on FP64 level f above the coarsest (mg_k_adj.hip): A, the
// transposed post-smoothing (ns sweeps) and prolongation into the coarse
// level (ubc = P^T ubar, rbc = 0); B, the transposed restriction and residual
// from the coarse rbc and the transposed pre-smoothing (ns sweeps).  ubar
// out of place (ubin -> ubout), rbar in place (rb).
void launch_tdown(Ctx *c, int ns, const Lvl<double> &f, const double *ubin, double *ubout, double *rb,
                  const Lvl<double> &g, double *ubc, double *rbc);
void launch_tup(Ctx *c, int ns, const Lvl<double> &f, const double *ubin, double *ubout, double *rb,
                const Lvl<double> &g, double *rbc);

}  // namespace vlp2d
