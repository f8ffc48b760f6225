// Register-streaming fused multigrid kernels for gfx950.
//
// One wavefront owns a tile of R output rows by TJ = 128 - 2H output
// columns of one level (R is chosen per launch so that the tiles fill the
// GPU's resident wavefront slots about once).  It stages a strip of 128
// columns (lane m holds the column pair cb + 2m, cb + 2m + 1) and walks it
// top to bottom, one staged row per step, LR = R + 2H rows in all.  Every
// row is loaded from HBM once per tile; all red-black Gauss-Seidel
// half-sweeps of the launch are pipelined over the rows (temporal blocking in
// registers): at step q the wavefront loads row q + PF (prefetch), applies
// half-sweep s to row q - s for s = 1..S, and runs the epilogue on the rows
// the last half-sweep has finished.  Vertical neighbours are registers of the
// same lane; horizontal neighbours of the other lane come through DPP
// wavefront shifts (wave_shr:1 / wave_shl:1).  No LDS; the only barrier
// keeps the kWaves wavefronts of a workgroup (adjacent strips) in step.
//
// Red-black bookkeeping.  The tile origin (gi0, cb) has both coordinates
// even, so node (gi0 + p, cb + 2m + e) has colour (p + e) & 1 (0 = red,
// updated first, VisualLaplacePoisson2D/Model/CPU/MultiGridSlidingIteration.cs:
// 53-57, GridIterator.cs:184-188).  Half-sweep s updates colour
// c_s = (s - 1) & 1, one node per lane and row.  A Gauss-Seidel update reads
// only the four neighbours, which are all of the other colour, so each
// half-sweep level keeps just the value it updated: U[s][p] is the colour-c_s
// node of row p after half-sweep s (element (p + c_s) & 1 of the pair), and
// the full state of row p after half-sweep s is {U[s][p], U[s-1][p]}.
// Levels -1 and 0 are the staged values (red and black).  A node that is not
// updated (domain boundary) carries U[s-2][p].
//
// Within one colour the update is order independent, and every formula is
// evaluated in the reference's operation order, so the result is bitwise the
// reference's multigrid.  The weighted-Jacobi smoother (an addition, SM_JAC)
// runs in the same pipeline with both elements of a row per sweep level
// (JacState).  Rows and columns within distance s of the strip
// edge hold partial results after half-sweep s; the halo H covers them.
//
// Every tile keeps Dirichlet nodes fixed with per-lane column predicates
// (EDGE: one select per half-sweep; a separate body without them for tiles
// whose columns are all interior measured no faster).  Row checks (Dirichlet
// rows, rows outside the domain, predicated loads) are needed only in the
// loop bodies of tiles at the top and bottom of the domain: each loop body of
// P rows runs in one of two compiled forms, with row checks (ROWS) or, when
// every row it touches is an interior row, without them and with
// unconditional loads.  Row checks depend on ROWS alone and column checks on
// EDGE alone.
//
// Kernels:
//   k_sdown : [zero or staged iterate] -> S = 2*NS half-sweeps -> residual
//             and half-weighting restriction into the coarse right-hand side
//             (MultiGridScheme.cs:134-169) -> write the tile.  H = S + 2.
//   k_sup   : staged iterate + additive bilinear prolongation of the coarse
//             correction (MultiGridScheme.cs:171-186) -> S half-sweeps ->
//             write the tile, with epilogue EPI 0 (none), 1 (partial sums of
//             the squared residual; H = S + 2), 2 (partial max of the update
//             |u - u_after_prolongation|, MultiGridScheme.cs:69,111) or 3
//             (FP32 correction e on level 0 of a mixed-precision step: the
//             FP64 refinement u64 += e, r32 = (float)(fn - Lap u64), partial
//             sums of r^2 and partial max |e|; e itself is not written;
//             H = S + 2) or 4 (the level-0 correction e of the first FP32
//             cycle of a two-cycle mixed step, continued into the second
//             cycle's descent on the same rows: NB pre-sweeps on e, residual
//             and half-weighting restriction into the coarse right-hand side
//             a.rc; the pre-smoothed iterate, not e, is written, and the
//             second cycle's ascent stages it; H = S + 2NB + 2).
// On a zero-initialised level the pre-smoothed iterate is a local function
// of the right-hand side (radius 2 NS1 for NS1 pre-sweeps).  There the
// descent may skip writing it (k_sdown STORE = false) and the ascent
// recomputes it in front of its own pipeline (k_sup N1 = NS1): 2 N1 more
// half-sweeps on the right-hand-side rows it streams anyway, H + 2 N1, and
// one write plus one read of the level less per cycle.
#pragma once
#include <algorithm>
#include <type_traits>

#include "mg_tiles.hpp"


namespace vlp2d {

constexpr int SCOLS = 128;  // staged columns per strip (64 lanes x 2)

// Wavefronts per workgroup: kWaves tiles that are neighbours in tile order
// (adjacent strips of one row band) run in one workgroup and step through
// their rows together (one s_barrier per step, no memory fence), so the
// cache lines their halos share are requested together (measured on MI355X
// at 8193^2: 4 waves fetch 7% less than 1 on the level-0 ascent).
#ifndef VLP2D_STREAM_WAVES
#define VLP2D_STREAM_WAVES 4
#endif
constexpr int kWaves = VLP2D_STREAM_WAVES;
__device__ __forceinline__ void group_step() {
  if constexpr (kWaves > 1) __builtin_amdgcn_s_barrier();
}

// Steps per unrolled loop body: a multiple of the colour period (2), the
// level rings (6) and the right-hand-side ring (S + 4 rows live).
// Prefetch distance: staged row q + PF is loaded at step q.  kPf is the
// default of the streaming kernels' PF parameter; a launcher may pick another
// per kernel (measured: 3 suits the latency-bound FP32 level-1 ascent and the
// first FP32 descent with the FP64 residual, 2 everything else).
#ifndef VLP2D_PF
#define VLP2D_PF 2
#endif
constexpr int kPf = VLP2D_PF;
__host__ __device__ constexpr int pf_ring(int pf) { return pf <= 2 ? 3 : 6; }  // prefetch ring slots (a divisor of every period)
constexpr int kNru = pf_ring(kPf);
static_assert(kPf >= 1 && kPf <= 5, "prefetch distance");
// Smoothers of the streaming kernels: SM_GS, red-black Gauss-Seidel (the
// reference's), 2 ns half-sweep levels; SM_JAC, weighted Jacobi (an
// addition), ns sweep levels.  A level shrinks the exact region of a strip
// by one row and column on each side, so the halo is the number of levels
// (+2 for a residual and restriction, +1 or 2 for a residual), rounded up
// to an even number (the tile origin stays on an even row and column).
constexpr int SM_GS = 0, SM_JAC = 1;
__host__ __device__ constexpr int sweep_levels(int ns, int sm) { return sm == SM_JAC ? ns : 2 * ns; }
__host__ __device__ constexpr int even_up(int h) { return (h + 1) / 2 * 2; }
__host__ __device__ constexpr int s_period(int ns, int sm = SM_GS, int pf = kPf) {
  return sweep_levels(ns, sm) + pf + 2 <= 6    ? 6
         : sweep_levels(ns, sm) + pf + 2 <= 12 ? 12
         : sweep_levels(ns, sm) + pf + 2 <= 18 ? 18
                                               : 24;
}
__host__ __device__ constexpr int s_down_halo(int ns, int sm = SM_GS) { return even_up(sweep_levels(ns, sm) + 2); }
// nd > 0 (epilogue 3 only): the ascent also runs the next refinement step's
// first descent (nd pre-sweeps from zero, residual, restriction) on the
// FP32 residual it produces, which needs that residual 2 nd + 1 nodes
// around the tile's output region, hence the FP64 iterate one node further
// and the final correction there: 2 nd + 2 more levels of halo.
__host__ __device__ constexpr int s_up_halo(int ns, int epi, int sm = SM_GS, int nd = 0) {
  return (epi == 3 && nd > 0) ? even_up(sweep_levels(ns, sm) + 2 * nd + 2)
         : (epi == 1 || epi == 3 || epi == 4)
             ? even_up(sweep_levels(ns, sm) + 1)
             : (sweep_levels(ns, sm) < 2 ? 2 : even_up(sweep_levels(ns, sm)));
}
// Output rows per tile R (a launch parameter): the staged rows R + 2H must
// be a whole number of loop bodies; the smallest such R >= want.
inline int s_rows_at_least(int want, int H, int P) {
  const int lr = ((std::max(want, 2) + 2 * H + P - 1) / P) * P;
  return lr - 2 * H;
}

struct STileGrid {
  int tiles_i, tiles_j, tiles, chunk;
};
inline STileGrid s_grid(int nx, int ny, int R, int H) {
  STileGrid g;
  const int TJ = SCOLS - 2 * H;
  g.tiles_i = (nx - 2) / R + 1;  // output rows [ti R, ti R + R) cover 1..nx-2
  g.tiles_j = (ny - 2) / TJ + 1;
  g.tiles = g.tiles_i * g.tiles_j;
  g.chunk = ((g.tiles + kWaves - 1) / kWaves + 7) / 8;  // workgroups per XCD
  return g;
}

// ---- lane shifts (DPP) --------------------------------------------------
// from_left(x): lane m receives lane m-1's x (lane 0 receives 0);
// from_right(x): lane m receives lane m+1's x (lane 63 receives 0).
// Called with all 64 lanes active.
__device__ inline int dpp_shr1(int x) { return __builtin_amdgcn_update_dpp(0, x, 0x138, 0xF, 0xF, true); }
__device__ inline int dpp_shl1(int x) { return __builtin_amdgcn_update_dpp(0, x, 0x130, 0xF, 0xF, true); }
__device__ inline float from_left(float x) { return __int_as_float(dpp_shr1(__float_as_int(x))); }
__device__ inline float from_right(float x) { return __int_as_float(dpp_shl1(__float_as_int(x))); }
__device__ inline double from_left(double x) {
  const long long b = __double_as_longlong(x);
  const int lo = dpp_shr1((int)b), hi = dpp_shr1((int)(b >> 32));
  return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}
__device__ inline double from_right(double x) {
  const long long b = __double_as_longlong(x);
  const int lo = dpp_shl1((int)b), hi = dpp_shl1((int)(b >> 32));
  return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}

template <int N> __host__ __device__ constexpr int md(int x) { return ((x % N) + N) % N; }

// compile-time loop: f(integral_constant<int, 0>) ... f(integral_constant<int, N-1>)
template <int I, int N, typename F> __device__ __forceinline__ void static_for(F &&f) {
  if constexpr (I < N) {
    f(std::integral_constant<int, I>{});
    static_for<I + 1, N>(f);
  }
}

// Point update of the reference (MultiGridSlidingIteration.cs:53-57):
// EQ (hx == hy): (uim + uip + ujm + ujp - f hx^2) / 4; otherwise
// coef ((uim + uip)/hx^2 + (ujm + ujp)/hy^2 - f).  Laplace levels carry f = 0,
// which leaves both forms bitwise unchanged (x - (+0) == x).
template <typename T, bool EQ, bool PW2>
__device__ __forceinline__ T s_gs(T uim, T uip, T ujm, T ujp, T f, const LevelView<T> &a) {
  if constexpr (EQ) {
    return (uim + uip + ujm + ujp - f * a.hx2) * T(0.25);
  } else {
    if constexpr (PW2) return a.coef * ((uim + uip) * a.ihx2 + (ujm + ujp) * a.ihy2 - f);
    else return a.coef * ((uim + uip) / a.hx2 + (ujm + ujp) / a.hy2 - f);
  }
}

// Residual rhs - Lap u of the reference (UtilsOpLap.cs:15-20).
template <typename T, bool PW2>
__device__ __forceinline__ T s_res(T u, T uim, T uip, T ujm, T ujp, T rhs, const LevelView<T> &a) {
  const T u2 = u * T(2);
  if constexpr (PW2) return rhs - ((uim + uip - u2) * a.ihx2 + (ujm + ujp - u2) * a.ihy2);
  else return rhs - ((uim + uip - u2) / a.hx2 + (ujm + ujp - u2) / a.hy2);
}

template <typename T> struct Pair {
  T x0, x1;
};

// Pair load / store at byte offset boff of a row (the row pointer is
// wavefront-uniform, the offset per lane).
template <typename T> __device__ __forceinline__ Pair<T> ld_pair(const T *row, unsigned boff) {
  const char *p = reinterpret_cast<const char *>(row) + boff;
  if constexpr (sizeof(T) == 8) {
    const double2 v = *reinterpret_cast<const double2 *>(p);
    return Pair<T>{v.x, v.y};
  } else {
    const float2 v = *reinterpret_cast<const float2 *>(p);
    return Pair<T>{v.x, v.y};
  }
}
template <typename T> __device__ __forceinline__ void st_pair(T *row, unsigned boff, T x0, T x1) {
  char *p = reinterpret_cast<char *>(row) + boff;
  if constexpr (sizeof(T) == 8) *reinterpret_cast<double2 *>(p) = make_double2(x0, x1);
  else *reinterpret_cast<float2 *>(p) = make_float2(x0, x1);
}

// Per-lane constants of a strip.
struct Lane {
  int m;          // lane
  int gj;         // column of element 0 (even); element 1 is gj + 1
  bool ld;        // the pair lies inside the row allocation
  bool in0, in1;  // interior columns
  bool oc;        // output column pair
  unsigned boff;  // byte offset of the pair within a row (0 when !ld)
};

template <typename T> __device__ inline Lane make_lane(int cb, int H, int TJ, int ny, int pitch) {
  Lane L;
  L.m = threadIdx.x & 63;
  L.gj = cb + 2 * L.m;
  L.ld = L.gj >= 0 && L.gj <= pitch - 2;
  L.in0 = L.gj >= 1 && L.gj <= ny - 2;
  L.in1 = L.gj + 1 >= 1 && L.gj + 1 <= ny - 2;
  L.oc = L.m >= H / 2 && L.m < H / 2 + TJ / 2;
  L.boff = L.ld ? (unsigned)L.gj * (unsigned)sizeof(T) : 0u;
  return L;
}

template <typename T> __device__ __forceinline__ const T *row_of(const T *base, int gi, int pitch) {
  return base + gi * pitch;
}
template <typename T> __device__ __forceinline__ T *row_of(T *base, int gi, int pitch) { return base + gi * pitch; }

// Store the final pair of row gi where the nodes are output nodes (interior
// nodes of the domain; every node of an interior tile is).
template <typename T, bool EDGE>
__device__ __forceinline__ void s_store(T *uout, int pitch, int gi, const Lane &L, T x0, T x1) {
  T *row = row_of(uout, gi, pitch);
  if constexpr (!EDGE) {
    if (L.oc) st_pair(row, L.boff, x0, x1);
  } else {
    const bool o0 = L.oc && L.in0, o1 = L.oc && L.in1;
    if (o0 && o1) st_pair(row, L.boff, x0, x1);
    else if (o0) row[L.gj] = x0;
    else if (o1) row[L.gj + 1] = x1;
  }
}

// ---------------------------------------------------------------------------
// Sweep state shared by both kernels.  Ring slots are compile-time: row p
// lives in slot p mod 6 of each level ring (a level-s row is read up to
// three steps after the next row of level s-1 is written), p mod P of the
// right-hand-side ring, p mod 3 of the prefetch ring (staged row q = q0 + J
// with q0 a multiple of P, itself a multiple of 6).  The rings start at zero;
// rows above the first staged row then read zeros, and their results lie in
// the halo.
// ---------------------------------------------------------------------------
template <typename T, int NS, int PFD = kPf> struct SweepState {
  static constexpr int S = 2 * NS;
  static constexpr int PF = PFD, NRU = pf_ring(PFD);
  static constexpr int P = s_period(NS, SM_GS, PF);
  static_assert(S + PF + 2 <= P, "right-hand-side ring too short for the pipeline");
  T U[S + 2][6];  // U[s + 1][p % 6]: level s in -1..S
  T F[P][2];      // right-hand side of row p (0 on Laplace levels and outside the domain)
  T RU[NRU][2];   // prefetched staged rows (raw)
  __device__ void zero() {
#pragma unroll
    for (int s = 0; s < S + 2; ++s)
#pragma unroll
      for (int k = 0; k < 6; ++k) U[s][k] = T(0);
#pragma unroll
    for (int k = 0; k < P; ++k) F[k][0] = F[k][1] = T(0);
#pragma unroll
    for (int k = 0; k < NRU; ++k) RU[k][0] = RU[k][1] = T(0);
  }
  // full value of element e of row p = q - d (J the step phase) after
  // half-sweep s
  template <int s, int J, int d, int e> __device__ __forceinline__ T full() {
    constexpr int p = J - d;  // row index mod P (only parity and mod 6 matter)
    constexpr int cs = md<2>(s - 1);
    constexpr int es = md<2>(p + cs);
    if constexpr (e == es) return U[s + 1][md<6>(p)];
    else return U[s][md<6>(p)];
  }
  // staged pair of row q (phase J): level -1 the red element, 0 the black
  template <int J> __device__ __forceinline__ void stage(T x0, T x1) {
    U[0][md<6>(J)] = (J & 1) == 0 ? x0 : x1;
    U[1][md<6>(J)] = (J & 1) == 0 ? x1 : x0;
  }
};

// Weighted-Jacobi sweep state: level s (0 staged, s after sweep s) keeps
// both elements of each row, since a Jacobi update reads the previous
// iterate at all four neighbours; the same ring slots as SweepState.
template <typename T, int NS, int PFD = kPf> struct JacState {
  static constexpr int S = NS;
  static constexpr int PF = PFD, NRU = pf_ring(PFD);
  static constexpr int P = s_period(NS, SM_JAC, PF);
  static_assert(S + PF + 2 <= P, "right-hand-side ring too short for the pipeline");
  T U[S + 1][6][2];
  T F[P][2];
  T RU[NRU][2];
  __device__ void zero() {
#pragma unroll
    for (int s = 0; s < S + 1; ++s)
#pragma unroll
      for (int k = 0; k < 6; ++k) U[s][k][0] = U[s][k][1] = T(0);
#pragma unroll
    for (int k = 0; k < P; ++k) F[k][0] = F[k][1] = T(0);
#pragma unroll
    for (int k = 0; k < NRU; ++k) RU[k][0] = RU[k][1] = T(0);
  }
  template <int s, int J, int d, int e> __device__ __forceinline__ T full() { return U[s][md<6>(J - d)][e]; }
  template <int J> __device__ __forceinline__ void stage(T x0, T x1) {
    U[0][md<6>(J)][0] = x0;
    U[0][md<6>(J)][1] = x1;
  }
};
template <typename T, int NS, int SM, int PF = kPf>
using SState = typename std::conditional<SM == SM_JAC, JacState<T, NS, PF>, SweepState<T, NS, PF>>::type;

// Apply half-sweeps SA..SB (default 1..S) at step J (row q = q0 + J):
// half-sweep s updates row q - s, using level s-1 rows q-s-1, q-s, q-s+1.
// On edge tiles nodes outside the domain interior keep their value.
template <typename T, int NS, int PF, bool EQ, bool PW2, bool EDGE, bool ROWS, int J, int SA = 1, int SB = 2 * NS>
__device__ __forceinline__ void s_sweeps(SweepState<T, NS, PF> &st, const LevelView<T> &a, const Lane &L, int gi0,
                                         int q) {
  constexpr int P = SweepState<T, NS, PF>::P;
  static_for<SA, SB + 1>([&](auto sc) {
    constexpr int s = decltype(sc)::value;
    constexpr int p = J - s;  // phase of the row
    constexpr int cs = md<2>(s - 1);
    constexpr int e = md<2>(p + cs);  // element holding colour cs in this row
    const T above = st.U[s][md<6>(p - 1)];
    const T below = st.U[s][md<6>(p + 1)];
    const T mid = st.U[s][md<6>(p)];
    T ujm, ujp;
    if constexpr (e == 0) {
      ujm = from_left(mid);
      ujp = mid;
    } else {
      ujm = mid;
      ujp = from_right(mid);
    }
    const T g = s_gs<T, EQ, PW2>(above, below, ujm, ujp, st.F[md<P>(p)][e], a);
    if constexpr (EDGE || ROWS) {
      bool ok = true;
      if constexpr (EDGE) ok = e == 0 ? L.in0 : L.in1;
      if constexpr (ROWS) {
        const int gi = gi0 + q - s;
        ok = ok && gi >= 1 && gi <= a.nx - 2;
      }
      st.U[s + 1][md<6>(p)] = ok ? g : st.U[s - 1][md<6>(p)];
    } else {
      st.U[s + 1][md<6>(p)] = g;
    }
  });
}

// Weighted-Jacobi sweeps SA..SB at step J: sweep s updates both elements of
// row q - s from level s-1 rows q-s-1, q-s, q-s+1, u + w (g - u) with g the
// Gauss-Seidel point formula (the oracle's jacobi_sweep).
template <typename T, int NS, int PF, bool EQ, bool PW2, bool EDGE, bool ROWS, int J, int SA, int SB>
__device__ __forceinline__ void j_sweeps(JacState<T, NS, PF> &st, const LevelView<T> &a, const Lane &L, int gi0,
                                         int q) {
  constexpr int P = JacState<T, NS, PF>::P;
  static_for<SA, SB + 1>([&](auto sc) {
    constexpr int s = decltype(sc)::value;
    constexpr int p = J - s;
    const T o0 = st.U[s - 1][md<6>(p)][0], o1 = st.U[s - 1][md<6>(p)][1];
    const T g0 = s_gs<T, EQ, PW2>(st.U[s - 1][md<6>(p - 1)][0], st.U[s - 1][md<6>(p + 1)][0], from_left(o1), o1,
                                  st.F[md<P>(p)][0], a);
    const T g1 = s_gs<T, EQ, PW2>(st.U[s - 1][md<6>(p - 1)][1], st.U[s - 1][md<6>(p + 1)][1], o0, from_right(o0),
                                  st.F[md<P>(p)][1], a);
    T n0 = o0 + a.w * (g0 - o0), n1 = o1 + a.w * (g1 - o1);
    if constexpr (EDGE || ROWS) {
      bool ok0 = true, ok1 = true;
      if constexpr (EDGE) {
        ok0 = L.in0;
        ok1 = L.in1;
      }
      if constexpr (ROWS) {
        const int gi = gi0 + q - s;
        const bool rin = gi >= 1 && gi <= a.nx - 2;
        ok0 = ok0 && rin;
        ok1 = ok1 && rin;
      }
      n0 = ok0 ? n0 : o0;
      n1 = ok1 ? n1 : o1;
    }
    st.U[s][md<6>(p)][0] = n0;
    st.U[s][md<6>(p)][1] = n1;
  });
}

// Sweep levels SA..SB of either smoother at step J.
template <bool EQ, bool PW2, bool EDGE, bool ROWS, int J, int SA, int SB, typename T, int NS, int PF>
__device__ __forceinline__ void sweeps(SweepState<T, NS, PF> &st, const LevelView<T> &a, const Lane &L, int gi0,
                                       int q) {
  s_sweeps<T, NS, PF, EQ, PW2, EDGE, ROWS, J, SA, SB>(st, a, L, gi0, q);
}
template <bool EQ, bool PW2, bool EDGE, bool ROWS, int J, int SA, int SB, typename T, int NS, int PF>
__device__ __forceinline__ void sweeps(JacState<T, NS, PF> &st, const LevelView<T> &a, const Lane &L, int gi0,
                                       int q) {
  j_sweeps<T, NS, PF, EQ, PW2, EDGE, ROWS, J, SA, SB>(st, a, L, gi0, q);
}

// Load the raw pair of staged row q + PF (u unless zero-initialised) and
// its right-hand side (PF: the prefetch distance of the state St).
// ROWS = false: every row this loop body touches is an interior row of the
// domain, so the load needs no predicate: lanes outside the row allocation
// read its columns 0..1 (L.boff = 0), values that reach no interior node.
// FBR: the right-hand-side load sits behind a uniform branch (skipped on
// Laplace levels).  The branch makes the compiler wait for the previous
// step's loads and stores before the step goes on; without it more memory
// operations stay in flight.  Measured on MI355X at 8193^2, the bandwidth-
// bound descents and level-0 ascents with a residual epilogue are faster with
// the branch (refinement ascent 509 against 541 us, FP64 ascent 360 against
// 371 us), the latency-bound ascents without it (the cycle-boundary kernel
// 177 against 218 us, the FP64 level-1 ascent 70 against 76 us).
template <bool ZI, bool EDGE, bool ROWS, int J, bool FBR, typename St, typename T>
__device__ __forceinline__ void s_prefetch(St &st, const LevelView<T> &a, const Lane &L, int gi0, int q, int LR) {
  constexpr int P = St::P, PF = St::PF, NRU = St::NRU;
  const int qq = q + PF, gi = gi0 + qq;
  Pair<T> u{T(0), T(0)}, f{T(0), T(0)};
  if constexpr (ROWS) {
    bool go = qq < LR && gi >= 0 && gi <= a.nx - 1;
    if constexpr (EDGE) go = go && L.ld;
    if (go) {
      if constexpr (!ZI) u = ld_pair(row_of(a.uin, gi, a.pitch), L.boff);
      if constexpr (FBR) {
        if (a.fpitch) f = ld_pair(row_of(a.f, gi, a.fpitch), L.boff);
      } else {
        f = ld_pair(row_of(a.f, gi, a.fpitch), L.boff);
      }
    }
  } else {
    if constexpr (!ZI) u = ld_pair(row_of(a.uin, gi, a.pitch), L.boff);
    if constexpr (FBR) {
      if (a.fpitch) f = ld_pair(row_of(a.f, gi, a.fpitch), L.boff);
    } else {
      f = ld_pair(row_of(a.f, gi, a.fpitch), L.boff);
    }
  }
  st.RU[md<NRU>(J + PF)][0] = u.x0;
  st.RU[md<NRU>(J + PF)][1] = u.x1;
  st.F[md<P>(J + PF)][0] = f.x0;
  st.F[md<P>(J + PF)][1] = f.x1;
}

// Tile of this workgroup: output rows [ti R, ti R + R), columns
// [tj TJ, tj TJ + TJ); staged origin (gi0, cb) = (ti R - H, tj TJ - H).
// An interior tile stages only interior nodes of the domain.
struct STile {
  int t, gi0, cb;
};
template <int H> __device__ __forceinline__ bool s_tile(int R, int tiles_i, int tiles_j, int chunk, int nx, int ny,
                                                        STile &tl) {
  constexpr int TJ = SCOLS - 2 * H;
  const int LR = R + 2 * H;
  {
    // workgroup -> XCD as tile_of_block, kWaves consecutive tiles per workgroup
    const int b = blockIdx.x;
    // the wavefront's index in its workgroup is uniform across the wave:
    // readfirstlane makes the tile, its rows and its row addresses scalar,
    // so the choice of loop body below is a scalar branch
    const int w = __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
    const int t = ((b & 7) * chunk + (b >> 3)) * kWaves + w;
    tl.t = t < tiles_i * tiles_j ? t : -1;
  }
  if (tl.t < 0) return false;
  int ti, tj;
  tile_coords(tl.t, tiles_i, tiles_j, ti, tj);
  tl.gi0 = ti * R - H;
  tl.cb = tj * TJ - H;
  return true;
}

// A wavefront without a tile still takes part in its workgroup's LR steps.
__device__ __forceinline__ void s_idle(int LR) {
  if constexpr (kWaves > 1)
    for (int q = 0; q < LR; ++q) group_step();
}

// ---------------------------------------------------------------------------
// FP64 residual as the right-hand side of a zero-initialised FP32 descent (the
// first descent of a mixed-precision solve, R64): staged row m gets
// f = (float)(fn - Lap u64), evaluated in FP64 from the iterate rows m-1..m+1
// and the right-hand-side row m in the order of UtilsOpLap.cs:15-20 (the
// formula of the refinement epilogue and of k_refine_stream), with the rows
// loaded two steps ahead into rings.  The tile's output nodes also get r32
// written and r^2 summed (the initial residual norm of the solve).
// ---------------------------------------------------------------------------
#ifndef VLP2D_R64D
#define VLP2D_R64D 2
#endif
constexpr int kR64D = VLP2D_R64D;  // steps between the FP64 loads of a row and its residual
static_assert(kR64D >= 1 && kR64D <= 3, "FP64 prefetch distance of the residual descent");
struct R64State {
  double X[6][2];  // iterate rows, slot m % 6 (row m lives from step m - PF - 1 - kR64D to m - PF + 1)
  double F[6][2];  // right-hand-side rows, slot m % 6 (step m - PF - kR64D to m - PF)
  double acc;      // sum of r^2 over the tile's output nodes
};

// Step J (row q): loads iterate row q + PF + 1 + kR64D and right-hand-side
// row q + PF + kR64D, then (J >= -PF) evaluates the residual of staged row
// m = q + PF into st.F (PF: the state's prefetch distance).  The pre-loop
// runs J = -PF - 2 - kR64D .. -1 with q = J.
template <typename T, int NS, int PF, bool EDGE, bool ROWS, bool PW2, int J>
__device__ __forceinline__ void r64_stage(SweepState<T, NS, PF> &st, R64State &rs, const LevelView<T> &a,
                                          const MixView &mx, const Lane &L, unsigned boff8, int gi0, int q, int R,
                                          int H) {
  constexpr int P = SweepState<T, NS, PF>::P;
  {
    const int mu = q + PF + 1 + kR64D, gu = gi0 + mu;
    const int mf = q + PF + kR64D, gf = gi0 + mf;
    Pair<double> x{0.0, 0.0}, f{0.0, 0.0};
    if constexpr (ROWS) {
      bool gx = gu >= 0 && gu <= a.nx - 1, gff = mf >= 0 && gf >= 0 && gf <= a.nx - 1;
      if constexpr (EDGE) {
        gx = gx && L.ld;
        gff = gff && L.ld;
      }
      if (gx) x = ld_pair(row_of(mx.u64in, gu, a.pitch), boff8);
      if (gff) f = ld_pair(row_of(mx.fn64, gf, a.pitch), boff8);
    } else {
      x = ld_pair(row_of(mx.u64in, gu, a.pitch), boff8);
      f = ld_pair(row_of(mx.fn64, gf, a.pitch), boff8);
    }
    rs.X[md<6>(J + PF + 1 + kR64D)][0] = x.x0;
    rs.X[md<6>(J + PF + 1 + kR64D)][1] = x.x1;
    rs.F[md<6>(J + PF + kR64D)][0] = f.x0;
    rs.F[md<6>(J + PF + kR64D)][1] = f.x1;
  }
  if constexpr (J + PF >= 0) {
    constexpr int jm = J + PF;
    const double u0 = rs.X[md<6>(jm)][0], u1 = rs.X[md<6>(jm)][1];
    const double a0 = rs.X[md<6>(jm - 1)][0], a1 = rs.X[md<6>(jm - 1)][1];
    const double c0 = rs.X[md<6>(jm + 1)][0], c1 = rs.X[md<6>(jm + 1)][1];
    const double l = from_left(u1), rt = from_right(u0);
    const double u20 = u0 * 2.0, u21 = u1 * 2.0;
    double res0, res1;
    if constexpr (PW2) {
      res0 = rs.F[md<6>(jm)][0] - ((a0 + c0 - u20) * mx.ihx2 + (l + u1 - u20) * mx.ihy2);
      res1 = rs.F[md<6>(jm)][1] - ((a1 + c1 - u21) * mx.ihx2 + (u0 + rt - u21) * mx.ihy2);
    } else {
      res0 = rs.F[md<6>(jm)][0] - ((a0 + c0 - u20) / mx.hx2 + (l + u1 - u20) / mx.hy2);
      res1 = rs.F[md<6>(jm)][1] - ((a1 + c1 - u21) / mx.hx2 + (u0 + rt - u21) / mx.hy2);
    }
    const float r0f = (float)res0, r1f = (float)res1;
    st.F[md<P>(jm)][0] = r0f;
    st.F[md<P>(jm)][1] = r1f;
    const int m = q + PF, gm = gi0 + m;
    bool orow = m >= H && m < H + R;
    if constexpr (ROWS) orow = orow && gm >= 1 && gm <= a.nx - 2;
    if (orow && L.oc) {
      const bool o0 = !EDGE || L.in0, o1 = !EDGE || L.in1;
      float *rrow = row_of(mx.r32out, gm, a.pitch);
      if (o0 && o1) st_pair(rrow, L.boff, r0f, r1f);
      else if (o0) rrow[L.gj] = r0f;
      else if (o1) rrow[L.gj + 1] = r1f;
      if (o0) rs.acc += res0 * res0;
      if (o1) rs.acc += res1 * res1;
    }
  }
}

// ---------------------------------------------------------------------------
// Descent
// ---------------------------------------------------------------------------
// SM: smoother (SM_GS, SM_JAC); FW: full-weighting restriction (an addition,
// (1/16)[1 2 1; 2 4 2; 1 2 1], the oracle's restrict_fw) instead of half
// weighting; PF: prefetch distance.
template <typename T, int NS, bool ZI, bool STORE, bool EQ, bool PW2, bool EDGE, bool R64 = false, int SM = SM_GS,
          bool FW = false, int PF = kPf>
__device__ __forceinline__ void sdown_body(const LevelView<T> &a, const STile &tl, T *__restrict__ rhsc, int ncx,
                                           int ncy, int pc, const MixView &mx, double *__restrict__ part) {
  static_assert(!R64 || (ZI && !STORE && sizeof(T) == 4 && SM == SM_GS && !FW),
                "the FP64-residual descent starts the FP32 correction");
  static_assert(SM == SM_GS || STORE, "a Jacobi-smoothed level is stored (the ascent recomputes Gauss-Seidel only)");
  using St = SState<T, NS, SM, PF>;
  constexpr int S = St::S, H = s_down_halo(NS, SM), P = St::P, TJ = SCOLS - 2 * H, NRU = St::NRU;
  const int R = a.rows, LR = R + 2 * H;  // LR is a multiple of P
  const int gi0 = tl.gi0, cb = tl.cb;
  const Lane L = make_lane<T>(cb, H, TJ, a.ny, a.pitch);
  St st;
  st.zero();
  T RR[3][2];  // residual of rows (slot p % 3); element 1 only on even rows unless FW
#pragma unroll
  for (int k = 0; k < 3; ++k) RR[k][0] = RR[k][1] = T(0);
  R64State rs;
  const unsigned boff8 = L.ld ? (unsigned)L.gj * 8u : 0u;
  if constexpr (R64) {
#pragma unroll
    for (int k = 0; k < 6; ++k) rs.X[k][0] = rs.X[k][1] = 0.0;
#pragma unroll
    for (int k = 0; k < 6; ++k) rs.F[k][0] = rs.F[k][1] = 0.0;
    rs.acc = 0.0;
    // iterate rows -1 .. PF + kR64D, right-hand-side rows 0 .. PF + kR64D - 1,
    // F of rows 0 .. PF - 1
    static_for<-PF - 2 - kR64D, 0>([&](auto jc) {
      constexpr int J = decltype(jc)::value;
      r64_stage<T, NS, PF, EDGE, true, PW2, J>(st, rs, a, mx, L, boff8, gi0, J, R, H);
    });
  } else {
    // rows 0 .. PF - 1 (the loop prefetches q + PF)
    static_for<-PF, 0>([&](auto jc) {
      constexpr int J = decltype(jc)::value;
      s_prefetch<ZI, EDGE, true, J, true>(st, a, L, gi0, J, LR);
    });
  }
  const int Jc = L.m + cb / 2;  // coarse column of element 0 (fine column 2 Jc)
  const bool oc_c = EDGE ? L.oc && Jc >= 1 && Jc <= ncy - 2 : L.oc;
  T *const rc = rhsc + Jc;
#pragma unroll 1
  for (int q0 = 0; q0 < LR; q0 += P) {
    // rows q0 - S - 3 .. q0 + P + PF + 4 cover everything this loop body
    // reads or writes; when all are interior rows it runs without row checks
    const bool rows_in = gi0 + q0 - (S + 3) >= 1 && gi0 + q0 + P + PF + 4 <= a.nx - 2;
    auto body = [&](auto rows_c) {
    constexpr bool ROWS = decltype(rows_c)::value;
    static_for<0, P>([&](auto jc) {
      constexpr int J = decltype(jc)::value;
      const int q = q0 + J;
      group_step();
      if constexpr (R64) r64_stage<T, NS, PF, EDGE, ROWS, PW2, J>(st, rs, a, mx, L, boff8, gi0, q, R, H);
      else s_prefetch<ZI, EDGE, ROWS, J, true>(st, a, L, gi0, q, LR);
      st.template stage<J>(st.RU[md<NRU>(J)][0], st.RU[md<NRU>(J)][1]);
      sweeps<EQ, PW2, EDGE, ROWS, J, 1, S>(st, a, L, gi0, q);
      // store row q - S (unless the ascent recomputes it)
      if constexpr (STORE) {
        const int p = q - S, gi = gi0 + p;
        bool go = p >= H && p < H + R;
        if constexpr (ROWS) go = go && gi >= 1 && gi <= a.nx - 2;
        if (go) s_store<T, EDGE>(a.uout, a.pitch, gi, L, st.template full<S, J, S, 0>(), st.template full<S, J, S, 1>());
      }
      // residual of row q - S - 1
      {
        constexpr int d = S + 1;
        constexpr int pr = J - d;
        const T f0 = st.F[md<P>(pr)][0], f1 = st.F[md<P>(pr)][1];
        const T x0 = st.template full<S, J, d, 0>(), x1 = st.template full<S, J, d, 1>();
        const T a0 = st.template full<S, J, d + 1, 0>(), b0 = st.template full<S, J, d - 1, 0>();
        RR[md<3>(pr)][0] = s_res<T, PW2>(x0, a0, b0, from_left(x1), x1, f0, a);
        if constexpr (FW || md<2>(pr) == 0) {
          const T a1 = st.template full<S, J, d + 1, 1>(), b1 = st.template full<S, J, d - 1, 1>();
          RR[md<3>(pr)][1] = s_res<T, PW2>(x1, a1, b1, x0, from_right(x0), f1, a);
        }
      }
      // restriction at even row q - S - 2 (MultiGridScheme.cs:158-167; full
      // weighting: the oracle's restrict_fw, corners summed in its order)
      {
        constexpr int px = J - S - 2;
        if constexpr (md<2>(px) == 0) {
          const int p = q - S - 2;
          const T r0 = RR[md<3>(px)][0], rjp = RR[md<3>(px)][1];
          const T rjm = from_left(rjp);
          const T rim = RR[md<3>(px - 1)][0], rip = RR[md<3>(px + 1)][0];
          const int I = (gi0 + p) / 2;
          bool go = p >= H && p < H + R;
          if constexpr (ROWS) go = go && I >= 1 && I <= ncx - 2;
          if constexpr (FW) {
            const T cpp = RR[md<3>(px + 1)][1], cmp = RR[md<3>(px - 1)][1];
            const T cpm = from_left(cpp), cmm = from_left(cmp);
            if (go && oc_c)
              rc[I * pc] = T(0.0625) * (r0 * T(4) + (rip + rim + rjp + rjm) * T(2) + (cpp + cpm + cmp + cmm));
          } else {
            if (go && oc_c) rc[I * pc] = T(0.125) * (r0 * T(4) + (rip + rim + rjp + rjm));
          }
        }
      }
    });
    };
    if (rows_in) body(std::false_type{});
    else body(std::true_type{});
  }
  if constexpr (R64) {
    const double v = wave_sum(rs.acc);
    if (L.m == 0) part[tl.t] = v;
  }
}

// STORE = false: the smoothed iterate is not written (the ascent recomputes
// it, k_sup with N1 = NS).  SM, FW, PF as sdown_body.
template <typename T, int NS, bool ZI, bool EQ, bool STORE = true, int SM = SM_GS, bool FW = false, int PF = kPf>
__global__ void __launch_bounds__(64 * kWaves) k_sdown(LevelView<T> a, T *__restrict__ rhsc, int ncx, int ncy, int pc) {
  constexpr int H = s_down_halo(NS, SM);
  STile tl;
  if (!s_tile<H>(a.rows, a.tiles_i, a.tiles_j, a.chunk, a.nx, a.ny, tl)) {
    s_idle(a.rows + 2 * H);
    return;
  }
  // every tile runs the column-predicated body
  if (a.pw2)
    sdown_body<T, NS, ZI, STORE, EQ, true, true, false, SM, FW, PF>(a, tl, rhsc, ncx, ncy, pc, MixView{}, nullptr);
  else
    sdown_body<T, NS, ZI, STORE, EQ, false, true, false, SM, FW, PF>(a, tl, rhsc, ncx, ncy, pc, MixView{}, nullptr);
}

// The first FP32 descent of a mixed-precision solve on level 0 (zero
// correction, pre-smoothed iterate not stored), its right-hand side the FP64
// residual of mx.u64in computed as the rows stream in (R64): writes r32 =
// (float)(fn - Lap u64) into mx.r32out on interior nodes, the restricted
// residual into rhsc, and one partial sum of r^2 per tile into part.  The
// FP64 steps of mx are powers of two exactly when the FP32 steps of a are.
// PF: prefetch distance.
template <int NS, bool EQ, int PF = kPf>
__global__ void __launch_bounds__(64 * kWaves) k_sdown_r64(LevelView<float> a, float *__restrict__ rhsc, int ncx,
                                                           int ncy, int pc, MixView mx, double *__restrict__ part) {
  constexpr int H = s_down_halo(NS);
  STile tl;
  if (!s_tile<H>(a.rows, a.tiles_i, a.tiles_j, a.chunk, a.nx, a.ny, tl)) {
    s_idle(a.rows + 2 * H);
    return;
  }
  if (a.pw2)
    sdown_body<float, NS, true, false, EQ, true, true, true, SM_GS, false, PF>(a, tl, rhsc, ncx, ncy, pc, mx, part);
  else
    sdown_body<float, NS, true, false, EQ, false, true, true, SM_GS, false, PF>(a, tl, rhsc, ncx, ncy, pc, mx, part);
}

// ---------------------------------------------------------------------------
// Ascent
// ---------------------------------------------------------------------------
template <typename T, int NS, int N1, int EPI, bool EQ, bool PW2, bool EDGE, bool ZS, int NB, int SM = SM_GS,
          int PF = kPf, int ND = 0>
__device__ __forceinline__ void sup_body(const LevelView<T> &a, const STile &tl, const T *__restrict__ uc, int ncx,
                                         int ncy, int pc, double *__restrict__ part, const MixView &mx) {
  // N1 > 0: the level was pre-smoothed with N1 sweeps by the descent, which
  // did not store it; the pipeline recomputes it (half-sweeps 1..S1 on the
  // staged iterate: zeros when ZS, else a.uin, the iterate the descent
  // started from; the same operations in the same order, hence the same
  // bits) and adds the prolongation to level S1 of each row.
  // half-sweeps: 1..S1 recompute, S1+1..ST post-smoothing, ST+1..SE the
  // next cycle's pre-smoothing (EPI 4)
  // (SM_JAC: NS Jacobi sweep levels, no recomputed or next-cycle sweeps)
  // ND > 0 (EPI 3): the next refinement step's first descent runs on the
  // FP32 residual rows this ascent produces, D = ST + 1 steps behind the
  // staged row: ND Gauss-Seidel sweeps from zero (state S2), residual and
  // half-weighting restriction into a.rc, the coarse level's right-hand
  // side (as k_sdown<T, ND, ZI, EQ, false> on that residual, the same
  // operations in the same order)
  constexpr int NT = N1 + NS + NB, S1 = 2 * N1, ST = S1 + sweep_levels(NS, SM), SE = ST + 2 * NB;
  using St = SState<T, NT, SM, PF>;
  constexpr int H = s_up_halo(NT, EPI, SM, ND), P = St::P, TJ = SCOLS - 2 * H, NRU = St::NRU;
  static_assert(ND == 0 || (EPI == 3 && SM == SM_GS), "the next step's descent follows the refinement epilogue");
  // residual radius around the output region the folded descent reads
  constexpr int RX = ND > 0 ? 2 * ND + 1 : 0, D = ST + 1, SD = 2 * ND;
  // (its right-hand side is written one step before the sweeps read it, not
  // prefetched: distance 1 sizes its ring)
  using St2 = SweepState<T, (ND > 0 ? ND : 1), 1>;
  static_assert(ND == 0 || P % St2::P == 0, "folded descent ring period");
  static_assert(St::S == SE, "sweep levels of the state");
  static_assert(N1 == 0 || EPI != 2, "the update metric needs the staged iterate");
  static_assert((NB > 0) == (EPI == 4), "pre-sweeps of the next cycle belong to epilogue 4");
  static_assert(SM == SM_GS || (N1 == 0 && NB == 0), "recomputed and next-cycle sweeps are Gauss-Seidel's");
  constexpr bool FBR = EPI == 1 || EPI == 3;  // see s_prefetch
  const int R = a.rows, LR = R + 2 * H;  // LR is a multiple of P
  const int gi0 = tl.gi0, cb = tl.cb;
  const Lane L = make_lane<T>(cb, H, TJ, a.ny, a.pitch);
  St st;
  st.zero();
  // coarse rows kk = 0, 1, ... of the strip: global row gi0/2 + kk, column
  // Jc = cb/2 + m (fine column gj = 2 Jc); ring slot kk % 3.  The right
  // coarse neighbour Jc + 1 comes from lane m + 1; lane 63 loads its own (CR),
  // so the staged strip is exact after prolongation up to its last column.
  T C[3], CR[3];
  const int Jc = cb / 2 + L.m;
  const bool last = L.m == 63;
  const bool ldc = EDGE ? Jc >= 0 && Jc <= ncy - 1 : true;
  const bool ldr = last && (EDGE ? Jc + 1 >= 0 && Jc + 1 <= ncy - 1 : true);
  const T *const ucl = uc + (ldc ? Jc : 0);
  const T *const ucr = uc + (ldr ? Jc + 1 : 0);
  // rows_c (ROWS): the coarse row may lie outside the coarse grid; else it is
  // a coarse-grid row, and the clamped columns are loaded unconditionally and
  // masked (no divergent branch)
  auto load_coarse = [&](int kk, T &v, T &vr, auto rows_c) {
    const int I = gi0 / 2 + kk;
    if constexpr (decltype(rows_c)::value) {
      v = T(0);
      vr = T(0);
      if (I >= 0 && I <= ncx - 1) {
        if (ldc) v = ucl[I * pc];
        if (ldr) vr = ucr[I * pc];
      }
    } else {
      const T x = ucl[I * pc], xr = ucr[I * pc];
      v = ldc ? x : T(0);
      vr = ldr ? xr : T(0);
    }
  };
  load_coarse(0, C[0], CR[0], std::true_type{});
  load_coarse(1, C[1], CR[1], std::true_type{});
  C[2] = CR[2] = T(0);
  // EPI 2 keeps the staged (post-prolongation) pair of each row until its
  // last half-sweep
  constexpr int NV = EPI == 2 ? P : 1;
  T V0[NV][2];
#pragma unroll
  for (int k = 0; k < NV; ++k) V0[k][0] = V0[k][1] = T(0);
  double acc = 0.0, emax = 0.0;
  static_assert(EPI != 3 || sizeof(T) == 4, "the refinement epilogue runs on the FP32 correction");
  // EPI 3: FP64 iterate rows (raw, then refined x = u64 + e) and FP64
  // right-hand side; row r is loaded at step r + ST - 2, refined at step
  // r + ST (when e of row r is final), used by the residuals of rows r-1..r+1.
  constexpr int N3 = EPI == 3 ? 3 : 1;
  double X64[N3][2], XR[N3][2], FN[N3][2];
#pragma unroll
  for (int k = 0; k < N3; ++k) X64[k][0] = X64[k][1] = XR[k][0] = XR[k][1] = FN[k][0] = FN[k][1] = 0.0;
  // EPI 3's FP64 rows: only lanes H/2 - 1 .. H/2 + TJ/2 (the output columns
  // and the neighbour column on each side, which the FP64 residual reads;
  // RX more columns on each side with ND) need their values; the others load
  // a needed lane's pair instead of their own halo columns (addresses
  // already requested: no extra traffic)
  constexpr int m64lo = H / 2 - (RX + 2) / 2, m64hi = H / 2 + TJ / 2 + RX / 2;
  const int m64 = L.m < m64lo ? m64lo : (L.m > m64hi ? m64hi : L.m);
  const int gj64 = cb + 2 * m64;
  const unsigned boff8 = gj64 >= 0 && gj64 <= a.pitch - 2 ? (unsigned)gj64 * 8u : 0u;
  // EPI 4: residual of rows (slot p % 3; element 1 only on even rows) and
  // the coarse output column of element 0 (fine column 2 Jc)
  constexpr int N4 = EPI == 4 ? 3 : 1;
  T RR[N4][2];
#pragma unroll
  for (int k = 0; k < N4; ++k) RR[k][0] = RR[k][1] = T(0);
  const bool oc_c = EDGE ? L.oc && Jc >= 1 && Jc <= ncy - 2 : L.oc;
  // ND: the folded descent's state (staged zeros, right-hand side = the FP32
  // residual of row q - D) and its residual rows (slot p % 3)
  St2 s2;
  s2.zero();
  T R2[3][2];
#pragma unroll
  for (int k = 0; k < 3; ++k) R2[k][0] = R2[k][1] = T(0);
  static_for<-PF, 0>([&](auto jc) {
    constexpr int J = decltype(jc)::value;
    s_prefetch<ZS, EDGE, true, J, FBR>(st, a, L, gi0, J, LR);
  });
#pragma unroll 1
  for (int q0 = 0; q0 < LR; q0 += P) {
    // rows q0 - SL .. q0 + P + PF + 4 cover everything this loop body reads
    // or writes (coarse rows included); when all are interior rows it runs
    // without row checks
    constexpr int SL = ND > 0 ? D + SD + 3 : SE + 3;
    const bool rows_in = gi0 + q0 - SL >= 1 && gi0 + q0 + P + PF + 4 <= a.nx - 2;
    auto body = [&](auto rows_c) {
    constexpr bool ROWS = decltype(rows_c)::value;
    static_for<0, P>([&](auto jc) {
      constexpr int J = decltype(jc)::value;
      const int q = q0 + J;
      group_step();
      s_prefetch<ZS, EDGE, ROWS, J, FBR>(st, a, L, gi0, q, LR);
      // staged row q (Gauss-Seidel: level -1 the red element, 0 the black)
      if constexpr (N1 > 0) {
        st.template stage<J>(st.RU[md<NRU>(J)][0], st.RU[md<NRU>(J)][1]);
        sweeps<EQ, PW2, EDGE, ROWS, J, 1, S1>(st, a, L, gi0, q);
      }
      // prolongation of row pr = q - S1, whose level S1 is now complete
      {
        constexpr int JP = md<P>(J - S1);  // phase of row pr (P is a multiple of 6)
        constexpr int kk = JP >> 1;          // coarse row of row pr, mod 3
        const int pr = q - S1, gi = gi0 + pr;
        if constexpr ((JP & 1) == 0)
          load_coarse(pr / 2 + 2, C[md<3>(kk + 2)], CR[md<3>(kk + 2)], std::integral_constant<bool, ROWS>{});
        const T c00 = C[md<3>(kk)];
        const T c01r = from_right(c00);
        const T c01 = last ? CR[md<3>(kk)] : c01r;
        T p0, p1;
        if constexpr ((JP & 1) == 0) {
          p0 = c00;
          p1 = (c00 + c01) * T(0.5);
        } else {
          const T c10 = C[md<3>(kk + 1)];
          const T c11r = from_right(c10);
          const T c11 = last ? CR[md<3>(kk + 1)] : c11r;
          p0 = (c00 + c10) * T(0.5);
          p1 = (c00 + c01 + c10 + c11) * T(0.25);
        }
        bool a0 = true, a1 = true;
        if constexpr (EDGE) {
          a0 = L.in0;
          a1 = L.in1;
        }
        if constexpr (ROWS) {
          const bool rin = gi >= 1 && gi <= a.nx - 2;
          a0 = a0 && rin;
          a1 = a1 && rin;
        }
        if constexpr (N1 == 0) {
          T v0 = st.RU[md<NRU>(J)][0], v1 = st.RU[md<NRU>(J)][1];
          v0 = a0 ? v0 + p0 : v0;
          v1 = a1 ? v1 + p1 : v1;
          st.template stage<J>(v0, v1);
          if constexpr (EPI == 2) {
            V0[md<NV>(J)][0] = v0;
            V0[md<NV>(J)][1] = v1;
          }
        } else {
          // level S1 holds colour 1 (black): element (pr + 1) & 1; the
          // red element is overwritten by half-sweep S1 + 1 without being
          // read (nodes that sweep leaves alone get no prolongation either)
          constexpr int eb = md<2>(JP + 1);
          T &v = st.U[S1 + 1][md<6>(JP)];
          if constexpr (eb == 0) {
            v = a0 ? v + p0 : v;
          } else {
            v = a1 ? v + p1 : v;
          }
        }
      }
      if constexpr (EPI == 3) {
        // prefetch FP64 rows: iterate row q - ST + 2, right-hand side row q - ST + 1
        // only the rows the FP64 residual of the output rows [H, H + R) reads
        // (RX more on each side with ND): iterate rows H - 1 - RX .. H + R + RX,
        // right-hand-side rows H - RX .. H + R - 1 + RX (the slots of the other
        // rows hold zeros no output uses)
        const int r = q - ST + 2, gr = gi0 + r;
        Pair<double> x{0.0, 0.0}, f{0.0, 0.0};
        const int rf = q - ST + 1, gf = gi0 + rf;
        const bool need_x = r >= H - 1 - RX && r <= H + R + RX, need_f = rf >= H - RX && rf < H + R + RX;
        if constexpr (ROWS) {
          bool go = need_x && gr >= 0 && gr <= a.nx - 1;
          if constexpr (EDGE) go = go && L.ld;
          if (go) x = ld_pair(row_of(mx.u64in, gr, a.pitch), boff8);
          bool gof = need_f && gf >= 0 && gf <= a.nx - 1;
          if constexpr (EDGE) gof = gof && L.ld;
          if (gof) f = ld_pair(row_of(mx.fn64, gf, a.pitch), boff8);
        } else {
          // both rows are domain rows
          if (need_x) x = ld_pair(row_of(mx.u64in, gr, a.pitch), boff8);
          if (need_f) f = ld_pair(row_of(mx.fn64, gf, a.pitch), boff8);
        }
        X64[md<3>(J - ST + 2)][0] = x.x0;
        X64[md<3>(J - ST + 2)][1] = x.x1;
        FN[md<3>(J - ST + 1)][0] = f.x0;
        FN[md<3>(J - ST + 1)][1] = f.x1;
      }
      sweeps<EQ, PW2, EDGE, ROWS, J, S1 + 1, ST>(st, a, L, gi0, q);
      // store row q - ST (+ update metric)
      {
        const int p = q - ST, gi = gi0 + p;
        const T x0 = st.template full<ST, J, ST, 0>(), x1 = st.template full<ST, J, ST, 1>();
        bool orow = p >= H && p < H + R;
        if constexpr (ROWS) orow = orow && gi >= 1 && gi <= a.nx - 2;
        if constexpr (EPI != 3 && EPI != 4) {
          if (orow) s_store<T, EDGE>(a.uout, a.pitch, gi, L, x0, x1);
        } else {
          // refined FP64 row q - ST: u64 + e on interior nodes
          constexpr int sx = md<3>(J - ST);
          double r0 = X64[sx][0], r1 = X64[sx][1];
          if constexpr (EDGE || ROWS) {
            bool a0 = true, a1 = true;
            if constexpr (EDGE) {
              a0 = L.in0;
              a1 = L.in1;
            }
            if constexpr (ROWS) {
              const bool rin = gi >= 1 && gi <= a.nx - 2;
              a0 = a0 && rin;
              a1 = a1 && rin;
            }
            r0 = a0 ? r0 + (double)x0 : r0;
            r1 = a1 ? r1 + (double)x1 : r1;
          } else {
            r0 = r0 + (double)x0;
            r1 = r1 + (double)x1;
          }
          XR[sx][0] = r0;
          XR[sx][1] = r1;
        }
        if constexpr (EPI == 2) {
          constexpr int pv = md<NV>(J - ST);
          if (orow && L.oc) {
            if (!EDGE || L.in0) acc = nanmax(acc, fabs((double)(x0 - V0[pv][0])));
            if (!EDGE || L.in1) acc = nanmax(acc, fabs((double)(x1 - V0[pv][1])));
          }
        }
      }
      if constexpr (EPI == 4) {
        // the next cycle's pre-smoothing of the rows just stored, then the
        // residual of row q - SE - 1 and the restriction at even row
        // q - SE - 2 (as k_sdown, MultiGridScheme.cs:134-169)
        sweeps<EQ, PW2, EDGE, ROWS, J, ST + 1, SE>(st, a, L, gi0, q);
        {
          // the next cycle's pre-smoothed iterate of row q - SE, which its
          // ascent stages (instead of recomputing it from this cycle's e)
          const int p = q - SE, gi = gi0 + p;
          bool orow = p >= H && p < H + R;
          if constexpr (ROWS) orow = orow && gi >= 1 && gi <= a.nx - 2;
          if (orow)
            s_store<T, EDGE>(a.uout, a.pitch, gi, L, st.template full<SE, J, SE, 0>(),
                             st.template full<SE, J, SE, 1>());
        }
        {
          constexpr int d = SE + 1;
          constexpr int pr = J - d;
          const T f0 = st.F[md<P>(pr)][0], f1 = st.F[md<P>(pr)][1];
          const T x0 = st.template full<SE, J, d, 0>(), x1 = st.template full<SE, J, d, 1>();
          const T a0 = st.template full<SE, J, d + 1, 0>(), b0 = st.template full<SE, J, d - 1, 0>();
          RR[md<3>(pr)][0] = s_res<T, PW2>(x0, a0, b0, from_left(x1), x1, f0, a);
          if constexpr (md<2>(pr) == 0) {
            const T a1 = st.template full<SE, J, d + 1, 1>(), b1 = st.template full<SE, J, d - 1, 1>();
            RR[md<3>(pr)][1] = s_res<T, PW2>(x1, a1, b1, x0, from_right(x0), f1, a);
          }
        }
        {
          constexpr int px = J - SE - 2;
          if constexpr (md<2>(px) == 0) {
            const int p = q - SE - 2;
            const T r0 = RR[md<3>(px)][0], rjp = RR[md<3>(px)][1];
            const T rjm = from_left(rjp);
            const T rim = RR[md<3>(px - 1)][0], rip = RR[md<3>(px + 1)][0];
            const int I = (gi0 + p) / 2;
            bool go = p >= H && p < H + R;
            if constexpr (ROWS) go = go && I >= 1 && I <= ncx - 2;
            if (go && oc_c) a.rc[I * pc + Jc] = T(0.125) * (r0 * T(4) + (rip + rim + rjp + rjm));
          }
        }
      }
      // refinement of row q - ST - 1: u64out, r32out = (float)(fn - Lap u64out)
      // (UtilsOpLap.cs:15-20 order, FP64), partial sum r^2 and max |e|
      if constexpr (EPI == 3) {
        constexpr int d = ST + 1;
        constexpr int pr = J - d;
        const int p = q - d, gi = gi0 + p;
        const double x0 = XR[md<3>(pr)][0], x1 = XR[md<3>(pr)][1];
        const double a0 = XR[md<3>(pr - 1)][0], a1 = XR[md<3>(pr - 1)][1];
        const double b0 = XR[md<3>(pr + 1)][0], b1 = XR[md<3>(pr + 1)][1];
        const double xl = from_left(x1), xr = from_right(x0);
        const T e0 = st.template full<ST, J, d, 0>(), e1 = st.template full<ST, J, d, 1>();
        bool orow = p >= H && p < H + R;
        if constexpr (ROWS) orow = orow && gi >= 1 && gi <= a.nx - 2;
        auto fp64_residual = [&](double &res0, double &res1) {
          const double u20 = x0 * 2.0, u21 = x1 * 2.0;
          if constexpr (PW2) {
            res0 = FN[md<3>(pr)][0] - ((a0 + b0 - u20) * mx.ihx2 + (xl + x1 - u20) * mx.ihy2);
            res1 = FN[md<3>(pr)][1] - ((a1 + b1 - u21) * mx.ihx2 + (x0 + xr - u21) * mx.ihy2);
          } else {
            res0 = FN[md<3>(pr)][0] - ((a0 + b0 - u20) / mx.hx2 + (xl + x1 - u20) / mx.hy2);
            res1 = FN[md<3>(pr)][1] - ((a1 + b1 - u21) / mx.hx2 + (x0 + xr - u21) / mx.hy2);
          }
        };
        double res0 = 0.0, res1 = 0.0;
        if constexpr (ND > 0) {
          // every lane: the folded descent reads the residual around the
          // output region too (values of lanes and rows it does not read are
          // never used)
          fp64_residual(res0, res1);
          s2.F[md<St2::P>(pr)][0] = (float)res0;
          s2.F[md<St2::P>(pr)][1] = (float)res1;
        }
        if (orow && L.oc) {
          if constexpr (ND == 0) fp64_residual(res0, res1);
          const bool o0 = !EDGE || L.in0, o1 = !EDGE || L.in1;
          double *urow = row_of(mx.u64out, gi, a.pitch);
          float *rrow = row_of(mx.r32out, gi, a.pitch);
          if (o0 && o1) {
            st_pair(urow, boff8, x0, x1);
            st_pair(rrow, L.boff, (float)res0, (float)res1);
          } else if (o0) {
            urow[L.gj] = x0;
            rrow[L.gj] = (float)res0;
          } else if (o1) {
            urow[L.gj + 1] = x1;
            rrow[L.gj + 1] = (float)res1;
          }
          if (o0) {
            acc += res0 * res0;
            emax = nanmax(emax, fabs((double)e0));
          }
          if (o1) {
            acc += res1 * res1;
            emax = nanmax(emax, fabs((double)e1));
          }
        }
      }
      // the next step's first descent (ND): row q - D was just staged in s2
      // (its right-hand side above, zeros as its start); sweeps, residual of
      // row q - D - SD - 1, restriction at even row q - D - SD - 2
      if constexpr (ND > 0) {
        constexpr int J2 = J - D;
        const int q2 = q - D;
        sweeps<EQ, PW2, EDGE, ROWS, J2, 1, SD>(s2, a, L, gi0, q2);
        {
          constexpr int d = SD + 1;
          constexpr int pr = J2 - d;
          const T f0 = s2.F[md<St2::P>(pr)][0], f1 = s2.F[md<St2::P>(pr)][1];
          const T x0 = s2.template full<SD, J2, d, 0>(), x1 = s2.template full<SD, J2, d, 1>();
          const T a0 = s2.template full<SD, J2, d + 1, 0>(), b0 = s2.template full<SD, J2, d - 1, 0>();
          R2[md<3>(pr)][0] = s_res<T, PW2>(x0, a0, b0, from_left(x1), x1, f0, a);
          if constexpr (md<2>(pr) == 0) {
            const T a1 = s2.template full<SD, J2, d + 1, 1>(), b1 = s2.template full<SD, J2, d - 1, 1>();
            R2[md<3>(pr)][1] = s_res<T, PW2>(x1, a1, b1, x0, from_right(x0), f1, a);
          }
        }
        {
          constexpr int px = J2 - SD - 2;
          if constexpr (md<2>(px) == 0) {
            const int p = q2 - SD - 2;
            const T r0 = R2[md<3>(px)][0], rjp = R2[md<3>(px)][1];
            const T rjm = from_left(rjp);
            const T rim = R2[md<3>(px - 1)][0], rip = R2[md<3>(px + 1)][0];
            const int I = (gi0 + p) / 2;
            bool go = p >= H && p < H + R;
            if constexpr (ROWS) go = go && I >= 1 && I <= ncx - 2;
            if (go && oc_c) a.rc[I * pc + Jc] = T(0.125) * (r0 * T(4) + (rip + rim + rjp + rjm));
          }
        }
      }
      // residual norm of row q - ST - 1
      if constexpr (EPI == 1) {
        constexpr int d = ST + 1;
        constexpr int pr = J - d;
        const int p = q - d, gi = gi0 + p;
        const T x0 = st.template full<ST, J, d, 0>(), x1 = st.template full<ST, J, d, 1>();
        const T a0 = st.template full<ST, J, d + 1, 0>(), b0 = st.template full<ST, J, d - 1, 0>();
        const T a1 = st.template full<ST, J, d + 1, 1>(), b1 = st.template full<ST, J, d - 1, 1>();
        const T xl = from_left(x1), xr = from_right(x0);
        bool orow = p >= H && p < H + R;
        if constexpr (ROWS) orow = orow && gi >= 1 && gi <= a.nx - 2;
        if (orow && L.oc) {
          const double r0 = (double)s_res<T, PW2>(x0, a0, b0, xl, x1, st.F[md<P>(pr)][0], a);
          const double r1 = (double)s_res<T, PW2>(x1, a1, b1, x0, xr, st.F[md<P>(pr)][1], a);
          if (!EDGE || L.in0) acc += r0 * r0;
          if (!EDGE || L.in1) acc += r1 * r1;
        }
      }
    });
    };
    if (rows_in) body(std::false_type{});
    else body(std::true_type{});
  }
  if constexpr (EPI == 1 || EPI == 3) {
    acc = wave_sum(acc);
    if (L.m == 0) part[tl.t] = acc;
  } else if constexpr (EPI == 2) {
    acc = wave_max(acc);
    if (L.m == 0) part[tl.t] = acc;
  }
  if constexpr (EPI == 3) {
    emax = wave_max(emax);
    if (L.m == 0) mx.part_max[tl.t] = emax;
  }
}

// EPI 3 requires the FP64 steps of mx to be powers of two exactly when the
// FP32 steps of a are (the host checks).
// N1 > 0 pairs with k_sdown<T, N1, ZS, EQ, false> on the same level: ZS, the
// descent started from zero; otherwise from a.uin, which the ascent then
// stages (it may not write a.uin: only EPI 3, which writes no iterate).
// SM: the smoother (SM_JAC with N1 = NB = 0 only); PF: prefetch distance;
// ND: the next step's first descent folded in (EPI 3, restriction into a.rc).
template <typename T, int NS, int EPI, bool EQ, int N1 = 0, bool ZS = true, int NB = 0, int SM = SM_GS, int PF = kPf,
          int ND = 0>
__global__ void __launch_bounds__(64 * kWaves) k_sup(LevelView<T> a, const T *__restrict__ uc, int ncx, int ncy, int pc,
                                            double *__restrict__ part, MixView mx) {
  constexpr int H = s_up_halo(N1 + NS + NB, EPI, SM, ND);
  constexpr bool Z = ZS && N1 > 0;
  static_assert(ZS || (N1 > 0 && EPI == 3), "a recomputed non-zero start is read back in place: EPI 3 only");
  STile tl;
  if (!s_tile<H>(a.rows, a.tiles_i, a.tiles_j, a.chunk, a.nx, a.ny, tl)) {
    s_idle(a.rows + 2 * H);
    return;
  }
  // Every tile runs the column-predicated body; its loop bodies drop the row
  // checks wherever all their rows are interior rows (measured faster than a
  // separate body for interior tiles, which also needs more registers).
  if (a.pw2) {
    sup_body<T, NS, N1, EPI, EQ, true, true, Z, NB, SM, PF, ND>(a, tl, uc, ncx, ncy, pc, part, mx);
  } else {
    sup_body<T, NS, N1, EPI, EQ, false, true, Z, NB, SM, PF, ND>(a, tl, uc, ncx, ncy, pc, part, mx);
  }
}

}  // namespace vlp2d
