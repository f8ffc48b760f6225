// Register-streaming kernels of the transposed V-cycle (FP64), the reverse
// sweep of the truncated-cycle adjoint.
//
// The forward V(nu1,nu2) cycle (MultiGridScheme.cs:93-115) is an affine map
// of (u, fn, g); its transpose applied to (ubar, rbar) runs the transposed
// operators in reverse order.  On one level l (above the small levels, which
// keep the one-operator kernels of mg_misc_kernels.hpp) it is two launches:
//
//   k_tdown (A): the transposed post-smoothing, 2 nu2 transposed half-sweeps
//     (colours black, red, black, ...), then the transposed prolongation
//     ubar_c = P^T ubar into the coarse level and rbar_c = 0 there;
//   k_tup (B): the transposed restriction resbar = R^T rbar_c, the transposed
//     residual ubar -= L^T resbar, rbar += resbar, then the transposed
//     pre-smoothing, 2 nu1 transposed half-sweeps.
//
// A transposed half-sweep of colour c: every node Q of the other colour
// (boundary nodes included) gathers a * ubar_P from its interior colour-c
// neighbours P (a = ax along i, ay along j); then every interior colour-c node
// P gives rbar_P += -b ubar_P and ubar_P = 0.  Every formula keeps the
// operation order of the one-operator kernels (k_adj_gather, k_adj_color,
// k_adj_prolongT, k_adj_restrictT, k_adj_residualT), so the two paths give
// the same bits.
//
// Tiling as in mg_stream.hpp: a wavefront owns R output rows x (128 - 2H)
// output columns, walks its staged strip top to bottom with all half-sweeps
// pipelined in registers, lane m holding the column pair cb + 2m, cb + 2m + 1.
// Unlike the forward kernels the outputs include the domain's boundary
// nodes, which accumulate the gradient with respect to the Dirichlet data.
// Each level keeps both nodes of a row (3 rows per level), since boundary
// nodes carry values of either colour.  ubar is updated out of place, so that
// overlapping halos read consistent data; rbar never feeds back into ubar, so
// a tile reads and writes it on its own output nodes only, in place.
#pragma once
#include "mg_stream.hpp"

namespace vlp2d {

struct TLevel {
  const double *ubin;  // ubar of the level (read)
  double *ubout;       // its new values (written on the tile's output nodes)
  double *rb;          // rbar of the level, read and written on the tile's output nodes
  int nx, ny, pitch;
  double ax, ay, b;           // smoother coefficients (u_P = ax (uW + uE) + ay (uS + uN) - b fn_P)
  double ix2, iy2;            // 1 / hx^2, 1 / hy^2 (residual)
  int rows, tiles_i, tiles_j, chunk;
};

// coarse operands: k_tdown writes ubc = P^T ubar and rbc = 0; k_tup reads rbc
struct TCoarse {
  double *ubc, *rbc;
  int ncx, ncy, pc;
};

__host__ __device__ constexpr int t_halo(int ns) { return 2 * ns + 2; }  // 2 ns half-sweeps + one stencil, even
__host__ __device__ constexpr int t_period(int ns) { return 2 * ns + kPf + 4 <= 6 ? 6 : 2 * ns + kPf + 4 <= 12 ? 12 : 18; }

// Pipeline state: level s in 0..S holds rows p (slot p % 3, both elements);
// RB holds the rbar accumulation of staged rows (slot p % P); RU the
// prefetched ubar rows.
template <int NS> struct TState {
  static constexpr int S = 2 * NS, P = t_period(NS);
  double X[S + 1][3][2];
  double RB[P][2];
  double RU[kNru][2];
  __device__ void zero() {
#pragma unroll
    for (int s = 0; s <= S; ++s)
#pragma unroll
      for (int k = 0; k < 3; ++k) X[s][k][0] = X[s][k][1] = 0.0;
#pragma unroll
    for (int k = 0; k < P; ++k) RB[k][0] = RB[k][1] = 0.0;
#pragma unroll
    for (int k = 0; k < kNru; ++k) RU[k][0] = RU[k][1] = 0.0;
  }
};

// Per-lane column facts of a strip (both elements).
struct TLane {
  int m, gj;
  bool ld;          // the pair lies inside the row allocation
  bool in0, in1;    // interior columns gj, gj + 1
  bool inm, inp;    // interior columns gj - 1, gj + 2 (the outer neighbours of the pair)
  bool oc0, oc1;    // output nodes of the tile (columns inside the grid)
  unsigned boff;
};
__device__ inline TLane t_lane(int cb, int H, int TJ, int ny, int pitch) {
  TLane L;
  L.m = threadIdx.x & 63;
  L.gj = cb + 2 * L.m;
  L.ld = L.gj >= 0 && L.gj <= pitch - 2;
  L.in0 = L.gj >= 1 && L.gj <= ny - 2;
  L.in1 = L.gj + 1 >= 1 && L.gj + 1 <= ny - 2;
  L.inm = L.gj - 1 >= 1 && L.gj - 1 <= ny - 2;
  L.inp = L.gj + 2 >= 1 && L.gj + 2 <= ny - 2;
  const bool oc = L.m >= H / 2 && L.m < H / 2 + TJ / 2;
  L.oc0 = oc && L.gj >= 0 && L.gj <= ny - 1;
  L.oc1 = oc && L.gj + 1 >= 0 && L.gj + 1 <= ny - 1;
  L.boff = L.ld ? (unsigned)L.gj * 8u : 0u;
  return L;
}

__device__ __forceinline__ bool t_row_in(int gi, int nx) { return gi >= 1 && gi <= nx - 2; }

// Transposed half-sweep s (colour (s & 1) ? black : red) of row p = q - s
// (phase J - s) from level s - 1 (rows p - 1 .. p + 1) into level s; rbar of
// row p accumulates -b ubar_P on its interior colour nodes.
template <int NS, int J, int s>
__device__ __forceinline__ void t_half(TState<NS> &st, const TLevel &a, const TLane &L, int gi0, int q) {
  constexpr int P = TState<NS>::P;
  constexpr int c = s & 1;  // smooth_T: black (1) then red (0) in every sweep
  constexpr int p = J - s;
  constexpr int su = md<3>(p - 1), s0 = md<3>(p), sd = md<3>(p + 1), sp = md<P>(p);
  const int gi = gi0 + q - s;
  const bool rin = t_row_in(gi, a.nx);
  const bool up_in = t_row_in(gi - 1, a.nx), dn_in = t_row_in(gi + 1, a.nx);
  const double xl = from_left(st.X[s - 1][s0][1]), xr = from_right(st.X[s - 1][s0][0]);
  static_for<0, 2>([&](auto ec) {
    constexpr int e = decltype(ec)::value;
    constexpr bool colour_c = md<2>(p + e) == c;  // node (p, e) has colour c (tile origin even)
    const bool cin = e == 0 ? L.in0 : L.in1;
    const double x = st.X[s - 1][s0][e];
    if constexpr (colour_c) {
      // interior P: rbar += -b ubar_P, ubar_P = 0; boundary nodes keep theirs
      const bool inner = rin && cin;
      st.RB[sp][e] = inner ? st.RB[sp][e] + -a.b * x : st.RB[sp][e];
      st.X[s][s0][e] = inner ? 0.0 : x;
    } else {
      // Q gathers from its interior neighbours (k_adj_gather's order)
      const double left = e == 0 ? xl : st.X[s - 1][s0][0], right = e == 0 ? st.X[s - 1][s0][1] : xr;
      const bool lin = e == 0 ? L.inm : L.in0;  // column j - 1 interior
      const bool rgt = e == 0 ? L.in1 : L.inp;  // column j + 1 interior
      double g = 0.0;
      g = up_in && cin ? g + a.ax * st.X[s - 1][su][e] : g;
      g = dn_in && cin ? g + a.ax * st.X[s - 1][sd][e] : g;
      g = lin && rin ? g + a.ay * left : g;
      g = rgt && rin ? g + a.ay * right : g;
      st.X[s][s0][e] = x + g;
    }
  });
}

// Tile geometry: output rows [ti R, ti R + R) and columns [tj TJ, tj TJ + TJ)
// cover every node 0 .. nx-1 x 0 .. ny-1 (tiles_i = (nx - 1) / R + 1, ...).
inline STileGrid t_grid(int nx, int ny, int R, int H) {
  STileGrid g;
  const int TJ = SCOLS - 2 * H;
  g.tiles_i = (nx - 1) / R + 1;
  g.tiles_j = (ny - 1) / TJ + 1;
  g.tiles = g.tiles_i * g.tiles_j;
  g.chunk = ((g.tiles + kWaves - 1) / kWaves + 7) / 8;
  return g;
}

// staged row q + kPf of ubin (RU) and, on output rows and columns, of rb (RB;
// zero on the halo, whose rbar no output depends on)
template <int NS, int J>
__device__ __forceinline__ void t_prefetch(TState<NS> &st, const TLevel &a, const TLane &L, int gi0, int q, int LR) {
  constexpr int P = TState<NS>::P, H = t_halo(NS);
  const int qq = q + kPf, gi = gi0 + qq;
  Pair<double> u{0.0, 0.0}, r{0.0, 0.0};
  if (qq < LR && gi >= 0 && gi <= a.nx - 1) {
    if (L.ld) u = ld_pair(row_of(a.ubin, gi, a.pitch), L.boff);
    if (qq >= H && qq < H + a.rows && (L.oc0 || L.oc1)) r = ld_pair(row_of(a.rb, gi, a.pitch), L.boff);
  }
  st.RU[md<kNru>(J + kPf)][0] = u.x0;
  st.RU[md<kNru>(J + kPf)][1] = u.x1;
  st.RB[md<P>(J + kPf)][0] = r.x0;
  st.RB[md<P>(J + kPf)][1] = r.x1;
}

// write row p (tile-relative) of ubar and rbar if it is an output row
__device__ __forceinline__ void t_store(const TLevel &a, const TLane &L, int gi0, int H, int R, int p, double x0,
                                        double x1, double r0, double r1) {
  const int gi = gi0 + p;
  if (!(p >= H && p < H + R && gi >= 0 && gi <= a.nx - 1)) return;
  double *ur = row_of(a.ubout, gi, a.pitch);
  double *rr = row_of(a.rb, gi, a.pitch);
  if (L.oc0 && L.oc1) {
    st_pair(ur, L.boff, x0, x1);
    st_pair(rr, L.boff, r0, r1);
  } else if (L.oc0) {
    ur[L.gj] = x0;
    rr[L.gj] = r0;
  } else if (L.oc1) {
    ur[L.gj + 1] = x1;
    rr[L.gj + 1] = r1;
  }
}

// ---------------------------------------------------------------------------
// A: transposed post-smoothing + transposed prolongation (k_adj_prolongT's
// order) into the coarse level, whose rbar is set to zero.
// ---------------------------------------------------------------------------
template <int NS>
__global__ void __launch_bounds__(64 * kWaves) k_tdown(TLevel a, TCoarse cg) {
  constexpr int S = 2 * NS, H = t_halo(NS), P = t_period(NS), TJ = SCOLS - 2 * H;
  STile tl;
  if (!s_tile<H>(a.rows, a.tiles_i, a.tiles_j, a.chunk, a.nx, a.ny, tl)) {
    s_idle(a.rows + 2 * H);
    return;
  }
  const int R = a.rows, LR = R + 2 * H;
  const int gi0 = tl.gi0, cb = tl.cb;
  const TLane L = t_lane(cb, H, TJ, a.ny, a.pitch);
  TState<NS> st;
  st.zero();
  static_for<-kPf, 0>([&](auto jc) {
    constexpr int J = decltype(jc)::value;
    t_prefetch<NS, J>(st, a, L, gi0, J, LR);
  });
  // coarse column of element 0 (fine column 2 Jc) and its output flag
  const int Jc = L.m + cb / 2;
  const bool occ = L.m >= H / 2 && L.m < H / 2 + TJ / 2 && Jc >= 0 && Jc <= cg.ncy - 1;
  const bool cin = Jc >= 1 && Jc <= cg.ncy - 2;
#pragma unroll 1
  for (int q0 = 0; q0 < LR; q0 += P) {
    static_for<0, P>([&](auto jc) {
      constexpr int J = decltype(jc)::value;
      const int q = q0 + J;
      group_step();
      t_prefetch<NS, J>(st, a, L, gi0, q, LR);
      st.X[0][md<3>(J)][0] = st.RU[md<kNru>(J)][0];
      st.X[0][md<3>(J)][1] = st.RU[md<kNru>(J)][1];
      static_for<1, S + 1>([&](auto sc) { t_half<NS, J, decltype(sc)::value>(st, a, L, gi0, q); });
      // final row q - S
      {
        constexpr int sf = md<3>(J - S), sr = md<P>(J - S);
        t_store(a, L, gi0, H, R, q - S, st.X[S][sf][0], st.X[S][sf][1], st.RB[sr][0], st.RB[sr][1]);
      }
      // transposed prolongation at even row pc = q - S - 1 (rows pc - 1 .. pc + 1 final)
      constexpr int jp = J - S - 1;
      if constexpr (md<2>(jp) == 0) {
        const int pc = q - S - 1;
        const int I = (gi0 + pc) / 2;
        if (pc >= H && pc < H + R && I >= 0 && I <= cg.ncx - 1) {
          constexpr int ku = md<3>(jp - 1), k0 = md<3>(jp), kd = md<3>(jp + 1);
          const double fu0 = st.X[S][ku][0], fu1 = st.X[S][ku][1], f00 = st.X[S][k0][0], f01 = st.X[S][k0][1];
          const double fd0 = st.X[S][kd][0], fd1 = st.X[S][kd][1];
          const double l = from_left(f01), ul = from_left(fu1), dl = from_left(fd1);
          double v = 0.0;
          if (I >= 1 && I <= cg.ncx - 2 && cin)
            v = f00 + 0.5 * (fu0 + fd0 + l + f01) + 0.25 * (ul + fu1 + dl + fd1);
          if (occ) {
            cg.ubc[I * cg.pc + Jc] = v;
            cg.rbc[I * cg.pc + Jc] = 0.0;
          }
        }
      }
    });
  }
}

// ---------------------------------------------------------------------------
// B: transposed restriction (k_adj_restrictT) + transposed residual
// (k_adj_residualT) + transposed pre-smoothing.
// ---------------------------------------------------------------------------
template <int NS>
__global__ void __launch_bounds__(64 * kWaves) k_tup(TLevel a, TCoarse cg) {
  constexpr int S = 2 * NS, H = t_halo(NS), P = t_period(NS), TJ = SCOLS - 2 * H;
  STile tl;
  if (!s_tile<H>(a.rows, a.tiles_i, a.tiles_j, a.chunk, a.nx, a.ny, tl)) {
    s_idle(a.rows + 2 * H);
    return;
  }
  const int R = a.rows, LR = R + 2 * H;
  const int gi0 = tl.gi0, cb = tl.cb;
  const TLane L = t_lane(cb, H, TJ, a.ny, a.pitch);
  TState<NS> st;
  st.zero();
  double RS[3][2];  // resbar of rows (slot p % 3)
  double RZ[3][2];  // ubar of staged rows before the residual (slot p % 3)
#pragma unroll
  for (int k = 0; k < 3; ++k) RS[k][0] = RS[k][1] = RZ[k][0] = RZ[k][1] = 0.0;
  // coarse rbar * 0.125 of coarse rows (slot I % 3) at columns Jc (element 0) and Jc + 1
  const int Jc = L.m + cb / 2;
  const bool last = L.m == 63;
  const bool ldc = Jc >= 0 && Jc <= cg.ncy - 1, ldr = Jc + 1 >= 0 && Jc + 1 <= cg.ncy - 1;
  const double *cl = cg.rbc + (ldc ? Jc : 0), *cr = cg.rbc + (ldr ? Jc + 1 : 0);
  // g(I, J) of k_adj_restrictT: rbc * 0.125 on interior coarse nodes, else 0
  const bool gJ0 = Jc >= 1 && Jc <= cg.ncy - 2, gJ1 = Jc + 1 >= 1 && Jc + 1 <= cg.ncy - 2;
  auto coarse_g = [&](int I, double &g0, double &g1) {
    g0 = g1 = 0.0;
    if (I >= 1 && I <= cg.ncx - 2) {
      // the shift runs on every lane (a lane that skipped it would feed its
      // left neighbour a zero); lane 63 loads its own right column
      const double v0 = cl[I * cg.pc];
      const double vr = from_right(v0);
      const double v1 = last ? cr[I * cg.pc] : vr;
      g0 = gJ0 ? v0 * 0.125 : 0.0;
      g1 = gJ1 ? v1 * 0.125 : 0.0;
    }
  };
  const double d2 = 2.0 * a.ix2 + 2.0 * a.iy2;
  static_for<-kPf, 0>([&](auto jc) {
    constexpr int J = decltype(jc)::value;
    t_prefetch<NS, J>(st, a, L, gi0, J, LR);
  });
#pragma unroll 1
  for (int q0 = 0; q0 < LR; q0 += P) {
    static_for<0, P>([&](auto jc) {
      constexpr int J = decltype(jc)::value;
      const int q = q0 + J;
      group_step();
      t_prefetch<NS, J>(st, a, L, gi0, q, LR);
      RZ[md<3>(J)][0] = st.RU[md<kNru>(J)][0];
      RZ[md<3>(J)][1] = st.RU[md<kNru>(J)][1];
      // resbar of staged row q (fine row i = gi0 + q)
      {
        const int i = gi0 + q;
        double r0 = 0.0, r1 = 0.0;
        if (t_row_in(i, a.nx)) {
          if constexpr (md<2>(J) == 0) {
            // even row: even column -> 4 g(i/2, j/2); odd column -> g(I, J0) + g(I, J1)
            double g0, g1;
            coarse_g(i / 2, g0, g1);
            const int I = i / 2;
            const bool irow = I >= 1 && I <= cg.ncx - 2;
            if (L.in0 && irow && gJ0) r0 = 0.0 + 4.0 * g0;
            if (L.in1) {
              // j = gj + 1 odd: J0 = Jc, J1 = Jc + 1
              const bool aa = Jc >= 1 && irow, cc = Jc + 1 <= cg.ncy - 2 && irow;
              double v = 0.0;
              if (aa) v = v + g0;
              if (cc) v = v + g1;
              r1 = v;
            }
          } else {
            // odd row: even column -> g(I0, J) + g(I1, J); odd column -> 0
            double a0, a1, b0, b1;
            coarse_g((i - 1) / 2, a0, a1);
            coarse_g((i + 1) / 2, b0, b1);
            if (L.in0) {
              const int I0 = (i - 1) / 2, I1 = (i + 1) / 2;
              const bool aa = I0 >= 1 && gJ0, cc = I1 <= cg.ncx - 2 && gJ0;
              double v = 0.0;
              if (aa) v = v + a0;
              if (cc) v = v + b0;
              r0 = v;
            }
            (void)a1;
            (void)b1;
          }
        }
        RS[md<3>(J)][0] = r0;
        RS[md<3>(J)][1] = r1;
      }
      // transposed residual of row q - 1 -> level 0 of that row
      {
        constexpr int jr = J - 1;
        constexpr int ku = md<3>(jr - 1), k0 = md<3>(jr), kd = md<3>(jr + 1), kp = md<P>(jr);
        const int i = gi0 + q - 1;
        const double sl = from_left(RS[k0][1]), sr = from_right(RS[k0][0]);
        const bool rin = t_row_in(i, a.nx);
        static_for<0, 2>([&](auto ec) {
          constexpr int e = decltype(ec)::value;
          const int j = L.gj + e;
          const double left = e == 0 ? sl : RS[k0][0], right = e == 0 ? RS[k0][1] : sr;
          double sm = 0.0;
          sm = i > 0 ? sm + RS[ku][e] * a.ix2 : sm;
          sm = i < a.nx - 1 ? sm + RS[kd][e] * a.ix2 : sm;
          sm = j > 0 ? sm + left * a.iy2 : sm;
          sm = j < a.ny - 1 ? sm + right * a.iy2 : sm;
          const bool inner = rin && (e == 0 ? L.in0 : L.in1);
          sm = inner ? sm - d2 * RS[k0][e] : sm;
          st.X[0][k0][e] = RZ[k0][e] - sm;
          st.RB[kp][e] = inner ? st.RB[kp][e] + RS[k0][e] : st.RB[kp][e];
        });
      }
      // transposed pre-smoothing: half-sweep s on row q - 1 - s
      static_for<1, S + 1>([&](auto sc) {
        constexpr int s = decltype(sc)::value;
        t_half<NS, J - 1, s>(st, a, L, gi0, q - 1);
      });
      {
        constexpr int sf = md<3>(J - 1 - S), sr = md<P>(J - 1 - S);
        t_store(a, L, gi0, H, R, q - 1 - S, st.X[S][sf][0], st.X[S][sf][1], st.RB[sr][0], st.RB[sr][1]);
      }
    });
  }
}

}  // namespace vlp2d
