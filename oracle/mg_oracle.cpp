// TEST INFRASTRUCTURE ONLY -- the oracle.
//
// CPU restatement of the reference CPU multigrid (MultiGridScheme<T> with its
// red-black Gauss-Seidel smoother) for T in {double, float, double-double},
// plus restatements of the build's own additions on the same arithmetic
// (FP32 V-cycles inside FP64 iterative refinement, the transposed V-cycle used
// by the adjoint).  Only tests/, the smoke check and the bench's CPU baseline
// leg load this library; the product library never links or calls it.
//
// Reference files followed (paths relative to the reference root, VLP/ is
// VisualLaplacePoisson2D/):
//   VLP/Model/CPU/MultiGridScheme.cs:23-61    level hierarchy, steps
//   VLP/Model/CPU/MultiGridScheme.cs:63-71    doIteration + update-norm test
//   VLP/Model/CPU/MultiGridScheme.cs:93-115   VCycle (snapshot at :111)
//   VLP/Model/CPU/MultiGridScheme.cs:134-143  residual
//   VLP/Model/CPU/MultiGridScheme.cs:145-169  half-weighting restriction
//   VLP/Model/CPU/MultiGridScheme.cs:171-186  additive bilinear prolongation
//   VLP/Model/CPU/MultiGridSlidingIteration.cs:29-70  RB-GS point formulas
//   VLP/Common/GridIterator.cs:11-33,61-78,170-222    loop bounds, red first
//   VLP/Common/UtilsOpLap.cs:15-20             5-point operator association
//   VLP/DD128_QD256/DD128.cs:52-203,478-483, ArithmeticDD128.cs:8-131  DD ops
//
// Build: -O3 -fopenmp -ffp-contract=off (no FMA contraction, IEEE order as
// written), see oracle/Makefile.

#include <cmath>
#include <cstdint>
#include <cstdlib>
#include <cstring>
#include <vector>
#include <algorithm>
#include <omp.h>

// ---------------------------------------------------------------------------
// Double-double arithmetic (Dekker / Knuth error-free transforms, the
// Hida-Li-Bailey QD "ieee" add and "sloppy" multiply; long division with three
// quotient digits).  Follows ArithmeticDD128.cs:8-131 and DD128.cs:52-203.
// ---------------------------------------------------------------------------
struct DD {
  double hi, lo;
  DD() : hi(0.0), lo(0.0) {}
  DD(double h) : hi(h), lo(0.0) {}
  DD(double h, double l) : hi(h), lo(l) {}
};

static inline DD dd_quick_two_sum(double a, double b) {
  double s = a + b;
  return DD(s, b - (s - a));
}
static inline void dd_two_sum(double a, double b, double &s, double &e) {
  s = a + b;
  double bb = s - a;
  e = (a - (s - bb)) + (b - bb);
}
static inline void dd_two_diff(double a, double b, double &s, double &e) {
  s = a - b;
  double bb = s - a;
  e = (a - (s - bb)) - (b + bb);
}
static inline void dd_split(double a, double &hi, double &lo) {
  const double splitter = 134217729.0;           // 2^27 + 1
  const double thresh = 6.69692879491417e+299;   // 2^996
  if (a > thresh || a < -thresh) {
    a *= 3.7252902984619140625e-09;              // 2^-28
    double t = splitter * a;
    hi = t - (t - a);
    lo = a - hi;
    hi *= 268435456.0;                           // 2^28
    lo *= 268435456.0;
  } else {
    double t = splitter * a;
    hi = t - (t - a);
    lo = a - hi;
  }
}
static inline void dd_two_prod(double a, double b, double &p, double &e) {
  p = a * b;
  double ah, al, bh, bl;
  dd_split(a, ah, al);
  dd_split(b, bh, bl);
  e = ((ah * bh - p) + ah * bl + al * bh) + al * bl;
}
static inline DD dd_renorm(double s1, double s2, double t1, double t2) {
  s2 += t1;
  DD r = dd_quick_two_sum(s1, s2);
  return dd_quick_two_sum(r.hi, r.lo + t2);
}
static inline DD operator+(DD a, DD b) {
  double s1, s2, t1, t2;
  dd_two_sum(a.hi, b.hi, s1, s2);
  dd_two_sum(a.lo, b.lo, t1, t2);
  return dd_renorm(s1, s2, t1, t2);
}
static inline DD operator-(DD a, DD b) {
  double s1, s2, t1, t2;
  dd_two_diff(a.hi, b.hi, s1, s2);
  dd_two_diff(a.lo, b.lo, t1, t2);
  return dd_renorm(s1, s2, t1, t2);
}
static inline DD operator-(DD a) { return DD(-a.hi, -a.lo); }
static inline DD operator*(DD a, DD b) {
  double p1, p2;
  dd_two_prod(a.hi, b.hi, p1, p2);
  return dd_quick_two_sum(p1, p2 + a.hi * b.lo + a.lo * b.hi);
}
static inline DD dd_mul_d(DD a, double b) {
  double p1, p2;
  dd_two_prod(a.hi, b, p1, p2);
  return dd_quick_two_sum(p1, p2 + a.lo * b);
}
static inline DD dd_add_d(DD a, double b) {
  double s1, s2;
  dd_two_sum(a.hi, b, s1, s2);
  return dd_quick_two_sum(s1, s2 + a.lo);
}
static inline DD operator/(DD a, DD b) {
  double q1 = a.hi / b.hi;
  if (std::isinf(q1)) return DD(q1);
  DD r = a - dd_mul_d(b, q1);
  double q2 = r.hi / b.hi;
  r = r - dd_mul_d(b, q2);
  double q3 = r.hi / b.hi;
  return dd_add_d(dd_quick_two_sum(q1, q2), q3);
}
static inline DD &operator+=(DD &a, DD b) { a = a + b; return a; }
static inline bool operator<(DD a, DD b) { return a.hi < b.hi || (a.hi == b.hi && a.lo < b.lo); }
static inline DD dd_abs(DD a) { return (a.hi < 0.0 || (a.hi == 0.0 && a.lo < 0.0)) ? -a : a; }

// Scalar traits so the multigrid template is written once.
template <class T> struct Num;
template <> struct Num<double> {
  static double from(double v) { return v; }
  static double to_d(double v) { return v; }
  static double lo(double) { return 0.0; }
  static double abs_d(double v) { return std::fabs(v); }
  static bool abs_lt(double v, double t) { return std::fabs(v) < t; }
};
template <> struct Num<float> {
  static float from(double v) { return (float)v; }
  static double to_d(float v) { return (double)v; }
  static double lo(float) { return 0.0; }
  static double abs_d(float v) { return (double)std::fabs(v); }
  static bool abs_lt(float v, double t) { return std::fabs(v) < (float)t; }
};
template <> struct Num<DD> {
  static DD from(double v) { return DD(v); }
  static double to_d(DD v) { return v.hi; }
  static double lo(DD v) { return v.lo; }
  static double abs_d(DD v) { return dd_abs(v).hi; }
  static bool abs_lt(DD v, double t) { return dd_abs(v) < DD(t); }
};

// ---------------------------------------------------------------------------
// Multigrid hierarchy (MultiGridScheme.cs:23-61)
// ---------------------------------------------------------------------------
template <class T> struct Level {
  int nx = 0, ny = 0;          // nodes = segments + 1
  std::vector<T> u, rhs, res;  // rhs of level 0 = fn (or zeros for Laplace)
  T hx, hy, hx2, hy2, coef;
  bool eq = true;              // |hx - hy| < 1e-15 (MultiGridSlidingIteration.cs:43)
  bool lap = false;            // level 0 without fn: Laplace point formulas
  bool par = false;            // GridIterator parallel above 64 segments (:40-41)
  inline size_t at(int i, int j) const { return (size_t)i * ny + j; }
};

template <class T> struct MG {
  int L = 0;
  std::vector<Level<T>> lv;
  // smoother: 0 red-black Gauss-Seidel (the reference's), 1 weighted Jacobi
  // with weight omega (an addition: the reference has no Jacobi smoother)
  int smoother = 0;
  double omega = 0.8;
  // restriction: 0 half weighting (the reference's), 1 full weighting (an
  // addition: the reference's full-weighting branch is dead code)
  int restriction = 0;
  // cycle: 0 V-cycle (MultiGridScheme.cs:65), 1 full multigrid (the
  // reference's FMG(), :117-132, which its doIteration leaves commented out)
  int cycle = 0;
  std::vector<T> tmp;  // Jacobi: the iterate before the sweep

  // fn: scheme right-hand side (= -f) at all nodes, or nullptr for Laplace.
  MG(int segX, int segY, T hx, T hy, const T *fn) {
    int lx = (int)std::floor(std::log2((double)segX) + 1e-12);
    int ly = (int)std::floor(std::log2((double)segY) + 1e-12);
    L = std::min(lx, ly);
    lv.resize(L);
    int sx = segX, sy = segY;
    T lngX = hx * Num<T>::from((double)segX);
    T lngY = hy * Num<T>::from((double)segY);
    for (int l = 0; l < L; ++l) {
      Level<T> &v = lv[l];
      v.nx = sx + 1;
      v.ny = sy + 1;
      size_t n = (size_t)v.nx * v.ny;
      v.u.assign(n, Num<T>::from(0.0));
      v.rhs.assign(n, Num<T>::from(0.0));
      if (l < L - 1) v.res.assign(n, Num<T>::from(0.0));
      v.hx = lngX / Num<T>::from((double)sx);
      v.hy = lngY / Num<T>::from((double)sy);
      v.hx2 = v.hx * v.hx;
      v.hy2 = v.hy * v.hy;
      T two = Num<T>::from(2.0);
      v.coef = Num<T>::from(1.0) / (two / v.hx2 + two / v.hy2);
      v.eq = Num<T>::abs_lt(v.hx - v.hy, 1e-15);
      v.lap = (l == 0) && (fn == nullptr);
      v.par = (v.nx - 1) > 64;
      sx /= 2;
      sy /= 2;
    }
    if (fn) std::copy(fn, fn + (size_t)lv[0].nx * lv[0].ny, lv[0].rhs.begin());
  }

  // --- smoother: one red half-sweep then one black half-sweep per iteration
  //     (GridIterator.cs:184-188; red = i+j even: j starts at 1 on odd rows).
  void half_sweep(int l, int color) {
    Level<T> &v = lv[l];
    const int nx = v.nx, ny = v.ny;
    T *u = v.u.data();
    const T *f = v.rhs.data();
    const T q = Num<T>::from(0.25);
#pragma omp parallel for schedule(static) if (v.par)
    for (int i = 1; i < nx - 1; ++i) {
      int j0 = (i % 2 == 1) ? (color == 0 ? 1 : 2) : (color == 0 ? 2 : 1);
      for (int j = j0; j < ny - 1; j += 2) {
        size_t c = (size_t)i * ny + j;
        T uim = u[c - ny], uip = u[c + ny], ujm = u[c - 1], ujp = u[c + 1];
        T r;
        if (v.eq) {
          // MultiGridSlidingIteration.cs:53-54
          if (v.lap) r = (uim + uip + ujm + ujp) * q;
          else r = (uim + uip + ujm + ujp - f[c] * v.hx2) * q;
        } else {
          // MultiGridSlidingIteration.cs:56-57
          if (v.lap) r = v.coef * ((uim + uip) / v.hx2 + (ujm + ujp) / v.hy2);
          else r = v.coef * ((uim + uip) / v.hx2 + (ujm + ujp) / v.hy2 - f[c]);
        }
        u[c] = r;
      }
    }
  }
  // --- weighted Jacobi sweep: every interior node from the previous
  //     iterate, u = u + omega (g - u), g the same point formula as the
  //     Gauss-Seidel update (MultiGridSlidingIteration.cs:53-57)
  void jacobi_sweep(int l) {
    Level<T> &v = lv[l];
    const int nx = v.nx, ny = v.ny;
    tmp.assign(v.u.begin(), v.u.end());
    const T *o = tmp.data();
    T *u = v.u.data();
    const T *f = v.rhs.data();
    const T q = Num<T>::from(0.25), w = Num<T>::from(omega);
#pragma omp parallel for schedule(static) if (v.par)
    for (int i = 1; i < nx - 1; ++i)
      for (int j = 1; j < ny - 1; ++j) {
        size_t c = (size_t)i * ny + j;
        T uim = o[c - ny], uip = o[c + ny], ujm = o[c - 1], ujp = o[c + 1];
        T g;
        if (v.eq) {
          if (v.lap) g = (uim + uip + ujm + ujp) * q;
          else g = (uim + uip + ujm + ujp - f[c] * v.hx2) * q;
        } else {
          if (v.lap) g = v.coef * ((uim + uip) / v.hx2 + (ujm + ujp) / v.hy2);
          else g = v.coef * ((uim + uip) / v.hx2 + (ujm + ujp) / v.hy2 - f[c]);
        }
        u[c] = o[c] + w * (g - o[c]);
      }
  }
  void smooth(int l, int n) {
    for (int s = 0; s < n; ++s) {
      if (smoother == 1) {
        jacobi_sweep(l);
        continue;
      }
      half_sweep(l, 0);
      half_sweep(l, 1);
    }
  }

  // --- residual (MultiGridScheme.cs:134-143, UtilsOpLap.cs:15-20)
  void residual_into(int l, T *re) {
    Level<T> &v = lv[l];
    const int nx = v.nx, ny = v.ny;
    const T *u = v.u.data(), *rh = v.rhs.data();
    const T two = Num<T>::from(2.0);
#pragma omp parallel for schedule(static) if (v.par)
    for (int i = 1; i < nx - 1; ++i)
      for (int j = 1; j < ny - 1; ++j) {
        size_t c = (size_t)i * ny + j;
        T u2 = u[c] * two;
        re[c] = rh[c] - ((u[c - ny] + u[c + ny] - u2) / v.hx2 + (u[c - 1] + u[c + 1] - u2) / v.hy2);
      }
  }

  // --- full-weighting restriction (1/16)[1 2 1; 2 4 2; 1 2 1]
  void restrict_fw(int l) {
    Level<T> &f = lv[l];
    Level<T> &c = lv[l + 1];
    const T *re = f.res.data();
    T *rh = c.rhs.data();
    const T w = Num<T>::from(1.0) / Num<T>::from(16.0);
    const T four = Num<T>::from(4.0), two = Num<T>::from(2.0);
    const int nyf = f.ny;
#pragma omp parallel for schedule(static) if (c.par)
    for (int i = 1; i < c.nx - 1; ++i)
      for (int j = 1; j < c.ny - 1; ++j) {
        size_t k = (size_t)(2 * i) * nyf + 2 * j;
        rh[c.at(i, j)] = w * (re[k] * four + (re[k + nyf] + re[k - nyf] + re[k + 1] + re[k - 1]) * two +
                              (re[k + nyf + 1] + re[k + nyf - 1] + re[k - nyf + 1] + re[k - nyf - 1]));
      }
  }

  // --- half-weighting restriction (MultiGridScheme.cs:145-169)
  void restrict_hw(int l) {
    Level<T> &f = lv[l];
    Level<T> &c = lv[l + 1];
    const T *re = f.res.data();
    T *rh = c.rhs.data();
    const T w = Num<T>::from(1.0) / Num<T>::from(8.0);
    const T four = Num<T>::from(4.0);
    const int nyf = f.ny;
#pragma omp parallel for schedule(static) if (c.par)
    for (int i = 1; i < c.nx - 1; ++i)
      for (int j = 1; j < c.ny - 1; ++j) {
        size_t k = (size_t)(2 * i) * nyf + 2 * j;
        rh[c.at(i, j)] = w * (re[k] * four + (re[k + nyf] + re[k - nyf] + re[k + 1] + re[k - 1]));
      }
  }

  // --- additive bilinear prolongation (MultiGridScheme.cs:171-186)
  void interpolate(int l) {
    Level<T> &f = lv[l];
    Level<T> &c = lv[l + 1];
    T *uf = f.u.data();
    const T *uc = c.u.data();
    const T half = Num<T>::from(0.5), quarter = Num<T>::from(0.25);
    const int nyf = f.ny, nyc = c.ny;
#pragma omp parallel for schedule(static) if (c.par)
    for (int i = 0; i < c.nx - 1; ++i)
      for (int j = 0; j < c.ny - 1; ++j) {
        T a = uc[(size_t)i * nyc + j], b = uc[(size_t)i * nyc + j + 1];
        T d = uc[(size_t)(i + 1) * nyc + j], e = uc[(size_t)(i + 1) * nyc + j + 1];
        size_t k = (size_t)(2 * i) * nyf + 2 * j;
        if (i > 0 && j > 0) uf[k] += a;
        if (j > 0) uf[k + nyf] += (a + d) * half;
        if (i > 0) uf[k + 1] += (a + b) * half;
        uf[k + nyf + 1] += (a + b + d + e) * quarter;
      }
  }

  // --- One iteration: a V-cycle or FMG. Returns the update norm
  //     max|u - snapshot| over interior nodes of level 0 (:69, :111).
  double vcycle(int nu1, int nu2, int nuc) { return cycle == 1 ? fmg(nu1, nu2, nuc) : vcycle_from(nu1, nu2, nuc, 0); }

  // --- FMG (MultiGridScheme.cs:117-132): the level-0 residual restricted to
  //     every level without smoothing, one sweep on the coarsest, then from
  //     coarse to fine: prolongation and a V-cycle starting at that level;
  //     a last V-cycle from level 0 gives the update norm.
  double fmg(int nu1, int nu2, int nuc) {
    for (int l = 0; l <= L - 2; ++l) {
      residual_into(l, lv[l].res.data());
      if (restriction == 1) restrict_fw(l);
      else restrict_hw(l);
      std::fill(lv[l + 1].u.begin(), lv[l + 1].u.end(), Num<T>::from(0.0));
    }
    smooth(L - 1, 1);
    for (int l = L - 2; l >= 0; --l) {
      interpolate(l);
      vcycle_from(nu1, nu2, nuc, l);
    }
    return vcycle_from(nu1, nu2, nuc, 0);
  }

  // --- V-cycle over levels s..L-1 (MultiGridScheme.cs:93-115, startLevel
  //     s); the level-0 snapshot is taken only when s == 0.
  double vcycle_from(int nu1, int nu2, int nuc, int s) {
    for (int l = s; l <= L - 2; ++l) {
      smooth(l, nu1);
      residual_into(l, lv[l].res.data());
      if (restriction == 1) restrict_fw(l);
      else restrict_hw(l);
      std::fill(lv[l + 1].u.begin(), lv[l + 1].u.end(), Num<T>::from(0.0));
    }
    smooth(L - 1, nuc);
    for (int l = L - 2; l >= s; --l) {
      interpolate(l);
      if (l == 0) {
        Level<T> &v = lv[0];
        for (int i = 1; i < v.nx - 1; ++i)
          for (int j = 1; j < v.ny - 1; ++j) v.res[v.at(i, j)] = v.u[v.at(i, j)];
      }
      smooth(l, nu2);
    }
    if (L < 2 || s > 0) return 0.0;
    Level<T> &v = lv[0];
    double m = 0.0;
    for (int i = 1; i < v.nx - 1; ++i)
      for (int j = 1; j < v.ny - 1; ++j) {
        double d = Num<T>::abs_d(v.u[v.at(i, j)] - v.res[v.at(i, j)]);
        if (d > m || std::isnan(d)) m = d;
      }
    return m;
  }

  // Sum of squares of the level-0 residual, accumulated in double.
  double residual_norm2() {
    Level<T> &v = lv[0];
    std::vector<T> r((size_t)v.nx * v.ny, Num<T>::from(0.0));
    residual_into(0, r.data());
    double s = 0.0;
    for (int i = 1; i < v.nx - 1; ++i)
      for (int j = 1; j < v.ny - 1; ++j) {
        double d = Num<T>::to_d(r[v.at(i, j)]);
        s += d * d;
      }
    return s;
  }
};

// ---------------------------------------------------------------------------
// Transposed V-cycle (adjoint), FP64.  The forward cycle is linear in
// (u_in, fn, boundary values); this applies its transpose operator by
// operator in reverse order.  A half-sweep u_P = ax (u_W+u_E) + ay (u_S+u_N)
// - b fn_P on colour c has the transpose: every node Q of the other colour
// (boundary included) gathers ax/ay * ubar_P from its colour-c neighbours,
// then fnbar_P += -b ubar_P and ubar_P = 0.
// ---------------------------------------------------------------------------
struct Adj {
  MG<double> &mg;
  std::vector<std::vector<double>> ub, rb;  // adjoints of u and rhs per level
  explicit Adj(MG<double> &m) : mg(m) {
    ub.resize(mg.L);
    rb.resize(mg.L);
    for (int l = 0; l < mg.L; ++l) {
      size_t n = (size_t)mg.lv[l].nx * mg.lv[l].ny;
      ub[l].assign(n, 0.0);
      rb[l].assign(n, 0.0);
    }
  }
  void coefs(int l, double &ax, double &ay, double &b) {
    Level<double> &v = mg.lv[l];
    if (v.eq) { ax = 0.25; ay = 0.25; b = 0.25 * v.hx2; }
    else { ax = v.coef / v.hx2; ay = v.coef / v.hy2; b = v.coef; }
  }
  void half_sweep_T(int l, int color) {
    Level<double> &v = mg.lv[l];
    const int nx = v.nx, ny = v.ny;
    double ax, ay, b;
    coefs(l, ax, ay, b);
    std::vector<double> &u = ub[l];
    std::vector<double> &r = rb[l];
    auto isc = [&](int i, int j) {
      return i > 0 && i < nx - 1 && j > 0 && j < ny - 1 && (((i + j) & 1) == color);
    };
    // gather into nodes of the other colour (and all boundary nodes)
    for (int i = 0; i < nx; ++i)
      for (int j = 0; j < ny; ++j) {
        if (((i + j) & 1) == color) continue;
        double s = 0.0;
        if (i > 0 && isc(i - 1, j)) s += ax * u[v.at(i - 1, j)];
        if (i < nx - 1 && isc(i + 1, j)) s += ax * u[v.at(i + 1, j)];
        if (j > 0 && isc(i, j - 1)) s += ay * u[v.at(i, j - 1)];
        if (j < ny - 1 && isc(i, j + 1)) s += ay * u[v.at(i, j + 1)];
        u[v.at(i, j)] += s;
      }
    for (int i = 1; i < nx - 1; ++i)
      for (int j = 1; j < ny - 1; ++j)
        if (((i + j) & 1) == color) {
          r[v.at(i, j)] += -b * u[v.at(i, j)];
          u[v.at(i, j)] = 0.0;
        }
  }
  void smooth_T(int l, int n) {
    for (int s = 0; s < n; ++s) {
      half_sweep_T(l, 1);
      half_sweep_T(l, 0);
    }
  }
  void cycle_T(int l, int nu1, int nu2, int nuc) {
    if (l == mg.L - 1) { smooth_T(l, nuc); return; }
    smooth_T(l, nu2);
    Level<double> &f = mg.lv[l];
    Level<double> &c = mg.lv[l + 1];
    // transpose of u_f += P u_c: ubar_c = P^T ubar_f (coarse interior only;
    // the coarse boundary is the constant zero)
    std::fill(ub[l + 1].begin(), ub[l + 1].end(), 0.0);
    std::fill(rb[l + 1].begin(), rb[l + 1].end(), 0.0);
    for (int I = 1; I < c.nx - 1; ++I)
      for (int J = 1; J < c.ny - 1; ++J) {
        int i = 2 * I, j = 2 * J;
        const std::vector<double> &u = ub[l];
        double s = u[f.at(i, j)] +
                   0.5 * (u[f.at(i - 1, j)] + u[f.at(i + 1, j)] + u[f.at(i, j - 1)] + u[f.at(i, j + 1)]) +
                   0.25 * (u[f.at(i - 1, j - 1)] + u[f.at(i - 1, j + 1)] + u[f.at(i + 1, j - 1)] + u[f.at(i + 1, j + 1)]);
        ub[l + 1][c.at(I, J)] = s;
      }
    cycle_T(l + 1, nu1, nu2, nuc);
    // transpose of rhs_c = R res_f: resbar_f = R^T rhsbar_c (fine interior)
    std::vector<double> resb((size_t)f.nx * f.ny, 0.0);
    for (int I = 1; I < c.nx - 1; ++I)
      for (int J = 1; J < c.ny - 1; ++J) {
        double g = rb[l + 1][c.at(I, J)] * 0.125;
        int i = 2 * I, j = 2 * J;
        resb[f.at(i, j)] += 4.0 * g;
        resb[f.at(i - 1, j)] += g;
        resb[f.at(i + 1, j)] += g;
        resb[f.at(i, j - 1)] += g;
        resb[f.at(i, j + 1)] += g;
      }
    // transpose of res = rhs - Lap(u): rhsbar += resbar, ubar -= Lap(resbar)
    double ix2 = 1.0 / f.hx2, iy2 = 1.0 / f.hy2;
    for (int i = 0; i < f.nx; ++i)
      for (int j = 0; j < f.ny; ++j) {
        size_t k = f.at(i, j);
        bool inner = i > 0 && i < f.nx - 1 && j > 0 && j < f.ny - 1;
        double s = 0.0;
        if (i > 0) s += resb[f.at(i - 1, j)] * ix2;
        if (i < f.nx - 1) s += resb[f.at(i + 1, j)] * ix2;
        if (j > 0) s += resb[f.at(i, j - 1)] * iy2;
        if (j < f.ny - 1) s += resb[f.at(i, j + 1)] * iy2;
        if (inner) s -= (2.0 * ix2 + 2.0 * iy2) * resb[k];
        ub[l][k] -= s;
        if (inner) rb[l][k] += resb[k];
      }
    smooth_T(l, nu1);
  }
};

// ---------------------------------------------------------------------------
// C API
// ---------------------------------------------------------------------------
extern "C" {

struct OrcHandle {
  int prec;  // 0 double, 1 float, 2 double-double
  void *mg;
};

static void set_threads(int nthreads) {
  if (nthreads > 0) omp_set_num_threads(nthreads);
}

// f: PDE right-hand side (u_xx + u_yy = -f) at all nodes, or NULL (Laplace).
// As in the reference constructor the scheme stores fn = -f on interior nodes.
void *orc_create(int prec, int segX, int segY, double hx, double hy, const double *f, int nthreads) {
  set_threads(nthreads);
  int nx = segX + 1, ny = segY + 1;
  size_t n = (size_t)nx * ny;
  OrcHandle *h = new OrcHandle;
  h->prec = prec;
  auto fill = [&](auto *dst) {
    using T = std::remove_pointer_t<decltype(dst)>;
    for (size_t k = 0; k < n; ++k) dst[k] = Num<T>::from(0.0);
    for (int i = 1; i < nx - 1; ++i)
      for (int j = 1; j < ny - 1; ++j) dst[(size_t)i * ny + j] = -Num<T>::from(f[(size_t)i * ny + j]);
  };
  if (prec == 0) {
    std::vector<double> fn;
    if (f) { fn.resize(n); fill(fn.data()); }
    h->mg = new MG<double>(segX, segY, hx, hy, f ? fn.data() : nullptr);
  } else if (prec == 1) {
    std::vector<float> fn;
    if (f) { fn.resize(n); fill(fn.data()); }
    h->mg = new MG<float>(segX, segY, (float)hx, (float)hy, f ? fn.data() : nullptr);
  } else {
    std::vector<DD> fn;
    if (f) { fn.resize(n); fill(fn.data()); }
    h->mg = new MG<DD>(segX, segY, DD(hx), DD(hy), f ? fn.data() : nullptr);
  }
  return h;
}

void orc_set_smoother(void *p, int kind, double omega) {
  OrcHandle *h = (OrcHandle *)p;
  auto set = [&](auto *mg) {
    mg->smoother = kind;
    mg->omega = omega;
  };
  if (h->prec == 0) set((MG<double> *)h->mg);
  else if (h->prec == 1) set((MG<float> *)h->mg);
  else set((MG<DD> *)h->mg);
}

void orc_set_restriction(void *p, int kind) {
  OrcHandle *h = (OrcHandle *)p;
  if (h->prec == 0) ((MG<double> *)h->mg)->restriction = kind;
  else if (h->prec == 1) ((MG<float> *)h->mg)->restriction = kind;
  else ((MG<DD> *)h->mg)->restriction = kind;
}

void orc_set_cycle(void *p, int kind) {
  OrcHandle *h = (OrcHandle *)p;
  if (h->prec == 0) ((MG<double> *)h->mg)->cycle = kind;
  else if (h->prec == 1) ((MG<float> *)h->mg)->cycle = kind;
  else ((MG<DD> *)h->mg)->cycle = kind;
}

void orc_destroy(void *p) {
  OrcHandle *h = (OrcHandle *)p;
  if (!h) return;
  if (h->prec == 0) delete (MG<double> *)h->mg;
  else if (h->prec == 1) delete (MG<float> *)h->mg;
  else delete (MG<DD> *)h->mg;
  delete h;
}

int orc_levels(void *p) {
  OrcHandle *h = (OrcHandle *)p;
  if (h->prec == 0) return ((MG<double> *)h->mg)->L;
  if (h->prec == 1) return ((MG<float> *)h->mg)->L;
  return ((MG<DD> *)h->mg)->L;
}

// Level-0 iterate (boundary values + initial guess), all nodes.
void orc_set_u(void *p, const double *u) {
  OrcHandle *h = (OrcHandle *)p;
  auto put = [&](auto *mg) {
    using T = std::remove_reference_t<decltype(mg->lv[0].u[0])>;
    for (size_t k = 0; k < mg->lv[0].u.size(); ++k) mg->lv[0].u[k] = Num<T>::from(u[k]);
  };
  if (h->prec == 0) put((MG<double> *)h->mg);
  else if (h->prec == 1) put((MG<float> *)h->mg);
  else put((MG<DD> *)h->mg);
}

void orc_get_u(void *p, double *hi, double *lo) {
  OrcHandle *h = (OrcHandle *)p;
  auto get = [&](auto *mg) {
    for (size_t k = 0; k < mg->lv[0].u.size(); ++k) {
      hi[k] = Num<std::remove_reference_t<decltype(mg->lv[0].u[0])>>::to_d(mg->lv[0].u[k]);
      if (lo) lo[k] = Num<std::remove_reference_t<decltype(mg->lv[0].u[0])>>::lo(mg->lv[0].u[k]);
    }
  };
  if (h->prec == 0) get((MG<double> *)h->mg);
  else if (h->prec == 1) get((MG<float> *)h->mg);
  else get((MG<DD> *)h->mg);
}

// One V(nu1,nu2) cycle with nuc sweeps on the coarsest level; returns the
// reference update norm max|u - u_before_post_smoothing| on level 0.
double orc_vcycle(void *p, int nu1, int nu2, int nuc, int nthreads) {
  set_threads(nthreads);
  OrcHandle *h = (OrcHandle *)p;
  if (h->prec == 0) return ((MG<double> *)h->mg)->vcycle(nu1, nu2, nuc);
  if (h->prec == 1) return ((MG<float> *)h->mg)->vcycle(nu1, nu2, nuc);
  return ((MG<DD> *)h->mg)->vcycle(nu1, nu2, nuc);
}

double orc_residual_norm2(void *p, int nthreads) {
  set_threads(nthreads);
  OrcHandle *h = (OrcHandle *)p;
  if (h->prec == 0) return ((MG<double> *)h->mg)->residual_norm2();
  if (h->prec == 1) return ((MG<float> *)h->mg)->residual_norm2();
  return ((MG<DD> *)h->mg)->residual_norm2();
}

// Plain V-cycles until ||r_k||_2 / ||r_0||_2 <= rtol (or max_cycles).
// hist[k] = relative residual after k cycles (hist[0] = 1). Returns cycles run.
int orc_plain_solve(void *p, int nu1, int nu2, int nuc, double rtol, int max_cycles, double *hist, int nthreads) {
  set_threads(nthreads);
  double r0 = std::sqrt(orc_residual_norm2(p, nthreads));
  if (hist) hist[0] = 1.0;
  if (r0 == 0.0) return 0;
  int k = 0;
  while (k < max_cycles) {
    orc_vcycle(p, nu1, nu2, nuc, nthreads);
    ++k;
    double rel = std::sqrt(orc_residual_norm2(p, nthreads)) / r0;
    if (hist) hist[k] = rel;
    if (rel <= rtol) break;
  }
  return k;
}

// FP32 V-cycles inside FP64 iterative refinement.
//   r = fn - Lap(u) (FP64, reference association);  rhs32 = (float) r
//   e = m FP32 V(nu1,nu2) cycles on Lap(e) = rhs32 from e = 0 (zero boundary)
//   u += (double) e on interior nodes
// m: step 0 runs `inner` cycles; step k > 0 the fewest m <= inner with
// rel_k rho^m <= rtol, rho = (rel_k / rel_{k-1})^(1/m_{k-1}) the contraction
// per cycle of the previous step (inner when rho >= 1); inner = 1 is one
// cycle per step.
// Stops when ||r||_2 / ||r_0||_2 <= rtol. u (all nodes, FP64) is updated in
// place. hist[k] = relative residual after k outer steps. Returns steps run.
int orc_mixed_solve(int segX, int segY, double hx, double hy, const double *f, double *u, int nu1, int nu2,
                    int nuc, double rtol, int max_outer, double *hist, int nthreads, int inner, int smoother,
                    double omega, int restriction) {
  set_threads(nthreads);
  void *p64 = orc_create(0, segX, segY, hx, hy, f, nthreads);
  MG<double> &d = *(MG<double> *)((OrcHandle *)p64)->mg;
  orc_set_u(p64, u);
  const int nx = d.lv[0].nx, ny = d.lv[0].ny;
  const size_t n = (size_t)nx * ny;
  std::vector<float> dummy(n, 0.0f);
  MG<float> s(segX, segY, (float)hx, (float)hy, dummy.data());  // level 0 in Poisson form
  s.smoother = smoother;
  s.omega = omega;
  s.restriction = restriction;
  std::vector<double> r(n, 0.0);
  auto resid = [&]() {
    std::fill(r.begin(), r.end(), 0.0);
    d.residual_into(0, r.data());
    double ss = 0.0;
    for (int i = 1; i < nx - 1; ++i)
      for (int j = 1; j < ny - 1; ++j) ss += r[(size_t)i * ny + j] * r[(size_t)i * ny + j];
    return std::sqrt(ss);
  };
  double r0 = resid();
  if (hist) hist[0] = 1.0;
  int k = 0, m_prev = 1;
  double rel = 1.0, rel_prev = 1.0;
  if (r0 > 0.0) {
    while (k < max_outer) {
      Level<float> &z = s.lv[0];
      for (size_t q = 0; q < n; ++q) { z.rhs[q] = (float)r[q]; z.u[q] = 0.0f; }
      // FP32 cycles of this step: step 0 runs max_inner; later steps the
      // fewest that the last step's contraction per cycle predicts reach
      // rtol, at most max_inner
      int m = std::max(inner, 1);
      if (k > 0 && m > 1) {
        const double rho = std::pow(rel / rel_prev, 1.0 / m_prev);
        if (rho < 1.0) {
          int mm = 1;
          for (double pr = rel * rho; mm < m && pr > rtol; pr *= rho) ++mm;
          m = mm;
        }
      }
      for (int c = 0; c < m; ++c) s.vcycle(nu1, nu2, nuc);
      m_prev = m;
      for (int i = 1; i < nx - 1; ++i)
        for (int j = 1; j < ny - 1; ++j) d.lv[0].u[(size_t)i * ny + j] += (double)z.u[(size_t)i * ny + j];
      ++k;
      rel_prev = rel;
      rel = resid() / r0;
      if (hist) hist[k] = rel;
      if (rel <= rtol) break;
    }
  }
  std::copy(d.lv[0].u.begin(), d.lv[0].u.end(), u);
  orc_destroy(p64);
  return k;
}

// Forward map of K plain FP64 V-cycles: u_K = Phi^K(u0; f, g), always in
// Poisson form (f may be all zeros).  u: in = u0 (boundary g + interior
// guess), out = u_K.
void orc_forward_cycles(int segX, int segY, double hx, double hy, const double *f, double *u, int K, int nu1,
                        int nu2, int nuc, int nthreads) {
  set_threads(nthreads);
  size_t n = (size_t)(segX + 1) * (segY + 1);
  std::vector<double> zero(n, 0.0);
  void *p = orc_create(0, segX, segY, hx, hy, f ? f : zero.data(), nthreads);
  orc_set_u(p, u);
  for (int k = 0; k < K; ++k) orc_vcycle(p, nu1, nu2, nuc, nthreads);
  orc_get_u(p, u, nullptr);
  orc_destroy(p);
}

// Transpose of the K-cycle map. ubar: adjoint of u_K (all nodes; normally
// nonzero on interior only).  Outputs (all nodes): fbar = dL/df (PDE
// right-hand side f, fn = -f), gbar = dL/d(boundary values) (boundary nodes,
// zero inside), u0bar = dL/d(initial interior guess) (interior, zero on the
// boundary).  Any output pointer may be NULL.
void orc_adjoint_cycles(int segX, int segY, double hx, double hy, int K, int nu1, int nu2, int nuc,
                        const double *ubar, double *fbar, double *gbar, double *u0bar, int nthreads) {
  set_threads(nthreads);
  size_t n = (size_t)(segX + 1) * (segY + 1);
  std::vector<double> zero(n, 0.0);
  MG<double> mg(segX, segY, hx, hy, zero.data());
  Adj a(mg);
  std::copy(ubar, ubar + n, a.ub[0].begin());
  for (int k = 0; k < K; ++k) a.cycle_T(0, nu1, nu2, nuc);
  const int nx = segX + 1, ny = segY + 1;
  for (int i = 0; i < nx; ++i)
    for (int j = 0; j < ny; ++j) {
      size_t q = (size_t)i * ny + j;
      bool inner = i > 0 && i < nx - 1 && j > 0 && j < ny - 1;
      if (fbar) fbar[q] = inner ? -a.rb[0][q] : 0.0;
      if (gbar) gbar[q] = inner ? 0.0 : a.ub[0][q];
      if (u0bar) u0bar[q] = inner ? a.ub[0][q] : 0.0;
    }
}

// ---- operator-level entry points (FP64), for pinning each operator alone --
// Restriction (half weighting, or full weighting when full != 0) of a fine residual array (nx*ny, fine grid of
// segX x segY) to the coarse right-hand side (coarse interior written, coarse
// boundary 0).
void orc_op_restrict(int segX, int segY, const double *resf, double *rhsc, int full) {
  MG<double> mg(segX, segY, 1.0, 1.0, nullptr);
  std::copy(resf, resf + mg.lv[0].res.size(), mg.lv[0].res.begin());
  if (full) mg.restrict_fw(0);
  else mg.restrict_hw(0);
  std::copy(mg.lv[1].rhs.begin(), mg.lv[1].rhs.end(), rhsc);
}
// uf += P uc (additive bilinear prolongation onto the fine interior).
void orc_op_prolong(int segX, int segY, const double *uc, double *uf) {
  MG<double> mg(segX, segY, 1.0, 1.0, nullptr);
  std::copy(uc, uc + mg.lv[1].u.size(), mg.lv[1].u.begin());
  std::copy(uf, uf + mg.lv[0].u.size(), mg.lv[0].u.begin());
  mg.interpolate(0);
  std::copy(mg.lv[0].u.begin(), mg.lv[0].u.end(), uf);
}
// res = fn - Lap_h u on the interior (fn = -f; f NULL -> 0), 0 elsewhere.
void orc_op_residual(int segX, int segY, double hx, double hy, const double *f, const double *u, double *res) {
  void *p = orc_create(0, segX, segY, hx, hy, f, 1);
  MG<double> &mg = *(MG<double> *)((OrcHandle *)p)->mg;
  std::copy(u, u + mg.lv[0].u.size(), mg.lv[0].u.begin());
  std::fill(res, res + mg.lv[0].u.size(), 0.0);
  mg.residual_into(0, res);
  orc_destroy(p);
}
// n red-black sweeps in place on u (level 0 formulas; f NULL -> Laplace).
void orc_op_smooth(int segX, int segY, double hx, double hy, const double *f, double *u, int n, int smoother,
                   double omega) {
  void *p = orc_create(0, segX, segY, hx, hy, f, 1);
  orc_set_smoother(p, smoother, omega);
  MG<double> &mg = *(MG<double> *)((OrcHandle *)p)->mg;
  std::copy(u, u + mg.lv[0].u.size(), mg.lv[0].u.begin());
  mg.smooth(0, n);
  std::copy(mg.lv[0].u.begin(), mg.lv[0].u.end(), u);
  orc_destroy(p);
}
// Double-double arithmetic: op 0 add, 1 sub, 2 mul, 3 div on (hi, lo) pairs.
void orc_dd_op(int op, const double *a, const double *b, double *out) {
  DD x(a[0], a[1]), y(b[0], b[1]), r;
  if (op == 0) r = x + y;
  else if (op == 1) r = x - y;
  else if (op == 2) r = x * y;
  else r = x / y;
  out[0] = r.hi;
  out[1] = r.lo;
}

}  // extern "C"
