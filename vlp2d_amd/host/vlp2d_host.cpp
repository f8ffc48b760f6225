// Host layer over the solver's C ABI: input files, the scheme object and the
// model loop of VLP2D's MultiGrid runs.  See vlp2d_host.h.
#include "vlp2d_host.hpp"

#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <limits>
#include <map>
#include <sstream>

namespace vlp2d {

// ---- scheme ------------------------------------------------------------------

void MultiGridSchemeHIP::check(int rc, const char *what) const {
  if (rc != VLP2D_OK) throw SolverError(rc, std::string(what) + ": " + vlp2d_mg_last_error(h_));
}

MultiGridSchemeHIP::MultiGridSchemeHIP(int segX, int segY, double hx, double hy, const std::string &f_rhs, double eps,
                                       int device, const vlp2d_mg_opts *opts)
    : segX_(segX), segY_(segY) {
  const int rc = vlp2d_mg_create_expr(segX, segY, hx, hy, f_rhs.c_str(), eps, device, opts, &h_);
  if (rc != VLP2D_OK) throw SolverError(rc, std::string("create: ") + vlp2d_mg_last_error(nullptr));
}

MultiGridSchemeHIP::~MultiGridSchemeHIP() { cleanup(); }

void MultiGridSchemeHIP::initBorders(const std::string &bottom, const std::string &top, const std::string &left,
                                     const std::string &right, const std::string &boundary, double &vmin,
                                     double &vmax) {
  double lo, hi;
  check(vlp2d_mg_set_boundary_expr(h_, bottom.c_str(), top.c_str(), left.c_str(), right.c_str(), boundary.c_str(),
                                   &lo, &hi),
        "boundary");
  if (lo < vmin) vmin = lo;
  if (hi > vmax) vmax = hi;
}

void MultiGridSchemeHIP::initInitialIterationMean(double value) {
  check(vlp2d_mg_init_guess(h_, VLP2D_INIT_MEAN, value), "initial guess");
}
void MultiGridSchemeHIP::initInitialIterationArithmeticMean() {
  check(vlp2d_mg_init_guess(h_, VLP2D_INIT_ARITHMETIC_MEAN, 0.0), "initial guess");
}
void MultiGridSchemeHIP::initInitialIterationLinearInterpolation() {
  check(vlp2d_mg_init_guess(h_, VLP2D_INIT_LINEAR, 0.0), "initial guess");
}
void MultiGridSchemeHIP::initInitialIterationWeightLinearInterpolation() {
  check(vlp2d_mg_init_guess(h_, VLP2D_INIT_WEIGHT_LINEAR, 0.0), "initial guess");
}

double MultiGridSchemeHIP::doIteration(int) {
  double metric = 0.0;
  check(vlp2d_mg_cycle(h_, &metric), "cycle");
  return metric;
}

std::vector<double> MultiGridSchemeHIP::getArray() const {
  std::vector<double> u((size_t)(segX_ + 1) * (segY_ + 1));
  check(vlp2d_mg_download(h_, u.data()), "download");
  return u;
}

void MultiGridSchemeHIP::pointsMinMax(double &vmin, double &vmax) const {
  check(vlp2d_mg_points_min_max(h_, &vmin, &vmax), "min/max");
}

void MultiGridSchemeHIP::calculateDifference(std::vector<double> *diff, const std::string &analytic, double &dmin,
                                             double &dmax) const {
  double *out = nullptr;
  if (diff) {
    diff->assign((size_t)(segX_ + 1) * (segY_ + 1), 0.0);
    out = diff->data();
  }
  check(vlp2d_mg_difference(h_, analytic.c_str(), &dmin, &dmax, out), "difference");
}

void MultiGridSchemeHIP::cancelIterations() { vlp2d_mg_cancel(h_); }

void MultiGridSchemeHIP::cleanup() {
  vlp2d_mg_destroy(h_);
  h_ = nullptr;
}

// ---- input files ---------------------------------------------------------------

namespace {

std::string trim_cr(std::string s) {
  while (!s.empty() && (s.back() == '\r' || s.back() == '\n')) s.pop_back();
  return s;
}

// every occurrence of key replaced by val (InputParser.cs:107, string.Replace)
std::string replace_all(std::string s, const std::string &key, const std::string &val) {
  if (key.empty()) return s;
  size_t p = 0;
  while ((p = s.find(key, p)) != std::string::npos) {
    s.replace(p, key.size(), val);
    p += val.size();
  }
  return s;
}

void copy_text(char *dst, const std::string &src, const char *what, std::vector<std::string> &errs) {
  if (src.size() >= VLP2D_EXPR_LEN) {
    errs.push_back(std::string(what) + " is longer than " + std::to_string(VLP2D_EXPR_LEN - 1) + " characters");
    dst[0] = 0;
    return;
  }
  std::memcpy(dst, src.c_str(), src.size() + 1);
}

bool parse_int(const std::string &s, int &v) {
  if (s.empty()) return false;
  char *e = nullptr;
  const long r = std::strtol(s.c_str(), &e, 10);
  if (*e != 0 || r < INT32_MIN || r > INT32_MAX) return false;
  v = (int)r;
  return true;
}

bool parse_double(const std::string &s, double &v) {
  if (s.empty()) return false;
  char *e = nullptr;
  v = std::strtod(s.c_str(), &e);
  return *e == 0;
}

// "a, b, c, d" -> four integers (InputParser.cs:241-262 requires at least 4)
bool parse_iters(const std::string &s, int out[4]) {
  std::stringstream ss(s);
  std::string item;
  int k = 0;
  while (std::getline(ss, item, ',')) {
    size_t b = item.find_first_not_of(" \t"), e = item.find_last_not_of(" \t");
    if (b == std::string::npos) return false;
    int v;
    if (!parse_int(item.substr(b, e - b + 1), v)) return false;
    if (k < 4) out[k] = v;
    ++k;
  }
  return k >= 4;
}

}  // namespace

vlp2d_input parse_input_file(const std::string &path) {
  std::ifstream f(path);
  if (!f) throw SolverError(VLP2D_E_ARG, "cannot open " + path);
  std::map<std::string, std::string> kv;
  std::string line;
  while (std::getline(f, line)) {
    line = trim_cr(line);
    const size_t eq = line.find('=');
    if (eq == std::string::npos) continue;
    const std::string val = line.substr(eq + 1);
    if (!val.empty()) kv[line.substr(0, eq)] = val;  // empty values stay absent
  }
  std::vector<std::string> errs;
  vlp2d_input in;
  std::memset(&in, 0, sizeof(in));
  auto get = [&](const char *k) -> std::string {
    auto it = kv.find(k);
    return it == kv.end() ? std::string() : it->second;
  };
  for (const char *k : {"xMax", "yMax", "segmentsX", "segmentsY"})
    if (get(k).empty()) errs.push_back(std::string("required parameter ") + k + " is missing");
  const bool edges = !get("funcLeft").empty() && !get("funcRight").empty() && !get("funcTop").empty() &&
                     !get("funcBottom").empty();
  if (!edges && get("funcBoundary").empty())
    errs.push_back("the boundary needs funcBoundary or all of funcLeft, funcRight, funcTop, funcBottom");
  if (!get("segmentsX").empty() && !parse_int(get("segmentsX"), in.seg_x)) errs.push_back("segmentsX is not an integer");
  if (!get("segmentsY").empty() && !parse_int(get("segmentsY"), in.seg_y)) errs.push_back("segmentsY is not an integer");
  in.eps_double = 1e-13;
  if (!get("epsilon").empty() && !parse_double(get("epsilon"), in.eps_double)) errs.push_back("epsilon is not a number");
  if (!get("epsilonDouble").empty() && !parse_double(get("epsilonDouble"), in.eps_double))
    errs.push_back("epsilonDouble is not a number");
  for (int k = 0; k < 4; ++k) in.mg_iters[k] = 5000;
  for (const char *k : {"MultiGrid", "MultiGridDouble", "MultiGridCUDA", "MultiGridCUDADouble"}) {
    if (get(k).empty()) continue;
    int it[4];
    if (!parse_iters(get(k), it)) errs.push_back(std::string(k) + " needs four integers");
    else std::memcpy(in.mg_iters, it, sizeof(it));  // later keys in this list take precedence
  }
  const std::string xm = get("xMax"), ym = get("yMax");
  auto fn = [&](const char *k) { return replace_all(replace_all(get(k), "xMax", xm), "yMax", ym); };
  copy_text(in.x_max, xm, "xMax", errs);
  copy_text(in.y_max, ym, "yMax", errs);
  copy_text(in.f_rhs, fn("funcRHS"), "funcRHS", errs);
  copy_text(in.f_left, fn("funcLeft"), "funcLeft", errs);
  copy_text(in.f_right, fn("funcRight"), "funcRight", errs);
  copy_text(in.f_top, fn("funcTop"), "funcTop", errs);
  copy_text(in.f_bottom, fn("funcBottom"), "funcBottom", errs);
  copy_text(in.f_boundary, fn("funcBoundary"), "funcBoundary", errs);
  copy_text(in.f_analytic, fn("funcAnalytic"), "funcAnalytic", errs);
  std::string name = path;
  const size_t slash = name.find_last_of("/\\");
  if (slash != std::string::npos) name = name.substr(slash + 1);
  const size_t dot = name.find_last_of('.');
  if (dot != std::string::npos) name = name.substr(0, dot);
  if (name.size() >= sizeof(in.name)) name.resize(sizeof(in.name) - 1);
  std::memcpy(in.name, name.c_str(), name.size() + 1);
  if (!errs.empty()) {
    std::string all;
    for (auto &e : errs) all += e + "\n";
    throw SolverError(VLP2D_E_ARG, all);
  }
  return in;
}

// ---- model -------------------------------------------------------------------

int pow2_segments(int segments) {
  int p = 1;
  while (p < segments) p <<= 1;
  return p;
}

namespace {

double eval_const(const char *text, const char *what) {
  const double zero = 0.0;
  double v = 0.0;
  if (vlp2d_expr_eval(text, &zero, &zero, 1, &v) != VLP2D_OK)
    throw SolverError(VLP2D_E_ARG, std::string(what) + ": " + vlp2d_mg_last_error(nullptr));
  return v;
}

}  // namespace

vlp2d_run_result run_model(const vlp2d_input &in, int init_kind, int device, const vlp2d_mg_opts *opts,
                           int max_cycles) {
  using clk = std::chrono::steady_clock;
  if (init_kind < VLP2D_INIT_MEAN || init_kind > VLP2D_INIT_WEIGHT_LINEAR)
    throw SolverError(VLP2D_E_ARG, "unknown initial guess");
  vlp2d_run_result r;
  std::memset(&r, 0, sizeof(r));
  // StepsRecalculator: power-of-two segments, step = xMax / segments
  r.seg_x = pow2_segments(in.seg_x);
  r.seg_y = pow2_segments(in.seg_y);
  r.hx = eval_const(in.x_max, "xMax") / (double)r.seg_x;
  r.hy = eval_const(in.y_max, "yMax") / (double)r.seg_y;
  r.eps = in.eps_double;
  const int max_iters = max_cycles > 0 ? max_cycles : in.mg_iters[init_kind];
  const auto t0 = clk::now();
  MultiGridSchemeHIP s(r.seg_x, r.seg_y, r.hx, r.hy, in.f_rhs, r.eps, device, opts);
  double bmin = std::numeric_limits<double>::max(), bmax = std::numeric_limits<double>::lowest();
  s.initBorders(in.f_bottom, in.f_top, in.f_left, in.f_right, in.f_boundary, bmin, bmax);
  switch (init_kind) {
    case VLP2D_INIT_MEAN: s.initInitialIterationMean((bmin + bmax) / 2.0); break;
    case VLP2D_INIT_ARITHMETIC_MEAN: s.initInitialIterationArithmeticMean(); break;
    case VLP2D_INIT_LINEAR: s.initInitialIterationLinearInterpolation(); break;
    default: s.initInitialIterationWeightLinearInterpolation(); break;
  }
  s.initAfterBoundariesAndInitialIterationInited();
  const auto t1 = clk::now();
  int iter = 0;
  double metric = std::numeric_limits<double>::infinity();
  for (; iter < max_iters; iter++) {
    metric = s.doIteration(iter);
    ++r.cycles;
    if (metric < r.eps) break;
  }
  const auto t2 = clk::now();
  r.converged = metric < r.eps;
  r.last_metric = metric;
  r.elapsed_iters = iter == 0 ? 1 : iter;
  r.init_seconds = std::chrono::duration<double>(t1 - t0).count();
  r.solve_seconds = std::chrono::duration<double>(t2 - t1).count();
  s.pointsMinMax(r.vmin, r.vmax);
  r.dmin = r.dmax = std::numeric_limits<double>::quiet_NaN();
  if (in.f_analytic[0]) s.calculateDifference(nullptr, in.f_analytic, r.dmin, r.dmax);
  return r;
}

}  // namespace vlp2d

// ---- C ABI ---------------------------------------------------------------------

namespace {
void put_err(char *err, int errlen, const std::string &m) {
  if (err && errlen > 0) {
    std::snprintf(err, (size_t)errlen, "%s", m.c_str());
  }
}
}  // namespace

extern "C" int vlp2d_input_parse(const char *path, vlp2d_input *out, char *err, int errlen) {
  if (!path || !out) {
    put_err(err, errlen, "path and out must be non-NULL");
    return VLP2D_E_ARG;
  }
  try {
    *out = vlp2d::parse_input_file(path);
    put_err(err, errlen, "");
    return VLP2D_OK;
  } catch (const vlp2d::SolverError &e) {
    put_err(err, errlen, e.what());
    return e.code();
  }
}

extern "C" int vlp2d_run_input(const vlp2d_input *in, int init_kind, int device, const vlp2d_mg_opts *opts,
                               int max_cycles, vlp2d_run_result *out, char *err, int errlen) {
  if (!in || !out) {
    put_err(err, errlen, "in and out must be non-NULL");
    return VLP2D_E_ARG;
  }
  try {
    *out = vlp2d::run_model(*in, init_kind, device, opts, max_cycles);
    put_err(err, errlen, "");
    return VLP2D_OK;
  } catch (const vlp2d::SolverError &e) {
    put_err(err, errlen, e.what());
    return e.code();
  } catch (const std::exception &e) {
    put_err(err, errlen, e.what());
    return VLP2D_E_STATE;
  }
}
