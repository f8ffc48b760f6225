// C++ host layer over the solver's C ABI: an IScheme-shaped scheme object,
// the input-file reader and the model loop of the reference's MultiGrid runs.
#pragma once
#include <stdexcept>
#include <string>
#include <vector>

#include "vlp2d_host.h"

namespace vlp2d {

// Failure of a solver call (the reference's constructor and doIteration throw).
class SolverError : public std::runtime_error {
 public:
  SolverError(int code, const std::string &msg) : std::runtime_error(msg), code_(code) {}
  int code() const { return code_; }

 private:
  int code_;
};

// The MultiGrid scheme on one GPU, with the calls of IScheme<double>
// (VisualLaplacePoisson2D/Model/CPU/IScheme.cs:16-45) that a MultiGrid run
// uses.  Functions are expression texts (see vlp2d_mg.h) instead of managed
// delegates; the grid lives on the device, so the host mirror un0 of
// Iterative1DScheme exists only as getArray()'s result.
class MultiGridSchemeHIP {
 public:
  // ctor (MultiGridSchemeCU.cs:32-134): f_rhs "" = Laplace
  MultiGridSchemeHIP(int segX, int segY, double hx, double hy, const std::string &f_rhs, double eps, int device,
                     const vlp2d_mg_opts *opts = nullptr);
  ~MultiGridSchemeHIP();
  MultiGridSchemeHIP(const MultiGridSchemeHIP &) = delete;
  MultiGridSchemeHIP &operator=(const MultiGridSchemeHIP &) = delete;

  void getArrayDimensions(int &dim1, int &dim2) const {
    dim1 = segX_ + 1;
    dim2 = segY_ + 1;
  }
  // initTopBottomBorders + initLeftRightBorders in the reference's order;
  // vmin / vmax are widened by every value written
  void initBorders(const std::string &bottom, const std::string &top, const std::string &left,
                   const std::string &right, const std::string &boundary, double &vmin, double &vmax);
  void initInitialIterationMean(double value);
  void initInitialIterationArithmeticMean();
  void initInitialIterationLinearInterpolation();
  void initInitialIterationWeightLinearInterpolation();
  void initAfterBoundariesAndInitialIterationInited() {}  // boundary and guess are already on the device
  // one V-cycle; returns the update norm max |u - u_before_post_smoothing|
  double doIteration(int iter);
  std::vector<double> getArray() const;
  void pointsMinMax(double &vmin, double &vmax) const;
  // u - analytic on nodes i, j >= 1; diff (may be null) receives the array
  void calculateDifference(std::vector<double> *diff, const std::string &analytic, double &dmin, double &dmax) const;
  void cancelIterations();
  void cleanup();
  vlp2d_mg_handle handle() const { return h_; }

 private:
  void check(int rc, const char *what) const;
  vlp2d_mg_handle h_ = nullptr;
  int segX_, segY_;
};

// The input file (InputParser.parseFile); throws SolverError on errors.
vlp2d_input parse_input_file(const std::string &path);

// One MultiGrid run of an input (VLPRectangleModel.allIterations_DoWork).
vlp2d_run_result run_model(const vlp2d_input &in, int init_kind, int device, const vlp2d_mg_opts *opts = nullptr,
                           int max_cycles = 0);

// Segments rounded up to a power of two (Utils.calculatePowOf2, Common/Utils.cs:326-332)
int pow2_segments(int segments);

}  // namespace vlp2d
