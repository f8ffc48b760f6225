// vlp2d_solve: runs the MultiGrid method on one of the reference's input
// files on the GPU and prints the run's figures as one JSON line.
//
//   vlp2d_solve <input.txt> [--init mean|amean|linear|wlinear] [--device N]
//               [--mixed] [--cycle nu1,nu2] [--max-cycles K]
//               [--smoother rbgs|jacobi[:omega]] [--restriction half|full] [--fmg]
//
// Without --cycle the scheme runs the reference's V(3,3) cycle with its update
// norm stop (epsilonDouble of the file).  --smoother jacobi (weighted Jacobi,
// omega default 0.8), --restriction full (full weighting) and --fmg (each
// iteration the reference's FMG, FP64) are additions to the reference's
// scheme.
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>

#include "vlp2d_host.hpp"

int main(int argc, char **argv) {
  if (argc < 2) {
    std::fprintf(stderr, "usage: %s <input.txt> [--init mean|amean|linear|wlinear] [--device N] [--mixed] "
                         "[--cycle nu1,nu2] [--max-cycles K] [--smoother rbgs|jacobi[:omega]] "
                         "[--restriction half|full] [--fmg]\n", argv[0]);
    return 2;
  }
  int init = VLP2D_INIT_MEAN, device = 0, max_cycles = 0;
  vlp2d_mg_opts o;
  vlp2d_mg_default_opts(&o);
  for (int k = 2; k < argc; ++k) {
    const std::string a = argv[k];
    auto next = [&]() -> const char * {
      if (k + 1 >= argc) {
        std::fprintf(stderr, "%s needs a value\n", a.c_str());
        std::exit(2);
      }
      return argv[++k];
    };
    if (a == "--init") {
      const std::string v = next();
      init = v == "mean" ? VLP2D_INIT_MEAN : v == "amean" ? VLP2D_INIT_ARITHMETIC_MEAN
           : v == "linear" ? VLP2D_INIT_LINEAR : v == "wlinear" ? VLP2D_INIT_WEIGHT_LINEAR : -1;
      if (init < 0) {
        std::fprintf(stderr, "unknown initial guess %s\n", v.c_str());
        return 2;
      }
    } else if (a == "--device") {
      device = std::atoi(next());
    } else if (a == "--mixed") {
      o.precision = VLP2D_PREC_MIXED;
    } else if (a == "--cycle") {
      if (std::sscanf(next(), "%d,%d", &o.nu1, &o.nu2) != 2) {
        std::fprintf(stderr, "--cycle needs nu1,nu2\n");
        return 2;
      }
    } else if (a == "--max-cycles") {
      max_cycles = std::atoi(next());
    } else if (a == "--smoother") {
      const std::string v = next();
      if (v == "rbgs") {
        o.smoother = VLP2D_SMOOTH_RBGS;
      } else if (v == "jacobi" || v.rfind("jacobi:", 0) == 0) {
        o.smoother = VLP2D_SMOOTH_WJACOBI;
        if (v.size() > 7) {
          char *end = nullptr;
          o.omega = std::strtod(v.c_str() + 7, &end);
          if (!end || *end || !(o.omega > 0.0 && o.omega <= 1.0)) {
            std::fprintf(stderr, "jacobi omega must be in (0, 1]: %s\n", v.c_str());
            return 2;
          }
        }
      } else {
        std::fprintf(stderr, "unknown smoother %s\n", v.c_str());
        return 2;
      }
    } else if (a == "--restriction") {
      const std::string v = next();
      if (v != "half" && v != "full") {
        std::fprintf(stderr, "unknown restriction %s\n", v.c_str());
        return 2;
      }
      o.restriction = v == "full" ? VLP2D_RESTRICT_FULL : VLP2D_RESTRICT_HALF;
    } else if (a == "--fmg") {
      o.cycle = VLP2D_CYCLE_FMG;
    } else {
      std::fprintf(stderr, "unknown option %s\n", a.c_str());
      return 2;
    }
  }
  try {
    const vlp2d_input in = vlp2d::parse_input_file(argv[1]);
    const vlp2d_run_result r = vlp2d::run_model(in, init, device, &o, max_cycles);
    std::printf("{\"input\": \"%s\", \"segments\": [%d, %d], \"steps\": [%.17g, %.17g], \"eps\": %.3g, "
                "\"cycles\": %d, \"elapsed_iters\": %d, \"budget\": %d, \"converged\": %s, \"last_metric\": %.6g, "
                "\"init_s\": %.6f, \"solve_s\": %.6f, \"min\": %.17g, \"max\": %.17g, \"diff_min\": %.6g, "
                "\"diff_max\": %.6g}\n",
                in.name, r.seg_x, r.seg_y, r.hx, r.hy, r.eps, r.cycles, r.elapsed_iters, in.mg_iters[init],
                r.converged ? "true" : "false", r.last_metric, r.init_seconds, r.solve_seconds, r.vmin, r.vmax,
                r.dmin, r.dmax);
    return r.converged ? 0 : 1;
  } catch (const std::exception &e) {
    std::fprintf(stderr, "%s\n", e.what());
    return 1;
  }
}
