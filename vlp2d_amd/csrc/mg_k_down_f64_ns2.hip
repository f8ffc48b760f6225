// Descent launchers (k_sdown instances) for double levels, 2 pre-sweeps.
// One translation unit per precision and sweep count, compiled in parallel.
#include "mg_launch.hpp"

namespace vlp2d {

template void launch_down_ns<double, 2>(Ctx *, Lvl<double> &, Lvl<double> &, bool, bool);

}  // namespace vlp2d
