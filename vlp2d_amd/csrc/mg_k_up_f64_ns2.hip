// Ascent launchers (k_sup instances) for double levels, 2 post-sweeps.
// One translation unit per precision and sweep count, compiled in parallel.
#include "mg_launch.hpp"

namespace vlp2d {

template int launch_up_ns<double, 2>(Ctx *, Lvl<double> &, const Lvl<double> &, int, bool, bool, const MixView &, int);

}  // namespace vlp2d
