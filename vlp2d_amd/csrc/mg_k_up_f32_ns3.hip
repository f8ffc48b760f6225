// Ascent launchers (k_sup instances) for float levels, 3 post-sweeps.
// One translation unit per precision and sweep count, compiled in parallel.
#include "mg_launch.hpp"

namespace vlp2d {

template void launch_up_down0_ns<3>(Ctx *, Lvl<float> &, Lvl<float> &, int);
template int launch_up_ns<float, 3>(Ctx *, Lvl<float> &, const Lvl<float> &, int, bool, bool, const MixView &, int);

}  // namespace vlp2d
