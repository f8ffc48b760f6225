// Descent launchers (k_sdown instances) for float levels, 3 pre-sweeps.
// One translation unit per precision and sweep count, compiled in parallel.
#include "mg_launch.hpp"

namespace vlp2d {

template void launch_down_ns<float, 3>(Ctx *, Lvl<float> &, Lvl<float> &, bool, bool);
template int launch_down_r64_ns<3>(Ctx *, Lvl<float> &, Lvl<float> &, const MixView &);

}  // namespace vlp2d
