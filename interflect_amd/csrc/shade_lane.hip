// Timed shading kernels with the per-lane traversal.
#include "ifx_shadek.h"

namespace ifx {
void launch_shade_lane(const ShadeLaunch& a) {
    if (a.staged) launch_shade_t<false, true, false>(a);
    else launch_shade_t<false, false, false>(a);
}
}  // namespace ifx
