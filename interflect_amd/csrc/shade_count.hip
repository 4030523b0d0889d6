// Counting instantiation of the shading kernels (per-lane traversal with the
// operation tally of CountStack).
#include "ifx_shadek.h"

namespace ifx {
void launch_shade_count(const ShadeLaunch& a) {
    if (a.staged) launch_shade_t<true, true, false>(a);
    else launch_shade_t<true, false, false>(a);
}
}  // namespace ifx
