// Timed shading kernels with wavefront-shared queries, scene in global memory.
#include "ifx_shadek.h"

namespace ifx {
void launch_shade_wave_glb(const ShadeLaunch& a) { launch_shade_t<false, false, true>(a); }
}  // namespace ifx
