// Timed shading kernels with wavefront-shared queries, scene staged in LDS.
#include "ifx_shadek.h"

namespace ifx {
void launch_shade_wave_lds(const ShadeLaunch& a) { launch_shade_t<false, true, true>(a); }
}  // namespace ifx
