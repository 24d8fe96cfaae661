// spectral_f32.hip -- fused twofft / cosft1 / cosft2 kernels and the
// load-mapped first passes, float.
#include "spectral_kern.hpp"

namespace nrfft {

const SpecRef& spec_lookup_f32(int logl) {
    static const SpecTable<float> tab;
    return tab.t[logl];
}

const KernelRef& map_lookup_f32(int lmap, int logl, int tilog) {
    static const MapTable<float> tab;
    static const KernelRef none{nullptr, 0, 0, 0};
    if (lmap <= LMAP_PLAIN || lmap >= LMAP_COUNT || logl < MAP_MIN_LOGL || logl > MAP_MAX_LOGL ||
        tilog < MAP_MIN_TILOG || tilog > MAP_MAX_TILOG)
        return none;
    return tab.t[lmap][logl][tilog - MAP_MIN_TILOG];
}

}  // namespace nrfft
