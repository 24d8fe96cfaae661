// spectral_f32.hip -- fused twofft / cosft1 / cosft2 kernels, float.
#include "spectral_kern.hpp"

namespace nrfft {

const SpecRef& spec_lookup_f32(int logl) {
    static const SpecTable<float> tab;
    return tab.t[logl];
}

}  // namespace nrfft
