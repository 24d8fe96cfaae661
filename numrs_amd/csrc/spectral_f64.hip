// spectral_f64.hip -- fused twofft / cosft1 / cosft2 kernels, double.
#include "spectral_kern.hpp"

namespace nrfft {

const SpecRef& spec_lookup_f64(int logl) {
    static const SpecTable<double> tab;
    return tab.t[logl];
}

}  // namespace nrfft
