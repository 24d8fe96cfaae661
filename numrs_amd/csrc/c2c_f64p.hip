// Instantiations of the C2C pass kernel for double, isign = 1.
#include "c2c_instances.hpp"
#include "internal.hpp"

namespace nrfft {
const KernelRef& c2c_lookup_f64p(int mode, int logl, int tilog) {
    static const C2CTable<double, 1> table;
    return table.t[mode][logl][tilog];
}
}  // namespace nrfft
