// Instantiations of the C2C pass kernel for float, isign = -1.
#include "c2c_instances.hpp"
#include "pass_dma.hpp"
#include "internal.hpp"

namespace nrfft {
const KernelRef& c2c_lookup_f32m(int mode, int logl, int tilog) {
    static const C2CTable<float, -1> table;
    return table.t[mode][logl][tilog];
}
// DMA-staged column passes, log2 L = 8, 9 (else fn = nullptr).
const DmaRef& dma_lookup_f32m(int logl) {
    static const DmaRef none{nullptr, 0, 0};
    static const DmaRef t8{dma_kernel<-1, 256>(), 256, dma_lds_bytes<256>()};
    static const DmaRef t9{dma_kernel<-1, 512>(), 512, dma_lds_bytes<512>()};
    return logl == 8 ? t8 : logl == 9 ? t9 : none;
}
// The single-pass 16384-point row transform (one transform per group of
// 1024 threads, one group per CU).
const KernelRef& c2c16k_lookup_f32m() {
    constexpr int L = 16384, NT = 1024;
    static const KernelRef k{reinterpret_cast<const void*>(&k_c2c<float, -1, L, 1, NT, MAP_J, MAP_J>), NT, 1,
                             kc2c_lds_bytes<float, -1, L, 1, NT, MAP_J, MAP_J>()};
    return k;
}
}  // namespace nrfft
