// spectral_kern.hpp -- registry of the fused twofft / cosft1 / cosft2 kernels
// (one table per precision, instantiated in spectral_f32.hip / spectral_f64.hip).
#pragma once
#include <utility>

#include "c2c_instances.hpp"
#include "spectral.hpp"

namespace nrfft {

// Fused: the FFT of a transform fits one group -- L <= 16384 in FP32 (139 KB
// of LDS), L <= 8192 in FP64.
template <typename T>
constexpr int spec_max_logl() {
    return sizeof(T) == 4 ? 14 : 13;
}
constexpr int SPEC_MAX_LOGL = 14;

struct SpecRef {
    const void* twofft;
    const void* cosft1;
    const void* cosft2[2];  // [0] forward, [1] inverse
    int nt, ti;
};

template <typename T, int LOGL>
SpecRef spec_ref_one() {
    constexpr int L = 1 << LOGL;
    constexpr int TI = ti_row<T, L>();
    constexpr int EPT = L < 16 ? L : 16;
    constexpr int NT = TI * L / EPT;
    return SpecRef{reinterpret_cast<const void*>(&k_twofft<T, L, TI, NT>),
                   reinterpret_cast<const void*>(&k_cosft1<T, L, TI, NT>),
                   {reinterpret_cast<const void*>(&k_cosft2<T, L, TI, NT, 1>),
                    reinterpret_cast<const void*>(&k_cosft2<T, L, TI, NT, -1>)},
                   NT,
                   TI};
}

template <typename T, int... LS>
void spec_fill(SpecRef* tab, std::integer_sequence<int, LS...>) {
    ((tab[LS + 1] = spec_ref_one<T, LS + 1>()), ...);
}

template <typename T>
struct SpecTable {
    SpecRef t[SPEC_MAX_LOGL + 1];
    SpecTable() {
        for (auto& e : t) e = SpecRef{nullptr, nullptr, {nullptr, nullptr}, 0, 0};
        spec_fill<T>(t, std::make_integer_sequence<int, spec_max_logl<T>()>{});
    }
};

const SpecRef& spec_lookup_f32(int logl);
const SpecRef& spec_lookup_f64(int logl);

}  // namespace nrfft
