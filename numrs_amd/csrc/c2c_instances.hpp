// c2c_instances.hpp -- instantiates k_c2c over (log2 L, tile width, mapping)
// for one element type and fills the lookup table the planner reads.
#pragma once
#include <utility>

#include "pass_c2c.hpp"

namespace nrfft {

constexpr int lds_limit_bytes() { return 160 * 1024; }

// Row tiles (MODE_JJ): TI*L = 4096 (f32) / 2048 (f64) points per group.
template <typename T, int L>
constexpr int ti_row() {
    constexpr int TE = sizeof(T) == 4 ? 4096 : 2048;
    if constexpr (L < 16) return sizeof(T) == 4 ? 256 : 128;
    else return L >= TE ? 1 : TE / L;
}

template <typename T, int S, int LOGL, int MODE, int TILOG>
KernelRef c2c_ref_one() {
    constexpr int L = 1 << LOGL;
    constexpr int EPT = L < 16 ? L : 16;
    constexpr int TI = MODE == MODE_JJ ? ti_row<T, L>() : (1 << TILOG);
    constexpr int NT = TI * L / EPT;
    constexpr int ML = (MODE == MODE_JJ || MODE == MODE_JT) ? MAP_J : MAP_T;
    constexpr int MS = (MODE == MODE_JJ || MODE == MODE_TJ) ? MAP_J : MAP_T;
    constexpr int NST = RadixPlan<L>::NST;
    constexpr int LDS = NST > 1 ? kc2c_lds_bytes<T, S, L, TI, NT, ML, MS>() : (int)sizeof(cpx<T>);
    constexpr bool ok = (MODE != MODE_JJ || TILOG == 0) && NT >= 1 && NT <= 1024 && LDS <= lds_limit_bytes();
    if constexpr (ok) {
        return KernelRef{reinterpret_cast<const void*>(&k_c2c<T, S, L, TI, NT, ML, MS>), NT, TI, LDS};
    } else {
        return KernelRef{nullptr, 0, 0, 0};
    }
}

template <typename T, int S, int MODE, int LOGL, int... TIS>
void c2c_fill_ti(KernelRef* row, std::integer_sequence<int, TIS...>) {
    ((row[TIS] = c2c_ref_one<T, S, LOGL, MODE, TIS>()), ...);
}

template <typename T, int S, int MODE, int... LS>
void c2c_fill(KernelRef (*tab)[C2C_NTI_], std::integer_sequence<int, LS...>) {
    (c2c_fill_ti<T, S, MODE, LS + 1>(tab[LS + 1], std::make_integer_sequence<int, C2C_NTI_>{}), ...);
}

// table[mode][logL][tilog]
template <typename T, int S>
struct C2CTable {
    KernelRef t[C2C_NMODES][C2C_MAX_LOGL_ + 1][C2C_NTI_];
    C2CTable() {
        for (auto& a : t)
            for (auto& b : a)
                for (auto& c : b) c = KernelRef{nullptr, 0, 0, 0};
        c2c_fill<T, S, MODE_JJ>(t[MODE_JJ], std::make_integer_sequence<int, C2C_MAX_LOGL_>{});
        c2c_fill<T, S, MODE_TT>(t[MODE_TT], std::make_integer_sequence<int, C2C_MAX_LOGL_>{});
        c2c_fill<T, S, MODE_JT>(t[MODE_JT], std::make_integer_sequence<int, C2C_MAX_LOGL_>{});
        c2c_fill<T, S, MODE_TJ>(t[MODE_TJ], std::make_integer_sequence<int, C2C_MAX_LOGL_>{});
    }
};

}  // namespace nrfft
