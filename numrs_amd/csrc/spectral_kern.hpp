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
    const void* r3rows;  // rlft3's fused last axis for rows of L points (16 <= L <= 2048), r3_ti rows a tile
    int r3_nt, r3_ti;
    bool cf1_persist;  // cosft1 is persistent over signals (launch at most the resident groups)
};

// The fused rlft3 last-axis kernel for rows of L points: TI = 4096 / L FP32
// (2048 / L FP64) rows a tile, at least 2.
template <typename T, int L>
constexpr int r3_ti() {
    constexpr int q = (sizeof(T) == 4 ? 4096 : 2048) / L;
    return q < 2 ? 2 : q;
}
template <typename T, int L>
const void* r3rows_ref() {
    if constexpr (L >= 16 && L <= 2048) return reinterpret_cast<const void*>(&k_rlft3_rows<T, L, r3_ti<T, L>(), r3_ti<T, L>() * L / 16>);
    else return nullptr;
}

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
                   TI,
                   r3rows_ref<T, L>(),
                   r3_ti<T, L>() * L / 16,
                   r3_ti<T, L>(),
                   cf1_persistent<T, L, TI, NT>()};
}

template <typename T, int... LS>
void spec_fill(SpecRef* tab, std::integer_sequence<int, LS...>) {
    ((tab[LS + 1] = spec_ref_one<T, LS + 1>()), ...);
}

template <typename T>
struct SpecTable {
    SpecRef t[SPEC_MAX_LOGL + 1];
    SpecTable() {
        for (auto& e : t) e = SpecRef{nullptr, nullptr, {nullptr, nullptr}, 0, 0, nullptr, 0, 0, false};
        spec_fill<T>(t, std::make_integer_sequence<int, spec_max_logl<T>()>{});
    }
};

const SpecRef& spec_lookup_f32(int logl);
const SpecRef& spec_lookup_f64(int logl);

// First four-step passes that read a DCT's or real transform's input
// straight from the caller's reals
// (k_c2c with a load map; isign +1, or -1 for the inverse maps; column mapping
// on both sides), for pass lengths 2^MAP_MIN_LOGL .. 2^MAP_MAX_LOGL and tile
// widths 8 and 16 (the planner's defaults for both precisions).
constexpr int MAP_MIN_LOGL = 7, MAP_MAX_LOGL = 11, MAP_MIN_TILOG = 3, MAP_MAX_TILOG = 4;

template <typename T, int LMAP, int LOGL, int TILOG>
KernelRef map_ref_one() {
    constexpr int L = 1 << LOGL;
    constexpr int TI = 1 << TILOG;
    constexpr int NT = TI * L / 16;
    constexpr int S = (LMAP == LMAP_CF2INV || LMAP == LMAP_RFINV) ? -1 : 1;
    constexpr int LDS = kc2c_lds_bytes<T, S, L, TI, NT, MAP_T, MAP_T>();
    if constexpr (NT <= 1024 && LDS <= lds_limit_bytes()) {
        return KernelRef{reinterpret_cast<const void*>(&k_c2c<T, S, L, TI, NT, MAP_T, MAP_T, LMAP>), NT, TI, LDS};
    } else {
        return KernelRef{nullptr, 0, 0, 0};
    }
}

template <typename T, int LMAP, int... LS>
void map_fill(KernelRef (*row)[MAP_MAX_TILOG - MAP_MIN_TILOG + 1], std::integer_sequence<int, LS...>) {
    ((row[LS + MAP_MIN_LOGL][0] = map_ref_one<T, LMAP, LS + MAP_MIN_LOGL, MAP_MIN_TILOG>()), ...);
    ((row[LS + MAP_MIN_LOGL][1] = map_ref_one<T, LMAP, LS + MAP_MIN_LOGL, MAP_MAX_TILOG>()), ...);
}

template <typename T>
struct MapTable {
    KernelRef t[LMAP_COUNT][MAP_MAX_LOGL + 1][MAP_MAX_TILOG - MAP_MIN_TILOG + 1];  // [load map][log2 L][tile width]
    MapTable() {
        for (auto& a : t)
            for (auto& b : a)
                for (auto& e : b) e = KernelRef{nullptr, 0, 0, 0};
        constexpr int NL = MAP_MAX_LOGL - MAP_MIN_LOGL + 1;
        map_fill<T, LMAP_EVENEXT>(t[LMAP_EVENEXT], std::make_integer_sequence<int, NL>{});
        map_fill<T, LMAP_EVENODD>(t[LMAP_EVENODD], std::make_integer_sequence<int, NL>{});
        map_fill<T, LMAP_CF2INV>(t[LMAP_CF2INV], std::make_integer_sequence<int, NL>{});
        map_fill<T, LMAP_RFINV>(t[LMAP_RFINV], std::make_integer_sequence<int, NL>{});
        map_fill<T, LMAP_PLACE>(t[LMAP_PLACE], std::make_integer_sequence<int, NL>{});
    }
};

const KernelRef& map_lookup_f32(int lmap, int logl, int tilog);
const KernelRef& map_lookup_f64(int lmap, int logl, int tilog);

}  // namespace nrfft
