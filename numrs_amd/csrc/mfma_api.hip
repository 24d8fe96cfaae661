// mfma_api.hip -- split-precision MFMA kernels: DFT-matrix operand tables and
// the kernel registry the planner reads.
#include <utility>

#include "../../include/nrfft.h"
#include "internal.hpp"
#include "mfma_fft.hpp"

namespace nrfft {

// Operand tables of F/4 (F_mp = exp(S 2 pi i mp / 16)) in the A layout of
// v_mfma_f32_16x16x32, one table per output half h: lane l, element e -> row
// q = l % 16 (output m = 8 h + q / 2, its real part for even q, imaginary for
// odd q), column k = 8 (l / 16) + e = 2 p + c (point p, input part c):
//   real rows [Re F_mp, -Im F_mp],  imaginary rows [Im F_mp, Re F_mp];
// values computed in FP64 and split into FP16 hi/lo and BF16 hi/mid/lo parts.
__global__ void k_mfma_tables(uint16_t* f16tab, uint16_t* bf16tab) {
    const int idx = blockIdx.x * blockDim.x + threadIdx.x;  // (sign, half, lane, e)
    if (idx >= 2 * 2 * 64 * 8) return;
    const int e = idx % 8, l = (idx / 8) % 64, ri = (idx / 512) % 2, si = idx / 1024;
    const double S = si ? -1.0 : 1.0;
    const int q = l % 16, m = 8 * ri + q / 2, k = 8 * (l / 16) + e, p = k / 2, c = k % 2;
    const double fr = cospi(m * p / 8.0), fi = S * sinpi(m * p / 8.0);
    double v;
    if (q % 2 == 0) v = c == 0 ? fr : -fi;
    else v = c == 0 ? fi : fr;
    v *= 0.25;
    // FP16 hi / lo
    const _Float16 h = (_Float16)v;
    const _Float16 lo = (_Float16)(v - (double)h);
    f16tab[((si * 2 + 0) * 2 + ri) * 512 + l * 8 + e] = __builtin_bit_cast(uint16_t, h);
    f16tab[((si * 2 + 1) * 2 + ri) * 512 + l * 8 + e] = __builtin_bit_cast(uint16_t, lo);
    // BF16 hi / mid / lo
    const __bf16 bh = (__bf16)(float)v;
    const double r1 = v - (double)(float)bh;
    const __bf16 bm = (__bf16)(float)r1;
    const double r2 = r1 - (double)(float)bm;
    const __bf16 bl = (__bf16)(float)r2;
    bf16tab[((si * 3 + 0) * 2 + ri) * 512 + l * 8 + e] = __builtin_bit_cast(uint16_t, bh);
    bf16tab[((si * 3 + 1) * 2 + ri) * 512 + l * 8 + e] = __builtin_bit_cast(uint16_t, bm);
    bf16tab[((si * 3 + 2) * 2 + ri) * 512 + l * 8 + e] = __builtin_bit_cast(uint16_t, bl);
}

int build_mfma_tables(DeviceCtx* c) {
    if (hipMalloc(&c->mfma_f16, MFMA_TAB_F16 * sizeof(uint16_t)) != hipSuccess ||
        hipMalloc(&c->mfma_bf16, MFMA_TAB_BF16 * sizeof(uint16_t)) != hipSuccess)
        return fail(NRFFT_E_NOMEM, "mfma tables");
    hipLaunchKernelGGL(k_mfma_tables, dim3(8), dim3(256), 0, 0, (uint16_t*)c->mfma_f16, (uint16_t*)c->mfma_bf16);
    if (hipGetLastError() != hipSuccess) return fail(NRFFT_E_HIP, "mfma table kernel");
    return NRFFT_OK;
}

constexpr int MFMA_MIN_LOGL = 8, MFMA_MAX_LOGL = 10;

template <int PREC, int S, int LOGL, int MODE>
KernelRef mfma_ref_one() {
    constexpr int L = 1 << LOGL;
    constexpr int ML = (MODE == MODE_JJ || MODE == MODE_JT) ? MAP_J : MAP_T;
    constexpr int MS = (MODE == MODE_JJ || MODE == MODE_TJ) ? MAP_J : MAP_T;
    if constexpr (MODE == MODE_TJ) {
        return KernelRef{nullptr, 0, 0, 0};
    } else {
        using E = Engine<float, S, L, 16, L, ML, MS>;
        return KernelRef{reinterpret_cast<const void*>(&k_c2c_mfma<PREC, S, L, L, ML, MS>), L, 16,
                         E::LDS_ELEMS * (int)sizeof(c32)};
    }
}

template <int PREC, int S, int LOGL>
void mfma_fill_modes(KernelRef* row) {
    row[MODE_JJ] = mfma_ref_one<PREC, S, LOGL, MODE_JJ>();
    row[MODE_TT] = mfma_ref_one<PREC, S, LOGL, MODE_TT>();
    row[MODE_JT] = mfma_ref_one<PREC, S, LOGL, MODE_JT>();
    row[MODE_TJ] = KernelRef{nullptr, 0, 0, 0};
}

template <int PREC, int S>
struct MfmaTable {
    KernelRef t[MFMA_MAX_LOGL + 1][C2C_NMODES];
    MfmaTable() {
        for (auto& a : t)
            for (auto& b : a) b = KernelRef{nullptr, 0, 0, 0};
        mfma_fill_modes<PREC, S, 8>(t[8]);
        mfma_fill_modes<PREC, S, 9>(t[9]);
        mfma_fill_modes<PREC, S, 10>(t[10]);
    }
};

const KernelRef& mfma_lookup(int prec, int sign, int mode, int logl) {
    static const KernelRef none{nullptr, 0, 0, 0};
    if (logl < MFMA_MIN_LOGL || logl > MFMA_MAX_LOGL || mode < 0 || mode >= C2C_NMODES) return none;
    static const MfmaTable<MP_F16X2, 1> fp;
    static const MfmaTable<MP_F16X2, -1> fm;
    static const MfmaTable<MP_BF16X3, 1> bp;
    static const MfmaTable<MP_BF16X3, -1> bm;
    if (prec == MP_F16X2) return sign > 0 ? fp.t[logl][mode] : fm.t[logl][mode];
    if (prec == MP_BF16X3) return sign > 0 ? bp.t[logl][mode] : bm.t[logl][mode];
    return none;
}

}  // namespace nrfft
