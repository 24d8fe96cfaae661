// pass_c2c.hpp -- one batched C2C FFT pass over a strided 3-level batch.
//
// Transform b = (b0, b1, b2) (b0 fastest), element i:
//   in [b0*si0 + b1*si1 + b2*si2 + i*esi]  ->  out[b0*so0 + b1*so1 + b2*so2 + k*eso]
// A work-group takes TI consecutive b0.  Optional four-step twiddle on the
// output: element k of transform b is multiplied by W_N^{S k (b0 >> tw_b0_shift)},
// N = 2^tw_logn.  Every element a group reads is loaded before any element
// is stored, and groups own disjoint element sets when in == out with equal
// in/out strides, so such passes may run in place.
#pragma once
#include "device_fft.hpp"

namespace nrfft {

struct PassDesc {
    int64_t nb0, nb1, nb2;  // batch counts
    int64_t si1, si2;       // 64-bit batch strides (group base)
    int64_t so1, so2;
    int32_t si0, esi;       // 32-bit strides inside a tile
    int32_t so0, eso;
    int32_t tw_logn;        // 0 = no four-step twiddle
    int32_t tw_b0_shift;
};

template <typename T, int S, int L, int TI, int NT, int ML, int MS>
__global__ __launch_bounds__(NT) void k_c2c(const cpx<T>* in, cpx<T>* out, PassDesc d, Twiddles<T> tw) {
    using E = Engine<T, S, L, TI, NT, ML, MS>;
    __shared__ cpx<T> lds[E::NST > 1 ? E::LDS_ELEMS : 1];
    const int64_t tiles0 = (d.nb0 + TI - 1) / TI;
    int64_t blk = blockIdx.x;
    const int64_t tile0 = blk % tiles0;
    blk /= tiles0;
    const int64_t b1 = blk % d.nb1;
    const int64_t b2 = blk / d.nb1;
    const int64_t b0base = tile0 * TI;
    const int64_t rem = d.nb0 - b0base;
    const int nvalid = rem < TI ? (int)rem : TI;
    const cpx<T>* ib = in + b0base * d.si0 + b1 * d.si1 + b2 * d.si2;
    cpx<T>* ob = out + b0base * d.so0 + b1 * d.so1 + b2 * d.so2;
    const int si0 = d.si0, esi = d.esi, so0 = d.so0, eso = d.eso;
    auto ld = [&](int t, int i) -> cpx<T> {
        if (TI == 1 || t < nvalid) return ib[t * si0 + i * esi];
        return cpx<T>{(T)0, (T)0};
    };
    if (d.tw_logn) {
        const int logn = d.tw_logn, sh = d.tw_b0_shift;
        auto st = [&](int t, int k, cpx<T> v) {
            if (TI == 1 || t < nvalid) {
                const uint32_t n2 = (uint32_t)((b0base + t) >> sh);
                v = cmul(v, twiddle_big<S>(tw, (uint32_t)k * n2, logn));
                ob[t * so0 + k * eso] = v;
            }
        };
        E::run(threadIdx.x, lds, tw.coarse, ld, st);
    } else {
        auto st = [&](int t, int k, cpx<T> v) {
            if (TI == 1 || t < nvalid) ob[t * so0 + k * eso] = v;
        };
        E::run(threadIdx.x, lds, tw.coarse, ld, st);
    }
}

constexpr int C2C_MAX_LOGL_ = 13;  // single-pass transforms up to 8192 points
constexpr int C2C_NTI_ = 7;        // tile widths 1..64
// Load/store mappings of a pass: J = along the transform, T = across transforms.
enum C2CMode { MODE_JJ = 0, MODE_TT = 1, MODE_JT = 2, MODE_TJ = 3, C2C_NMODES = 4 };

// Kernel registry entry.
struct KernelRef {
    const void* fn;   // host stub of the kernel (nullptr = not instantiated)
    int nt;           // threads per group
    int ti;           // transforms per group
    int lds_bytes;    // static LDS
};

}  // namespace nrfft
