// mfma_fft.hpp -- split-precision MFMA form of the FP32 Stockham pass.
//
// Every radix-16 stage is a 16x16 complex DFT-matrix product on the matrix
// cores.  For 16 vectors x_n (16 complex points each) of one group:
//   [Re Y; Im Y] = [Fr -Fi; Fi Fr] [Re X; Im X],  F_mk = exp(S 2 pi i mk/16)/4
// as two v_mfma_f32_16x16x32 products (rows of Re Y, rows of Im Y; K = 32
// covers Re X and Im X).  Operands are split into 16-bit parts and the
// products accumulate in FP32:
//   F16x2 : x = hi + lo (FP16),  F X ~ Fhi Xhi + Fhi Xlo + Flo Xhi    (3 MFMA)
//   BF16x3: x = hi + mid + lo,   six products of combined order <= 2 (6 MFMA)
// F16 needs range control: each tile is scaled by a power of two so that its
// largest magnitude is < 1, and the 1/4 in F keeps a stage's outputs within
// 4x of its inputs; the exact inverse scale is applied at the final store.
// Twiddles stay FP32 (tables generated in FP64) and are applied by the stage
// that produces a value, for the stage that consumes it.  The last radix-2/4/8
// stage (when log2 L is not a multiple of 4) runs on the VALU.
//
// Operand lane layout of v_mfma_f32_16x16x32_{f16,bf16} (checked on MI355X):
//   A: lane l holds A[l % 16][8 (l / 16) + e],  e = 0..7
//   B: lane l holds B[8 (l / 16) + e][l % 16]
//   D: lane l holds D[4 (l / 16) + e][l % 16],  e = 0..3
// So in a group, lane (n = l % 16, kb = l / 16) supplies 8 consecutive input
// points of vector n (kb & 1 selects points 0-7 / 8-15, kb >> 1 re / im) and
// receives outputs 4 kb .. 4 kb + 3 of vector n.
#pragma once
#include "pass_c2c.hpp"

namespace nrfft {

typedef _Float16 h8v __attribute__((ext_vector_type(8)));
typedef __bf16 b8v __attribute__((ext_vector_type(8)));
typedef float f4v __attribute__((ext_vector_type(4)));

enum MfmaPrec { MP_F16X2 = 1, MP_BF16X3 = 2 };

// DFT-matrix operand tables (built in FP64 on the GPU at device init):
//   f16 [sign][part 0..1][r/i][lane][8]   bf16 [sign][part 0..2][r/i][lane][8]
constexpr int MFMA_TAB_F16 = 2 * 2 * 2 * 64 * 8;
constexpr int MFMA_TAB_BF16 = 2 * 3 * 2 * 64 * 8;

template <int PREC>
struct MfmaTraits;
template <>
struct MfmaTraits<MP_F16X2> {
    static constexpr int PARTS = 2;
    using V = h8v;
    using E = _Float16;
};
template <>
struct MfmaTraits<MP_BF16X3> {
    static constexpr int PARTS = 3;
    using V = b8v;
    using E = __bf16;
};

__device__ __forceinline__ f4v mfma(h8v a, h8v b, f4v c) {
    return __builtin_amdgcn_mfma_f32_16x16x32_f16(a, b, c, 0, 0, 0);
}
__device__ __forceinline__ f4v mfma(b8v a, b8v b, f4v c) {
    return __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c, 0, 0, 0);
}

// x (8 FP32 values) -> 16-bit parts.
__device__ __forceinline__ void split_parts(const float* x, h8v* p) {
#pragma unroll
    for (int e = 0; e < 8; ++e) {
        const _Float16 h = (_Float16)x[e];
        p[0][e] = h;
        p[1][e] = (_Float16)(x[e] - (float)h);
    }
}
__device__ __forceinline__ void split_parts(const float* x, b8v* p) {
#pragma unroll
    for (int e = 0; e < 8; ++e) {
        const __bf16 h = (__bf16)x[e];
        const float r1 = x[e] - (float)h;
        const __bf16 m = (__bf16)r1;
        p[0][e] = h;
        p[1][e] = m;
        p[2][e] = (__bf16)(r1 - (float)m);
    }
}

// D = A B with split operands; products added smallest first.
template <int PREC>
__device__ __forceinline__ f4v split_product(const typename MfmaTraits<PREC>::V* a,
                                             const typename MfmaTraits<PREC>::V* b) {
    f4v d = {0.f, 0.f, 0.f, 0.f};
    if constexpr (PREC == MP_F16X2) {
        d = mfma(a[1], b[0], d);
        d = mfma(a[0], b[1], d);
        d = mfma(a[0], b[0], d);
    } else {
        d = mfma(a[2], b[0], d);
        d = mfma(a[0], b[2], d);
        d = mfma(a[1], b[1], d);
        d = mfma(a[1], b[0], d);
        d = mfma(a[0], b[1], d);
        d = mfma(a[0], b[0], d);
    }
    return d;
}

template <int PREC, int L, int NT>
constexpr int kmfma_waves = waves_per_eu_target<2 * 17 * L * 4, NT, true>();

// Tile: 16 transforms (TI = 16) of length L, LDS planes re/im in the padded
// MAP_T layout of the VALU engine: position p of transform t at p * 17 + t,
// so every access below is a per-lane base plus compile-time offsets.
template <int PREC, int S, int L, int NT, int ML, int MS>
__global__ __attribute__((amdgpu_flat_work_group_size(1, NT), amdgpu_waves_per_eu(kmfma_waves<PREC, L, NT>))) void
k_c2c_mfma(const c32* in, c32* out, PassDesc d, Twiddles<float> tw, const void* amat) {
    using Tr = MfmaTraits<PREC>;
    using V = typename Tr::V;
    constexpr int TI = 16;
    constexpr int LOGL = ilog2c(L);
    constexpr int A = LOGL / 4;             // MFMA stages
    constexpr int REM = 1 << (LOGL % 4);    // trailing VALU radix (1 = none)
    constexpr int NB = L / 16;              // vectors per transform in an MFMA stage
    constexpr int G = TI * L / 256;         // groups per MFMA stage
    constexpr int W = NT / 64;              // waves
    constexpr int GPW = G / W;              // groups per wave
    static_assert(A >= 1 && G % W == 0 && GPW >= 1, "tile must split into whole groups per wave");
    static_assert(NB % 16 == 0 || ML == MAP_T, "row groups need 16 | L/16");

    __shared__ float pre[17 * L];
    __shared__ float pim[17 * L];
    __shared__ float red[W];

    const int tid0 = threadIdx.x;
    const int64_t tiles0 = (d.nb0 + TI - 1) / TI;
    const int64_t ntiles = tiles0 * d.nb1 * d.nb2;
    const int si0 = d.si0, esi = d.esi, so0 = d.so0, eso = d.eso;
    const int64_t Gd = gridDim.x;
    int64_t tile = blockIdx.x;
    if (d.xcd_remap && (Gd % 8) == 0) tile = (tile % 8) * (Gd / 8) + tile / 8;
    if (tile >= ntiles) return;

    // DFT-matrix operands for this lane.
    V ar[Tr::PARTS], ai[Tr::PARTS];
    {
        const uint16_t* base = static_cast<const uint16_t*>(amat) + (S > 0 ? 0 : Tr::PARTS * 2 * 64 * 8);
        const int l = tid0 & 63;
#pragma unroll
        for (int q = 0; q < Tr::PARTS; ++q) {
            const uint16_t* pr = base + ((q * 2 + 0) * 64 + l) * 8;
            const uint16_t* pi = base + ((q * 2 + 1) * 64 + l) * 8;
#pragma unroll
            for (int e = 0; e < 8; ++e) {
                ar[q][e] = __builtin_bit_cast(typename Tr::E, pr[e]);
                ai[q][e] = __builtin_bit_cast(typename Tr::E, pi[e]);
            }
        }
    }

    auto lpos = [](int t, int p) -> int { return p * 17 + t; };

    bool first = true;
    for (;;) {
        const int tid = fresh_tid(tid0);
        const int lane = tid & 63, wave = tid >> 6;
        const int n = lane & 15, kb = lane >> 4;
        const int64_t t0 = tile % tiles0;
        const int64_t rest = tile / tiles0;
        const int64_t b1 = rest % d.nb1, b2 = rest / d.nb1;
        const int64_t b0base = t0 * TI;
        const int64_t remn = d.nb0 - b0base;
        const int nvalid = remn < TI ? (int)remn : TI;
        const c32* ib = in + b0base * si0 + b1 * d.si1 + b2 * d.si2;
        c32* ob = out + b0base * so0 + b1 * d.so1 + b2 * d.so2;

        // ---- stage 0 operands straight from HBM ------------------------
        float x[GPW][8];
#pragma unroll
        for (int g = 0; g < GPW; ++g) {
            const int gid = wave * GPW + g;
            int t, j;
            if constexpr (ML == MAP_T) {
                t = n;
                j = gid;
            } else {
                t = gid / (NB / 16);
                j = (gid % (NB / 16)) * 16 + n;
            }
            const bool ok = t < nvalid;
            const float* src = reinterpret_cast<const float*>(ib + t * si0) + (kb >> 1);
#pragma unroll
            for (int e = 0; e < 8; ++e) {
                const int r = 8 * (kb & 1) + e;
                x[g][e] = ok ? src[2 * (int64_t)(j + r * NB) * esi] : 0.f;
            }
        }
        // Tile scale: power of two putting the largest magnitude below 1.
        float inv_scale = 1.f;
        if constexpr (PREC == MP_F16X2) {
            float m = 0.f;
#pragma unroll
            for (int g = 0; g < GPW; ++g)
#pragma unroll
                for (int e = 0; e < 8; ++e) m = fmaxf(m, fabsf(x[g][e]));
#pragma unroll
            for (int off = 32; off >= 1; off >>= 1) m = fmaxf(m, __shfl_xor(m, off));
            if (!first) __syncthreads();  // red[] and planes read by the previous tile
            if (lane == 0) red[wave] = m;
            __syncthreads();
            m = 0.f;
#pragma unroll
            for (int w = 0; w < W; ++w) m = fmaxf(m, red[w]);
            int ex = 0;
            if (m > 0.f) (void)frexpf(m, &ex);  // m = f 2^ex, f in [0.5, 1)
            const float sc = ldexpf(1.f, -ex);
            inv_scale = ldexpf(1.f, ex);
#pragma unroll
            for (int g = 0; g < GPW; ++g)
#pragma unroll
                for (int e = 0; e < 8; ++e) x[g][e] *= sc;
        } else {
            if (!first) __syncthreads();
        }

        // ---- MFMA stages ------------------------------------------------
        int ns = 1;
#pragma unroll
        for (int s = 0; s < A; ++s) {
            const bool last_mfma = (s == A - 1);
            const bool to_global = last_mfma && REM == 1;
            // Row groups for the first stage of row-loaded tiles and for a
            // final store of row-stored tiles; column groups otherwise.
            const bool rowgrp = (s == 0 && ML == MAP_J) || (to_global && MS == MAP_J);
            if (s > 0) {
                __syncthreads();  // previous stage's LDS writes visible
#pragma unroll
                for (int g = 0; g < GPW; ++g) {
                    const int gid = wave * GPW + g;
                    int t, j;
                    if (rowgrp) {
                        t = gid / (NB / 16);
                        j = (gid % (NB / 16)) * 16 + n;
                    } else {
                        t = n;
                        j = gid;
                    }
                    const float* src = ((kb >> 1) ? pim : pre) + lpos(t, j + 8 * (kb & 1) * NB);
#pragma unroll
                    for (int e = 0; e < 8; ++e) x[g][e] = src[e * NB * 17];
                }
                __syncthreads();  // all reads done before this stage's writes
            }
            f4v yr[GPW], yi[GPW];
#pragma unroll
            for (int g = 0; g < GPW; ++g) {
                V b[Tr::PARTS];
                split_parts(x[g], b);
                yr[g] = split_product<PREC>(ar, b);
                yi[g] = split_product<PREC>(ai, b);
            }
            // outputs: lane (n, kb) holds Y_m of its vector, m = 4 kb + e
            const int ns_next = ns * 16;
#pragma unroll
            for (int g = 0; g < GPW; ++g) {
                const int gid = wave * GPW + g;
                int t, j;
                if (rowgrp) {
                    t = gid / (NB / 16);
                    j = (gid % (NB / 16)) * 16 + n;
                } else {
                    t = n;
                    j = gid;
                }
#pragma unroll
                for (int e = 0; e < 4; ++e) {
                    const int m = 4 * kb + e;
                    const int p = (j / ns) * ns * 16 + (j % ns) + m * ns;
                    c32 v{yr[g][e], yi[g][e]};
                    if (to_global) {
                        if (t < nvalid) {
                            // p = j + m * NB: natural output index
                            v = cscale(v, inv_scale * (float)(1 << (2 * A)));
                            if (d.tw_logn) {
                                const uint32_t n2 = uint32_t((b0base + t) >> d.tw_b0_shift);
                                const uint32_t mask = (1u << d.tw_logn) - 1u;
                                v = cmul(v, twiddle_big<S>(tw, (uint32_t(p) * n2) & mask, d.tw_logn));
                            }
                            ob[t * so0 + p * eso] = v;
                        }
                    } else {
                        // twiddle of the consuming stage (radix R', NS' = ns_next)
                        const int Rn = (s + 1 < A) ? 16 : REM;
                        const int NBn = L / Rn;
                        const int jn = p % NBn, rn = p / NBn;
                        const int kn = jn % ns_next;
                        const int idx = (rn * kn) * (TW_COARSE / (ns_next * Rn));
                        v = cmul(v, dirw<S>(tw.coarse[idx]));
                        const int a = lpos(t, p);
                        pre[a] = v.x;
                        pim[a] = v.y;
                    }
                }
            }
            ns = ns_next;
        }

        // ---- trailing VALU radix-REM stage --------------------------------
        if constexpr (REM > 1) {
            __syncthreads();
            constexpr int NBR = L / REM;
            constexpr int BPT = TI * NBR / NT;  // butterflies per thread
            static_assert(TI * NBR % NT == 0, "butterflies must split evenly");
            const float out_scale = inv_scale * (float)(1 << (2 * A));
#pragma unroll
            for (int b = 0; b < BPT; ++b) {
                const int q = tid + b * NT;
                int t, j;
                if constexpr (MS == MAP_T) {
                    t = q % TI;
                    j = q / TI;
                } else {
                    t = q / NBR;
                    j = q % NBR;
                }
                c32 v[REM];
#pragma unroll
                for (int r = 0; r < REM; ++r) {
                    const int a = lpos(t, j + r * NBR);
                    v[r] = c32{pre[a], pim[a]};
                }
                dft<REM, S>(v);
                if (t < nvalid) {
                    if (d.tw_logn) {
                        const uint32_t n2 = uint32_t((b0base + t) >> d.tw_b0_shift);
                        const uint32_t mask = (1u << d.tw_logn) - 1u;
                        c32 f[REM];
                        twiddle_run<REM, S>(tw, (uint32_t(j) * n2) & mask, (uint32_t(NBR) * n2) & mask, d.tw_logn, f);
#pragma unroll
                        for (int r = 0; r < REM; ++r) ob[t * so0 + (j + r * NBR) * eso] = cmul(cscale(v[r], out_scale), f[r]);
                    } else {
#pragma unroll
                        for (int r = 0; r < REM; ++r) ob[t * so0 + (j + r * NBR) * eso] = cscale(v[r], out_scale);
                    }
                }
            }
        }

        const int64_t next = tile + Gd;
        if (next >= ntiles) break;
        first = false;
        tile = next;
    }
}

}  // namespace nrfft
