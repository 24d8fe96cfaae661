// mfma_fft.hpp -- split-precision MFMA form of the FP32 Stockham pass.
//
// Stage 0 (the 2/4/8 radix, or the first radix 16 when log2 L is a multiple
// of 4) runs on the VALU through the stage engine; every later radix-16 stage
// is a 16x16 complex DFT-matrix product on the matrix cores.  For 16 vectors
// x_n (16 complex points each) of one group, with F_mp = exp(S 2 pi i mp/16)/4,
// two v_mfma_f32_16x16x32 products (halves h = 0, 1) give outputs
// m = 8h .. 8h + 7 as interleaved (Re Y_m, Im Y_m) rows:
//   A_h[2r][2p] = Re F_mp,  A_h[2r][2p+1] = -Im F_mp,
//   A_h[2r+1][2p] = Im F_mp, A_h[2r+1][2p+1] = Re F_mp,   m = 8h + r,
// applied to B = [x.re_0, x.im_0, x.re_1, x.im_1, ...].  With real and
// imaginary parts interleaved along K and along the output rows, a lane's
// eight B values are four consecutive complex points and its accumulators
// hold whole complex outputs, so every global and LDS access moves complex
// values.  Operands are split into 16-bit parts and the products accumulate
// in FP32:
//   F16x2 : x = hi + lo (FP16),  F X ~ Fhi Xhi + Fhi Xlo + Flo Xhi    (3 MFMA)
//   BF16x3: x = hi + mid + lo,   six products of combined order <= 2 (6 MFMA)
// F16 needs range control: each tile is scaled by a power of two so that its
// largest magnitude is < 1, and the 1/4 in F keeps a stage's outputs within
// 4x of its inputs; the exact inverse scale is applied at the final store.
// Twiddles stay FP32 (tables generated in FP64): the consuming stage reads its
// four contiguous per-stage table entries from the group's LDS copy of the
// stage tables, as the VALU engine does.
//
// Operand lane layout of v_mfma_f32_16x16x32_{f16,bf16} (checked on MI355X):
//   A: lane l holds A[l % 16][8 (l / 16) + e],  e = 0..7
//   B: lane l holds B[8 (l / 16) + e][l % 16]
//   D: lane l holds D[4 (l / 16) + e][l % 16],  e = 0..3
// So in a group, lane (n = l % 16, kb = l / 16) supplies points 4 kb .. 4 kb + 3
// of vector n and receives outputs 2 kb, 2 kb + 1 (half 0) and 8 + 2 kb,
// 9 + 2 kb (half 1) of vector n.
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
constexpr int kmfma_waves =
    waves_per_eu_target<(16 * L + 16 * L / NRFFT_LDS_PAD + 1 + Engine<float, 1, L, 16, NT, MAP_T, MAP_T>::TW_ELEMS) * 8,
                        NT, false>();

// Radix-16 MFMA stages s .. NST-1 of one tile whose stage s - 1 output is in
// LDS (a barrier has published it).  Groups of 16 vectors: 16 transforms
// with the same butterfly (column layout) or 16 consecutive butterflies of one
// transform (row layout).
template <int PREC, int S, int L, int NT, int ML, int MS, int s>
__device__ __forceinline__ void mfma_stages(cpx<float>* lds, const cpx<float>* ltw, unsigned lane, unsigned wave,
                                            int nvalid, const Twiddles<float>& tw,
                                            const typename MfmaTraits<PREC>::V* a0,
                                            const typename MfmaTraits<PREC>::V* a1, c32* ob, int so0, int eso,
                                            const PassDesc& d, int64_t b0base, float out_scale) {
    using E = Engine<float, S, L, 16, NT, ML, MS>;
    using P = RadixPlan<L>;
    using Tr = MfmaTraits<PREC>;
    using V = typename Tr::V;
    constexpr unsigned NST = P::NST, NB = L / 16, NS = P::ns(s), W = NT / 64, GPW = (L / 16) / W;
    constexpr bool LAST = s == NST - 1;
    constexpr bool COLG = E::LAYOUT == MAP_T;
    static_assert(P::radix(s) == 16, "MFMA stages are radix 16");
    static_assert(COLG || NB % 16 == 0, "row groups need 16 | L/16");
    const unsigned n = lane & 15u, kb = lane >> 4;
    auto vec = [&](unsigned g, unsigned& t, unsigned& j) {
        const unsigned gid = wave * GPW + g;
        if constexpr (COLG) {
            t = n;
            j = gid;
        } else {
            t = gid / (NB / 16);
            j = (gid % (NB / 16)) * 16 + n;
        }
    };
    // inputs: points 4 kb + e of butterfly j, twiddled by row j mod NS of the
    // stage table (its LDS copy, rows E::twrow(s) entries apart)
    const cpx<float>* stab = ltw + E::twoff(s) + 4 * kb;
    cpx<float> x[GPW][4];
#pragma unroll
    for (unsigned g = 0; g < GPW; ++g) {
        unsigned t, j;
        vec(g, t, j);
        const cpx<float>* row = stab + (j & (NS - 1)) * E::twrow(s);
        if constexpr (E::template affine<NB>) {
            const cpx<float>* src = lds + E::lidx(t, j + 4 * kb * NB);
#pragma unroll
            for (int e = 0; e < 4; ++e) x[g][e] = cmul(src[E::template loff<NB>(e)], dirw<S>(row[e]));
        } else {
#pragma unroll
            for (int e = 0; e < 4; ++e) x[g][e] = cmul(lds[E::lidx(t, j + (4 * kb + e) * NB)], dirw<S>(row[e]));
        }
    }
    if constexpr (!LAST) lds_barrier();  // every read of this stage done before its writes
#pragma unroll
    for (unsigned g = 0; g < GPW; ++g) {
        unsigned t, j;
        vec(g, t, j);
        V b[Tr::PARTS];
        split_parts(reinterpret_cast<const float*>(x[g]), b);
        const f4v y0 = split_product<PREC>(a0, b);  // outputs 2kb, 2kb + 1
        const f4v y1 = split_product<PREC>(a1, b);  // outputs 8 + 2kb, 9 + 2kb
        cpx<float> o[4] = {{y0[0], y0[1]}, {y0[2], y0[3]}, {y1[0], y1[1]}, {y1[2], y1[3]}};
        if constexpr (LAST) {
            // NS = NB: output m of butterfly j is natural index j + m NB
            if ((int)t < nvalid) {
#pragma unroll
                for (int e = 0; e < 4; ++e) o[e] = cpx<float>{o[e].x * out_scale, o[e].y * out_scale};
                if (d.tw_logn) {
                    const uint32_t n2 = uint32_t((b0base + t) >> d.tw_b0_shift);
                    const uint32_t mask = (1u << d.tw_logn) - 1u;
                    const cpx<float> step = twiddle_big<S>(tw, (NB * n2) & mask, d.tw_logn);
                    const cpx<float> w0 = twiddle_big<S>(tw, ((j + 2 * kb * NB) * n2) & mask, d.tw_logn);
                    const cpx<float> w2 = twiddle_big<S>(tw, ((j + (8 + 2 * kb) * NB) * n2) & mask, d.tw_logn);
                    o[0] = cmul(o[0], w0);
                    o[1] = cmul(o[1], cmul(w0, step));
                    o[2] = cmul(o[2], w2);
                    o[3] = cmul(o[3], cmul(w2, step));
                }
                c32* obt = ob + (int64_t)t * so0 + (int64_t)(j + 2 * kb * NB) * eso;
                const int64_t st = (int64_t)NB * eso;
                if (d.nt_out) {  // the plan's result, not re-read by it
                    using V2 = float __attribute__((ext_vector_type(2)));
                    __builtin_nontemporal_store(V2{o[0].x, o[0].y}, reinterpret_cast<V2*>(obt));
                    __builtin_nontemporal_store(V2{o[1].x, o[1].y}, reinterpret_cast<V2*>(obt + st));
                    __builtin_nontemporal_store(V2{o[2].x, o[2].y}, reinterpret_cast<V2*>(obt + 8 * st));
                    __builtin_nontemporal_store(V2{o[3].x, o[3].y}, reinterpret_cast<V2*>(obt + 9 * st));
                } else {
                    obt[0] = o[0];
                    obt[st] = o[1];
                    obt[8 * st] = o[2];
                    obt[9 * st] = o[3];
                }
            }
        } else {
            const unsigned i0 = (j / NS) * NS * 16 + (j & (NS - 1)) + 2 * kb * NS;
            if constexpr (E::template affine<NS>) {
                cpx<float>* dst = lds + E::lidx(t, i0);
                dst[E::template loff<NS>(0)] = o[0];
                dst[E::template loff<NS>(1)] = o[1];
                dst[E::template loff<NS>(8)] = o[2];
                dst[E::template loff<NS>(9)] = o[3];
            } else {
                lds[E::lidx(t, i0)] = o[0];
                lds[E::lidx(t, i0 + NS)] = o[1];
                lds[E::lidx(t, i0 + 8 * NS)] = o[2];
                lds[E::lidx(t, i0 + 9 * NS)] = o[3];
            }
        }
    }
    if constexpr (!LAST) {
        lds_barrier();
        mfma_stages<PREC, S, L, NT, ML, MS, s + 1>(lds, ltw, lane, wave, nvalid, tw, a0, a1, ob, so0, eso, d,
                                                    b0base, out_scale);
    }
}

// Tile: TI = 16 transforms of length L (256..1024), NT = L threads; persistent
// over tiles like k_c2c, with the same PassDesc addressing.
template <int PREC, int S, int L, int NT, int ML, int MS>
__global__ __attribute__((amdgpu_flat_work_group_size(1, NT), amdgpu_waves_per_eu(kmfma_waves<PREC, L, NT>))) void
k_c2c_mfma(const c32* in, c32* out, PassDesc d, Twiddles<float> tw, const void* amat) {
    using Tr = MfmaTraits<PREC>;
    using V = typename Tr::V;
    using E = Engine<float, S, L, 16, NT, ML, MS>;
    constexpr int TI = 16, W = NT / 64;
    constexpr int A = RadixPlan<L>::NST - 1;  // MFMA stages
    static_assert(A >= 1 && E::EPT == 16, "tile shape");
    static_assert(((L / 16) % W) == 0, "groups must split evenly over waves");

    __shared__ cpx<float> lds[E::LDS_ELEMS];
    __shared__ __align__(16) cpx<float> ltw[E::TW_ELEMS];  // stage twiddles, stages 1 .. NST-1
    __shared__ float red[W];

    const int tid0 = threadIdx.x;
    const int64_t tiles0 = (d.nb0 + TI - 1) / TI;
    const int64_t ntiles = tiles0 * d.nb1 * d.nb2;
    const int si0 = d.si0, esi = d.esi, so0 = d.so0, eso = d.eso;
    const int64_t Gd = gridDim.x;
    int64_t tile = blockIdx.x;
    if (d.xcd_remap && (Gd % 8) == 0) tile = (tile % 8) * (Gd / 8) + tile / 8;
    if (tile >= ntiles) return;

    // DFT-matrix operands for this lane (halves 0 and 1).
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

    E::stage_tw_to_lds(ltw, tw.stage, tid0);  // published by the first stage-0 exchange barrier
    bool first = true;
    for (;;) {
        const int tid = fresh_tid(tid0);
        const int lane = tid & 63, wave = tid >> 6;
        const int64_t t0 = tile % tiles0;
        const int64_t rest = tile / tiles0;
        const int64_t b1 = rest % d.nb1, b2 = rest / d.nb1;
        const int64_t b0base = t0 * TI;
        const int64_t remn = d.nb0 - b0base;
        const int nvalid = remn < TI ? (int)remn : TI;
        const c32* ib = in + b0base * si0 + b1 * d.si1 + b2 * d.si2;
        c32* ob = out + b0base * so0 + b1 * d.so1 + b2 * d.so2;

        // ---- stage 0 on the VALU, operands straight from HBM ----------------
        cpx<float> v[16];
        E::load0(v, tid, [&](int t, int i) -> c32 {
            if (t < nvalid) return ib[t * si0 + i * esi];
            return c32{0.f, 0.f};
        });
        float inv_scale = 1.f;
        if constexpr (PREC == MP_F16X2) {
            // tile scale: a power of two putting the largest magnitude below 1
            float m = 0.f;
#pragma unroll
            for (int e = 0; e < 16; ++e) m = fmaxf(m, fmaxf(fabsf(v[e].x), fabsf(v[e].y)));
#pragma unroll
            for (int off = 32; off >= 1; off >>= 1) m = fmaxf(m, __shfl_xor(m, off));
            if (!first) lds_barrier();  // red[] and LDS read by the previous tile
            if (lane == 0) red[wave] = m;
            lds_barrier();
            m = 0.f;
#pragma unroll
            for (int w = 0; w < W; ++w) m = fmaxf(m, red[w]);
            int ex = 0;
            if (m > 0.f) (void)frexpf(m, &ex);  // m = f 2^ex, f in [0.5, 1)
            const float sc = ldexpf(1.f, -ex);
            inv_scale = ldexpf(1.f, ex);
#pragma unroll
            for (int e = 0; e < 16; ++e) v[e] = cscale(v[e], sc);
        } else {
            if (!first) lds_barrier();
        }
        E::template compute<0>(v, tid, tw.stage);
        E::template store_lds<0>(v, tid, lds);
        lds_barrier();

        // ---- radix-16 stages on the matrix cores ----------------------------
        mfma_stages<PREC, S, L, NT, ML, MS, 1>(lds, ltw, (unsigned)lane, (unsigned)wave, nvalid, tw, ar, ai, ob, so0,
                                               eso, d, b0base, inv_scale * (float)(1 << (2 * A)));

        const int64_t next = tile + Gd;
        if (next >= ntiles) break;
        first = false;
        tile = next;
    }
}

}  // namespace nrfft
