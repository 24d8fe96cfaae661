// pass_dma.hpp -- FP32 column pass (MODE_TT: loads and stores both mapped
// across transforms) whose next tile travels HBM -> LDS on the LDS-DMA path
// (global_load_lds_dwordx4) while the current tile is being computed.
//
// A tile is TI = 16 adjacent transforms (b0 .. b0+15) of length L: L rows of
// 16 complex FP32 = 128 contiguous bytes each, row i at element offset
// i * esi.  One DMA instruction moves 64 lanes x 16 B = 8 rows into LDS
// (lane l: row l / 8, bytes 16 (l mod 8) .. +15), so the staging buffer holds
// the tile unpadded as [i][t].  Per tile:
//   wait for this tile's DMA; barrier
//   stage-0 inputs from the staging buffer, stage 0, outputs to the exchange
//   buffer; barrier (the staging buffer is free)
//   issue the next tile's DMA
//   store the previous tile's outputs (held in registers)
//   stages 1.. through the exchange buffer (stage twiddles from an LDS copy,
//   so no global load queues behind the DMA), four-step twiddle; the
//   outputs stay in registers until the next iteration's stores.
// The four-step twiddle's table loads are issued before the next tile's DMA
// (vector memory operations complete in order), and the barriers wait for
// LDS operations only (lds_barrier), so the DMA stays in flight through the
// stages.  Full tiles only (nb0 % 16 == 0), si0 == 1, even esi / si1 / si2,
// a tile's output span below 2^31 bytes, 16-byte aligned input: the planner
// checks these and otherwise uses k_c2c.
#pragma once
#include "pass_c2c.hpp"

namespace nrfft {

typedef __attribute__((address_space(3))) void lds_void_t;

// NRFFT_DMA_STAGED = 1: the next tile lands in a staging buffer of its own
// while all stages of the current tile run (one group per CU at L = 512);
// 0: it lands in the exchange buffer once the last stage has read its inputs
// (no extra LDS, so as many groups per CU as k_c2c).
#ifndef NRFFT_DMA_STAGED
#define NRFFT_DMA_STAGED 1
#endif
constexpr bool DMA_STAGED = NRFFT_DMA_STAGED != 0;

template <int L>
constexpr int dma_lds_bytes() {
    using E = Engine<float, 1, L, DMA_TI, L, MAP_T, MAP_T>;
    return ((DMA_STAGED ? DMA_TI * L : 0) + E::LDS_ELEMS + E::TW_ELEMS) * (int)sizeof(c32);
}

// NRFFT_DMA_DEFER_STORES = 1 (staged form only): a tile's outputs are
// stored one tile later, right after the next DMA is issued, instead of at
// the end of their own tile (where the wait for the next tile's DMA at the
// top of the loop also waits for them).  Measured on MI355X for the
// 512-point pass over 16 transforms x 2^20: 69.5 us deferred against 62.8
// us at the end of the tile (the same pass with its stores removed takes 44
// us), so the default stores at the end of the tile.
#ifndef NRFFT_DMA_DEFER_STORES
#define NRFFT_DMA_DEFER_STORES 0
#endif
constexpr bool DMA_DEFER = NRFFT_DMA_DEFER_STORES != 0 && DMA_STAGED;

// Exchanges for stages s .. LAST: stages s .. LAST-1 computed, stage LAST's
// inputs only loaded (stage s-1 has been computed on v).
template <class E, int s, int LAST>
__device__ __forceinline__ void dma_exchanges(c32* v, int tid, c32* lds, const c32* ltw) {
    if constexpr (s <= LAST) {
        lds_barrier();
        E::template store_lds<s - 1>(v, tid, lds);
        lds_barrier();
        E::template load_lds<s>(v, tid, lds);
        if constexpr (s < LAST) {
            E::template compute<s, LT_ALL>(v, tid, ltw, ltw);
            dma_exchanges<E, s + 1, LAST>(v, tid, lds, ltw);
        }
    }
}

template <int S, int L>
__global__ __attribute__((amdgpu_flat_work_group_size(1, L), amdgpu_waves_per_eu(DMA_STAGED ? 2 : 4))) void
k_c2c_dma(const c32* in, c32* out, PassDesc d, Twiddles<float> tw) {
    constexpr int TI = DMA_TI, NT = L;  // 16 values per thread
    using E = Engine<float, S, L, TI, NT, MAP_T, MAP_T>;
    using P = typename E::P;
    static_assert(E::EPT == 16 && E::NST > 1, "radix-16 engine with at least one exchange");
    constexpr int NW = NT / 64;
    constexpr int NI = L / 8 / NW;  // DMA instructions per wave per tile
    static_assert(L % (8 * NW) == 0, "tile rows split evenly over the waves");
    constexpr bool STAGED = DMA_STAGED;
    __shared__ __align__(16) c32 stg[STAGED ? TI * L : 1];
    __shared__ __align__(16) c32 lds[E::LDS_ELEMS];
    static_assert(E::LDS_ELEMS >= TI * L, "raw tile fits the exchange buffer");
    __shared__ __align__(16) c32 ltw[E::TW_ELEMS];

    const int tid = threadIdx.x;
    const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
    const int lane = tid & 63;
    const int64_t tiles0 = d.nb0 / TI;
    const int64_t ntiles = tiles0 * d.nb1 * d.nb2;
    const int esi = d.esi, so0 = d.so0, eso = d.eso;
    const int64_t G = gridDim.x;
    int64_t tile = blockIdx.x;
    if (d.xcd_remap && (G % 8) == 0) tile = (tile % 8) * (G / 8) + tile / 8;
    if (tile >= ntiles) return;

    struct Ctx {
        const c32* ib;
        c32* ob;
        int64_t b0base;
    };
    auto ctx_of = [&](int64_t tl) {
        Ctx c;
        const int64_t t0 = tl % tiles0;
        const int64_t rest = tl / tiles0;
        const int64_t b1 = rest % d.nb1, b2 = rest / d.nb1;
        c.b0base = t0 * TI;
        c.ib = in + c.b0base + b1 * d.si1 + b2 * d.si2;
        c.ob = out + c.b0base * so0 + b1 * d.so1 + b2 * d.so2;
        return c;
    };
    // Wave w moves the 8-row blocks k * NW + w, k < NI.
    auto dma = [&](const c32* ib) {
        const int r_in = lane >> 3, c_in = 2 * (lane & 7);
#pragma unroll
        for (int k = 0; k < NI; ++k) {
            const int row0 = (k * NW + wave) * 8;
            const c32* g = ib + (row0 + r_in) * esi + c_in;
            if constexpr (STAGED)
                __builtin_amdgcn_global_load_lds((const void*)g, (lds_void_t*)(stg + row0 * TI), 16, 0, 0);
            else
                __builtin_amdgcn_global_load_lds((const void*)g, (lds_void_t*)(lds + row0 * TI), 16, 0, 0);
        }
    };

    constexpr int LR = ilog2c(E::RLAST);
    constexpr int BPTL = E::EPT / E::RLAST;
    auto fw_load = [&](c32 (*fa)[LR + 1], c32 (*fb)[LR + 1], const Ctx& c, int ltid) {
        const int logn = d.tw_logn, sh = d.tw_b0_shift;
        const uint32_t mask = (1u << logn) - 1u;
#pragma unroll
        for (int b = 0; b < BPTL; ++b) {
            int t, j;
            E::template bmap<MAP_T, E::NBLAST>(ltid + b * NT, t, j);
            const uint32_t n2 = uint32_t((c.b0base + t) >> sh);
            const uint32_t m1 = (uint32_t(E::NBLAST) * n2) & mask;
            twiddle_parts(tw, (uint32_t(j) * n2) & mask, logn, fa[b][0], fb[b][0]);
#pragma unroll
            for (int q = 0; q < LR; ++q) twiddle_parts(tw, (m1 << q) & mask, logn, fa[b][1 + q], fb[b][1 + q]);
        }
    };

    // Stage-0 inputs of the current tile, from the raw [i][t] copy the DMA left.
    auto load_raw = [&](c32* v, const c32* src, int ltid) {
        constexpr int R0 = P::radix(0), NB0 = L / R0, BPT0 = 16 / R0;
#pragma unroll
        for (int b = 0; b < BPT0; ++b) {
            int t, j;
            E::template bmap<MAP_T, NB0>(ltid + b * NT, t, j);
#pragma unroll
            for (int r = 0; r < R0; ++r) v[b * R0 + r] = src[(j + r * NB0) * TI + t];
        }
    };
    constexpr int R = E::RLAST, NBL = E::NBLAST;

    // Output k = j + r NBL of transform t: tile base ob (scalar registers)
    // plus the 32-bit byte offset (t so0 + (j + r NBL) eso) 8 (the planner
    // keeps the output span below 2^31 bytes), so a store costs one 32-bit
    // vector add.
    auto out_at = [&](c32* ob, int r, uint32_t vbyte) -> c32* {
        return reinterpret_cast<c32*>(reinterpret_cast<char*>(ob) + (vbyte + uint32_t(r * NBL * eso) * 8u));
    };
    // Last-stage outputs v (four-step twiddle fw applied when set) -> o, in
    // store_last order; then o stored to the tile at ob.
    auto finish = [&](const c32* v, c32 (*fw)[LR + 1], c32* o, int ltid) {
        if (d.tw_logn) {
            int bb = 0;
            E::store_last(v, ltid, [&](int, int, const c32* w) {
                const int b = bb++;
                c32 f[R];
                f[0] = fw[b][0];
#pragma unroll
                for (int r = 1; r < R; ++r) {
                    const int lb = r & (-r);
                    f[r] = cmul(f[r - lb], fw[b][1 + ilog2c(lb)]);
                }
#pragma unroll
                for (int r = 0; r < R; ++r) o[b * R + r] = cmul(w[r], f[r]);
            });
        } else {
#pragma unroll
            for (int e = 0; e < 16; ++e) o[e] = v[e];
        }
    };
    auto store_out = [&](const c32* o, c32* ob, int ltid) {
        int bb = 0;
        E::store_last(o, ltid, [&](int t, int j, const c32*) {
            const int b = bb++;
            const uint32_t vbyte = uint32_t(t * so0 + j * eso) * 8u;
#pragma unroll
            for (int r = 0; r < R; ++r) *out_at(ob, r, vbyte) = o[b * R + r];
        });
    };

    E::stage_tw_to_lds(ltw, tw.stage, tid);  // published by the first loop barrier
    Ctx c = ctx_of(tile);
    dma(c.ib);
    c32 o[16];              // the previous tile's outputs (deferred stores)
    c32* ob_prev = nullptr;
    for (;;) {
        const int64_t next = tile + G;
        const bool has_next = next < ntiles;
        // Staged: 2 waves per SIMD leave 256 VGPRs, so loop-invariant indices
        // stay in registers; otherwise they are recomputed per tile.
        const int ltid = STAGED ? tid : fresh_tid(tid);
        // This tile's DMA (and the previous tile's stores) have completed.
        asm volatile("s_waitcnt vmcnt(0)" ::: "memory");
        lds_barrier();  // every wave's DMA landed; previous tile's exchange reads done

        c32 v[16];
        if constexpr (STAGED) load_raw(v, stg, ltid);
        else load_raw(v, lds, ltid);
        // Twiddle table loads after the raw-tile reads (which the compiler
        // guards with a vmcnt wait for the DMA) and before the next DMA.
        asm volatile("" ::: "memory");
        c32 fa[BPTL][LR + 1], fb[BPTL][LR + 1];
        if (d.tw_logn) fw_load(fa, fb, c, ltid);
        E::template compute<0>(v, ltid, tw.stage);
        if constexpr (!STAGED) lds_barrier();  // raw tile read before the exchange overwrites it
        E::template store_lds<0>(v, ltid, lds);
        lds_barrier();  // (staging buffer fully read;) stage-0 outputs visible
        // Products before the next DMA: a wait for the table loads after a
        // conditional DMA issue would have to wait for the DMA as well.
        c32 fw[BPTL][LR + 1];
        if (d.tw_logn) {
#pragma unroll
            for (int b = 0; b < BPTL; ++b)
#pragma unroll
                for (int q = 0; q <= LR; ++q) fw[b][q] = twiddle_join<S>(fa[b][q], fb[b][q]);
        }
        Ctx cn = c;
        if constexpr (STAGED) {
            if (has_next) {
                cn = ctx_of(next);
                dma(cn.ib);
            }
            if constexpr (DMA_DEFER) {
                if (ob_prev) store_out(o, ob_prev, ltid);
            }
            E::template load_lds<1>(v, ltid, lds);
            E::template compute<1, LT_ALL>(v, ltid, tw.stage, ltw);
            E::template middle<2, LT_ALL>(v, ltid, lds, tw.stage, true, ltw);
        } else {
            // Stages 1 .. NST-2 through the exchange buffer, then the last
            // stage's inputs; once every wave has read them the buffer takes
            // the next tile while the last stage computes and stores.
            E::template load_lds<1>(v, ltid, lds);
            if constexpr (E::NST > 2) {
                E::template compute<1, LT_ALL>(v, ltid, tw.stage, ltw);
                dma_exchanges<E, 2, E::NST - 1>(v, ltid, lds, ltw);
            }
            lds_barrier();
            if (has_next) {
                cn = ctx_of(next);
                dma(cn.ib);
            }
            E::template compute<E::NST - 1, LT_ALL>(v, ltid, tw.stage, ltw);
        }

        finish(v, fw, o, ltid);
        if constexpr (DMA_DEFER) {
            ob_prev = c.ob;
        } else {
            store_out(o, c.ob, ltid);
        }
        if (!has_next) break;
        tile = next;
        c = cn;
    }
    if constexpr (DMA_DEFER) store_out(o, ob_prev, tid);
    // Drain: no DMA is left in flight (the last tile issued none).
}

// The DMA-staged column pass the planner uses.
template <int S, int L>
constexpr const void* dma_kernel() {
    return reinterpret_cast<const void*>(&k_c2c_dma<S, L>);
}

}  // namespace nrfft
