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
#include <type_traits>

#include "device_fft.hpp"
#include "real_fft.hpp"

namespace nrfft {

struct PassDesc {
    int64_t nb0, nb1, nb2;  // batch counts
    int64_t si1, si2;       // 64-bit batch strides (group base)
    int64_t so1, so2;
    int32_t si0, esi;       // 32-bit strides inside a tile
    int32_t so0, eso;
    int32_t tw_logn;        // 0 = no four-step twiddle
    int32_t tw_b0_shift;
    int32_t xcd_remap;      // 1: give each XCD a contiguous run of tiles
    int32_t nt_out;         // 1: non-temporal output stores (the plan does not re-read them)
    int64_t map_arg;        // load map argument (LMAP_PLACE: the response length m)
};

// Persistent pass kernel: group g walks tiles g, g + G, g + 2G, ... (G =
// gridDim.x, sized to the groups that fit on the chip) and prefetches the
// next tile's stage-0 inputs into registers before computing the current
// one, so the global loads of one tile overlap the LDS/VALU work of the
// previous one.  With xcd_remap, group ids are renumbered so the groups of
// one XCD (blockIdx mod 8 on MI355X) take adjacent tiles and share their L2
// lines.
// Register prefetch of the next tile (1) costs occupancy that the no-prefetch
// form (0) turns into more resident groups; 0 measured faster on MI355X.
#ifndef NRFFT_PREFETCH
#define NRFFT_PREFETCH 0
#endif

// Waves per SIMD a kernel is compiled for: as many groups as its LDS lets
// share a CU (MI355X: 160 KB LDS, 4 SIMDs), capped at 4 with register
// prefetch (128 VGPRs per lane) and at 8 otherwise.
template <int LDSB, int NT, bool PF>
constexpr int waves_per_eu_target() {
    constexpr int per_cu = LDSB > 0 ? (163840 / LDSB) : 16;
    constexpr int w = (per_cu * (NT < 64 ? 1 : NT / 64)) / 4;
    constexpr int cap = PF ? 4 : 8;
    return w < 1 ? 1 : (w > cap ? cap : w);
}
// Stages 1 .. kc2c_ltm read their twiddles from an LDS copy: the most
// stages whose copy fits beside the exchange buffer without lowering the
// groups a CU holds (LDS- or wave-limited).  An LDS read replaces an L2 load
// whose latency each exchange otherwise exposes.  NRFFT_PASS_LDS_TW = 0
// reads every stage's twiddles from the global table.
#ifndef NRFFT_PASS_LDS_TW
#define NRFFT_PASS_LDS_TW 1
#endif
// Kernels whose one exchange runs across lanes (Engine::DPP_X) use no LDS:
// no exchange tile and no LDS twiddle copy (stage 1 reads the global table).
template <typename T, int S, int L, int TI, int NT, int ML, int MS>
constexpr bool kc2c_uses_lds() {
    using E = Engine<T, S, L, TI, NT, ML, MS>;
    return E::NST > 1 && !E::DPP_X;
}
template <typename T, int S, int L, int TI, int NT, int ML, int MS>
constexpr int kc2c_ltm() {
    using E = Engine<T, S, L, TI, NT, ML, MS>;
    return (NRFFT_PASS_LDS_TW && kc2c_uses_lds<T, S, L, TI, NT, ML, MS>()) ? engine_ltm<E>(E::LDS_ELEMS * E::CSIZE) : 0;
}
template <typename T, int S, int L, int TI, int NT, int ML, int MS>
constexpr int kc2c_lds_bytes() {
    using E = Engine<T, S, L, TI, NT, ML, MS>;
    constexpr int ltm = kc2c_ltm<T, S, L, TI, NT, ML, MS>();
    return kc2c_uses_lds<T, S, L, TI, NT, ML, MS>() ? (E::LDS_ELEMS + (ltm > 0 ? E::tw_elems(ltm) : 0)) * (int)sizeof(cpx<T>) : 0;
}
// Register prefetch (NRFFT_PREFETCH).  Not for the 1024-thread FP32 pass B
// of 2048 points alone either: its 4 waves per SIMD leave 128 VGPRs, and
// the next tile's 32 would spill 136 bytes per lane.
template <typename T, int NT, int ML, int MS>
constexpr bool kc2c_pf() {
    return NRFFT_PREFETCH != 0;
}
// The cross-lane (no-LDS) kernels are compiled for 4 waves per SIMD: at 8
// their 16 values and the exchange's temporaries spill.
template <typename T, int S, int L, int TI, int NT, int ML, int MS>
constexpr int kc2c_waves = (Engine<T, S, L, TI, NT, ML, MS>::NST > 1 && !kc2c_uses_lds<T, S, L, TI, NT, ML, MS>())
                               ? 4
                               : waves_per_eu_target<kc2c_lds_bytes<T, S, L, TI, NT, ML, MS>(), NT, kc2c_pf<T, NT, ML, MS>()>();

// Thread index laundered through a register move the compiler cannot hoist,
// so per-tile address arithmetic is recomputed inside the persistent loop
// rather than kept live across it.
__device__ __forceinline__ int fresh_tid(int tid) {
    int r;
    asm volatile("v_mov_b32 %0, %1" : "=v"(r) : "v"(tid));
    return r;
}

// Load maps of a column pass (LMAP) over transforms of N = L * esi complex
// points whose input is a real array of d.si1 reals per signal: complex
// element m of a signal is the pair (x_{2m}, x_{2m+1}) of a sequence x read
// straight from the reals y, so no packed copy of x is written first.
//   LMAP_PLAIN   the complex input as it lies (every other pass);
//   LMAP_EVENEXT x_d = y_{min(d, 2N - d)}: the 2N-point even extension of the
//                N + 1 reals y_0 .. y_N (the DCT-I's sequence);
//   LMAP_EVENODD x_i = y_{2i} (i < N), x_{2N-1-i} = y_{2i+1}: the even samples
//                ascending, then the odd ones descending (the DCT-II's).
//   LMAP_CF2INV  z_m of the cosft2 inverse (map_cf2inv): the rotation and
//                realft-inverse untangle of the 2N reals Y, isign -1.
//   LMAP_RFINV   z_m of the realft inverse (rfinv_core): the inverse
//                untangle of the packed spectrum d (N complex), isign -1.
//   LMAP_PLACE   x_e = placed_response(r, e, 2N, m): convlv's response of m
//                = d.map_arg reals padded to 2N (Convolve.rs:41-63).
// The first-pass tile holds every element i of its TI transforms, so with
// si0 == 1 the element is m = b0 + t + i * esi.
enum LoadMap {
    LMAP_PLAIN = 0,
    LMAP_EVENEXT = 1,
    LMAP_EVENODD = 2,
    LMAP_CF2INV = 3,
    LMAP_RFINV = 4,
    LMAP_PLACE = 5,
    LMAP_COUNT = 6
};
// NRFFT_LMAP_PAIR = 1: load-mapped passes run each tile beside its mirror
// on one XCD (ctx_of in k_c2c).  64 x 2^20 FP32 on MI355X: cosft1 0.657 ->
// 0.620 ms (read requests past L2 3.6x -> 1.22x the bytes of y), cosft2
// forward 0.380 -> 0.352 ms (2.0x -> 1.3x).
#ifndef NRFFT_LMAP_PAIR
#define NRFFT_LMAP_PAIR 1
#endif

// Positions in y of x_{2m} and x_{2m+1} under map MAP (N complex points).
template <int MAP>
__device__ __forceinline__ void map_pair(uint32_t m, uint32_t N, uint32_t& p0, uint32_t& p1) {
    const bool lo = 2 * m < N;
    if constexpr (MAP == LMAP_EVENEXT) {
        p0 = lo ? 2 * m : 2 * N - 2 * m;
        p1 = lo ? p0 + 1 : p0 - 1;
    } else {
        p0 = lo ? 4 * m : 4 * N - 4 * m - 1;
        p1 = lo ? p0 + 2 : p0 - 2;
    }
}

// z_m, 0 <= m < N, of the cosft2 inverse (Cos_FT2.rs:116-173 inverse branch
// read with NR's indices; k_cosft2_ipre computes the same per bin pair):
// with n = 2N (logn = log2 n), the pair p = min(m, N - m) is
//   V_k = exp(-i pi k / 2n) (Y_k + i Y_{n-k}),  k = p and N - p,
// untangled with conj W_n^p (isign -1) into z_p and z_{N-p}; z_0 =
// dc_inverse(Y_0, sqrt2 Y_N) and z_{N/2} pairs with itself.
// cf2inv_core: the general element (0 < p < N/2) given r4 = W_{4n}^p;
// W_{4n}^N = exp(i pi / 4) and W_n^p = (W_{4n}^p)^4 need no table.
template <typename T>
__device__ __forceinline__ cpx<T> cf2inv_core(const T* ys, uint32_t p, uint32_t N, cpx<T> r4, bool lo) {
    const uint32_t n = 2 * N;
    const T h = (T)0.70710678118654752440;
    const cpx<T> w8{h, h};
    const cpx<T> r2 = cmul(r4, r4);
    const cpx<T> wn = cmul(r2, r2);
    const cpx<T> va = cmul(conj(r4), cpx<T>{ys[p], ys[n - p]});
    const cpx<T> vb = cmul(conj(cmul(w8, conj(r4))), cpx<T>{ys[N - p], ys[N + p]});
    cpx<T> za, zb;
    untangle_pair(va, vb, conj(wn), (T)0.5, za, zb);
    return cpx<T>{lo ? za.x : zb.x, lo ? za.y : zb.y};  // per component: a struct select went through scratch
}
template <typename T>
__device__ __forceinline__ cpx<T> map_cf2inv(const T* ys, uint32_t m, uint32_t N, int logn, const Twiddles<T>& tw) {
    const uint32_t n = 2 * N, nq = N / 2;
    if (m == 0) return dc_inverse(cpx<T>{ys[0], (T)1.41421356237309504880 * ys[N]});
    if (m == nq) {
        cpx<T> za, zb;
        const cpx<T> vm = cmul(conj(twiddle_big<1>(tw, nq, logn + 2)), cpx<T>{ys[nq], ys[n - nq]});
        untangle_pair(vm, vm, conj(twiddle_big<1>(tw, nq, logn)), (T)0.5, za, zb);
        return zb;
    }
    const bool lo = m < nq;
    const uint32_t p = lo ? m : N - m;
    return cf2inv_core(ys, p, N, twiddle_big<1>(tw, p, logn + 2), lo);
}

// z_m of the realft inverse (Real_FT.rs:119-213 read with NR's indices;
// k_untangle<T, -1> computes the same per bin pair) for 0 < m < N, m != N/2:
// the pair p = min(m, N - m) of the packed spectrum d untangled with
// w = W_n^{-p} (n = 2N) and c2 = 1/2, z_p and z_{N-p}; z_0 = dc_inverse(d_0)
// and z_{N/2} = d_{N/2}.
template <typename T>
__device__ __forceinline__ cpx<T> rfinv_core(const cpx<T>* d, uint32_t p, uint32_t N, cpx<T> w, bool lo) {
    cpx<T> x, y;
    untangle_pair(d[p], d[N - p], w, (T)0.5, x, y);
    return cpx<T>{lo ? x.x : y.x, lo ? x.y : y.y};
}

// exp(pi i r / 16), r < 16: the row step of the realft inverse's W_n (rows
// N/16 apart; index r 16/R0 for a stage-0 radix R0).
template <typename T>
__device__ __forceinline__ cpx<T> w32(int r) {
    constexpr double c[16][2] = {
        {1.0, 0.0},
        {0.980785280403230449126182, 0.195090322016128267848285},
        {0.923879532511286756128183, 0.382683432365089771728460},
        {0.831469612302545237078788, 0.555570233019602224742831},
        {0.707106781186547524400844, 0.707106781186547524400844},
        {0.555570233019602224742831, 0.831469612302545237078788},
        {0.382683432365089771728460, 0.923879532511286756128183},
        {0.195090322016128267848285, 0.980785280403230449126182},
        {0.0, 1.0},
        {-0.195090322016128267848285, 0.980785280403230449126182},
        {-0.382683432365089771728460, 0.923879532511286756128183},
        {-0.555570233019602224742831, 0.831469612302545237078788},
        {-0.707106781186547524400844, 0.707106781186547524400844},
        {-0.831469612302545237078788, 0.555570233019602224742831},
        {-0.923879532511286756128183, 0.382683432365089771728460},
        {-0.980785280403230449126182, 0.195090322016128267848285}};
    return cpx<T>{(T)c[r][0], (T)c[r][1]};
}

// exp(2 pi i r / 128), r < 16: the step between the rows a first-pass
// butterfly loads, in units of the cosft2 inverse's rotation W_{4n} (rows
// N/16 apart; index r 16/R0 for a stage-0 radix R0, rows N/R0 apart).
template <typename T>
__device__ __forceinline__ cpx<T> w128(int r) {
    constexpr double c[16][2] = {
        {1.00000000000000000000, 0.00000000000000000000}, {0.99879545620517240501, 0.04906767432741801493},
        {0.99518472667219692873, 0.09801714032956060363}, {0.98917650996478101444, 0.14673047445536174793},
        {0.98078528040323043058, 0.19509032201612824808}, {0.97003125319454397424, 0.24298017990326387094},
        {0.95694033573220882438, 0.29028467725446233105}, {0.94154406518302080631, 0.33688985339222005111},
        {0.92387953251128673848, 0.38268343236508978178}, {0.90398929312344333820, 0.42755509343028208491},
        {0.88192126434835504956, 0.47139673682599764204}, {0.85772861000027211809, 0.51410274419322166128},
        {0.83146961230254523567, 0.55557023301960217765}, {0.80320753148064494287, 0.59569930449243335691},
        {0.77301045336273699338, 0.63439328416364548779}, {0.74095112535495910588, 0.67155895484701833009}};
    return cpx<T>{(T)c[r][0], (T)c[r][1]};
}

template <typename T, int S, int L, int TI, int NT, int ML, int MS, int LMAP = LMAP_PLAIN>
__global__ __attribute__((amdgpu_flat_work_group_size(1, NT), amdgpu_waves_per_eu(kc2c_waves<T, S, L, TI, NT, ML, MS>))) void
k_c2c(const cpx<T>* in, cpx<T>* out, PassDesc d, Twiddles<T> tw) {
    static_assert(LMAP == LMAP_PLAIN || ML == MAP_T, "load maps belong to column passes");
    using E = Engine<T, S, L, TI, NT, ML, MS>;
    using V2 = T __attribute__((ext_vector_type(2)));
    constexpr int EPT = E::EPT;
    constexpr int LTM = kc2c_ltm<T, S, L, TI, NT, ML, MS>();
    __shared__ cpx<T> lds[kc2c_uses_lds<T, S, L, TI, NT, ML, MS>() ? E::LDS_ELEMS : 1];
    __shared__ __align__(16) cpx<T> ltw[LTM > 0 ? E::tw_elems(LTM) : 1];
    const int tid = threadIdx.x;
    const int64_t tiles0 = (d.nb0 + TI - 1) / TI;
    const int64_t ntiles = tiles0 * d.nb1 * d.nb2;
    const int si0 = d.si0, esi = d.esi, so0 = d.so0, eso = d.eso;
    int64_t G = gridDim.x;
    int64_t tile = blockIdx.x;
    if (d.xcd_remap && (G % 8) == 0) tile = (tile % 8) * (G / 8) + tile / 8;
    if (tile >= ntiles) return;

    struct Ctx {
        const cpx<T>* ib;
        const T* yb;  // LMAP: the signal's reals
        cpx<T>* ob;
        int64_t b0base;
        int nvalid;
    };
    auto ctx_of = [&](int64_t tl) {
        Ctx c;
        int64_t t0 = tl % tiles0;
        if constexpr (LMAP != LMAP_PLAIN && NRFFT_LMAP_PAIR) {
            // Tile q and its mirror tiles0 - 1 - q (which reads the other
            // half of q's lines of y) at tile indices j and j + 8: groups
            // are dealt to the 8 XCDs round-robin (tile index mod 8 = XCD
            // for grids that are multiples of 8), so the pair runs at the
            // same time on one XCD and the second read hits its L2.  XCD
            // j mod 8 takes a contiguous run of q (tiles0 / 16 of them), so
            // the lines that neighbouring tiles share (a signal of cosft1's
            // N + 1 reals starts at any 4-byte offset) stay on it too.
            if (tiles0 % 16 == 0) {
                const int64_t q = (t0 & 7) * (tiles0 >> 4) + (t0 >> 4);
                t0 = (t0 & 8) ? tiles0 - 1 - q : q;
            }
        }
        const int64_t rest = tl / tiles0;
        const int64_t b1 = rest % d.nb1, b2 = rest / d.nb1;
        c.b0base = t0 * TI;
        const int64_t rem = d.nb0 - c.b0base;
        c.nvalid = rem < TI ? (int)rem : TI;
        c.ib = in + c.b0base * si0 + b1 * d.si1 + b2 * d.si2;
        c.yb = reinterpret_cast<const T*>(in) + b1 * d.si1;  // LMAP: si1 counts reals, nb2 == 1
        c.ob = out + c.b0base * so0 + b1 * d.so1 + b2 * d.so2;
        return c;
    };
    auto load = [&](cpx<T>* v, const Ctx& c, int ltid) {
        if constexpr (LMAP == LMAP_CF2INV || LMAP == LMAP_RFINV) {
            // Butterfly (t, j) of stage 0 loads rows i = j + r L/R0, r < R0:
            // elements m = m0 + r N/R0 with m0 < N/R0, so r < R0/2 is the
            // lower half (p = m), and one table twiddle per butterfly gives
            // every element's: CF2INV W_{4n}^m = W_{4n}^{m0} exp(2 pi i r /
            // (8 R0)), the upper half W_{4n}^{N-m} = exp(i pi / 4)
            // conj(W_{4n}^m); RFINV W_n^m = W_n^{m0} exp(pi i r / R0), the
            // lower half w = conj(W_n^m), the upper W_n^{-(N-m)} = -W_n^m.
            // m = 0 and m = N/2 (m0 = 0) are special.
            constexpr int R0 = E::P::radix(0), NB0 = L / R0, BPT0 = EPT / R0;
            static_assert(16 % R0 == 0, "stage-0 radix divides 16");
            const uint32_t N = uint32_t(L) * uint32_t(esi);
            const int logn = __builtin_ctz(N) + 1;
            const uint32_t step = uint32_t(NB0) * uint32_t(esi);
            const T h = (T)0.70710678118654752440;
#pragma unroll
            for (int b = 0; b < BPT0; ++b) {
                int t, j;
                E::template bmap<MAP_T, NB0>(ltid + b * NT, t, j);
                if (TI == 1 || t < c.nvalid) {
                    const uint32_t m0 = uint32_t(c.b0base) + uint32_t(t) + uint32_t(j) * uint32_t(esi);
                    const cpx<T> w0 = twiddle_big<1>(tw, m0, LMAP == LMAP_CF2INV ? logn + 2 : logn);
                    const cpx<T>* d = reinterpret_cast<const cpx<T>*>(c.yb);  // RFINV: the packed spectrum
#pragma unroll
                    for (int r = 0; r < R0; ++r) {
                        const uint32_t m = m0 + uint32_t(r) * step;
                        const bool lo = r < R0 / 2;
                        if constexpr (LMAP == LMAP_CF2INV) {
                            if ((r == 0 || r == R0 / 2) && m0 == 0) {
                                v[b * R0 + r] = map_cf2inv(c.yb, m, N, logn, tw);
                            } else {
                                const cpx<T> wm = cmul(w0, w128<T>(r * (16 / R0)));
                                const cpx<T> r4 = lo ? wm : cmul(cpx<T>{h, h}, conj(wm));
                                v[b * R0 + r] = cf2inv_core(c.yb, lo ? m : N - m, N, r4, lo);
                            }
                        } else {
                            if (r == 0 && m0 == 0) {
                                v[b * R0 + r] = dc_inverse(d[0]);
                            } else if (r == R0 / 2 && m0 == 0) {
                                v[b * R0 + r] = d[N / 2];
                            } else {
                                const cpx<T> wm = cmul(w0, w32<T>(r * (16 / R0)));
                                const cpx<T> w = lo ? conj(wm) : cpx<T>{-wm.x, -wm.y};
                                v[b * R0 + r] = rfinv_core(d, lo ? m : N - m, N, w, lo);
                            }
                        }
                    }
                } else {
#pragma unroll
                    for (int r = 0; r < R0; ++r) v[b * R0 + r] = cpx<T>{(T)0, (T)0};
                }
            }
            return;
        }
        E::load0(v, ltid, [&](int t, int i) -> cpx<T> {
            if (TI == 1 || t < c.nvalid) {
                if constexpr (LMAP == LMAP_PLAIN) {
                    return c.ib[t * si0 + i * esi];
                } else {
                    // 32-bit indices: N <= 2^26 points, 4N reals at most.
                    const uint32_t N = uint32_t(L) * uint32_t(esi);
                    const uint32_t m = uint32_t(c.b0base) + uint32_t(t) + uint32_t(i) * uint32_t(esi);
                    if constexpr (LMAP == LMAP_PLACE) {
                        const int64_t e = 2 * int64_t(m), n2 = 2 * int64_t(N);
                        return cpx<T>{placed_response(c.yb, e, n2, d.map_arg), placed_response(c.yb, e + 1, n2, d.map_arg)};
                    } else {
                        uint32_t p0, p1;
                        map_pair<LMAP>(m, N, p0, p1);
                        return cpx<T>{c.yb[p0], c.yb[p1]};
                    }
                }
            }
            return cpx<T>{(T)0, (T)0};
        });
    };

    // Published by the first exchange barrier, before stage 1 reads it.
    if constexpr (LTM > 0) E::template stage_tw_to_lds<LTM>(ltw, tw.stage, tid);
    constexpr bool PF = kc2c_pf<T, NT, ML, MS>();
    Ctx cn{};
    cpx<T> vn[PF ? EPT : 1];
    if constexpr (PF) {
        cn = ctx_of(tile);
        load(vn, cn, fresh_tid(tid));
    }
    // The tile loop, with plain or non-temporal result stores (NTS): a
    // runtime choice between two stores of the same value would be merged
    // into one plain store by the compiler.
    auto tiles = [&](auto nts) {
        constexpr bool NTS = decltype(nts)::value;
        bool first = true;
        for (;;) {
            cpx<T> v[EPT];
            const int64_t next = tile + G;
            const bool has_next = next < ntiles;
            Ctx c;
            if constexpr (PF) {
#pragma unroll
                for (int e = 0; e < EPT; ++e) v[e] = vn[e];
                c = cn;
                if (has_next) {
                    cn = ctx_of(next);
                    load(vn, cn, fresh_tid(tid));
                }
            } else {
                c = ctx_of(tile);
                load(v, c, fresh_tid(tid));
            }
            const int ltid = fresh_tid(tid);
            E::template compute<0>(v, ltid, tw.stage);
            E::template middle<1, LTM, true>(v, ltid, lds, tw.stage, !first, ltw);
            if (d.tw_logn) {
                const int logn = d.tw_logn, sh = d.tw_b0_shift;
                const uint32_t mask = (1u << logn) - 1u;
                E::store_last(v, ltid, [&](int t, int j, const cpx<T>* w) {
                    if (TI == 1 || t < c.nvalid) {
                        const uint32_t n2 = uint32_t((c.b0base + t) >> sh);
                        cpx<T> f[E::RLAST];
                        twiddle_run<E::RLAST, S>(tw, (uint32_t(j) * n2) & mask, (uint32_t(E::NBLAST) * n2) & mask, logn, f);
#pragma unroll
                        for (int r = 0; r < E::RLAST; ++r) c.ob[t * so0 + (j + r * E::NBLAST) * eso] = cmul(w[r], f[r]);
                    }
                });
            } else {
                E::store_last(v, ltid, [&](int t, int j, const cpx<T>* w) {
                    if (TI == 1 || t < c.nvalid) {
#pragma unroll
                        for (int r = 0; r < E::RLAST; ++r) {
                            cpx<T>* o = &c.ob[t * so0 + (j + r * E::NBLAST) * eso];
                            if constexpr (NTS)
                                __builtin_nontemporal_store(V2{w[r].x, w[r].y}, reinterpret_cast<V2*>(o));
                            else
                                *o = w[r];
                        }
                    }
                });
            }
            if (!has_next) break;
            first = false;
            tile = next;
        }
    };
    if constexpr (ML == MAP_J && MS == MAP_T) {  // a four-step pass B: may store the plan's result
        if (d.nt_out)
            tiles(std::true_type{});
        else
            tiles(std::false_type{});
    } else {
        tiles(std::false_type{});
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

// DMA-staged FP32 column pass (k_c2c_dma): TI = 16, nt = L threads.
// NRFFT_PASS_DMA = 0 builds the library without it (k_c2c everywhere).
#ifndef NRFFT_PASS_DMA
#define NRFFT_PASS_DMA 1
#endif
constexpr int DMA_TI = 16;
struct DmaRef {
    const void* fn;
    int nt;
    int lds_bytes;
};

}  // namespace nrfft
