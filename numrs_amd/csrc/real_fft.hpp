// real_fft.hpp -- realft / convlv / correl kernels (numrs Real_FT.rs,
// Convolve.rs, Correlation.rs) on the Stockham engine.
//
// A real signal of n = 2L samples is read as L complex values
// z_j = x_{2j} + i x_{2j+1} (the same bytes), transformed with an L-point
// complex FFT and "untangled" into the packed real spectrum
//   [F_0, F_{n/2}, Re F_1, Im F_1, ..., Re F_{L-1}, Im F_{L-1}],
//   F_k = sum_j x_j e^{+2 pi i jk/n}                           (NR realft)
// The untangle pairs bins k and L-k; it is fused into the kernel that runs
// the FFT (the FFT result is left in LDS and one thread handles each pair).
// convlv/correl hold one operand's spectrum in registers while the other's
// is computed in the same LDS buffer, apply the spectrum operation, untangle
// back and run the inverse FFT: R2C -> op -> C2R in one launch.
#pragma once
#include "device_fft.hpp"

namespace nrfft {

// (A at k, B at L-k) -> (out_k, out_{L-k}) with c2 = -isign/2 and
// w = exp(isign i pi k / L):
//   h1 = (A + conj B)/2,  h2 = c2 i (A - conj B),
//   out_k = h1 + w h2,  out_{L-k} = conj(h1 - w h2).
// isign = +1 turns the L-point FFT of z into F (Real_FT.rs:23-117 read with
// NR's indices); isign = -1 turns a packed spectrum into the z' whose
// inverse L-point FFT is (n/2) x (Real_FT.rs:119-213).
template <typename T>
__device__ __forceinline__ void untangle_pair(cpx<T> A, cpx<T> B, cpx<T> w, T c2, cpx<T>& ok, cpx<T>& olk) {
    const T half = (T)0.5;
    cpx<T> h1{half * (A.x + B.x), half * (A.y - B.y)};
    cpx<T> h2{-c2 * (A.y + B.y), c2 * (A.x - B.x)};
    cpx<T> wh = cmul(w, h2);
    ok = h1 + wh;
    olk = conj(h1 - wh);
}

// W_M^p for the index p of a loop over a signal's elements or pairs: W_M^threadIdx (per thread, once)
// times W_M^(p - threadIdx), the same for the whole block (one cache line per
// wave) -- a direct table lookup of consecutive p would step through the
// fine table 2^(26 - log2 M) entries at a time, one cache line per lane.
template <int S, typename T>
struct RowTw {
    cpx<T> wt;
    int logm;
    __device__ __forceinline__ RowTw(const Twiddles<T>& tw, int logm_) : logm(logm_) {
        wt = twiddle_big<S>(tw, uint32_t(threadIdx.x), logm);
    }
    __device__ __forceinline__ cpx<T> at(const Twiddles<T>& tw, int64_t p) const {
        return cmul(wt, twiddle_big<S>(tw, uint32_t(p - threadIdx.x), logm));
    }
};

// Bin pairs per thread in the multi-pass untangle kernels: a block takes
// PAIR_U * blockDim.x consecutive pairs p = base + threadIdx.x + u blockDim.x,
// issues every load of its pairs (data and block-uniform twiddle factors)
// before the first store, then finishes them -- one pair per thread left too
// few loads in flight to cover HBM latency.
// Indices inside one signal are 32-bit (n <= 2^26): 64-bit index arithmetic
// on eight address streams spilled registers.  Blocks are 256 threads.
constexpr int PAIR_U = 4;
constexpr int PAIR_NT = 256;
__device__ __forceinline__ int pair_first() { return int(blockIdx.x) * PAIR_NT * PAIR_U; }
__device__ __forceinline__ int pair_stride() { return int(gridDim.x) * PAIR_NT * PAIR_U; }
__device__ __forceinline__ int pair_at(int p0, int u) { return p0 + int(threadIdx.x) + u * PAIR_NT; }

// Grid of the PAIR_U kernels: y over signals, x over spans of PAIR_NT *
// PAIR_U items of one signal, at most PAIR_BLOCKS blocks, so a block walks
// several spans and its per-thread twiddle factor (one scattered table read
// per thread) is paid once per block, not once per span.
constexpr int64_t PAIR_BLOCKS = 8192;
inline dim3 grid_pairs(int64_t items, int64_t nsig) {
    const int64_t gy = nsig < 1 ? 1 : (nsig > 65535 ? 65535 : nsig);
    const int64_t cap = PAIR_BLOCKS / gy > 0 ? PAIR_BLOCKS / gy : 1;
    int64_t gx = (items + int64_t(PAIR_NT) * PAIR_U - 1) / (int64_t(PAIR_NT) * PAIR_U);
    gx = gx < 1 ? 1 : (gx > cap ? cap : gx);
    return dim3((unsigned)gx, (unsigned)gy);
}

// Slot 0 of the packed array: forward (Z_0 -> [F_0, F_{n/2}]) and inverse.
template <typename T>
__device__ __forceinline__ cpx<T> dc_forward(cpx<T> z0) { return {z0.x + z0.y, z0.x - z0.y}; }
template <typename T>
__device__ __forceinline__ cpx<T> dc_inverse(cpx<T> f) {
    const T half = (T)0.5;
    return {half * (f.x + f.y), half * (f.x - f.y)};
}

// Spectrum operations of convlv (Convolve.rs:90-138, NR semantics: every bin,
// DC and Nyquist included) and correl (Correlation.rs:68-110).
enum SpecOp { OP_CONV = 0, OP_DECONV = 1, OP_CORREL = 2 };

template <typename T>
__device__ __forceinline__ cpx<T> spec_apply(int op, cpx<T> a, cpx<T> b, T inv_no2) {
    if (op == OP_CONV) return cscale(cmul(a, b), inv_no2);
    if (op == OP_CORREL) return cscale(cmulc(a, b), inv_no2);
    // Deconvolution: a conj(b) / |b|^2 / no2, zero where |b|^2 < 1e-12
    // (Convolve.rs:117-128).
    const T mag2 = b.x * b.x + b.y * b.y;
    if (mag2 < (T)1e-12) return cpx<T>{(T)0, (T)0};
    const cpx<T> q = cmulc(a, b);
    return {q.x / mag2 * inv_no2, q.y / mag2 * inv_no2};
}
// DC / Nyquist slot: both parts are real bins.
template <typename T>
__device__ __forceinline__ T spec_apply_real(int op, T a, T b, T inv_no2) {
    if (op == OP_CONV || op == OP_CORREL) return a * b * inv_no2;
    const T mag2 = b * b;
    if (mag2 < (T)1e-12) return (T)0;
    return a * b / mag2 * inv_no2;
}

// Response placement of Convolve.rs:41-63 (prepare_response_function),
// element e of the n-long padded response, mid = (m-1)/2.
template <typename T>
__device__ __forceinline__ T placed_response(const T* r, int64_t e, int64_t n, int64_t m) {
    const int64_t mid = (m - 1) / 2;
    if (e < mid) return r[m - mid + e];
    if (e < m) return r[e];
    if (e < n - mid) return (T)0;
    return r[e - (n - mid)];
}

// Window of the placed response's support in complex indices (z_i = x_{2i}
// + i x_{2i+1}, taken circularly): nonzero reals lie at e < m and e >= n -
// mid, so z_i can be nonzero only for lo <= i < hi with lo = floor((n -
// mid)/2) - L <= 0 and hi = ceil(m/2).  win_start is lo rounded down to a
// multiple of 64; the window [win_start, win_start + 256) covers the support
// when win_start + 256 >= hi.
__host__ __device__ inline int resp_win_lo(int64_t L, int64_t m) { return int((2 * L - (m - 1) / 2) / 2 - L); }
__host__ __device__ inline int resp_win_hi(int64_t m) { return int((m + 1) / 2); }
__host__ __device__ inline int resp_win_start(int64_t L, int64_t m) {
    const int lo = resp_win_lo(L, m);
    return -(((-lo) + 63) / 64) * 64;
}
__host__ __device__ inline bool resp_win_fits(int64_t L, int64_t m) {
    return resp_win_start(L, m) + 256 >= resp_win_hi(m) && -resp_win_lo(L, m) < L;
}

// Slot of bin k of the windowed response's spectrum in the LDS tile (Q =
// L / 256): bits 1-4 of k XOR-ed with (k / Q) mod 16.  The window step's
// writes (lanes across k1 and g at k = k1 + Q (g + 16 kb)) and the pair
// loop's reads (consecutive k) then both touch distinct banks; with the
// tile's padded layout the writes were 4-way conflicts.
template <int Q>
__device__ __forceinline__ int resp_slot(int k) { return k ^ (((k / Q) & 15) << 1); }

// Spectrum Z_k, k < L, of the placed response (Convolve.rs:41-63) of one
// signal of n = 2L reals read as L complex values, left in LDS at
// lds[resp_slot<L / 256>(k)] (then a barrier), for a response whose support
// fits the window [js, js + 256) (resp_win_fits; js a multiple of 64):
//   Z_{k1 + Q k2} = W_256^{js k2} sum_{j < 256} (z_{js+j} W_L^{(js+j) k1}) W_256^{j k2},
// Q = L / 256, W_M = exp(+2 pi i / M) (the forward transform).  The 256
// window values are gathered once into zw (LDS); each k1 < Q is then one
// 256-point DFT on the 16 lanes k1*16 .. k1*16+15 (one wave): lane g takes
// j = 16 j2 + g, j2 < 16, with the twiddles W_L^{(js+g) k1 + 16 k1 j2} built
// from five table values (twiddle_run's scheme); radix 16 over j2, twiddle
// W_256^{g ka}, a 16 x 16 transpose through LDS inside the wave, radix 16
// over j1, and the window factor W_256^{js k2} = i^{(js/64) k2}, which
// depends on k2 mod 4 = ka mod 4 only.  This replaces log2 L stages over an
// input that is zero outside 2 m of its 2 L reals.  Needs one transform per
// group, NT = L / 16 threads, 4096 <= L <= 8192 (NT >= 256 gathers the
// window); the transposes use Q * 256 = L values of the LDS buffer.
template <class E, typename T, int L>
__device__ __forceinline__ void window_response_spectrum(const T* r, int64_t m, int js, cpx<T>* lds, cpx<T>* zw,
                                                         int tid, const Twiddles<T>& tw) {
    constexpr int Q = L / 256;
    constexpr int64_t N = 2 * int64_t(L);
    constexpr uint32_t CS = TW_COARSE / L;  // W_L^e = coarse[e CS]
    static_assert(L % 256 == 0 && L >= 4096 && L <= TW_COARSE && E::LDS_ELEMS >= Q * 256, "window spectrum layout");
    {
        const int lo = resp_win_lo(L, m), hi = resp_win_hi(m);
        const int i = js + tid;
        cpx<T> z{(T)0, (T)0};
        if (tid < 256 && i >= lo && i < hi) {
            const int ii = i < 0 ? i + L : i;
            z = cpx<T>{placed_response(r, 2 * int64_t(ii), N, m), placed_response(r, 2 * int64_t(ii) + 1, N, m)};
        }
        if (tid < 256) zw[tid] = z;
    }
    const int k1 = tid >> 4, g = tid & 15;
    // W_L^{m0 + j2 m1}, j2 < 16: m0 = (js + g) k1, m1 = 16 k1 (mod L).
    const uint32_t mask = uint32_t(L - 1);
    const uint32_t m0 = uint32_t((js + g + L) * k1) & mask, m1 = uint32_t(16 * k1) & mask;
    cpx<T> p[4], w[16];
#pragma unroll
    for (int q = 0; q < 4; ++q) p[q] = tw.coarse[((m1 << q) & mask) * CS];
    w[0] = tw.coarse[m0 * CS];
    lds_barrier();  // zw published
    // Each w[j2] is used as soon as it is formed; only the parents of later
    // entries stay live, which keeps this step's registers low while the
    // caller holds the signal spectrum.
    cpx<T> v[16];
    v[0] = cmul(zw[g], w[0]);
#pragma unroll
    for (int j2 = 1; j2 < 16; ++j2) {
        const int lb = j2 & (-j2);
        w[j2] = cmul(w[j2 - lb], p[ilog2c(lb)]);
        v[j2] = cmul(zw[16 * j2 + g], w[j2]);
    }
    dft16<1>(v);  // v[ka] = sum_j2 a_{16 j2 + g} W_16^{j2 ka}
    // Row g of the 256-point stage-1 table: W_256^{g ka}, ka < 16.
    // The window factor i^{(js/64) k2} of output k2 = ka + 16 kb goes with
    // column ka here: it is the same for the whole group (js is), so it costs
    // a uniform select instead of a per-lane choice after the transpose.
    const cpx<T>* row = tw.stage + 2 * 256 + StageTable<256>::off(1) + g * 16;
    const int jsq = js / 64;
#pragma unroll
    for (int ka = 1; ka < 16; ++ka) {
        const cpx<T> y = cmul(v[ka], row[ka]);
        const int rk = (jsq * ka) & 3;
        v[ka] = rk == 0 ? y : rk == 1 ? cpx<T>{-y.y, y.x} : rk == 2 ? cpx<T>{-y.x, -y.y} : cpx<T>{y.y, -y.x};
    }
    // Transpose [j1 = g][ka] -> [ka][j1] in this group's 16 x 16 region, the
    // column index XOR-ed with the row so that 16 lanes writing one column or
    // reading one row touch 16 distinct banks; the 16 lanes are in one wave,
    // whose LDS operations complete in order.
    cpx<T>* tr = lds + k1 * 256;
#pragma unroll
    for (int ka = 0; ka < 16; ++ka) tr[g * 16 + (ka ^ g)] = v[ka];
    asm volatile("s_waitcnt lgkmcnt(0)" ::: "memory");
#pragma unroll
    for (int j1 = 0; j1 < 16; ++j1) v[j1] = tr[j1 * 16 + (g ^ j1)];
    dft16<1>(v);  // v[kb] = Z[k1 + Q (g + 16 kb)]
    lds_barrier();  // every group has read its transpose region
#pragma unroll
    for (int kb = 0; kb < 16; ++kb) lds[resp_slot<Q>(k1 + Q * (g + 16 * kb))] = v[kb];
    lds_barrier();
}

}  // namespace nrfft

namespace nrfft {

// ---------------------------------------------------------------------------
// realft: TI signals of n = 2L per group, in place or out of place.
//   DIR = +1: n reals -> packed spectrum;  DIR = -1: packed -> (n/2) x.
// ---------------------------------------------------------------------------
template <typename T, int L, int TI, int NT, int DIR>
__global__ __launch_bounds__(NT) void k_realft(const cpx<T>* in, cpx<T>* out, int64_t nsig, Twiddles<T> tw) {
    using E = Engine<T, DIR, L, TI, NT, MAP_J, MAP_J>;
    constexpr int NP = L / 2;  // pairs per signal
    constexpr int LOG2L = ilog2c(L) + 1;
    constexpr int LTM = engine_ltm<E>(E::LDS_ELEMS * E::CSIZE);
    __shared__ cpx<T> lds[E::LDS_ELEMS];
    __shared__ __align__(16) cpx<T> ltw[LTM > 0 ? E::tw_elems(LTM) : 1];
    const int tid = threadIdx.x;
    const int64_t s0 = int64_t(blockIdx.x) * TI;
    const int nvalid = (nsig - s0) < TI ? int(nsig - s0) : TI;
    const cpx<T>* ib = in + s0 * L;
    cpx<T>* ob = out + s0 * L;
    const T c2 = (T)(-0.5 * DIR);
    if constexpr (LTM > 0) E::template stage_tw_to_lds<LTM>(ltw, tw.stage, tid);  // read after the first exchange
    // Untangle twiddle of pair p = (tid + u NT) mod NP: W^(tid mod NP) (two
    // table loads, once) times W^((u NT) mod NP), the same for every thread
    // (an LDS table read after the first barrier).
    constexpr int PPT = (TI * NP + NT - 1) / NT;
    __shared__ cpx<T> uw[PPT];
    if (tid < PPT) uw[tid] = twiddle_big<DIR>(tw, uint32_t((int64_t(tid) * NT) % NP), LOG2L);
    const cpx<T> wt = twiddle_big<DIR>(tw, uint32_t(tid % NP), LOG2L);
    auto wpair = [&](int u) -> cpx<T> { return NP <= NT ? wt : cmul(wt, uw[u]); };
    if constexpr (DIR > 0) {
        E::template run_to_lds<LTM>(tid, lds, tw.stage, [&](int t, int i) -> cpx<T> {
            return (TI == 1 || t < nvalid) ? ib[t * L + i] : cpx<T>{(T)0, (T)0};
        }, false, ltw);
#pragma unroll
        for (int u = 0; u < PPT; ++u) {
            const int q = tid + u * NT;
            if ((TI * NP) % NT != 0 && q >= TI * NP) break;
            const int t = q / NP, p = q % NP;
            if (TI > 1 && t >= nvalid) continue;
            if (p == 0) {
                ob[t * L] = dc_forward(lds[E::lidx(t, 0)]);
                if (L > 1) ob[t * L + NP] = lds[E::lidx(t, NP)];
            } else {
                cpx<T> a, b;
                untangle_pair(lds[E::lidx(t, p)], lds[E::lidx(t, L - p)], wpair(u), c2, a, b);
                ob[t * L + p] = a;
                ob[t * L + L - p] = b;
            }
        }
    } else {
        lds_barrier();  // uw published
#pragma unroll
        for (int u = 0; u < PPT; ++u) {
            const int q = tid + u * NT;
            if ((TI * NP) % NT != 0 && q >= TI * NP) break;
            const int t = q / NP, p = q % NP;
            const bool ok = TI == 1 || t < nvalid;
            if (p == 0) {
                lds[E::lidx(t, 0)] = ok ? dc_inverse(ib[t * L]) : cpx<T>{(T)0, (T)0};
                lds[E::lidx(t, NP)] = ok ? ib[t * L + NP] : cpx<T>{(T)0, (T)0};
            } else {
                cpx<T> a{(T)0, (T)0}, b{(T)0, (T)0};
                if (ok) untangle_pair(ib[t * L + p], ib[t * L + L - p], wpair(u), c2, a, b);
                lds[E::lidx(t, p)] = a;
                lds[E::lidx(t, L - p)] = b;
            }
        }
        lds_barrier();
        E::template run_from_lds<LTM>(tid, lds, tw.stage, [&](int t, int k, cpx<T> v) {
            if (TI == 1 || t < nvalid) ob[t * L + k] = v;
        }, ltw);
    }
}

// ---------------------------------------------------------------------------
// convlv / deconvolution / correl: one launch, TI signals of n = 2L per
// group.  `second` is the per-signal response (m values, placed as
// Convolve.rs:41-63) for OP_CONV / OP_DECONV, or data2 (n values) for
// OP_CORREL.  ans may alias data.
// ---------------------------------------------------------------------------
// Waves per SIMD k_convlv is compiled for.  Capping the FP32 kernels at 4
// (128 VGPRs) lets two 512-thread groups share a CU at L = 8192 (they need
// 153-168 VGPRs otherwise): with the windowed response formed first, 4096 x
// 2^14 takes 0.258 ms on MI355X with the cap (3 VGPRs spilled) against
// 0.346 ms without; also faster at L = 512.  At L = 32..128 the cap costs
// 33-41 spilled VGPRs and 25-35%, at L = 1024..4096 it measured equal, so
// those keep the compiler's choice.
template <typename T, int L>
constexpr int conv_waves() {
    return (sizeof(T) == 4 && (L == 512 || L == 8192)) ? 4 : 1;
}

// Whether a (T, L, TI, NT) k_convlv has a windowed-response form (WIN = 1):
// one transform per group, NT = L / 16 threads, 4096 <= L <= TW_COARSE.
template <typename T, int L, int TI, int NT>
constexpr bool conv_has_window() {
    return TI == 1 && NT * 16 == L && L >= 4096 && L <= TW_COARSE;
}

// WIN = 1: OP_CONV / OP_DECONV with a response whose placed support fits the
// 256-value window (resp_win_fits), the response spectrum from that window;
// WIN = 0: every other case, the response (or data2) through a full FFT.  Two
// instantiations, so the full FFT's registers do not count against the
// windowed form, which holds the signal spectrum across its response step.
template <typename T, int L, int TI, int NT, int WIN>
__global__ __launch_bounds__(NT) __attribute__((amdgpu_waves_per_eu(conv_waves<T, L>()))) void k_convlv(const cpx<T>* data, const T* second, cpx<T>* ans, int64_t nsig,
                                               int64_t m, int op, Twiddles<T> tw) {
    using EF = Engine<T, 1, L, TI, NT, MAP_J, MAP_J>;
    using EI = Engine<T, -1, L, TI, NT, MAP_J, MAP_J>;
    static_assert(EF::LDS_ELEMS == EI::LDS_ELEMS, "same layout both ways");
    constexpr int NP = L / 2;
    constexpr int PPT = (TI * NP + NT - 1) / NT;
    constexpr int LOG2N = ilog2c(L) + 1;
    constexpr int64_t N = 2 * int64_t(L);
    // Twiddles of every stage but the last come from an LDS copy (a few KB):
    // shorter latency than the L2 and fewer registers held while they load.
    constexpr int LTM = EF::NST >= 2 ? EF::NST - 2 : 0;
    __shared__ cpx<T> lds[EF::LDS_ELEMS];
    __shared__ __align__(16) cpx<T> ltw[EF::tw_elems(LTM) > 0 ? EF::tw_elems(LTM) : 1];
    // The response's spectrum from its 256-value window (a short response).
    constexpr bool WINDOW = WIN != 0;
    static_assert(!WINDOW || conv_has_window<T, L, TI, NT>(), "windowed form needs one 8192/4096-point signal per group");
    __shared__ cpx<T> zw[WINDOW ? 256 : 1];
    const int tid = threadIdx.x;
    const int64_t s0 = int64_t(blockIdx.x) * TI;
    const int nvalid = (nsig - s0) < TI ? int(nsig - s0) : TI;
    const cpx<T>* db = data + s0 * L;
    cpx<T>* ob = ans + s0 * L;
    const T inv_no2 = (T)1 / (T)L;
    EF::template stage_tw_to_lds<LTM>(ltw, tw.stage, tid);  // published by the first exchange barrier
    // Untangle twiddles W_N^p of bin pair p = (tid + u NT) mod NP: the
    // per-thread W^(tid mod NP) (two table loads, once) times the factor
    // W^((u NT) mod NP) shared by every thread (an LDS table of PPT entries),
    // instead of two table loads per pair after a barrier.
    __shared__ cpx<T> uw[PPT];
    if (tid < PPT) uw[tid] = twiddle_big<1>(tw, uint32_t((int64_t(tid) * NT) % NP), LOG2N);
    const cpx<T> wt = twiddle_big<1>(tw, uint32_t(tid % NP), LOG2N);
    auto wpair = [&](int u) -> cpx<T> { return NP <= NT ? wt : cmul(wt, uw[u]); };

    // The two forward transforms share the tile: the first one's spectrum is
    // untangled pair by pair into registers and held while the second runs.
    // With one signal per group the second operand (the response, or data2)
    // goes first: its FFT or window step needs more registers than the
    // signal's plain loads and holds nothing meanwhile (FP32 L = 8192: 8
    // bytes of spill per lane instead of 88).  With several signals per group
    // the placed response's loader for TI transforms weighs more, and the
    // signal goes first.
    constexpr bool SECOND_FIRST = TI == 1;
    auto second_spectrum = [&]() {
        if constexpr (WINDOW) {
            // A short response: its spectrum from the window it occupies
            // (the launch checked op != OP_CORREL and resp_win_fits).
            window_response_spectrum<EF, T, L>(second + s0 * m, m, resp_win_start(L, m), lds, zw, tid, tw);
        } else if (op == OP_CORREL) {
            const cpx<T>* sb = reinterpret_cast<const cpx<T>*>(second) + s0 * L;
            EF::template run_to_lds<LTM>(tid, lds, tw.stage, [&](int t, int i) -> cpx<T> {
                return (TI == 1 || t < nvalid) ? sb[t * L + i] : cpx<T>{(T)0, (T)0};
            }, false, ltw);
        } else {
            const T* rsp = second + s0 * m;
            EF::template run_to_lds<LTM>(tid, lds, tw.stage, [&](int t, int i) -> cpx<T> {
                if (TI > 1 && t >= nvalid) return cpx<T>{(T)0, (T)0};
                const T* r = rsp + t * m;
                return cpx<T>{placed_response(r, 2 * int64_t(i), N, m), placed_response(r, 2 * int64_t(i) + 1, N, m)};
            }, false, ltw);
        }
    };
    auto signal_load = [&](int t, int i) -> cpx<T> {
        return (TI == 1 || t < nvalid) ? db[t * L + i] : cpx<T>{(T)0, (T)0};
    };
    auto signal_spectrum = [&]() {
        EF::template run_to_lds<LTM>(tid, lds, tw.stage, [&](int t, int i) -> cpx<T> {
            return (TI == 1 || t < nvalid) ? db[t * L + i] : cpx<T>{(T)0, (T)0};
        }, false, ltw);
    };

    // 1. The first spectrum, held untangled (bins p and L - p of pair p).
    // When the signal comes second its stage-0 loads are issued here, so
    // their HBM latency overlaps the untangle of the held spectrum.
    if constexpr (SECOND_FIRST) second_spectrum(); else signal_spectrum();
    cpx<T> sv[SECOND_FIRST ? EF::EPT : 1];
    if constexpr (SECOND_FIRST) EF::load0(sv, tid, signal_load);
    auto hidx = [](int t, int i) -> int {
        if constexpr (WINDOW) return resp_slot<L / 256>(i);
        else return EF::lidx(t, i);
    };
    cpx<T> ha[PPT], hb[PPT];
#pragma unroll
    for (int u = 0; u < PPT; ++u) {
        const int q = tid + u * NT;
        if ((TI * NP) % NT != 0 && q >= TI * NP) continue;
        const int t = TI == 1 ? 0 : q / NP, p = TI == 1 ? q : q % NP;
        if (p == 0) {
            ha[u] = dc_forward(lds[hidx(t, 0)]);
            hb[u] = lds[hidx(t, NP)];
        } else {
            untangle_pair(lds[hidx(t, p)], lds[hidx(t, L - p)], wpair(u), (T)-0.5, ha[u], hb[u]);
        }
    }
    lds_barrier();  // every value is read before the next FFT overwrites the tile

    // 2. The other spectrum into the tile.
    if constexpr (SECOND_FIRST) EF::template run_to_lds_loaded<LTM>(sv, tid, lds, tw.stage, false, ltw);
    else second_spectrum();

    // 3. Each pair (p, L - p): untangle, spectrum operation (Convolve.rs:90-138,
    // Correlation.rs:68-110) with the signal's bins first, inverse untangle,
    // back in place.  A pair is read and written by one thread only.
#pragma unroll
    for (int u = 0; u < PPT; ++u) {
        const int q = tid + u * NT;
        if ((TI * NP) % NT != 0 && q >= TI * NP) continue;
        const int t = TI == 1 ? 0 : q / NP, p = TI == 1 ? q : q % NP;
        if (p == 0) {
            const cpx<T> t0 = dc_forward(lds[EF::lidx(t, 0)]);
            const cpx<T> th = lds[EF::lidx(t, NP)];
            const cpx<T> x0 = SECOND_FIRST ? t0 : ha[u], r0 = SECOND_FIRST ? ha[u] : t0;
            const cpx<T> xh = SECOND_FIRST ? th : hb[u], rh = SECOND_FIRST ? hb[u] : th;
            const cpx<T> y0{spec_apply_real(op, x0.x, r0.x, inv_no2), spec_apply_real(op, x0.y, r0.y, inv_no2)};
            lds[EF::lidx(t, 0)] = dc_inverse(y0);
            // Bin L/2 pairs with itself: untangle leaves it unchanged both ways.
            lds[EF::lidx(t, NP)] = spec_apply(op, xh, rh, inv_no2);
        } else {
            cpx<T> ta, tb, oa, ob2;
            const cpx<T> w = wpair(u);
            untangle_pair(lds[EF::lidx(t, p)], lds[EF::lidx(t, L - p)], w, (T)-0.5, ta, tb);
            const cpx<T> xa = SECOND_FIRST ? ta : ha[u], ra = SECOND_FIRST ? ha[u] : ta;
            const cpx<T> xb = SECOND_FIRST ? tb : hb[u], rb = SECOND_FIRST ? hb[u] : tb;
            // Inverse untangle: conj(w) = exp(-i pi k / L).
            untangle_pair(spec_apply(op, xa, ra, inv_no2), spec_apply(op, xb, rb, inv_no2), conj(w), (T)0.5, oa, ob2);
            lds[EF::lidx(t, p)] = oa;
            lds[EF::lidx(t, L - p)] = ob2;
        }
    }
    lds_barrier();

    // 4. Inverse FFT: (n/2) * (result / no2) = result.
    EI::template run_from_lds<LTM>(tid, lds, tw.stage, [&](int t, int k, cpx<T> v) {
        if (TI == 1 || t < nvalid) ob[t * L + k] = v;
    }, ltw);
}

}  // namespace nrfft
