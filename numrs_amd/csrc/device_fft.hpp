// device_fft.hpp -- register/LDS building blocks of the Stockham autosort FFT
// for gfx950 (CDNA4, 64-wide wavefronts).
//
// A work-group owns a tile of TI independent transforms of length L (a power
// of two).  Each of its NT threads holds EPT = TI*L/NT complex values in
// registers.  Stage s applies radix R_s butterflies with the Stockham
// indexing (Ns = product of the earlier radices, butterfly j in [0, L/R)):
//   read   X[j + r*L/R]            r = 0..R-1
//   scale  by W_{Ns*R}^{S*r*(j mod Ns)}
//   DFT_R
//   write  Y[(j/Ns)*Ns*R + (j mod Ns) + r*Ns]
// so the first stage reads, and the last stage writes, unit-stride runs: the
// global load and store of a pass are coalesced without a separate
// transpose.  Stages exchange data through LDS; the thread->butterfly mapping
// may change between stages (MAP_J: consecutive threads take consecutive j of
// one transform, MAP_T: consecutive threads take the same j of consecutive
// transforms), which is how a pass reads rows and writes columns, or the
// reverse, with both sides coalesced.
//
// Sign convention (numrs four1, FFT_1.rs:5-44): S = isign, kernel
// exp(+S * 2 pi i jk / L), no normalisation.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace nrfft {

// Work-group barrier that orders LDS only: wait for this wave's LDS
// operations, then s_barrier; the asm statement's memory clobber keeps the
// compiler from moving memory accesses across it.  __syncthreads() is a
// release / acquire fence on every address space, which on gfx950 waits for
// all of the wave's outstanding global loads and stores (vmcnt(0)) first; a
// fence on the local address space alone still waits for outstanding LDS-DMA
// loads.  This form lets global stores drain and prefetches (register or
// LDS-DMA) stay in flight across the barrier.
// Valid wherever a barrier guards LDS exchanges and no thread of the group
// reads global memory that another thread of the group writes in the same
// kernel (every kernel here: groups own disjoint output elements, and a
// group's loads of an element are consumed before the barrier that precedes
// any store to it).  LDS-DMA data is waited for explicitly (vmcnt) before
// the barrier that publishes it.
__device__ __forceinline__ void lds_barrier() {
    asm volatile("s_waitcnt lgkmcnt(0)\n\ts_barrier" ::: "memory");
}

template <typename T>
struct alignas(2 * sizeof(T)) cpx {
    T x, y;
};
using c32 = cpx<float>;
using c64 = cpx<double>;

template <typename T>
__device__ __forceinline__ cpx<T> operator+(cpx<T> a, cpx<T> b) { return {a.x + b.x, a.y + b.y}; }
template <typename T>
__device__ __forceinline__ cpx<T> operator-(cpx<T> a, cpx<T> b) { return {a.x - b.x, a.y - b.y}; }
template <typename T>
__device__ __forceinline__ cpx<T> cmul(cpx<T> a, cpx<T> b) {
    return {a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x};
}
// a * conj(b)
template <typename T>
__device__ __forceinline__ cpx<T> cmulc(cpx<T> a, cpx<T> b) {
    return {a.x * b.x + a.y * b.y, a.y * b.x - a.x * b.y};
}
template <typename T>
__device__ __forceinline__ cpx<T> cscale(cpx<T> a, T s) { return {a.x * s, a.y * s}; }
template <typename T>
__device__ __forceinline__ cpx<T> conj(cpx<T> a) { return {a.x, -a.y}; }

// a * (S i)
template <int S, typename T>
__device__ __forceinline__ cpx<T> mul_si(cpx<T> a) {
    if constexpr (S > 0) return {-a.y, a.x};
    else return {a.y, -a.x};
}
// Twiddle tables hold exp(+2 pi i m / M); direction S uses the conjugate for S < 0.
template <int S, typename T>
__device__ __forceinline__ cpx<T> dirw(cpx<T> w) {
    if constexpr (S > 0) return w;
    else return {w.x, -w.y};
}

// ---------------------------------------------------------------------------
// Small DFTs in registers, natural order in and out:
//   v[k] <- sum_j v[j] exp(S 2 pi i jk / R)
// ---------------------------------------------------------------------------
template <int S, typename T>
__device__ __forceinline__ void dft2(cpx<T>* v) {
    cpx<T> a = v[0], b = v[1];
    v[0] = a + b;
    v[1] = a - b;
}

template <int S, typename T>
__device__ __forceinline__ void dft4(cpx<T>& x0, cpx<T>& x1, cpx<T>& x2, cpx<T>& x3) {
    cpx<T> a0 = x0 + x2, a1 = x0 - x2, a2 = x1 + x3, a3 = mul_si<S>(x1 - x3);
    x0 = a0 + a2;
    x2 = a0 - a2;
    x1 = a1 + a3;
    x3 = a1 - a3;
}

template <int S, typename T>
__device__ __forceinline__ void dft4(cpx<T>* v) { dft4<S>(v[0], v[1], v[2], v[3]); }

// exp(S i pi m / 8) * a for the m used by the 8- and 16-point kernels.
template <int S, int M, typename T>
__device__ __forceinline__ cpx<T> rot16(cpx<T> a) {
    constexpr double C1 = 0.92387953251128675612818318939678829;  // cos(pi/8)
    constexpr double S1 = 0.38268343236508977172845998403039887;  // sin(pi/8)
    constexpr double H = 0.70710678118654752440084436210484904;   // cos(pi/4)
    constexpr double sg = (double)S;
    if constexpr (M == 0) return a;
    else if constexpr (M == 4) return mul_si<S>(a);
    else if constexpr (M == 8) return {-a.x, -a.y};
    else if constexpr (M == 2) {
        const T h = (T)H;
        return {h * (a.x - (T)sg * a.y), h * ((T)sg * a.x + a.y)};
    } else if constexpr (M == 6) {
        const T h = (T)H;
        return {h * (-a.x - (T)sg * a.y), h * ((T)sg * a.x - a.y)};
    } else {
        constexpr double c = (M == 1) ? C1 : (M == 3) ? S1 : (M == 5) ? -S1 : (M == 7) ? -C1 : (M == 9) ? -C1 : 0.0;
        constexpr double s = (M == 1) ? S1 : (M == 3) ? C1 : (M == 5) ? C1 : (M == 7) ? S1 : (M == 9) ? -S1 : 0.0;
        const T cc = (T)c, ss = (T)(sg * s);
        return {a.x * cc - a.y * ss, a.x * ss + a.y * cc};
    }
}

template <int S, typename T>
__device__ __forceinline__ void dft8(cpx<T>* v) {
    cpx<T> e0 = v[0], e1 = v[2], e2 = v[4], e3 = v[6];
    cpx<T> o0 = v[1], o1 = v[3], o2 = v[5], o3 = v[7];
    dft4<S>(e0, e1, e2, e3);
    dft4<S>(o0, o1, o2, o3);
    o1 = rot16<S, 2>(o1);
    o2 = rot16<S, 4>(o2);
    o3 = rot16<S, 6>(o3);
    v[0] = e0 + o0;
    v[4] = e0 - o0;
    v[1] = e1 + o1;
    v[5] = e1 - o1;
    v[2] = e2 + o2;
    v[6] = e2 - o2;
    v[3] = e3 + o3;
    v[7] = e3 - o3;
}

// 16 = 4 x 4: DFT4 down the columns j1 of x[4 j1 + j2], twiddle by
// w16^(j2 k1), DFT4 along j2, output X[k1 + 4 k2].
template <int S, typename T>
__device__ __forceinline__ void dft16(cpx<T>* v) {
#pragma unroll
    for (int j2 = 0; j2 < 4; ++j2) dft4<S>(v[j2], v[4 + j2], v[8 + j2], v[12 + j2]);
    // v[j2 + 4 k1] = Y[j2][k1]
    v[5] = rot16<S, 1>(v[5]);
    v[6] = rot16<S, 2>(v[6]);
    v[7] = rot16<S, 3>(v[7]);
    v[9] = rot16<S, 2>(v[9]);
    v[10] = rot16<S, 4>(v[10]);
    v[11] = rot16<S, 6>(v[11]);
    v[13] = rot16<S, 3>(v[13]);
    v[14] = rot16<S, 6>(v[14]);
    v[15] = rot16<S, 9>(v[15]);
#pragma unroll
    for (int k1 = 0; k1 < 4; ++k1) dft4<S>(v[4 * k1], v[4 * k1 + 1], v[4 * k1 + 2], v[4 * k1 + 3]);
    // v[4 k1 + k2] = X[k1 + 4 k2]; transpose the 4x4 index to natural order.
    cpx<T> t[16];
#pragma unroll
    for (int k1 = 0; k1 < 4; ++k1)
#pragma unroll
        for (int k2 = 0; k2 < 4; ++k2) t[k1 + 4 * k2] = v[4 * k1 + k2];
#pragma unroll
    for (int i = 0; i < 16; ++i) v[i] = t[i];
}

template <int R, int S, typename T>
__device__ __forceinline__ void dft(cpx<T>* v) {
    if constexpr (R == 1) {
    } else if constexpr (R == 2) dft2<S>(v);
    else if constexpr (R == 4) dft4<S>(v);
    else if constexpr (R == 8) dft8<S>(v);
    else if constexpr (R == 16) dft16<S>(v);
    else static_assert(R == 2, "radix must be 2, 4, 8 or 16");
}

// ---------------------------------------------------------------------------
// Radix plan: log2(L) bits covered by radix-16 stages plus one radix-2/4/8
// stage for the remainder, taken first (NRFFT_SMALL_RADIX_FIRST = 1, 10-15%
// faster at L = 512..2048 on MI355X) or last.  Padded LDS addresses are affine
// in the butterfly slot r whenever the stage's stride is a multiple of 16,
// and are computed per element otherwise.  L < 16 is a single stage.
__host__ __device__ constexpr int ilog2c(long long v) { return v <= 1 ? 0 : 1 + ilog2c(v >> 1); }

// Build switches: order of the small radix, and the LDS index scheme (padding
// with affine offsets, or an XOR swizzle -- equal speed on MI355X).
#ifndef NRFFT_SMALL_RADIX_FIRST
#define NRFFT_SMALL_RADIX_FIRST 1
#endif
#ifndef NRFFT_LDS_SWIZZLE
#define NRFFT_LDS_SWIZZLE 0
#endif
// Cross-lane (DPP) exchange for 256-point row-mapped passes.
#ifndef NRFFT_DPP_EXCHANGE
#define NRFFT_DPP_EXCHANGE 1
#endif
// One pad slot per NRFFT_LDS_PAD complex values of a padded LDS tile.
#ifndef NRFFT_LDS_PAD
#define NRFFT_LDS_PAD 32
#endif

__host__ __device__ constexpr int plan_radix(int L, int LOG, int NST, int s) {
    return (L < 16) ? L
                    : (NRFFT_SMALL_RADIX_FIRST ? (s == 0 ? ((LOG % 4) == 0 ? 16 : (1 << (LOG % 4))) : 16)
                                               : (s < NST - 1 ? 16 : ((LOG % 4) == 0 ? 16 : (1 << (LOG % 4)))));
}

template <int L>
struct RadixPlan {
    static constexpr int LOG = ilog2c(L);
    static constexpr int NST = (L < 16) ? 1 : (LOG + 3) / 4;
    static constexpr int radix(int s) { return plan_radix(L, LOG, NST, s); }
    static constexpr int ns(int s) {
        int p = 1;
        for (int i = 0; i < s; ++i) p *= radix(i);
        return p;
    }
    static constexpr int EPT_MIN = (L < 16) ? L : 16;
};

enum { MAP_J = 0, MAP_T = 1 };

// ---------------------------------------------------------------------------
// Cross-lane exchange: a 16 x 16 transpose across the 16 lanes of each DPP
// row (lanes 16k .. 16k+15 of a wave), in registers -- lane l's v[r] and lane
// r's v[l] trade places.  Four steps, one per bit m = 8, 4, 2, 1 of the lane
// and register indices: an element moves to (lane ^ m, register ^ m) when
// the two indices differ in bit m (after all four every element has swapped
// its lane and register index).  m = 8 and 4 use DPP row_ror:8 and row_shl /
// row_shr:4 with bank masks (the receiving lanes are whole 4-lane banks, so
// each step is one DPP move per float and register); m = 2 and 1 use
// quad_perm plus a per-lane select.
// ---------------------------------------------------------------------------
// DPP moves of one 32-bit lane value (FP64 values move as two halves).
template <int CTRL, int BANKS>
__device__ __forceinline__ float dpp_update(float old, float src) {
    return __int_as_float(__builtin_amdgcn_update_dpp(__float_as_int(old), __float_as_int(src), CTRL, 0xF, BANKS, false));
}
template <int CTRL, int BANKS>
__device__ __forceinline__ double dpp_update(double old, double src) {
    const long long o = __double_as_longlong(old), v = __double_as_longlong(src);
    const int lo = __builtin_amdgcn_update_dpp(int(o), int(v), CTRL, 0xF, BANKS, false);
    const int hi = __builtin_amdgcn_update_dpp(int(o >> 32), int(v >> 32), CTRL, 0xF, BANKS, false);
    return __longlong_as_double((long long)(unsigned)lo | ((long long)hi << 32));
}
template <int CTRL, typename T>
__device__ __forceinline__ T dpp_move(T src) {
    return dpp_update<CTRL, 0xF>(src, src);
}
constexpr int DPP_ROW_SHL4 = 0x104, DPP_ROW_SHR4 = 0x114, DPP_ROW_ROR8 = 0x128;
constexpr int DPP_QUAD_XOR1 = 1 | (0 << 2) | (3 << 4) | (2 << 6);  // quad_perm [1, 0, 3, 2]
constexpr int DPP_QUAD_XOR2 = 2 | (3 << 2) | (0 << 4) | (1 << 6);  // quad_perm [2, 3, 0, 1]

template <int M, typename T>
__device__ __forceinline__ void row_transpose_step(cpx<T>* v, int lane) {
#pragma unroll
    for (int r = 0; r < 16; ++r) {
        if (r & M) continue;
        const cpx<T> a = v[r], b = v[r | M];
        if constexpr (M == 8) {
            // lanes 8-15 (banks 2, 3) take the partner's v[r | 8] into v[r];
            // lanes 0-7 (banks 0, 1) its v[r] into v[r | 8]
            v[r] = {dpp_update<DPP_ROW_ROR8, 0xC>(a.x, b.x), dpp_update<DPP_ROW_ROR8, 0xC>(a.y, b.y)};
            v[r | M] = {dpp_update<DPP_ROW_ROR8, 0x3>(b.x, a.x), dpp_update<DPP_ROW_ROR8, 0x3>(b.y, a.y)};
        } else if constexpr (M == 4) {
            // banks 1, 3 (lane bit 2 set) read lane - 4, banks 0, 2 lane + 4
            v[r] = {dpp_update<DPP_ROW_SHR4, 0xA>(a.x, b.x), dpp_update<DPP_ROW_SHR4, 0xA>(a.y, b.y)};
            v[r | M] = {dpp_update<DPP_ROW_SHL4, 0x5>(b.x, a.x), dpp_update<DPP_ROW_SHL4, 0x5>(b.y, a.y)};
        } else {
            constexpr int Q = M == 2 ? DPP_QUAD_XOR2 : DPP_QUAD_XOR1;
            const bool hi = (lane & M) != 0;
            const cpx<T> pb{dpp_move<Q>(b.x), dpp_move<Q>(b.y)};  // partner's v[r | M]
            const cpx<T> pa{dpp_move<Q>(a.x), dpp_move<Q>(a.y)};  // partner's v[r]
            v[r] = hi ? pb : a;
            v[r | M] = hi ? b : pa;
        }
    }
}
template <typename T>
__device__ __forceinline__ void row_transpose16(cpx<T>* v) {
    const int lane = int(threadIdx.x) & 15;
    row_transpose_step<8>(v, lane);
    row_transpose_step<4>(v, lane);
    row_transpose_step<2>(v, lane);
    row_transpose_step<1>(v, lane);
}

// Twiddle tables (device memory, built in FP64 on the GPU):
//   coarse[m] = exp(2 pi i m / 2^13),       m < 2^13
//   fine[m]   = exp(2 pi i m / 2^26),       m < 2^13
// exp(2 pi i M / 2^26) = coarse[M >> 13] * fine[M & 8191] for M < 2^26.
//   stage: per-stage rows for every single-group length L = 2^1..2^14; the
//   stage-s table of L starts at 2L + off_s and holds NS_s rows of R_s entries,
//   row k = (exp(2 pi i r k / (NS_s R_s)), r = 0..R_s-1), so one thread's
//   twiddles are contiguous (wide loads) and consecutive k are adjacent.
constexpr int TW_LOG_COARSE = 13;
constexpr int TW_COARSE = 1 << TW_LOG_COARSE;
constexpr int TW_LOG_VIRT = 26;
constexpr int TW_STAGE_MAX_LOGL = 14;
constexpr int TW_STAGE_ELEMS = 4 << TW_STAGE_MAX_LOGL;

template <typename T>
struct Twiddles {
    const cpx<T>* coarse;
    const cpx<T>* fine;
    const cpx<T>* stage;
};

// exp(S 2 pi i m / 2^logn), 0 <= m < 2^logn <= 2^26.
template <int S, typename T>
__device__ __forceinline__ cpx<T> twiddle_big(const Twiddles<T>& tw, uint32_t m, int logn) {
    uint32_t M = m << (TW_LOG_VIRT - logn);
    cpx<T> w = cmul(tw.coarse[M >> TW_LOG_COARSE], tw.fine[M & (TW_COARSE - 1)]);
    return dirw<S>(w);
}

// The two table values whose product twiddle_big forms (loaded now,
// multiplied later with twiddle_join: same arithmetic, same result).
template <typename T>
__device__ __forceinline__ void twiddle_parts(const Twiddles<T>& tw, uint32_t m, int logn, cpx<T>& a, cpx<T>& b) {
    uint32_t M = m << (TW_LOG_VIRT - logn);
    a = tw.coarse[M >> TW_LOG_COARSE];
    b = tw.fine[M & (TW_COARSE - 1)];
}
template <int S, typename T>
__device__ __forceinline__ cpx<T> twiddle_join(cpx<T> a, cpx<T> b) {
    return dirw<S>(cmul(a, b));
}

// out[r] = W^{S (m0 + r m1)} for r < R (W = exp(2 pi i / 2^logn)), from table
// values of W^{m0} and W^{2^q m1}; out[r] = out[r - lowbit(r)] * W^{lowbit(r) m1},
// so every value is at most popcount(r) + 1 products away from the tables.
template <int R, int S, typename T>
__device__ __forceinline__ void twiddle_run(const Twiddles<T>& tw, uint32_t m0, uint32_t m1, int logn, cpx<T>* out) {
    constexpr int LR = ilog2c(R);
    cpx<T> p[LR > 0 ? LR : 1];
#pragma unroll
    for (int q = 0; q < LR; ++q) p[q] = twiddle_big<S>(tw, (m1 << q) & ((1u << logn) - 1u), logn);
    out[0] = twiddle_big<S>(tw, m0, logn);
#pragma unroll
    for (int r = 1; r < R; ++r) {
        const int lb = r & (-r);
        out[r] = cmul(out[r - lb], p[ilog2c(lb)]);
    }
}

// ---------------------------------------------------------------------------
// The stage engine.
// ---------------------------------------------------------------------------
// Stage limit meaning "every stage" for the LDS twiddle copy (Engine::compute).
constexpr int LT_ALL = 64;

template <int L>
struct StageTable {
    // offset of stage s inside L's block (which starts at 2L)
    static constexpr int off(int s) {
        int o = 0;
        for (int q = 0; q < s; ++q) o += RadixPlan<L>::ns(q + 1);
        return o;
    }
};

template <typename T, int S, int L, int TI, int NT, int MAPL, int MAPS>
struct Engine {
    using P = RadixPlan<L>;
    static constexpr int EPT = TI * L / NT;
    static constexpr int THREADS = NT;
    static constexpr int CSIZE = (int)sizeof(cpx<T>);
    static_assert(L <= (1 << TW_STAGE_MAX_LOGL), "per-stage twiddle tables cover L <= 2^14 only");
    static_assert(TI * L % NT == 0, "tile must split evenly over threads");
    static constexpr int NST = P::NST;
    static_assert(EPT % P::EPT_MIN == 0 || L < 16, "EPT must cover the largest radix");
    // A 256-point transform on 16 lanes with the row mapping on both
    // sides: its one exchange is a 16 x 16 transpose inside a DPP row
    // (row_transpose16), no LDS (NRFFT_DPP_EXCHANGE).
    static constexpr bool DPP_X = NRFFT_DPP_EXCHANGE && L == 256 && NST == 2 &&
                                  P::radix(0) == 16 && P::radix(1) == 16 && EPT == 16 && MAPL == MAP_J &&
                                  MAPS == MAP_J;
    // LDS layout follows the mapping used after the first stage.
    static constexpr int LAYOUT = (NST == 1) ? MAPL : MAPS;
    static constexpr int map_of(int s) { return s == 0 ? MAPL : MAPS; }

    // Padded LDS index of element i of tile transform t: one pad slot per PAD
    // complex values (NRFFT_LDS_PAD, 32 by default), so strides that are
    // multiples of PAD land in distinct banks (MAP_T layout [i][TI]: lanes
    // across t are contiguous, lanes along i are (PAD + 1) TI / PAD apart;
    // MAP_J layout [t][L]: lanes along i are contiguous, stride-PAD runs are
    // PAD + 1 apart).  Measured on MI355X (64 x 2^20 FP32): PAD 32 0.441 ms,
    // 16 0.452 ms, 8 0.491 ms (the 2048-point pass B 53.8 / 56.4 / 63.1 us);
    // convlv 4096 x 2^14 0.306 / 0.318 / 0.318 ms.
    // With NRFFT_LDS_SWIZZLE the MAP_T layout is unpadded and XOR-swizzles the
    // column with a hash of the row instead (no affine offsets).
    static constexpr int LSI = (LAYOUT == MAP_T) ? TI : 1;  // raw distance between i and i + 1
    static constexpr bool SWZ = NRFFT_LDS_SWIZZLE && LAYOUT == MAP_T && TI > 1;
    static constexpr int PAD = NRFFT_LDS_PAD;
    static constexpr int LDS_ELEMS = SWZ ? TI * L : TI * L + (TI * L) / PAD + 1;
    static __device__ __forceinline__ int lidx(int t, int i) {
        if constexpr (SWZ) {
            const unsigned h = unsigned(i) + (unsigned(i) >> 4) + (unsigned(i) >> 8);
            return i * TI + (t ^ int(h & unsigned(TI - 1)));
        } else {
            const int raw = (LAYOUT == MAP_T) ? i * TI + t : t * L + i;
            return raw + raw / PAD;
        }
    }
    // lidx(t, i0 + r * STRIDE) = lidx(t, i0) + loff<STRIDE>(r) whenever the raw
    // step STRIDE * LSI is a multiple of 16 (then the pad term is affine too).
    template <int STRIDE>
    static constexpr bool affine = !SWZ && ((STRIDE * LSI) % PAD) == 0;
    template <int STRIDE>
    static constexpr int loff(int r) {
        return r * (STRIDE * LSI + (STRIDE * LSI) / PAD);
    }

    template <int MAP, int NB>
    static __device__ __forceinline__ void bmap(int q, int& t, int& j) {
        const unsigned uq = unsigned(q);
        if constexpr (MAP == MAP_J) {
            t = int(uq / unsigned(NB));
            j = int(uq % unsigned(NB));
        } else {
            t = int(uq % unsigned(TI));
            j = int(uq / unsigned(TI));
        }
    }

    // Stage-0 inputs straight from global memory through a load functor
    // ld(t, i) -> cpx<T> (element i of tile transform t).
    template <class LD>
    static __device__ __forceinline__ void load0(cpx<T>* v, int tid, LD&& ld) {
        constexpr int R = P::radix(0), NB = L / R, BPT = EPT / R;
#pragma unroll
        for (int b = 0; b < BPT; ++b) {
            int t, j;
            bmap<MAPL, NB>(tid + b * NT, t, j);
#pragma unroll
            for (int r = 0; r < R; ++r) v[b * R + r] = ld(t, j + r * NB);
        }
    }

    // Stage-s inputs from an LDS buffer holding the previous stage's output
    // (or any natural-order data when s == 0).
    template <int s>
    static __device__ __forceinline__ void load_lds(cpx<T>* v, int tid, const cpx<T>* lds) {
        constexpr int R = P::radix(s), NB = L / R, BPT = EPT / R;
#pragma unroll
        for (int b = 0; b < BPT; ++b) {
            int t, j;
            bmap<map_of(s), NB>(tid + b * NT, t, j);
            if constexpr (affine<NB>) {
                const cpx<T>* base = lds + lidx(t, j);
#pragma unroll
                for (int r = 0; r < R; ++r) v[b * R + r] = base[loff<NB>(r)];
            } else {
#pragma unroll
                for (int r = 0; r < R; ++r) v[b * R + r] = lds[lidx(t, j + r * NB)];
            }
        }
    }

    // LDS copy of the stage twiddles of stages 1..NST-1 (stage 0 has none):
    // the rows of stage s start at twoff(s), R_s + 2 entries apart, so the
    // 16-byte row reads of up to 8 consecutive rows hit distinct banks.
    static constexpr int twrow(int s) { return P::radix(s) + 2; }
    static constexpr int twoff(int s) {
        int o = 0;
        for (int q = 1; q < s; ++q) o += P::ns(q) * twrow(q);
        return o;
    }
    static constexpr int TW_ELEMS = twoff(NST);
    // Entries of the LDS copy that holds stages 1 .. min(LTM, NST-1) only.
    static constexpr int tw_elems(int ltm) { return twoff((ltm < NST - 1 ? ltm : NST - 1) + 1); }
    template <int LTM = LT_ALL, int s = 1>
    static __device__ __forceinline__ void stage_tw_to_lds(cpx<T>* ltw, const cpx<T>* stage_tw, int tid) {
        if constexpr (s < NST && s <= LTM) {
            constexpr int R = P::radix(s), NS = P::ns(s);
            const cpx<T>* src = stage_tw + 2 * L + StageTable<L>::off(s);
            for (int e = tid; e < NS * R; e += NT) ltw[twoff(s) + (e / R) * twrow(s) + (e % R)] = src[e];
            stage_tw_to_lds<LTM, s + 1>(ltw, stage_tw, tid);
        }
    }

    // Stages 1 .. LTM read their twiddles from the LDS copy ltw, the others
    // from the global stage table.
    template <int s, int LTM = 0>
    static __device__ __forceinline__ void compute(cpx<T>* v, int tid, const cpx<T>* __restrict__ stage_tw,
                                                   const cpx<T>* __restrict__ ltw = nullptr) {
        constexpr int R = P::radix(s), NB = L / R, BPT = EPT / R, NS = P::ns(s);
        constexpr bool LT = s >= 1 && s <= LTM;
#pragma unroll
        for (int b = 0; b < BPT; ++b) {
            if constexpr (NS > 1) {
                int t, j;
                bmap<map_of(s), NB>(tid + b * NT, t, j);
                // Row k = j mod NS of this stage's table: W_{NS R}^{r k}, r < R.
                const cpx<T>* row = LT ? ltw + twoff(s) + (j % NS) * twrow(s)
                                       : stage_tw + 2 * L + StageTable<L>::off(s) + (j % NS) * R;
#pragma unroll
                for (int r = 1; r < R; ++r) v[b * R + r] = cmul(v[b * R + r], dirw<S>(row[r]));
            }
            dft<R, S>(v + b * R);
        }
    }

    // Stage-s outputs to LDS at their Stockham positions.
    template <int s>
    static __device__ __forceinline__ void store_lds(const cpx<T>* v, int tid, cpx<T>* lds) {
        constexpr int R = P::radix(s), NB = L / R, BPT = EPT / R, NS = P::ns(s);
#pragma unroll
        for (int b = 0; b < BPT; ++b) {
            int t, j;
            bmap<map_of(s), NB>(tid + b * NT, t, j);
            const int i0 = (j / NS) * NS * R + (j % NS);
            if constexpr (affine<NS>) {
                cpx<T>* base = lds + lidx(t, i0);
#pragma unroll
                for (int r = 0; r < R; ++r) base[loff<NS>(r)] = v[b * R + r];
            } else {
#pragma unroll
                for (int r = 0; r < R; ++r) lds[lidx(t, i0 + r * NS)] = v[b * R + r];
            }
        }
    }

    // Last-stage outputs (natural order k = j + r L/R) through a store
    // functor st(t, j, r, value) -- the functor sees the butterfly (t, j) and
    // the output index k = j + r * (L / R_last).
    static constexpr int RLAST = P::radix(NST - 1);
    static constexpr int NBLAST = L / RLAST;
    template <class ST>
    static __device__ __forceinline__ void store_last(const cpx<T>* v, int tid, ST&& st) {
        constexpr int s = NST - 1;
        constexpr int R = RLAST, NB = NBLAST, BPT = EPT / R;
#pragma unroll
        for (int b = 0; b < BPT; ++b) {
            int t, j;
            bmap<map_of(s), NB>(tid + b * NT, t, j);
            st(t, j, v + b * R);
        }
    }

    // Stages 1..NST-1 (stage 0 has been computed on v).  The leading barrier
    // of each exchange protects the LDS buffer against the previous stage's
    // readers (WAR); the trailing one publishes the writes (RAW).
    // DPP_OK: the caller lets the exchange run across lanes where DPP_X
    // allows it (then this call touches no LDS and has no barrier, so the
    // caller must not rely on it to publish other LDS data).
    template <int s = 1, int LTM = 0, bool DPP_OK = false>
    static __device__ __forceinline__ void middle(cpx<T>* v, int tid, cpx<T>* lds, const cpx<T>* stage_tw,
                                                  bool lead_barrier, const cpx<T>* ltw = nullptr) {
        if constexpr (s < NST) {
            if constexpr (DPP_OK && DPP_X && s == 1) {
                row_transpose16(v);
            } else {
                if (lead_barrier) lds_barrier();
                store_lds<s - 1>(v, tid, lds);
                lds_barrier();
                load_lds<s>(v, tid, lds);
            }
            compute<s, LTM>(v, tid, stage_tw, ltw);
            middle<s + 1, LTM, DPP_OK>(v, tid, lds, stage_tw, true, ltw);
        }
    }

    // Store functor adaptor for element-wise stores st(t, k, value).
    template <class ST>
    struct PerElem {
        ST st;
        __device__ __forceinline__ void operator()(int t, int j, const cpx<T>* w) const {
#pragma unroll
            for (int r = 0; r < RLAST; ++r) st(t, j + r * NBLAST, w[r]);
        }
    };
    template <class ST>
    static __device__ __forceinline__ PerElem<ST> per_elem(ST st) { return PerElem<ST>{st}; }

    // Whole transform: load0 -> stages -> store_last (element-wise store).
    template <class LD, class ST>
    static __device__ __forceinline__ void run(int tid, cpx<T>* lds, const cpx<T>* stage_tw, LD&& ld, ST&& st) {
        cpx<T> v[EPT];
        load0(v, tid, ld);
        compute<0>(v, tid, stage_tw);
        middle<1>(v, tid, lds, stage_tw, false);
        store_last(v, tid, per_elem(st));
    }

    // Transform whose output is left in LDS in natural order (index lidx(t, k)).
    // lead_barrier: a barrier is needed before the first LDS write (the buffer
    // was read earlier by other threads).
    template <int LTM = 0, class LD>
    static __device__ __forceinline__ void run_to_lds(int tid, cpx<T>* lds, const cpx<T>* stage_tw, LD&& ld,
                                                      bool lead_barrier = false, const cpx<T>* ltw = nullptr) {
        cpx<T> v[EPT];
        load0(v, tid, ld);
        compute<0>(v, tid, stage_tw);
        middle<1, LTM>(v, tid, lds, stage_tw, lead_barrier, ltw);
        if constexpr (NST > 1) lds_barrier();
        else if (lead_barrier) lds_barrier();
        store_lds<NST - 1>(v, tid, lds);
        lds_barrier();
    }

    // run_to_lds on stage-0 inputs the caller loaded with load0 earlier (so
    // their latency can overlap other work).
    template <int LTM = 0>
    static __device__ __forceinline__ void run_to_lds_loaded(cpx<T>* v, int tid, cpx<T>* lds, const cpx<T>* stage_tw,
                                                             bool lead_barrier = false, const cpx<T>* ltw = nullptr) {
        compute<0>(v, tid, stage_tw);
        middle<1, LTM>(v, tid, lds, stage_tw, lead_barrier, ltw);
        if constexpr (NST > 1) lds_barrier();
        else if (lead_barrier) lds_barrier();
        store_lds<NST - 1>(v, tid, lds);
        lds_barrier();
    }

    // Transform whose natural-order input is already in LDS (a barrier must
    // separate the writes of that input from this call).
    template <int LTM = 0, class ST>
    static __device__ __forceinline__ void run_from_lds(int tid, cpx<T>* lds, const cpx<T>* stage_tw, ST&& st,
                                                        const cpx<T>* ltw = nullptr) {
        cpx<T> v[EPT];
        load_lds<0>(v, tid, lds);
        compute<0>(v, tid, stage_tw);
        middle<1, LTM>(v, tid, lds, stage_tw, true, ltw);
        store_last(v, tid, per_elem(st));
    }
};

// Largest LTM (stages 1 .. LTM read twiddles from an LDS copy) whose copy
// fits beside the kernel's other LDS bytes without lowering the groups a CU
// holds (MI355X: 160 KB LDS, at most 32 waves per CU); 0 = none.
constexpr int groups_per_cu(int lds_bytes, int nt) {
    const int by_waves = 32 / (nt < 64 ? 1 : nt / 64);
    const int by_lds = lds_bytes > 0 ? 163840 / lds_bytes : by_waves;
    return by_lds < by_waves ? by_lds : by_waves;
}
template <class E>
constexpr int engine_ltm(int other_lds_bytes) {
    if (E::NST <= 1) return 0;
    const int g0 = groups_per_cu(other_lds_bytes, E::THREADS);
    for (int ltm = E::NST - 1; ltm >= 1; --ltm)
        if (groups_per_cu(other_lds_bytes + E::tw_elems(ltm) * E::CSIZE, E::THREADS) >= g0) return ltm;
    return 0;
}

}  // namespace nrfft
