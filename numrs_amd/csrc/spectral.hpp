// spectral.hpp -- twofft, cosft1, cosft2 and rlft3 kernels on the Stockham
// engine and the real-FFT untangle (numrs FFT_2.rs, Cos_FT.rs, Cos_FT2.rs,
// Real_FT3.rs; Numerical Recipes semantics, 0-based).
//
// twofft: two real sequences d1, d2 of n points are one complex sequence
//   z = d1 + i d2; after an n-point FFT (isign +1)
//     F1[k] = (Z_k + conj Z_{n-k}) / 2,   F2[k] = (Z_k - conj Z_{n-k}) / (2i).
// cosft1 (DCT-I, n + 1 points): the even extension x of y (2n reals,
//   x_j = y_j for j <= n, x_{2n-j} = y_j) has a real spectrum, and
//   Y_k = F_k(x) / 2; x is read as n complex values and the 2n-point real
//   spectrum is untangled from the n-point FFT.
// cosft2 (n points; DCT-II forward, DCT-III inverse): v_i = y_{2i},
//   v_{n-1-i} = y_{2i+1} (i < n/2); with V the isign +1 real spectrum of v
//   and u_k = exp(i pi k / 2n) V_k:  Y_k = Re u_k,  Y_{n-k} = Im u_k.  The
//   inverse rebuilds V_k = exp(-i pi k / 2n) (Y_k + i Y_{n-k}) (Y_n = 0) and
//   returns unpermute(realft^-1(V)), which is NR's (n/2)-scaled inverse.
// rlft3: fourn over (nn1, nn2, nn3/2) of the data read as complex values,
//   then the realft untangle between bin (k1, k2, k3) and its mirror
//   (-k1, -k2, nn3/2 - k3); the k3 = 0 column pairs with the speq plane
//   (k3 = nn3/2).  The middle column k3 = nn3/4 pairs with itself through a
//   twiddle of +-i, for which the untangle is the identity, so it is skipped.
#pragma once
#include "real_fft.hpp"

namespace nrfft {

template <typename T>
__device__ __forceinline__ T cos_quarter_pi() {
    return (T)0.70710678118654752440;
}

// NRFFT_CF2_LDS = 0: the fused cosft2 inverse with one signal per group
// scatters its unpermuted output element by element from the last stage
// (4-byte stores 16 bytes apart) instead of through LDS in 16-byte groups
// (4096 x 2^14 FP32: 0.212 ms against 0.160-0.165).
#ifndef NRFFT_CF2_LDS
#define NRFFT_CF2_LDS 1
#endif
constexpr bool CF2_LDS = NRFFT_CF2_LDS != 0;

// NRFFT_CF1_PERSIST = 1: the one-group cosft1 whose tile takes more than
// half the LDS (one group per CU: FP32 n = 16384, FP64 n = 8192) runs
// persistent over signals with the next signal's loads in flight during the
// finishing (2: during the FFT as well, a second register set).  MI355X,
// 4096 x 2^14 FP32 0.270 -> 0.252 ms, FP64 4096 x 2^13 0.244 -> 0.228; with
// two or more groups per CU it measured 7-10% slower, and the second
// register set spilled (0.378 ms).  (k_twofft at one group per CU the same
// way: 4096 x 2^14 FP32 0.386 against 0.356 ms -- its stores bound it.)
#ifndef NRFFT_CF1_PERSIST
#define NRFFT_CF1_PERSIST 1
#endif
constexpr bool CF1_PERSIST = NRFFT_CF1_PERSIST != 0;
template <typename T, int L, int TI, int NT>
constexpr bool cf1_persistent() {
    using E = Engine<T, 1, L, TI, NT, MAP_J, MAP_J>;
    return CF1_PERSIST && TI == 1 && E::LDS_ELEMS * E::CSIZE > 80 * 1024;
}

// Four consecutive reals at p: one 16-byte store (FP32) or two (FP64) when
// p is 16-byte aligned, else element by element.
template <typename T>
__device__ __forceinline__ void st_group4(T* p, T a, T b, T c, T d, bool al16) {
    if (al16) {
        if constexpr (sizeof(T) == 4) {
            using V4 = T __attribute__((ext_vector_type(4)));
            *reinterpret_cast<V4*>(p) = V4{a, b, c, d};
        } else {
            using V2 = T __attribute__((ext_vector_type(2)));
            reinterpret_cast<V2*>(p)[0] = V2{a, b};
            reinterpret_cast<V2*>(p)[1] = V2{c, d};
        }
    } else {
        p[0] = a;
        p[1] = b;
        p[2] = c;
        p[3] = d;
    }
}

// cosft2 storage position of v_i (n = 2L values).
__device__ __forceinline__ int64_t cf2_pos(int64_t i, int64_t L) { return i < L ? 2 * i : 2 * (2 * L - 1 - i) + 1; }

// cosft1 even-extension value x_j of y (n = L).
template <typename T>
__device__ __forceinline__ T cf1_ext(const T* ys, int64_t j, int64_t L) {
    return ys[j <= L ? j : 2 * L - j];
}

// ---------------------------------------------------------------------------
// Per-pair finishing steps shared by the fused and the multi-pass kernels.
// ---------------------------------------------------------------------------
// twofft: bins k and n-k of both spectra from Z_k, Z_{n-k}.
template <typename T>
__device__ __forceinline__ void twofft_pair(cpx<T> zk, cpx<T> zr, cpx<T>& g1, cpx<T>& g2) {
    const T half = (T)0.5;
    g1 = cpx<T>{half * (zk.x + zr.x), half * (zk.y - zr.y)};
    g2 = cpx<T>{half * (zk.y + zr.y), -half * (zk.x - zr.x)};
}

// cosft2 forward: Y_k and Y_{n-k} from V_k (LOGN = log2 n).
template <typename T>
__device__ __forceinline__ void cf2_fwd_out(const Twiddles<T>& tw, cpx<T> vk, int64_t k, int logn, T* os, int64_t n) {
    const cpx<T> u = cmul(twiddle_big<1>(tw, (uint32_t)k, logn + 2), vk);
    os[k] = u.x;
    os[n - k] = u.y;
}

// Same with the rotation twiddle exp(i pi k / 2n) given.
template <typename T>
__device__ __forceinline__ void cf2_fwd_out_w(cpx<T> rot, cpx<T> vk, int64_t k, T* os, int64_t n) {
    const cpx<T> u = cmul(rot, vk);
    os[k] = u.x;
    os[n - k] = u.y;
}

// Twiddles W_M^p of the bin pairs p = tid + u NT, u < PPT, of a group that
// holds one signal: W_M^tid (per thread, loaded once) times W_M^(u NT)
// (the same for every thread: a small LDS table written before the FFT's
// first barrier) -- instead of two table loads per pair after a barrier.
template <typename T, int PPT>
struct PairTw {
    cpx<T> wt;
    __device__ __forceinline__ void init(cpx<T>* tab, const Twiddles<T>& tw, int tid, int nt, int logm) {
        if (tid < PPT) tab[tid] = twiddle_big<1>(tw, uint32_t(tid * nt), logm);
        wt = twiddle_big<1>(tw, uint32_t(tid), logm);
    }
    __device__ __forceinline__ cpx<T> at(const cpx<T>* tab, int u) const { return cmul(wt, tab[u]); }
};

// cosft2 inverse: V_k from Y_k, Y_{n-k}.
template <typename T>
__device__ __forceinline__ cpx<T> cf2_inv_in(const Twiddles<T>& tw, const T* ys, int64_t k, int logn, int64_t n) {
    return cmul(twiddle_big<-1>(tw, (uint32_t)k, logn + 2), cpx<T>{ys[k], ys[n - k]});
}

// ---------------------------------------------------------------------------
// Fused kernels: TI transforms per group, the FFT on the engine, the
// finishing step on the result left in LDS (one thread per bin pair).
// ---------------------------------------------------------------------------
template <typename T, int L, int TI, int NT>
__global__ __launch_bounds__(NT) void k_twofft(const T* d1, const T* d2, cpx<T>* f1, cpx<T>* f2, int64_t nsig,
                                               Twiddles<T> tw) {
    using E = Engine<T, 1, L, TI, NT, MAP_J, MAP_J>;
    constexpr int LTM = engine_ltm<E>(E::LDS_ELEMS * E::CSIZE);
    __shared__ cpx<T> lds[E::LDS_ELEMS];
    __shared__ __align__(16) cpx<T> ltw[LTM > 0 ? E::tw_elems(LTM) : 1];
    const int tid = threadIdx.x;
    const int64_t s0 = int64_t(blockIdx.x) * TI;
    const int nvalid = (nsig - s0) < TI ? int(nsig - s0) : TI;
    const T* a = d1 + s0 * L;
    const T* b = d2 + s0 * L;
    if constexpr (LTM > 0) E::template stage_tw_to_lds<LTM>(ltw, tw.stage, tid);  // read after the first exchange
    E::template run_to_lds<LTM>(tid, lds, tw.stage, [&](int t, int i) -> cpx<T> {
        return (TI == 1 || t < nvalid) ? cpx<T>{a[t * L + i], b[t * L + i]} : cpx<T>{(T)0, (T)0};
    }, false, ltw);
    constexpr int NP = L / 2 + 1;
    for (int q = tid; q < TI * NP; q += NT) {
        const int t = q / NP, k = q % NP;
        if (TI > 1 && t >= nvalid) continue;
        const int r = (L - k) & (L - 1);
        cpx<T> g1, g2;
        twofft_pair(lds[E::lidx(t, k)], lds[E::lidx(t, r)], g1, g2);
        cpx<T>* o1 = f1 + (s0 + t) * L;
        cpx<T>* o2 = f2 + (s0 + t) * L;
        o1[k] = g1;
        o2[k] = g2;
        if (r != k) {
            o1[r] = conj(g1);
            o2[r] = conj(g2);
        }
    }
}

// cosft1, n = L: n + 1 reals per transform (stride L + 1).
template <typename T, int L, int TI, int NT>
__global__ __launch_bounds__(NT) void k_cosft1(const T* y, T* out, int64_t nsig, Twiddles<T> tw) {
    using E = Engine<T, 1, L, TI, NT, MAP_J, MAP_J>;
    constexpr int64_t S = L + 1;
    constexpr int LOG2N2 = ilog2c(L) + 1;  // log2 of the 2n-point real transform
    constexpr int LTM = engine_ltm<E>(E::LDS_ELEMS * E::CSIZE);
    __shared__ cpx<T> lds[E::LDS_ELEMS];
    __shared__ __align__(16) cpx<T> ltw[LTM > 0 ? E::tw_elems(LTM) : 1];
    const int tid = threadIdx.x;
    const int64_t s0 = int64_t(blockIdx.x) * TI;
    const int nvalid = (nsig - s0) < TI ? int(nsig - s0) : TI;
    const T* yb = y + s0 * S;
    if constexpr (LTM > 0) E::template stage_tw_to_lds<LTM>(ltw, tw.stage, tid);  // read after the first exchange
    constexpr int PPT1 = (L / 2 + 1 + NT - 1) / NT;
    __shared__ cpx<T> uw[PPT1];
    PairTw<T, PPT1> ptw;
    if constexpr (TI == 1) ptw.init(uw, tw, tid, NT, LOG2N2);
    constexpr int NP = L / 2 + 1;
    const T half = (T)0.5;
    if constexpr (cf1_persistent<T, L, TI, NT>()) {
        // Persistent over signals blockIdx.x + k gridDim.x, the next
        // signal's stage-0 loads in flight while this one finishes (signals
        // are disjoint, so in place is safe).
        const int64_t G = gridDim.x;
        auto load = [&](cpx<T>* v, int64_t sg) {
            const T* ys = y + sg * S;
            E::load0(v, tid, [&](int, int m) -> cpx<T> {
                return cpx<T>{cf1_ext(ys, 2 * m, L), cf1_ext(ys, 2 * m + 1, L)};
            });
        };
        auto finish = [&](int64_t sg) {
            T* os = out + sg * S;
            int u = 0;
            for (int p = tid; p < NP; p += NT, ++u) {
                if (p == 0) {
                    const cpx<T> z0 = lds[E::lidx(0, 0)];
                    os[0] = half * (z0.x + z0.y);
                    os[L] = half * (z0.x - z0.y);
                } else {
                    cpx<T> fa, fb;
                    untangle_pair(lds[E::lidx(0, p)], lds[E::lidx(0, L - p)], ptw.at(uw, u), (T)-0.5, fa, fb);
                    os[p] = half * fa.x;
                    if (L - p != p) os[L - p] = half * fb.x;
                }
            }
        };
        auto step = [&](cpx<T>* vc, cpx<T>* vn, int64_t sg) {
            if (sg + G < nsig) load(vn, sg + G);
            E::template run_to_lds_loaded<LTM>(vc, tid, lds, tw.stage, true, ltw);
            finish(sg);
        };
        if constexpr (NRFFT_CF1_PERSIST == 2) {
            cpx<T> va[E::EPT], vb[E::EPT];
            if (s0 < nsig) load(va, s0);
            for (int64_t sg = s0; sg < nsig; sg += 2 * G) {
                step(va, vb, sg);
                if (sg + G < nsig) step(vb, va, sg + G);
            }
        } else {
            // One register set: the next loads are issued after the FFT and
            // overlap the finishing only.
            cpx<T> v[E::EPT];
            if (s0 < nsig) load(v, s0);
            for (int64_t sg = s0; sg < nsig; sg += G) {
                E::template run_to_lds_loaded<LTM>(v, tid, lds, tw.stage, true, ltw);
                if (sg + G < nsig) load(v, sg + G);
                finish(sg);
            }
        }
        return;
    }
    E::template run_to_lds<LTM>(tid, lds, tw.stage, [&](int t, int m) -> cpx<T> {
        if (!(TI == 1 || t < nvalid)) return cpx<T>{(T)0, (T)0};
        const T* ys = yb + t * S;
        return cpx<T>{cf1_ext(ys, 2 * m, L), cf1_ext(ys, 2 * m + 1, L)};
    }, false, ltw);
    int u = 0;
    for (int q = tid; q < TI * NP; q += NT, ++u) {
        const int t = q / NP, p = q % NP;
        if (TI > 1 && t >= nvalid) continue;
        T* os = out + (s0 + t) * S;
        if (p == 0) {
            const cpx<T> z0 = lds[E::lidx(t, 0)];
            os[0] = half * (z0.x + z0.y);
            os[L] = half * (z0.x - z0.y);
        } else {
            cpx<T> fa, fb;
            const cpx<T> w = TI == 1 ? ptw.at(uw, u) : twiddle_big<1>(tw, (uint32_t)p, LOG2N2);
            untangle_pair(lds[E::lidx(t, p)], lds[E::lidx(t, L - p)], w, (T)-0.5, fa, fb);
            os[p] = half * fa.x;
            if (L - p != p) os[L - p] = half * fb.x;
        }
    }
}

// cosft2, n = 2L reals per transform; DIR = +1 forward, -1 inverse.
template <typename T, int L, int TI, int NT, int DIR>
__global__ __launch_bounds__(NT) void k_cosft2(const T* y, T* out, int64_t nsig, Twiddles<T> tw) {
    using E = Engine<T, DIR, L, TI, NT, MAP_J, MAP_J>;
    constexpr int64_t N = 2 * L;
    constexpr int LOGN = ilog2c(L) + 1;
    constexpr int LTM = engine_ltm<E>(E::LDS_ELEMS * E::CSIZE);
    __shared__ cpx<T> lds[E::LDS_ELEMS];
    __shared__ __align__(16) cpx<T> ltw[LTM > 0 ? E::tw_elems(LTM) : 1];
    const int tid = threadIdx.x;
    const int64_t s0 = int64_t(blockIdx.x) * TI;
    const int nvalid = (nsig - s0) < TI ? int(nsig - s0) : TI;
    const T* yb = y + s0 * N;
    constexpr int NP = L / 2 + 1;
    if constexpr (LTM > 0) E::template stage_tw_to_lds<LTM>(ltw, tw.stage, tid);  // read after the first exchange
    // One signal per group: untangle twiddles W_n^p and rotations W_4n^p,
    // W_4n^(L-p) = W_8 conj(W_4n^p) from per-thread and shared factors.
    constexpr int PPT2 = (NP + NT - 1) / NT;
    __shared__ cpx<T> uwn[PPT2], uw4[PPT2];
    PairTw<T, PPT2> pn, p4;
    cpx<T> w8{(T)0, (T)0};
    if constexpr (TI == 1) {
        pn.init(uwn, tw, tid, NT, LOGN);
        p4.init(uw4, tw, tid, NT, LOGN + 2);
        w8 = twiddle_big<1>(tw, uint32_t(L), LOGN + 2);
    }
    if constexpr (DIR > 0) {
        // The permuted gather stays element by element: through LDS in
        // 16-byte groups (as the inverse stores) it measured slower, 4096 x
        // 2^14 FP32 0.164 ms against 0.155.
        E::template run_to_lds<LTM>(tid, lds, tw.stage, [&](int t, int m) -> cpx<T> {
            if (!(TI == 1 || t < nvalid)) return cpx<T>{(T)0, (T)0};
            const T* ys = yb + t * N;
            return cpx<T>{ys[cf2_pos(2 * m, L)], ys[cf2_pos(2 * m + 1, L)]};
        }, false, ltw);
        int u = 0;
        for (int q = tid; q < TI * NP; q += NT, ++u) {
            const int t = q / NP, p = q % NP;
            if (TI > 1 && t >= nvalid) continue;
            T* os = out + (s0 + t) * N;
            if (p == 0) {
                const cpx<T> z0 = lds[E::lidx(t, 0)];
                os[0] = z0.x + z0.y;
                os[L] = (z0.x - z0.y) * cos_quarter_pi<T>();
            } else if constexpr (TI == 1) {
                cpx<T> va, vb;
                untangle_pair(lds[E::lidx(t, p)], lds[E::lidx(t, L - p)], pn.at(uwn, u), (T)-0.5, va, vb);
                const cpx<T> r4 = p4.at(uw4, u);
                cf2_fwd_out_w(r4, va, p, os, N);
                if (L - p != p) cf2_fwd_out_w(cmul(w8, conj(r4)), vb, L - p, os, N);
            } else {
                cpx<T> va, vb;
                untangle_pair(lds[E::lidx(t, p)], lds[E::lidx(t, L - p)], twiddle_big<1>(tw, (uint32_t)p, LOGN),
                              (T)-0.5, va, vb);
                cf2_fwd_out(tw, va, p, LOGN, os, N);
                if (L - p != p) cf2_fwd_out(tw, vb, L - p, LOGN, os, N);
            }
        }
    } else {
        if constexpr (TI == 1) lds_barrier();  // uwn, uw4 published
        int u = 0;
        for (int q = tid; q < TI * NP; q += NT, ++u) {
            const int t = q / NP, p = q % NP;
            const bool ok = TI == 1 || t < nvalid;
            const T* ys = yb + t * N;
            if (p == 0) {
                const T sqrt2 = (T)1.41421356237309504880;
                lds[E::lidx(t, 0)] = ok ? dc_inverse(cpx<T>{ys[0], sqrt2 * ys[L]}) : cpx<T>{(T)0, (T)0};
            } else {
                cpx<T> za{(T)0, (T)0}, zb{(T)0, (T)0};
                if constexpr (TI == 1) {
                    // V_k = conj(W_4n^k) (Y_k + i Y_{n-k}) for k = p and L - p
                    const cpx<T> r4 = p4.at(uw4, u);
                    const cpx<T> va = cmul(conj(r4), cpx<T>{ys[p], ys[N - p]});
                    const cpx<T> vb = cmul(conj(cmul(w8, conj(r4))), cpx<T>{ys[L - p], ys[N - (L - p)]});
                    untangle_pair(va, vb, conj(pn.at(uwn, u)), (T)0.5, za, zb);
                } else if (ok)
                    untangle_pair(cf2_inv_in(tw, ys, p, LOGN, N), cf2_inv_in(tw, ys, L - p, LOGN, N),
                                  twiddle_big<-1>(tw, (uint32_t)p, LOGN), (T)0.5, za, zb);
                lds[E::lidx(t, p)] = za;
                lds[E::lidx(t, L - p)] = zb;
            }
        }
        lds_barrier();
        if constexpr (TI == 1 && CF2_LDS) {
            // The inverse FFT back into LDS in natural order, then the
            // unpermute as 16-byte groups y[4q .. 4q+3] = z[q].x, z[L-1-q].y,
            // z[q].y, z[L-1-q].x (q < L/2) -- whole lines per wave instead of
            // 4-byte stores 16 bytes apart.
            E::template run_to_lds<LTM>(tid, lds, tw.stage, [&](int t, int m) { return lds[E::lidx(t, m)]; }, true,
                                        ltw);
            T* os = out + s0 * N;
            const bool al16 = (reinterpret_cast<uintptr_t>(os) & 15) == 0;
            for (int q = tid; q < L / 2; q += NT) {
                const cpx<T> a = lds[E::lidx(0, q)], b = lds[E::lidx(0, L - 1 - q)];
                st_group4(os + 4 * q, a.x, b.y, a.y, b.x, al16);
            }
        } else {
            E::template run_from_lds<LTM>(tid, lds, tw.stage, [&](int t, int m, cpx<T> v) {
                if (TI == 1 || t < nvalid) {
                    T* os = out + (s0 + t) * N;
                    os[cf2_pos(2 * m, L)] = v.x;
                    os[cf2_pos(2 * m + 1, L)] = v.y;
                }
            }, ltw);
        }
    }
}

// cosft2 with n = 2 (no FFT): Y0 = y0 + y1, Y1 = (y0 - y1) cos(pi/4); the
// inverse y_j = Y0 / 2 +- Y1 cos(pi/4).
template <typename T, int DIR>
__global__ void k_cosft2_n2(const T* y, T* out, int64_t nsig) {
    const int64_t s = blockIdx.x * int64_t(blockDim.x) + threadIdx.x;
    if (s >= nsig) return;
    const T a = y[2 * s], b = y[2 * s + 1];
    const T c = cos_quarter_pi<T>();
    if (DIR > 0) {
        out[2 * s] = a + b;
        out[2 * s + 1] = (a - b) * c;
    } else {
        out[2 * s] = (T)0.5 * a + b * c;
        out[2 * s + 1] = (T)0.5 * a - b * c;
    }
}

// ---------------------------------------------------------------------------
// Multi-pass kernels (transforms longer than one tile): pack -> C2C plan ->
// finish, one thread per element or bin pair.
// ---------------------------------------------------------------------------
// The multi-pass helpers below run on a 2-D grid: signals over blockIdx.y
// (stride gridDim.y), the elements of one signal over the x grid (stride
// gridDim.x * blockDim.x), so no index is divided by a runtime count.
__device__ __forceinline__ int64_t x_first() { return blockIdx.x * int64_t(blockDim.x) + threadIdx.x; }
__device__ __forceinline__ int64_t x_stride() { return int64_t(gridDim.x) * blockDim.x; }

// twofft beyond one group: Z1, Z2 are the n/2-point spectra of data1, data2
// read as complex values; each realft untangle (W = exp(2 pi i / n)) gives
// the bins k <= n/2 of its real input, written as the full n-bin spectrum
// (F[n - k] = conj F[k]).  Grid y runs over (signal, array) pairs, so a
// thread keeps two load and four store streams.  Pairs (p, N - p), N = n/2,
// p < N/2; the thread of p = 0 also writes DC / Nyquist and the self-paired
// bin N/2 (F = Z there).
template <typename T>
__global__ __launch_bounds__(PAIR_NT) void k_twofft_real(const cpx<T>* z1, const cpx<T>* z2, cpx<T>* f1, cpx<T>* f2,
                                                         int64_t n64, int64_t nsig, int logn, Twiddles<T> tw) {
    const int n = int(n64), N = n / 2, nq = N / 2;
    const RowTw<1, T> wr(tw, logn);
    for (int64_t y = blockIdx.y; y < 2 * nsig; y += gridDim.y) {
        const int64_t s = y >> 1;
        const cpx<T>* a = ((y & 1) ? z2 : z1) + s * int64_t(N);
        cpx<T>* o = ((y & 1) ? f2 : f1) + s * n64;
        for (int p0 = pair_first(); p0 < nq; p0 += pair_stride()) {
            cpx<T> x[PAIR_U], z[PAIR_U], w[PAIR_U];
#pragma unroll
            for (int u = 0; u < PAIR_U; ++u) {
                const int p = pair_at(p0, u);
                if (p < nq) {
                    x[u] = a[p];
                    z[u] = a[(N - p) & (N - 1)];
                    w[u] = wr.at(tw, p);
                }
            }
#pragma unroll
            for (int u = 0; u < PAIR_U; ++u) {
                const int p = pair_at(p0, u);
                if (p >= nq) continue;
                if (p == 0) {
                    o[0] = cpx<T>{x[u].x + x[u].y, (T)0};
                    o[N] = cpx<T>{x[u].x - x[u].y, (T)0};
                    const cpx<T> m = a[nq];
                    o[nq] = m;
                    o[N + nq] = conj(m);
                } else {
                    cpx<T> g, h;
                    untangle_pair(x[u], z[u], w[u], (T)-0.5, g, h);
                    o[p] = g;
                    o[N - p] = h;
                    o[n - p] = conj(g);
                    o[N + p] = conj(h);
                }
            }
        }
    }
}

template <typename T>
__global__ void k_cosft1_pack(const T* y, cpx<T>* z, int64_t L, int64_t nsig) {
    for (int64_t s = blockIdx.y; s < nsig; s += gridDim.y) {
        const T* ys = y + s * (L + 1);
        cpx<T>* zs = z + s * L;
        for (int64_t m = x_first(); m < L; m += x_stride())
            zs[m] = cpx<T>{cf1_ext(ys, 2 * m, L), cf1_ext(ys, 2 * m + 1, L)};
    }
}

// The pairs p < L/2 of a signal; the thread of p = 0 also finishes the
// self-paired bin p = L/2, so the spans of a signal divide evenly among blocks.
template <typename T>
__global__ __launch_bounds__(PAIR_NT) void k_cosft1_post(const cpx<T>* z, T* out, int64_t L64, int64_t nsig,
                                                         int log2n2, Twiddles<T> tw) {
    const int L = int(L64), nq = L / 2;
    const T half = (T)0.5;
    const RowTw<1, T> wp(tw, log2n2);
    for (int64_t s = blockIdx.y; s < nsig; s += gridDim.y) {
        const cpx<T>* zs = z + s * L64;
        T* os = out + s * (L64 + 1);
        for (int p0 = pair_first(); p0 < nq; p0 += pair_stride()) {
            cpx<T> a[PAIR_U], b[PAIR_U], w[PAIR_U];
#pragma unroll
            for (int u = 0; u < PAIR_U; ++u) {
                const int p = pair_at(p0, u);
                if (p < nq) {
                    a[u] = zs[p];
                    b[u] = zs[(L - p) & (L - 1)];
                    w[u] = wp.at(tw, p);
                }
            }
#pragma unroll
            for (int u = 0; u < PAIR_U; ++u) {
                const int p = pair_at(p0, u);
                if (p >= nq) continue;
                if (p == 0) {
                    os[0] = half * (a[u].x + a[u].y);
                    os[L] = half * (a[u].x - a[u].y);
                    const cpx<T> m = zs[nq];
                    cpx<T> fa, fb;
                    untangle_pair(m, m, twiddle_big<1>(tw, uint32_t(nq), log2n2), (T)-0.5, fa, fb);
                    os[nq] = half * fa.x;
                } else {
                    cpx<T> fa, fb;
                    untangle_pair(a[u], b[u], w[u], (T)-0.5, fa, fb);
                    os[p] = half * fa.x;
                    os[L - p] = half * fb.x;
                }
            }
        }
    }
}

// cosft2's even/odd permutation, 16 bytes of the real array per thread: the
// four reals y[4q .. 4q+3] are z[q].x, z[L-1-q].y, z[q].y, z[L-1-q].x
// (y[2i] = x_i and y[2(2L-1-i)+1] = x_i for the packed reals x of z), so
// every lane reads or writes one contiguous aligned group and whole cache
// lines come from one wave.  q < L/2.
template <typename T>
__global__ __launch_bounds__(PAIR_NT) void k_cosft2_fpack(const T* y, cpx<T>* z, int64_t L64, int64_t nsig) {
    const int L = int(L64), nq = L / 2;
    for (int64_t s = blockIdx.y; s < nsig; s += gridDim.y) {
        const T* ys = y + s * 2 * L64;
        cpx<T>* zs = z + s * L64;
        for (int q0 = pair_first(); q0 < nq; q0 += pair_stride()) {
            T v[PAIR_U][4];
#pragma unroll
            for (int u = 0; u < PAIR_U; ++u) {
                const int q = pair_at(q0, u);
                if (q < nq) {
#pragma unroll
                    for (int k = 0; k < 4; ++k) v[u][k] = ys[4 * q + k];
                }
            }
#pragma unroll
            for (int u = 0; u < PAIR_U; ++u) {
                const int q = pair_at(q0, u);
                if (q >= nq) continue;
                zs[q] = cpx<T>{v[u][0], v[u][2]};
                zs[L - 1 - q] = cpx<T>{v[u][3], v[u][1]};
            }
        }
    }
}

template <typename T>
__global__ __launch_bounds__(PAIR_NT) void k_cosft2_fpost(const cpx<T>* z, T* out, int64_t L64, int64_t nsig, int logn,
                                                          Twiddles<T> tw) {
    const int L = int(L64), nq = L / 2, n = 2 * L;
    // exp(i pi (L - p) / 2n) = W_8 conj(exp(i pi p / 2n))
    const cpx<T> w8 = twiddle_big<1>(tw, (uint32_t)L, logn + 2);
    const RowTw<1, T> wn(tw, logn), w4(tw, logn + 2);
    for (int64_t s = blockIdx.y; s < nsig; s += gridDim.y) {
        const cpx<T>* zs = z + s * L64;
        T* os = out + s * 2 * L64;
        for (int p0 = pair_first(); p0 < nq; p0 += pair_stride()) {
            cpx<T> a[PAIR_U], b[PAIR_U], wu[PAIR_U], r4[PAIR_U];
#pragma unroll
            for (int u = 0; u < PAIR_U; ++u) {
                const int p = pair_at(p0, u);
                if (p < nq) {
                    a[u] = zs[p];
                    b[u] = zs[(L - p) & (L - 1)];
                    wu[u] = wn.at(tw, p);
                    r4[u] = w4.at(tw, p);
                }
            }
#pragma unroll
            for (int u = 0; u < PAIR_U; ++u) {
                const int p = pair_at(p0, u);
                if (p >= nq) continue;
                if (p == 0) {
                    os[0] = a[u].x + a[u].y;
                    os[L] = (a[u].x - a[u].y) * cos_quarter_pi<T>();
                    const cpx<T> m = zs[nq];
                    cpx<T> va, vb;
                    untangle_pair(m, m, twiddle_big<1>(tw, uint32_t(nq), logn), (T)-0.5, va, vb);
                    cf2_fwd_out_w(twiddle_big<1>(tw, uint32_t(nq), logn + 2), va, nq, os, n);
                } else {
                    cpx<T> va, vb;
                    untangle_pair(a[u], b[u], wu[u], (T)-0.5, va, vb);
                    cf2_fwd_out_w(r4[u], va, p, os, n);
                    cf2_fwd_out_w(cmul(w8, conj(r4[u])), vb, L - p, os, n);
                }
            }
        }
    }
}

template <typename T>
__global__ __launch_bounds__(PAIR_NT) void k_cosft2_ipre(const T* y, cpx<T>* z, int64_t L64, int64_t nsig, int logn,
                                                         Twiddles<T> tw) {
    const int L = int(L64), nq = L / 2, n = 2 * L;
    const cpx<T> w8 = twiddle_big<1>(tw, (uint32_t)L, logn + 2);
    const RowTw<1, T> wn(tw, logn), w4(tw, logn + 2);
    for (int64_t s = blockIdx.y; s < nsig; s += gridDim.y) {
        const T* ys = y + s * 2 * L64;
        cpx<T>* zs = z + s * L64;
        for (int p0 = pair_first(); p0 < nq; p0 += pair_stride()) {
            // Y_p, Y_{n-p}, Y_{L-p}, Y_{L+p} (p = 0: Y_0 and Y_L)
            T y0[PAIR_U], y1[PAIR_U], y2[PAIR_U], y3[PAIR_U];
            cpx<T> wu[PAIR_U], r4[PAIR_U];
#pragma unroll
            for (int u = 0; u < PAIR_U; ++u) {
                const int p = pair_at(p0, u);
                if (p < nq) {
                    y0[u] = ys[p];
                    y1[u] = ys[(n - p) & (n - 1)];
                    y2[u] = ys[L - p];
                    y3[u] = ys[L + p];
                    wu[u] = wn.at(tw, p);
                    r4[u] = w4.at(tw, p);
                }
            }
#pragma unroll
            for (int u = 0; u < PAIR_U; ++u) {
                const int p = pair_at(p0, u);
                if (p >= nq) continue;
                if (p == 0) {
                    zs[0] = dc_inverse(cpx<T>{y0[u], (T)1.41421356237309504880 * y2[u]});
                    // the self-paired bin p = L/2 (k = L - p = p)
                    const cpx<T> rm = twiddle_big<1>(tw, uint32_t(nq), logn + 2);
                    const cpx<T> vm = cmul(conj(rm), cpx<T>{ys[nq], ys[n - nq]});
                    cpx<T> za, zb;
                    untangle_pair(vm, vm, conj(twiddle_big<1>(tw, uint32_t(nq), logn)), (T)0.5, za, zb);
                    zs[nq] = zb;
                } else {
                    // V_k = exp(-i pi k / 2n) (Y_k + i Y_{n-k}) for k = p and L - p
                    const cpx<T> va = cmul(conj(r4[u]), cpx<T>{y0[u], y1[u]});
                    const cpx<T> vb = cmul(conj(cmul(w8, conj(r4[u]))), cpx<T>{y2[u], y3[u]});
                    cpx<T> za, zb;
                    untangle_pair(va, vb, conj(wu[u]), (T)0.5, za, zb);
                    zs[p] = za;
                    zs[L - p] = zb;
                }
            }
        }
    }
}

template <typename T>
__global__ __launch_bounds__(PAIR_NT) void k_cosft2_iunpack(const cpx<T>* z, T* out, int64_t L64, int64_t nsig) {
    const int L = int(L64), nq = L / 2;
    for (int64_t s = blockIdx.y; s < nsig; s += gridDim.y) {
        const cpx<T>* zs = z + s * L64;
        T* os = out + s * 2 * L64;
        for (int q0 = pair_first(); q0 < nq; q0 += pair_stride()) {
            cpx<T> a[PAIR_U], b[PAIR_U];
#pragma unroll
            for (int u = 0; u < PAIR_U; ++u) {
                const int q = pair_at(q0, u);
                if (q < nq) {
                    a[u] = zs[q];
                    b[u] = zs[L - 1 - q];
                }
            }
#pragma unroll
            for (int u = 0; u < PAIR_U; ++u) {
                const int q = pair_at(q0, u);
                if (q >= nq) continue;
                os[4 * q] = a[u].x;
                os[4 * q + 1] = b[u].y;
                os[4 * q + 2] = a[u].y;
                os[4 * q + 3] = b[u].x;
            }
        }
    }
}

// rlft3: speq[s][i1][i2] = data[s][i1][i2][0] (complex; forward, after fourn).
template <typename T>
__global__ void k_rlft3_copycol(const cpx<T>* data, cpx<T>* speq, int64_t plane, int64_t h, int64_t nsig) {
    for (int64_t q = blockIdx.x * int64_t(blockDim.x) + threadIdx.x; q < plane * nsig;
         q += int64_t(gridDim.x) * blockDim.x)
        speq[q] = data[q * h];
}

// rlft3 untangle, one thread per (s, i1, i2, k3), k3 < max(1, h/2); every
// extent is a power of two, so the index splits into bit fields.
template <typename T, int DIR>
__global__ void k_rlft3_untangle(cpx<T>* data, cpx<T>* speq, int64_t nn1, int64_t nn2, int64_t h, int64_t nsig,
                                 int lognn3, Twiddles<T> tw) {
    const int64_t nk = h > 1 ? h / 2 : 1;
    const int64_t plane = nn1 * nn2;
    const int64_t total = plane * nk * nsig;
    const int lk = __builtin_ctzll((unsigned long long)nk), l2 = __builtin_ctzll((unsigned long long)nn2),
              l1 = __builtin_ctzll((unsigned long long)nn1);
    const T c2 = (T)(-0.5 * DIR);
    // k3 = q mod nk with q = (block base) + threadIdx.x: when nk <= blockDim.x
    // (both powers of two) k3 = threadIdx.x mod nk for every q of the thread;
    // otherwise W^k3 = W^threadIdx.x W^(k3 - threadIdx.x), the second factor
    // the same for the whole block.
    const cpx<T> wt = twiddle_big<DIR>(tw, uint32_t(nk <= (int64_t)blockDim.x ? (threadIdx.x & (nk - 1)) : threadIdx.x),
                                       lognn3);
    auto k3tw = [&](int64_t k3) -> cpx<T> {
        if (nk <= (int64_t)blockDim.x) return wt;
        return cmul(wt, twiddle_big<DIR>(tw, uint32_t(k3 - threadIdx.x), lognn3));
    };
    for (int64_t q = blockIdx.x * int64_t(blockDim.x) + threadIdx.x; q < total; q += int64_t(gridDim.x) * blockDim.x) {
        const int64_t k3 = q & (nk - 1);
        const int64_t r = q >> lk;
        const int64_t i2 = r & (nn2 - 1), i1 = (r >> l2) & (nn1 - 1), s = r >> (l1 + l2);
        const int64_t j1 = (nn1 - i1) & (nn1 - 1), j2 = (nn2 - i2) & (nn2 - 1);
        cpx<T>* ds = data + s * plane * h;
        cpx<T>* ss = speq + s * plane;
        cpx<T>& a = ds[(i1 * nn2 + i2) * h + k3];
        cpx<T>& b = k3 == 0 ? ss[j1 * nn2 + j2] : ds[(j1 * nn2 + j2) * h + (h - k3)];
        const cpx<T> w = k3 == 0 ? cpx<T>{(T)1, (T)0} : k3tw(k3);
        cpx<T> oa, ob;
        untangle_pair(a, b, w, c2, oa, ob);
        a = oa;
        b = ob;
    }
}

// rlft3 forward, last axis fused with the untangle.  The rows (i1, i2) of
// each volume pair with their mirrors m = (-i1, -i2); the pairs are
// enumerated per volume as i1 = 0 with i2 <= nn2/2, 0 < i1 < nn1/2 with
// every i2, i1 = nn1/2 with i2 <= nn2/2 (nn1 nn2 / 2 + 2 pairs, four rows
// their own mirror; nn1 = 1, the 2-D case: i1 = 0 only, nn2 / 2 + 1 pairs,
// two rows their own mirror).  A tile holds the H = nn3/2-point rows of TI/2
// pairs, slot t the row and slot t + TI/2 its mirror, read from z (the
// other axes already transformed, out of place); after the FFTs each row r
// of the tile, with m its partner, writes
//   k = 0        data[r][0] = untangle(Z_r[0], Z_m[0], 1).a,
//                speq[r]    = untangle(Z_m[0], Z_r[0], 1).b,
//   0 < k < H/2  data[r][k] = untangle(Z_r[k], Z_m[H-k], W^k).a,
//   k = H/2      data[r][k] = Z_r[k],
//   H/2 < k      data[r][k] = untangle(Z_m[H-k], Z_r[k], W^(H-k)).b,
// W = W_nn3 (isign +1): the pairs k_rlft3_untangle forms.  A row that is
// its own mirror is written by its first slot only.
template <typename T, int H, int TI, int NT>
__global__ __launch_bounds__(NT) void k_rlft3_rows(const cpx<T>* z, cpx<T>* data, cpx<T>* speq, int64_t npairs,
                                                   int64_t nn1, int64_t nn2, Twiddles<T> tw) {
    using E = Engine<T, 1, H, TI, NT, MAP_J, MAP_J>;
    constexpr int HALF = TI / 2;
    static_assert(TI >= 2 && TI % 2 == 0 && H >= 16 && NT >= TI, "rlft3 row tile");
    constexpr int LTM = engine_ltm<E>(E::LDS_ELEMS * E::CSIZE);
    constexpr int LOGNN3 = ilog2c(H) + 1;
    __shared__ cpx<T> lds[E::LDS_ELEMS];
    __shared__ __align__(16) cpx<T> ltw[LTM > 0 ? E::tw_elems(LTM) : 1];
    __shared__ int64_t rows[TI];  // global row of each slot, -1 = none
    const int tid = threadIdx.x;
    if (tid < TI) {
        const int64_t pi = int64_t(blockIdx.x) * HALF + (tid % HALF);
        int64_t row = -1;
        if (pi < npairs) {
            const int64_t plane = nn1 * nn2, h2 = nn2 / 2 + 1, per = nn1 >= 2 ? plane / 2 + 2 : h2;
            const int64_t s = pi / per;
            int64_t q = pi - s * per, i1, i2;
            if (q < h2) {
                i1 = 0;
                i2 = q;
            } else if ((q -= h2) < (nn1 / 2 - 1) * nn2) {
                i1 = 1 + q / nn2;
                i2 = q % nn2;
            } else {
                i1 = nn1 / 2;
                i2 = q - (nn1 / 2 - 1) * nn2;
            }
            if (tid >= HALF) {
                i1 = (nn1 - i1) & (nn1 - 1);
                i2 = (nn2 - i2) & (nn2 - 1);
            }
            row = s * plane + i1 * nn2 + i2;
        }
        rows[tid] = row;
    }
    if constexpr (LTM > 0) E::template stage_tw_to_lds<LTM>(ltw, tw.stage, tid);  // read after the first exchange
    lds_barrier();  // rows published
    E::template run_to_lds<LTM>(tid, lds, tw.stage, [&](int t, int i) -> cpx<T> {
        const int64_t row = rows[t];
        return row >= 0 ? z[row * H + i] : cpx<T>{(T)0, (T)0};
    }, false, ltw);
    const T c2 = (T)-0.5;
    for (int w = tid; w < TI * H; w += NT) {
        const int t = w / H, k = w % H;
        const int tm = t < HALF ? t + HALF : t - HALF;  // the partner's slot
        const int64_t r = rows[t];
        if (r < 0 || (t >= HALF && r == rows[tm])) continue;
        const cpx<T> a = lds[E::lidx(t, k)];
        cpx<T> o, oa, ob;
        if (k == 0) {
            const cpx<T> b = lds[E::lidx(tm, 0)];
            const cpx<T> one{(T)1, (T)0};
            untangle_pair(a, b, one, c2, oa, ob);
            o = oa;
            untangle_pair(b, a, one, c2, oa, ob);
            speq[r] = ob;
        } else if (k == H / 2) {
            o = a;
        } else if (k < H / 2) {
            untangle_pair(a, lds[E::lidx(tm, H - k)], twiddle_big<1>(tw, uint32_t(k), LOGNN3), c2, oa, ob);
            o = oa;
        } else {
            untangle_pair(lds[E::lidx(tm, H - k)], a, twiddle_big<1>(tw, uint32_t(H - k), LOGNN3), c2, oa, ob);
            o = ob;
        }
        data[r * H + k] = o;
    }
}

}  // namespace nrfft
