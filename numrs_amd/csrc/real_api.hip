// real_api.hip -- realft / convlv / correl: kernel instantiations, planning,
// execution and the extern "C" entry points of include/nrfft.h.
#include <cstring>
#include <memory>
#include <new>
#include <utility>
#include <vector>

#include "../../include/nrfft.h"
#include "c2c_instances.hpp"
#include "internal.hpp"
#include "real_fft.hpp"

namespace nrfft {

#define HIPCHK(x)                                                                                      \
    do {                                                                                               \
        hipError_t e_ = (x);                                                                           \
        if (e_ != hipSuccess) return fail(NRFFT_E_HIP, std::string(#x) + ": " + hipGetErrorString(e_)); \
    } while (0)

constexpr int REAL_MAX_LOGL = 13;  // fused single-group path: n/2 <= 8192
// NRFFT_REALFT_16K = 1: FP32 realft also runs fused at n/2 = 16384 (one
// group per CU, 135 KB of LDS); 0: the four-step C2C and an untangle pass.
#ifndef NRFFT_REALFT_16K
#define NRFFT_REALFT_16K 1
#endif
// NRFFT_CONV_16K = 1: FP32 convlv / correl likewise at n/2 = 16384 (the
// full-FFT form: both spectra and the inverse in one group).
#ifndef NRFFT_CONV_16K
#define NRFFT_CONV_16K 1
#endif
// NRFFT_SPEC_RAW = 0: convlv / correl beyond one group untangle both forward
// spectra in passes of their own before the spectrum operation.
#ifndef NRFFT_SPEC_RAW
#define NRFFT_SPEC_RAW 1
#endif

// ---------------------------------------------------------------------------
// Simple kernels of the multi-pass path (n/2 > 8192) and of small sizes.
// ---------------------------------------------------------------------------
// In-place untangle of packed spectra: bin pairs (p, L - p), p < L/2, PAIR_U
// per thread with every load before the first store; p = 0 is the packed
// DC / Nyquist slot.  The self-paired bin L/2 meets the twiddle W^{L/2} = i
// (forward) or -i (inverse), for which the untangle is the identity.
template <typename T, int DIR>
__global__ __launch_bounds__(PAIR_NT) void k_untangle(cpx<T>* data, int64_t L64, int64_t nsig, int logn,
                                                      Twiddles<T> tw) {
    const int L = int(L64), nq = L / 2;
    const RowTw<DIR, T> wr(tw, logn);
    for (int64_t s = blockIdx.y; s < nsig; s += gridDim.y) {
        cpx<T>* d = data + s * L64;
        for (int p0 = pair_first(); p0 < nq; p0 += pair_stride()) {
            cpx<T> a[PAIR_U], b[PAIR_U], w[PAIR_U];
#pragma unroll
            for (int u = 0; u < PAIR_U; ++u) {
                const int p = pair_at(p0, u);
                if (p < nq) {
                    a[u] = d[p];
                    b[u] = d[(L - p) & (L - 1)];
                    w[u] = wr.at(tw, p);
                }
            }
#pragma unroll
            for (int u = 0; u < PAIR_U; ++u) {
                const int p = pair_at(p0, u);
                if (p >= nq) continue;
                if (p == 0) {
                    d[0] = DIR > 0 ? dc_forward(a[u]) : dc_inverse(a[u]);
                } else {
                    cpx<T> x, y;
                    untangle_pair(a[u], b[u], w[u], (T)(-0.5 * DIR), x, y);
                    d[p] = x;
                    d[L - p] = y;
                }
            }
        }
    }
}

// Padded responses (Convolve.rs:41-63) for nsig signals: grid y over
// signals, x over the n slots of one signal.
template <typename T>
__global__ void k_place(const T* r, T* out, int64_t n, int64_t m, int64_t nsig) {
    for (int64_t s = blockIdx.y; s < nsig; s += gridDim.y) {
        const T* rs = r + s * m;
        T* os = out + s * n;
        for (int64_t e = blockIdx.x * int64_t(blockDim.x) + threadIdx.x; e < n; e += int64_t(gridDim.x) * blockDim.x)
            os[e] = placed_response(rs, e, n, m);
    }
}

// a <- inverse-untangle(op(a, b) / no2) on forward-untangled packed spectra:
// bin pairs (p, L - p), p < L/2, PAIR_U per thread with every load first;
// the thread of p = 0 also takes the packed DC / Nyquist slot and the
// self-paired bin L/2.  RAW: a and b are the L-point C2C outputs before the
// forward untangle, which each pair applies first (k_untangle<T, 1>'s
// arithmetic: w = W_n^p, c2 = -1/2; the packed slot dc_forward; bin L/2 is
// its own untangle), so the two forward untangle passes are not run.
template <typename T, bool RAW>
__global__ __launch_bounds__(PAIR_NT) void k_spec(cpx<T>* a, const cpx<T>* b, int64_t L64, int64_t nsig, int op,
                                                  int logn, Twiddles<T> tw) {
    const int L = int(L64), np = L / 2;
    const T inv_no2 = (T)1 / (T)L;
    const RowTw<1, T> wr(tw, logn);
    for (int64_t s = blockIdx.y; s < nsig; s += gridDim.y) {
        cpx<T>* x = a + s * L64;
        const cpx<T>* y = b + s * L64;
        for (int p0 = pair_first(); p0 < np; p0 += pair_stride()) {
            cpx<T> xa[PAIR_U], xb[PAIR_U], ya[PAIR_U], yb[PAIR_U], w[PAIR_U];
#pragma unroll
            for (int u = 0; u < PAIR_U; ++u) {
                const int p = pair_at(p0, u);
                if (p < np) {
                    xa[u] = x[p];
                    xb[u] = x[(L - p) & (L - 1)];
                    ya[u] = y[p];
                    yb[u] = y[(L - p) & (L - 1)];
                    w[u] = wr.at(tw, p);
                }
            }
#pragma unroll
            for (int u = 0; u < PAIR_U; ++u) {
                const int p = pair_at(p0, u);
                if (p >= np) continue;
                if constexpr (RAW) {
                    if (p == 0) {
                        xa[u] = dc_forward(xa[u]);
                        ya[u] = dc_forward(ya[u]);
                    } else {
                        cpx<T> f1, f2;
                        untangle_pair(xa[u], xb[u], w[u], (T)-0.5, f1, f2);
                        xa[u] = f1;
                        xb[u] = f2;
                        untangle_pair(ya[u], yb[u], w[u], (T)-0.5, f1, f2);
                        ya[u] = f1;
                        yb[u] = f2;
                    }
                }
                if (p == 0) {
                    const cpx<T> y0{spec_apply_real(op, xa[u].x, ya[u].x, inv_no2),
                                    spec_apply_real(op, xa[u].y, ya[u].y, inv_no2)};
                    x[0] = dc_inverse(y0);
                    x[np] = spec_apply(op, x[np], y[np], inv_no2);
                } else {
                    const cpx<T> sa = spec_apply(op, xa[u], ya[u], inv_no2);
                    const cpx<T> sb = spec_apply(op, xb[u], yb[u], inv_no2);
                    cpx<T> o1, o2;
                    untangle_pair(sa, sb, conj(w[u]), (T)0.5, o1, o2);
                    x[p] = o1;
                    x[L - p] = o2;
                }
            }
        }
    }
}

// n = 2 realft: [x0, x1] <-> [x0 + x1, x0 - x1] (inverse returns (n/2) x = x).
template <typename T, int DIR>
__global__ void k_real_n2(const cpx<T>* in, cpx<T>* out, int64_t nsig) {
    const int64_t s = blockIdx.x * int64_t(blockDim.x) + threadIdx.x;
    if (s < nsig) out[s] = DIR > 0 ? dc_forward(in[s]) : dc_inverse(in[s]);
}

// correl for n <= 32: linear correlation at lags 0..n-1 (Correlation.rs:37-50).
template <typename T>
__global__ void k_correl_direct(const T* d1, const T* d2, T* ans, int64_t n, int64_t npairs) {
    const int64_t q = blockIdx.x * int64_t(blockDim.x) + threadIdx.x;
    if (q >= n * npairs) return;
    const int64_t s = q / n, lag = q % n;
    const T* a = d1 + s * n;
    const T* b = d2 + s * n;
    T sum = (T)0;
    for (int64_t i = 0; i < n - lag; ++i) sum += a[i + lag] * b[i];
    ans[q] = sum;
}

// Direct circular convolution / deconvolution for n = 2 (n/2 = 1 has no FFT):
// spectrum [x0 + x1, x0 - x1], both bins real.
template <typename T>
__global__ void k_convlv_n2(const T* data, const T* second, T* ans, int64_t m, int64_t nsig, int op) {
    const int64_t s = blockIdx.x * int64_t(blockDim.x) + threadIdx.x;
    if (s >= nsig) return;
    const T* x = data + 2 * s;
    T r0, r1;
    if (op == OP_CORREL) {
        r0 = second[2 * s];
        r1 = second[2 * s + 1];
    } else {
        r0 = placed_response(second + s * m, 0, 2, m);
        r1 = placed_response(second + s * m, 1, 2, m);
    }
    const T f0 = x[0] + x[1], f1 = x[0] - x[1];
    const T g0 = r0 + r1, g1 = r0 - r1;
    const T y0 = spec_apply_real(op, f0, g0, (T)1), y1 = spec_apply_real(op, f1, g1, (T)1);
    ans[2 * s] = (T)0.5 * (y0 + y1);
    ans[2 * s + 1] = (T)0.5 * (y0 - y1);
}

// ---------------------------------------------------------------------------
// Registry of the fused kernels.
// ---------------------------------------------------------------------------
struct RealRef {
    const void* fwd;
    const void* inv;
    const void* conv;
    const void* conv_win;  // windowed-response convlv, or nullptr
    int nt, ti;
};

template <typename T, int LOGL>
RealRef real_ref_one() {
    constexpr int L = 1 << LOGL;
    constexpr int TI = ti_row<T, L>();
    constexpr int EPT = L < 16 ? L : 16;
    constexpr int NT = TI * L / EPT;
    const void* win = nullptr;
    if constexpr (conv_has_window<T, L, TI, NT>()) win = reinterpret_cast<const void*>(&k_convlv<T, L, TI, NT, 1>);
    return RealRef{reinterpret_cast<const void*>(&k_realft<T, L, TI, NT, 1>),
                   reinterpret_cast<const void*>(&k_realft<T, L, TI, NT, -1>),
                   reinterpret_cast<const void*>(&k_convlv<T, L, TI, NT, 0>), win, NT, TI};
}

template <typename T, int... LS>
void real_fill(RealRef* tab, std::integer_sequence<int, LS...>) {
    ((tab[LS + 1] = real_ref_one<T, LS + 1>()), ...);
}

template <typename T>
struct RealTable {
    RealRef t[REAL_MAX_LOGL + 2];  // [REAL_MAX_LOGL + 1]: FP32 realft at n/2 = 16384 only
    RealTable() {
        t[0] = RealRef{nullptr, nullptr, nullptr, nullptr, 0, 0};
        real_fill<T>(t, std::make_integer_sequence<int, REAL_MAX_LOGL>{});
        t[REAL_MAX_LOGL + 1] = RealRef{nullptr, nullptr, nullptr, nullptr, 0, 0};
        if constexpr (sizeof(T) == 4 && NRFFT_REALFT_16K) {
            constexpr int L = 2 << REAL_MAX_LOGL, NT = L / 16;
            const void* conv = nullptr;
            if constexpr (NRFFT_CONV_16K) conv = reinterpret_cast<const void*>(&k_convlv<T, L, 1, NT, 0>);
            t[REAL_MAX_LOGL + 1] = RealRef{reinterpret_cast<const void*>(&k_realft<T, L, 1, NT, 1>),
                                           reinterpret_cast<const void*>(&k_realft<T, L, 1, NT, -1>), conv, nullptr,
                                           NT, 1};
        }
    }
};

static const RealRef& real_lookup(int dtype, int logl) {
    static const RealTable<float> tf;
    static const RealTable<double> td;
    return dtype == NRFFT_F32 ? tf.t[logl] : td.t[logl];
}

// Kinds of LK_REAL launches.
enum RealSpecial {
    RS_FUSED_REALFT = 1,  // args: nsig, L
    RS_FUSED_CONV = 2,    // args: nsig, L, m
    RS_UNTANGLE = 3,      // args: nsig, L, dir, element offset (complex values) into src and dst
    RS_PLACE = 4,         // args: nsig, n, m
    RS_SPEC = 5,          // args: nsig, L, raw (1: both spectra before their forward untangle)
    RS_N2 = 6,            // args: nsig, dir(0 = both directions by isign)
    RS_CORREL_DIRECT = 7, // args: npairs, n
    RS_CONV_N2 = 8,       // args: nsig, m
};

static Launch real_launch(int special, int src, int dst, std::initializer_list<int64_t> args) {
    Launch l;
    l.kind = LK_REAL;
    l.special = special;
    l.src = src;
    l.dst = dst;
    int i = 0;
    for (int64_t a : args) l.args[i++] = a;
    return l;
}

// Forward realft of N = 2^logn complex points per signal as the four-step's
// two passes and the untangle chunk by chunk: each chunk is untangled right
// after its second pass, which stores plainly, so the untangle re-reads the
// chunk from the caches (MI355X, 64 x 2^21 FP32: 0.663 -> 0.621 ms; FP64
// 32 x 2^21 0.614 -> 0.598 ms; with non-temporal second-pass stores the
// chunked form took 0.675 ms).  NRFFT_REALFT_CHUNKED = 0: one untangle
// launch after all chunks; 1: chunked with non-temporal stores.
#ifndef NRFFT_REALFT_CHUNKED
#define NRFFT_REALFT_CHUNKED 2
#endif
static int plan_realft_fwd_chunked(Plan& p, int logn, int64_t batch, size_t* tmp_elems) {
    FourStep fs;
    if (int rc = fourstep_setup(p, logn, &fs)) return rc;
    const int64_t N = int64_t(1) << logn;
    const int64_t chunk = (p.tune.chunk > 0 && p.tune.chunk < batch) ? p.tune.chunk : batch;
    *tmp_elems = std::max<size_t>(*tmp_elems, size_t(N) * size_t(chunk));
    for (int64_t c0 = 0; c0 < batch; c0 += chunk) {
        const int64_t cb = std::min<int64_t>(chunk, batch - c0);
        Launch la, lb;
        if (int rc = fourstep_launches(p, fs, cb, B_IN, B_OUT, &la, &lb)) return rc;
        la.off_src = c0 * N;
        lb.off_dst = c0 * N;
        if (NRFFT_REALFT_CHUNKED == 2) lb.desc.nt_out = 0;
        p.launches.push_back(la);
        p.launches.push_back(lb);
        p.launches.push_back(real_launch(RS_UNTANGLE, B_OUT, B_OUT, {cb, N, 1, c0 * N}));
    }
    return NRFFT_OK;
}

int build_plan_real(Plan& p) {
    const size_t n = p.nn.at(0);
    const int64_t batch = (int64_t)p.batch;
    if (p.kind == PK_CORREL && n <= 32) {
        p.launches.push_back(real_launch(RS_CORREL_DIRECT, B_IN, B_OUT, {batch, (int64_t)n}));
        return NRFFT_OK;
    }
    if (!is_pow2(n) || n < 2) return fail(NRFFT_E_NOT_POW2, "n must be a power of two >= 2");
    const int64_t L = (int64_t)n / 2;
    const int logl = ilog2((size_t)L);
    if (p.kind == PK_REALFT) {
        if (L == 1) {
            p.launches.push_back(real_launch(RS_N2, B_IN, B_OUT, {batch}));
        } else if (logl <= REAL_MAX_LOGL || (logl == REAL_MAX_LOGL + 1 && real_lookup(p.dtype, logl).fwd)) {
            p.launches.push_back(real_launch(RS_FUSED_REALFT, B_IN, B_OUT, {batch, L}));
        } else {
            // Forward: C2C(+1) then untangle; inverse: untangle^-1 then C2C(-1).
            // The direction-dependent order is resolved at execution time: the
            // plan holds both sequences, tagged by args[3] (1 = fwd only, -1 = inv only).
            size_t tmp = 0;
            use_c2c_tune(p, logl);
            Plan fw = p, iv = p;
            fw.launches.clear();
            iv.launches.clear();
            // From 2^FUSED_3PASS_MIN_LOGN the C2C is the three-pass plan and
            // the untangle one whole-batch pass after it.
            const bool chunked = NRFFT_REALFT_CHUNKED && logl < FUSED_3PASS_MIN_LOGN;
            int rc = chunked ? plan_realft_fwd_chunked(fw, logl, batch, &tmp)
                             : plan_c2c_1d(fw, logl, batch, B_IN, B_OUT, &tmp);
            if (rc) return rc;
            if (!chunked) fw.launches.push_back(real_launch(RS_UNTANGLE, B_OUT, B_OUT, {batch, L, 1}));
            if (map_fits(iv, logl, LMAP_RFINV)) {
                // The inverse untangle as the C2C's first-pass load map, chunk
                // by chunk, in place.
                if ((rc = plan_c2c_mapped(iv, logl, batch, B_IN, B_OUT, LMAP_RFINV, 2 * L, 0, -1, 0, &tmp))) return rc;
            } else {
                iv.launches.push_back(real_launch(RS_UNTANGLE, B_IN, B_OUT, {batch, L, -1}));
                rc = plan_c2c_1d(iv, logl, batch, B_OUT, B_OUT, &tmp);
                if (rc) return rc;
            }
            for (auto& l : fw.launches) {
                l.args[7] = 1;
                p.launches.push_back(l);
            }
            for (auto& l : iv.launches) {
                l.args[7] = -1;
                p.launches.push_back(l);
            }
            p.tmp_elems = tmp;
        }
        return NRFFT_OK;
    }
    // convlv / correl
    const int64_t m = (int64_t)p.extra;
    if (L == 1) {
        p.launches.push_back(real_launch(RS_CONV_N2, B_IN, B_OUT, {batch, m}));
        return NRFFT_OK;
    }
    if (logl <= REAL_MAX_LOGL || (logl == REAL_MAX_LOGL + 1 && real_lookup(p.dtype, logl).conv)) {
        p.launches.push_back(real_launch(RS_FUSED_CONV, B_IN, B_OUT, {batch, L, m}));
        return NRFFT_OK;
    }
    // Multi-pass: spectra of the signal (into out) and of the response
    // (into TMP2), spectrum op + inverse untangle, inverse C2C.
    size_t tmp = 0;
    use_c2c_tune(p, logl);
    int rc;
    if (p.kind == PK_CONVLV && map_fits(p, logl, LMAP_PLACE)) {
        // The response's C2C reads the m taps through the placement map:
        // no padded copy of the response is written.
        rc = plan_c2c_mapped(p, logl, batch, B_AUX, B_TMP2, LMAP_PLACE, m, m, 1, 0, &tmp);
    } else if (p.kind == PK_CONVLV) {
        p.launches.push_back(real_launch(RS_PLACE, B_AUX, B_TMP2, {batch, (int64_t)n, m}));
        rc = plan_c2c_1d(p, logl, batch, B_TMP2, B_TMP2, &tmp);
    } else {
        rc = plan_c2c_1d(p, logl, batch, B_AUX, B_TMP2, &tmp);
    }
    if (rc) return rc;
    // The forward untangles of both spectra happen inside the spectrum
    // operation's pairs (RS_SPEC with args[2] = 1), not as passes of their own.
    if (!NRFFT_SPEC_RAW) p.launches.push_back(real_launch(RS_UNTANGLE, B_TMP2, B_TMP2, {batch, L, 1}));
    if ((rc = plan_c2c_1d(p, logl, batch, B_IN, B_OUT, &tmp))) return rc;
    if (!NRFFT_SPEC_RAW) p.launches.push_back(real_launch(RS_UNTANGLE, B_OUT, B_OUT, {batch, L, 1}));
    // Both forward spectra use isign = +1 whatever isign the call carries
    // (for convlv isign selects convolution or deconvolution).
    for (auto& l : p.launches)
        if (l.kind == LK_C2C) l.args[6] = 1;
    p.launches.push_back(real_launch(RS_SPEC, B_OUT, B_OUT, {batch, L, NRFFT_SPEC_RAW ? 1 : 0}));
    // The inverse C2C of the last step runs with isign = -1 whatever isign
    // the call carries (isign selects conv/deconv); args[6] marks it.
    const size_t before = p.launches.size();
    if ((rc = plan_c2c_1d(p, logl, batch, B_OUT, B_OUT, &tmp))) return rc;
    for (size_t i = before; i < p.launches.size(); ++i) p.launches[i].args[6] = -1;
    // TMP2 (response spectrum) follows the C2C scratch in the workspace.
    p.tmp2_offset = std::max<size_t>(tmp, size_t(batch) * size_t(L));
    p.tmp_elems = p.tmp2_offset + size_t(batch) * size_t(L);
    return NRFFT_OK;
}

static unsigned grid_for(int64_t work, int nt) {
    int64_t g = (work + nt - 1) / nt;
    if (g > 65536) g = 65536;
    if (g < 1) g = 1;
    return (unsigned)g;
}

template <typename T>
static int exec_real_t(const Plan& p, const Launch& l, void* const* bufs, int isign, hipStream_t stream,
                       DeviceCtx* ctx) {
    Twiddles<T> tw = ctx_twiddles<T>(ctx);
    void* src = bufs[l.src];
    void* dst = bufs[l.dst];
    const int64_t* a = l.args;
    switch (l.special) {
        case RS_FUSED_REALFT: {
            const RealRef& r = real_lookup(p.dtype, ilog2((size_t)a[1]));
            int64_t nsig = a[0];
            void* args[] = {&src, &dst, &nsig, &tw};
            unsigned grid = (unsigned)((nsig + r.ti - 1) / r.ti);
            HIPCHK(hipLaunchKernel(isign > 0 ? r.fwd : r.inv, dim3(grid), dim3(r.nt), args, 0, stream));
            return NRFFT_OK;
        }
        case RS_FUSED_CONV: {
            const RealRef& r = real_lookup(p.dtype, ilog2((size_t)a[1]));
            int64_t nsig = a[0], m = a[2];
            int op = p.kind == PK_CORREL ? OP_CORREL : (isign > 0 ? OP_CONV : OP_DECONV);
            void* aux = bufs[B_AUX];
            void* args[] = {&src, &aux, &dst, &nsig, &m, &op, &tw};
            unsigned grid = (unsigned)((nsig + r.ti - 1) / r.ti);
            const bool win = r.conv_win != nullptr && op != OP_CORREL && resp_win_fits(a[1], m);
            HIPCHK(hipLaunchKernel(win ? r.conv_win : r.conv, dim3(grid), dim3(r.nt), args, 0, stream));
            return NRFFT_OK;
        }
        case RS_UNTANGLE: {
            src = static_cast<cpx<T>*>(src) + a[3];
            dst = static_cast<cpx<T>*>(dst) + a[3];
            if (src != dst) HIPCHK(hipMemcpyAsync(dst, src, size_t(a[0] * a[1]) * sizeof(cpx<T>),
                                                  hipMemcpyDeviceToDevice, stream));
            const int logn = ilog2((size_t)a[1]) + 1;
            if (a[2] > 0)
                hipLaunchKernelGGL((k_untangle<T, 1>), grid_pairs(a[1] / 2, a[0]), dim3(PAIR_NT), 0, stream,
                                   (cpx<T>*)dst, a[1], a[0], logn, tw);
            else
                hipLaunchKernelGGL((k_untangle<T, -1>), grid_pairs(a[1] / 2, a[0]), dim3(PAIR_NT), 0, stream,
                                   (cpx<T>*)dst, a[1], a[0], logn, tw);
            HIPCHK(hipGetLastError());
            return NRFFT_OK;
        }
        case RS_PLACE:
            hipLaunchKernelGGL((k_place<T>), grid_pairs(a[1], a[0]), dim3(PAIR_NT), 0, stream, (const T*)src, (T*)dst,
                               a[1], a[2], a[0]);
            HIPCHK(hipGetLastError());
            return NRFFT_OK;
        case RS_SPEC: {
            int op = p.kind == PK_CORREL ? OP_CORREL : (isign > 0 ? OP_CONV : OP_DECONV);
            const int logn = ilog2((size_t)a[1]) + 1;
            if (a[2])
                hipLaunchKernelGGL((k_spec<T, true>), grid_pairs(a[1] / 2, a[0]), dim3(PAIR_NT), 0, stream,
                                   (cpx<T>*)dst, (const cpx<T>*)bufs[B_TMP2], a[1], a[0], op, logn, tw);
            else
                hipLaunchKernelGGL((k_spec<T, false>), grid_pairs(a[1] / 2, a[0]), dim3(PAIR_NT), 0, stream,
                                   (cpx<T>*)dst, (const cpx<T>*)bufs[B_TMP2], a[1], a[0], op, logn, tw);
            HIPCHK(hipGetLastError());
            return NRFFT_OK;
        }
        case RS_N2:
            if (isign > 0)
                hipLaunchKernelGGL((k_real_n2<T, 1>), dim3(grid_for(a[0], 256)), dim3(256), 0, stream,
                                   (const cpx<T>*)src, (cpx<T>*)dst, a[0]);
            else
                hipLaunchKernelGGL((k_real_n2<T, -1>), dim3(grid_for(a[0], 256)), dim3(256), 0, stream,
                                   (const cpx<T>*)src, (cpx<T>*)dst, a[0]);
            HIPCHK(hipGetLastError());
            return NRFFT_OK;
        case RS_CORREL_DIRECT:
            hipLaunchKernelGGL((k_correl_direct<T>), dim3(grid_for(a[0] * a[1], 256)), dim3(256), 0, stream,
                               (const T*)src, (const T*)bufs[B_AUX], (T*)dst, a[1], a[0]);
            HIPCHK(hipGetLastError());
            return NRFFT_OK;
        case RS_CONV_N2: {
            int op = p.kind == PK_CORREL ? OP_CORREL : (isign > 0 ? OP_CONV : OP_DECONV);
            hipLaunchKernelGGL((k_convlv_n2<T>), dim3(grid_for(a[0], 256)), dim3(256), 0, stream, (const T*)src,
                               (const T*)bufs[B_AUX], (T*)dst, a[1], a[0], op);
            HIPCHK(hipGetLastError());
            return NRFFT_OK;
        }
        default:
            return fail(NRFFT_E_INVALID, "unknown launch");
    }
}

int exec_special(const Plan& p, const Launch& l, void* const* bufs, int isign, hipStream_t stream, DeviceCtx* ctx) {
    if (p.dtype == NRFFT_F32) return exec_real_t<float>(p, l, bufs, isign, stream, ctx);
    return exec_real_t<double>(p, l, bufs, isign, stream, ctx);
}

// ---------------------------------------------------------------------------
// Host-pointer calls.
// ---------------------------------------------------------------------------
int host_real(int kind, int dtype, const void* in, const void* aux, size_t aux_bytes, void* out, size_t n, size_t m,
              size_t batch, int isign) {
    std::shared_ptr<Plan> p;
    int rc = cached_plan(kind, dtype, {n}, batch, m, &p);
    if (rc) return rc;
    const size_t es = dtype == NRFFT_F32 ? 4 : 8;
    const size_t bytes = n * batch * es;
    hipStream_t s = hipStreamPerThread;
    DevBuf dbuf, dout, daux, dtmp;
    if ((rc = dev_alloc(dbuf, bytes, s))) return rc;
    if ((rc = dev_alloc(dout, bytes, s))) return rc;
    if (aux_bytes && (rc = dev_alloc(daux, aux_bytes, s))) return rc;
    if ((rc = dev_alloc(dtmp, p->tmp_elems * csize(dtype), s))) return rc;
    HIPCHK(hipMemcpyAsync(dbuf.p, in, bytes, hipMemcpyHostToDevice, s));
    if (aux_bytes) HIPCHK(hipMemcpyAsync(daux.p, aux, aux_bytes, hipMemcpyHostToDevice, s));
    if ((rc = exec_plan(*p, dbuf.p, dout.p, dtmp.p, daux.p, isign, s))) return rc;
    HIPCHK(hipMemcpyAsync(out, dout.p, bytes, hipMemcpyDeviceToHost, s));
    HIPCHK(hipStreamSynchronize(s));
    return NRFFT_OK;
}

}  // namespace nrfft

using namespace nrfft;

static int check_realft(const void* data, size_t n, int isign) {
    if (isign != 1 && isign != -1) return fail(NRFFT_E_ISIGN, "isign must be 1 or -1");
    if (n == 0) return fail(NRFFT_E_EMPTY, "n == 0");
    if (!data) return fail(NRFFT_E_INVALID, "null data");
    // Real_FT.rs:5 asserts n % 2 == 0; four1 underneath needs n/2 a power of two.
    if (n % 2 != 0 || !is_pow2(n)) return fail(NRFFT_E_NOT_POW2, "n must be an even power of two");
    return NRFFT_OK;
}

// Convolve.rs:13-21 order of checks, then the power-of-two requirement.
static int check_convlv(const void* data, size_t n, const void* respns, size_t m, int isign, const void* ans) {
    if (n == 0 || m == 0) return fail(NRFFT_E_EMPTY, "Input arrays cannot be empty");
    if (m > n) return fail(NRFFT_E_RESP_TOO_LONG, "Response function longer than data");
    if (isign != 1 && isign != -1) return fail(NRFFT_E_ISIGN, "isign must be 1 (convolution) or -1 (deconvolution)");
    if (!data || !respns || !ans) return fail(NRFFT_E_INVALID, "null pointer");
    if (!is_pow2(n) || n < 2) return fail(NRFFT_E_NOT_POW2, "n must be a power of two >= 2");
    return NRFFT_OK;
}

// Correlation.rs:10-16; n > 32 must be a power of two.
static int check_correl(const void* d1, const void* d2, size_t n, const void* ans) {
    if (n == 0) return fail(NRFFT_E_EMPTY, "Input arrays cannot be empty");
    if (!d1 || !d2 || !ans) return fail(NRFFT_E_INVALID, "null pointer");
    if (n > 32 && !is_pow2(n)) return fail(NRFFT_E_NOT_POW2, "n > 32 must be a power of two");
    return NRFFT_OK;
}

extern "C" {

int nrfft_realft_f64(double* data, size_t n, int isign) {
    g_err.clear();
    int rc = check_realft(data, n, isign);
    return rc ? rc : host_real(PK_REALFT, NRFFT_F64, data, nullptr, 0, data, n, 0, 1, isign);
}
int nrfft_realft_f32(float* data, size_t n, int isign) {
    g_err.clear();
    int rc = check_realft(data, n, isign);
    return rc ? rc : host_real(PK_REALFT, NRFFT_F32, data, nullptr, 0, data, n, 0, 1, isign);
}

}  // extern "C"

namespace nrfft {
// RealFTProcessor::process_batch (Real_FT.rs:365-369) for `count` host
// buffers of n reals each: gathered into one contiguous batch, transformed
// by one batched plan execution, scattered back.
template <typename T>
static int realft_batch(T* const* data, size_t n, size_t count, int isign, int dtype) {
    g_err.clear();
    if (count == 0) return NRFFT_OK;
    if (!data) return fail(NRFFT_E_INVALID, "null buffer list");
    for (size_t i = 0; i < count; ++i)
        if (int rc = check_realft(data[i], n, isign)) return rc;
    if (count == 1) return host_real(PK_REALFT, dtype, data[0], nullptr, 0, data[0], n, 0, 1, isign);
    std::vector<T> buf;
    try {
        buf.resize(n * count);
    } catch (const std::bad_alloc&) {
        return fail(NRFFT_E_NOMEM, "host staging");
    }
    for (size_t i = 0; i < count; ++i) std::memcpy(buf.data() + i * n, data[i], n * sizeof(T));
    if (int rc = host_real(PK_REALFT, dtype, buf.data(), nullptr, 0, buf.data(), n, 0, count, isign)) return rc;
    for (size_t i = 0; i < count; ++i) std::memcpy(data[i], buf.data() + i * n, n * sizeof(T));
    return NRFFT_OK;
}
}  // namespace nrfft

extern "C" {

int nrfft_realft_batch_f64(double* const* data, size_t n, size_t count, int isign) {
    return realft_batch<double>(data, n, count, isign, NRFFT_F64);
}
int nrfft_realft_batch_f32(float* const* data, size_t n, size_t count, int isign) {
    return realft_batch<float>(data, n, count, isign, NRFFT_F32);
}

int nrfft_convlv_f64(const double* data, size_t n, const double* respns, size_t m, int isign, double* ans) {
    g_err.clear();
    int rc = check_convlv(data, n, respns, m, isign, ans);
    return rc ? rc : host_real(PK_CONVLV, NRFFT_F64, data, respns, m * 8, ans, n, m, 1, isign);
}
int nrfft_convlv_f32(const float* data, size_t n, const float* respns, size_t m, int isign, float* ans) {
    g_err.clear();
    int rc = check_convlv(data, n, respns, m, isign, ans);
    return rc ? rc : host_real(PK_CONVLV, NRFFT_F32, data, respns, m * 4, ans, n, m, 1, isign);
}
int nrfft_convlv_batch_f64(const double* data, size_t n, const double* respns, size_t m, size_t batch, int isign,
                           double* ans) {
    g_err.clear();
    int rc = check_convlv(data, n, respns, m, isign, ans);
    if (rc || batch == 0) return rc;
    return host_real(PK_CONVLV, NRFFT_F64, data, respns, m * batch * 8, ans, n, m, batch, isign);
}
int nrfft_convlv_batch_f32(const float* data, size_t n, const float* respns, size_t m, size_t batch, int isign,
                           float* ans) {
    g_err.clear();
    int rc = check_convlv(data, n, respns, m, isign, ans);
    if (rc || batch == 0) return rc;
    return host_real(PK_CONVLV, NRFFT_F32, data, respns, m * batch * 4, ans, n, m, batch, isign);
}

int nrfft_correl_f64(const double* d1, const double* d2, size_t n, double* ans) {
    g_err.clear();
    int rc = check_correl(d1, d2, n, ans);
    return rc ? rc : host_real(PK_CORREL, NRFFT_F64, d1, d2, n * 8, ans, n, 0, 1, 1);
}
int nrfft_correl_f32(const float* d1, const float* d2, size_t n, float* ans) {
    g_err.clear();
    int rc = check_correl(d1, d2, n, ans);
    return rc ? rc : host_real(PK_CORREL, NRFFT_F32, d1, d2, n * 4, ans, n, 0, 1, 1);
}

int nrfft_plan_realft(nrfft_plan* plan, int dtype, size_t n, size_t batch) {
    g_err.clear();
    int rc = check_realft(plan, n, 1);
    return rc ? rc : make_plan(plan, PK_REALFT, dtype, {n}, batch, 0);
}
int nrfft_execute_realft(nrfft_plan plan, void* d_data, int isign, void* stream) {
    g_err.clear();
    if (!plan || plan->plan.kind != PK_REALFT) return fail(NRFFT_E_INVALID, "not a realft plan");
    if (isign != 1 && isign != -1) return fail(NRFFT_E_ISIGN, "isign must be 1 or -1");
    if (!d_data) return fail(NRFFT_E_INVALID, "null buffer");
    return exec_plan(plan->plan, d_data, d_data, plan->tmp, nullptr, isign, (hipStream_t)stream);
}

int nrfft_plan_convlv(nrfft_plan* plan, int dtype, size_t n, size_t m, size_t batch) {
    g_err.clear();
    int rc = check_convlv(plan, n, plan, m, 1, plan);
    return rc ? rc : make_plan(plan, PK_CONVLV, dtype, {n}, batch, m);
}
int nrfft_execute_convlv(nrfft_plan plan, const void* d_data, const void* d_respns, void* d_ans, int isign,
                         void* stream) {
    g_err.clear();
    if (!plan || plan->plan.kind != PK_CONVLV) return fail(NRFFT_E_INVALID, "not a convlv plan");
    if (isign != 1 && isign != -1) return fail(NRFFT_E_ISIGN, "isign must be 1 (convolution) or -1 (deconvolution)");
    if (!d_data || !d_respns || !d_ans) return fail(NRFFT_E_INVALID, "null buffer");
    return exec_plan(plan->plan, d_data, d_ans, plan->tmp, d_respns, isign, (hipStream_t)stream);
}

int nrfft_plan_correl(nrfft_plan* plan, int dtype, size_t n, size_t batch) {
    g_err.clear();
    int rc = check_correl(plan, plan, n, plan);
    return rc ? rc : make_plan(plan, PK_CORREL, dtype, {n}, batch, 0);
}
int nrfft_execute_correl(nrfft_plan plan, const void* d1, const void* d2, void* d_ans, void* stream) {
    g_err.clear();
    if (!plan || plan->plan.kind != PK_CORREL) return fail(NRFFT_E_INVALID, "not a correl plan");
    if (!d1 || !d2 || !d_ans) return fail(NRFFT_E_INVALID, "null buffer");
    return exec_plan(plan->plan, d1, d_ans, plan->tmp, d2, 1, (hipStream_t)stream);
}

}  // extern "C"
