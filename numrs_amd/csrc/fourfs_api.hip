// fourfs_api.hip -- out-of-core n-D FFT (numrs Four_FS.rs, NR fourfs): an
// in-place transform of a host array of any size (an mmap-ed file works)
// streamed through HBM with at most `device_bytes` of device memory.
//
//   last axis (contiguous rows)  chunks of whole rows -> batched C2C plan
//   other axes (stride inner)    column bands of w columns gathered by the
//                                CPU into a pinned staging buffer, one DMA to
//                                [L][w], an axis-0 pass (column mapping,
//                                element stride w), scattered back the same
//                                way; whole slabs are batched when they fit
//   1-D longer than the budget   four-step N = N1 N2 over the host array:
//                                column bands of [N1][N2] + the W_N^{k1 n2}
//                                twiddle, then chunks of rows transformed and
//                                transposed on the device into `scratch` at
//                                natural positions, then copied back
// Results equal fourn's (C order, unnormalised, kernel exp(+isign 2 pi i ..)).
#include <cstring>
#include <string>
#include <vector>

#include "../../include/nrfft.h"
#include "internal.hpp"

namespace nrfft {

#define HIPCHK(x)                                                                                      \
    do {                                                                                               \
        hipError_t e_ = (x);                                                                           \
        if (e_ != hipSuccess) return fail(NRFFT_E_HIP, std::string(#x) + ": " + hipGetErrorString(e_)); \
    } while (0)

int col_tilog(int dtype);

// x[k1][c] *= W_N^{isign k1 (c0 + c)} for a band [n1][w] of [N1][N2].
template <typename T, int S>
__global__ void k_fs_twiddle(cpx<T>* x, int64_t n1, int64_t w, int64_t c0, int logn, Twiddles<T> tw) {
    const uint32_t mask = uint32_t((uint64_t(1) << logn) - 1);
    for (int64_t q = blockIdx.x * int64_t(blockDim.x) + threadIdx.x; q < n1 * w; q += int64_t(gridDim.x) * blockDim.x) {
        const int64_t k1 = q / w, c = q % w;
        const uint32_t m = uint32_t((uint64_t(k1) * uint64_t(c0 + c)) & mask);
        x[q] = cmul(x[q], twiddle_big<S>(tw, m, logn));
    }
}

// out[c][r] = in[r][c] for an [rows][cols] array (32 x 32 tiles through LDS).
template <typename T>
__global__ void k_fs_transpose(const cpx<T>* in, cpx<T>* out, int64_t rows, int64_t cols) {
    __shared__ cpx<T> tile[32][33];
    const int64_t tc = int64_t(blockIdx.x) * 32, tr = int64_t(blockIdx.y) * 32;
    const int tx = threadIdx.x, ty = threadIdx.y;  // 32 x 8
    for (int k = ty; k < 32; k += 8) {
        const int64_t r = tr + k, c = tc + tx;
        if (r < rows && c < cols) tile[k][tx] = in[r * cols + c];
    }
    lds_barrier();
    for (int k = ty; k < 32; k += 8) {
        const int64_t c = tc + k, r = tr + tx;
        if (r < rows && c < cols) out[c * rows + r] = tile[tx][k];
    }
}

// One transform pass along the first axis of [L][w] (x slabs of L*w apart), in place.
static int plan_axis0(Plan& p, int64_t L, int64_t w, int64_t slabs) {
    PassDesc d{};
    d.nb0 = w;
    d.si0 = 1;
    d.so0 = 1;
    d.esi = (int32_t)w;
    d.eso = (int32_t)w;
    d.nb1 = slabs;
    d.si1 = L * w;
    d.so1 = L * w;
    d.nb2 = 1;
    const int logl = ilog2((size_t)L);
    const int t = pick_tilog(p.dtype, MODE_TT, logl, w, col_tilog(p.dtype));
    if (t < 0) return fail(NRFFT_E_UNSUPPORTED, "no column kernel");
    return plan_add_c2c_pass(p, MODE_TT, logl, t, d, B_IN, B_OUT);
}

struct FsCtx {
    int dtype;
    size_t es;  // bytes per complex value
    hipStream_t s;
    DevBuf buf, ws;
    char* pinned = nullptr;  // staging for strided gathers / scatters, allocated on first use
    size_t pinned_bytes = 0;
    ~FsCtx() {
        if (pinned) (void)hipHostFree(pinned);
    }
};

static int ensure_pinned(FsCtx& c, size_t bytes) {
    if (bytes <= c.pinned_bytes) return NRFFT_OK;
    if (c.pinned) (void)hipHostFree(c.pinned);
    c.pinned = nullptr;
    c.pinned_bytes = 0;
    if (hipHostMalloc((void**)&c.pinned, bytes, hipHostMallocDefault) != hipSuccess) {
        c.pinned = nullptr;
        return fail(NRFFT_E_NOMEM, "pinned staging buffer");
    }
    c.pinned_bytes = bytes;
    return NRFFT_OK;
}

// Strided band [rows][w] of a host array with row pitch `pitch` (complex
// values) <-> the device buffer, through the pinned staging buffer: the CPU
// gathers / scatters rows, one DMA moves the band.
static int band_in(FsCtx& c, const char* src, int64_t pitch, int64_t w, int64_t rows) {
    const size_t rb = size_t(w) * c.es;
    if (int rc = ensure_pinned(c, size_t(rows) * rb)) return rc;
    for (int64_t r = 0; r < rows; ++r) std::memcpy(c.pinned + size_t(r) * rb, src + size_t(r * pitch) * c.es, rb);
    HIPCHK(hipMemcpyAsync(c.buf.p, c.pinned, size_t(rows) * rb, hipMemcpyHostToDevice, c.s));
    HIPCHK(hipStreamSynchronize(c.s));  // staging free for the next band
    return NRFFT_OK;
}
static int band_out(FsCtx& c, const void* dsrc, char* dst, int64_t pitch, int64_t w, int64_t rows) {
    const size_t rb = size_t(w) * c.es;
    if (int rc = ensure_pinned(c, size_t(rows) * rb)) return rc;
    HIPCHK(hipMemcpyAsync(c.pinned, dsrc, size_t(rows) * rb, hipMemcpyDeviceToHost, c.s));
    HIPCHK(hipStreamSynchronize(c.s));
    for (int64_t r = 0; r < rows; ++r) std::memcpy(dst + size_t(r * pitch) * c.es, c.pinned + size_t(r) * rb, rb);
    return NRFFT_OK;
}

static Plan fs_plan(int dtype) {
    Plan p;
    p.kind = PK_C2C;
    p.dtype = dtype;
    p.tune = Tune{};
    return p;
}

// Batched contiguous transforms of length L over `rows` rows of host data.
static int fs_rows(FsCtx& c, char* host, int64_t L, int64_t rows, int64_t max_rows, int isign) {
    for (int64_t r0 = 0; r0 < rows; r0 += max_rows) {
        const int64_t b = std::min(max_rows, rows - r0);
        Plan p = fs_plan(c.dtype);
        size_t tmp = 0;
        int rc = plan_c2c_1d(p, ilog2((size_t)L), b, B_IN, B_OUT, &tmp);
        if (rc) return rc;
        const size_t bytes = size_t(b * L) * c.es;
        HIPCHK(hipMemcpyAsync(c.buf.p, host + size_t(r0 * L) * c.es, bytes, hipMemcpyHostToDevice, c.s));
        if ((rc = exec_plan(p, c.buf.p, c.buf.p, c.ws.p, nullptr, isign, c.s))) return rc;
        HIPCHK(hipMemcpyAsync(host + size_t(r0 * L) * c.es, c.buf.p, bytes, hipMemcpyDeviceToHost, c.s));
    }
    return NRFFT_OK;
}

// Transforms along an axis of length L <= 8192 with element stride `inner`,
// for `outer` slabs of L * inner values; optional four-step twiddle (1-D
// split, logn > 0) on the band [L][w] after the transform.
static int fs_strided(FsCtx& c, char* host, int64_t L, int64_t inner, int64_t outer, size_t budget, int isign,
                      int tw_logn) {
    const int64_t slab = L * inner;
    const int64_t cap = int64_t(budget / c.es);  // complex values the device buffer holds
    DeviceCtx* ctx = nullptr;
    int rc = get_ctx(&ctx);
    if (rc) return rc;
    if (slab <= cap && tw_logn == 0) {
        const int64_t k = std::max<int64_t>(1, cap / slab);
        for (int64_t o0 = 0; o0 < outer; o0 += k) {
            const int64_t ks = std::min(k, outer - o0);
            Plan p = fs_plan(c.dtype);
            if ((rc = plan_axis0(p, L, inner, ks))) return rc;
            const size_t bytes = size_t(ks * slab) * c.es;
            HIPCHK(hipMemcpyAsync(c.buf.p, host + size_t(o0 * slab) * c.es, bytes, hipMemcpyHostToDevice, c.s));
            if ((rc = exec_plan(p, c.buf.p, c.buf.p, nullptr, nullptr, isign, c.s))) return rc;
            HIPCHK(hipMemcpyAsync(host + size_t(o0 * slab) * c.es, c.buf.p, bytes, hipMemcpyDeviceToHost, c.s));
        }
        return NRFFT_OK;
    }
    // column bands of w (a power of two) columns
    int64_t w = 1;
    while (w * 2 <= inner && L * w * 2 <= cap) w *= 2;
    if (L * w > cap) return fail(NRFFT_E_NOMEM, "device budget below one column of the transform");
    for (int64_t o = 0; o < outer; ++o) {
        char* sb = host + size_t(o * slab) * c.es;
        for (int64_t c0 = 0; c0 < inner; c0 += w) {
            const int64_t wb = std::min(w, inner - c0);
            Plan p = fs_plan(c.dtype);
            if ((rc = plan_axis0(p, L, wb, 1))) return rc;
            if ((rc = band_in(c, sb + size_t(c0) * c.es, inner, wb, L))) return rc;
            if ((rc = exec_plan(p, c.buf.p, c.buf.p, nullptr, nullptr, isign, c.s))) return rc;
            if (tw_logn) {
                const unsigned g = (unsigned)std::min<int64_t>(65536, (L * wb + 255) / 256);
                if (c.dtype == NRFFT_F32) {
                    Twiddles<float> tw = ctx_twiddles<float>(ctx);
                    if (isign > 0)
                        hipLaunchKernelGGL((k_fs_twiddle<float, 1>), dim3(g), dim3(256), 0, c.s, (cpx<float>*)c.buf.p,
                                           L, wb, c0, tw_logn, tw);
                    else
                        hipLaunchKernelGGL((k_fs_twiddle<float, -1>), dim3(g), dim3(256), 0, c.s,
                                           (cpx<float>*)c.buf.p, L, wb, c0, tw_logn, tw);
                } else {
                    Twiddles<double> tw = ctx_twiddles<double>(ctx);
                    if (isign > 0)
                        hipLaunchKernelGGL((k_fs_twiddle<double, 1>), dim3(g), dim3(256), 0, c.s,
                                           (cpx<double>*)c.buf.p, L, wb, c0, tw_logn, tw);
                    else
                        hipLaunchKernelGGL((k_fs_twiddle<double, -1>), dim3(g), dim3(256), 0, c.s,
                                           (cpx<double>*)c.buf.p, L, wb, c0, tw_logn, tw);
                }
                HIPCHK(hipGetLastError());
            }
            if ((rc = band_out(c, c.buf.p, sb + size_t(c0) * c.es, inner, wb, L))) return rc;
        }
    }
    return NRFFT_OK;
}

static int fourfs_host(int dtype, void* data, void* scratch, const size_t* nn, size_t ndim, int isign,
                       size_t device_bytes) {
    // Four_FS.rs:328-339 order: dimensions, isign, sizes; then powers of two.
    if (!nn || ndim == 0) return fail(NRFFT_E_DIMS, "Invalid dimensions");
    if (isign != 1 && isign != -1) return fail(NRFFT_E_ISIGN, "isign must be 1 or -1");
    for (size_t i = 0; i < ndim; ++i)
        if (nn[i] <= 1) return fail(NRFFT_E_DIMS, "Invalid dimension size");
    for (size_t i = 0; i < ndim; ++i)
        if (!is_pow2(nn[i])) return fail(NRFFT_E_NOT_POW2, "dimensions must be powers of two");
    if (!data) return fail(NRFFT_E_INVALID, "null data");
    FsCtx c;
    c.dtype = dtype;
    c.es = csize(dtype);
    c.s = hipStreamPerThread;
    const size_t budget = device_bytes ? device_bytes : (size_t(1) << 30);
    // half for the data chunk, half for a four-step plan's workspace
    const size_t half = budget / 2;
    if (half < 64 * c.es) return fail(NRFFT_E_NOMEM, "device budget too small");
    int rc;
    if ((rc = dev_alloc(c.buf, half, c.s))) return rc;
    if ((rc = dev_alloc(c.ws, half, c.s))) return rc;
    char* host = static_cast<char*>(data);
    int64_t total = 1;
    for (size_t i = 0; i < ndim; ++i) total *= (int64_t)nn[i];
    const int64_t cap = int64_t(half / c.es);
    if (ndim == 1 && total > cap) {
        // 1-D four-step over the host array
        const int logn = ilog2((size_t)total);
        if (logn > 26) return fail(NRFFT_E_UNSUPPORTED, "1-D out-of-core transforms up to 2^26 points");
        if (!scratch) return fail(NRFFT_E_INVALID, "a 1-D transform larger than the device budget needs a scratch buffer");
        const int l1 = std::min(C2C_MAX_LOGL_, logn / 2);
        const int64_t N1 = int64_t(1) << l1, N2 = total >> l1;
        // a row chunk and its transposed copy share the chunk buffer
        if (2 * N2 > cap) return fail(NRFFT_E_NOMEM, "device budget below two rows of the four-step split");
        // columns of [N1][N2] with the twiddle
        if ((rc = fs_strided(c, host, N1, N2, 1, half, isign, logn))) return rc;
        // rows: transform chunks, transpose on the device, store at natural positions in scratch
        const int64_t rows_cap = std::max<int64_t>(1, cap / (2 * N2));
        char* sc = static_cast<char*>(scratch);
        for (int64_t r0 = 0; r0 < N1; r0 += rows_cap) {
            const int64_t b = std::min(rows_cap, N1 - r0);
            Plan p = fs_plan(dtype);
            size_t tmp = 0;
            if ((rc = plan_c2c_1d(p, ilog2((size_t)N2), b, B_IN, B_OUT, &tmp))) return rc;
            if (tmp * c.es > half) return fail(NRFFT_E_NOMEM, "device budget below the row plan's workspace");
            const size_t bytes = size_t(b * N2) * c.es;
            char* tbuf = static_cast<char*>(c.buf.p) + bytes;  // second half of the chunk buffer
            HIPCHK(hipMemcpyAsync(c.buf.p, host + size_t(r0 * N2) * c.es, bytes, hipMemcpyHostToDevice, c.s));
            if ((rc = exec_plan(p, c.buf.p, c.buf.p, c.ws.p, nullptr, isign, c.s))) return rc;
            dim3 grid((unsigned)((N2 + 31) / 32), (unsigned)((b + 31) / 32));
            if (dtype == NRFFT_F32)
                hipLaunchKernelGGL((k_fs_transpose<float>), grid, dim3(32, 8), 0, c.s, (const cpx<float>*)c.buf.p,
                                   (cpx<float>*)tbuf, b, N2);
            else
                hipLaunchKernelGGL((k_fs_transpose<double>), grid, dim3(32, 8), 0, c.s, (const cpx<double>*)c.buf.p,
                                   (cpx<double>*)tbuf, b, N2);
            HIPCHK(hipGetLastError());
            // tbuf is [N2][b]: row k2 goes to scratch[k2 * N1 + r0 .. + b]
            if ((rc = band_out(c, tbuf, sc + size_t(r0) * c.es, N1, b, N2))) return rc;
        }
        HIPCHK(hipStreamSynchronize(c.s));
        std::memcpy(host, scratch, size_t(total) * c.es);
        return NRFFT_OK;
    }
    // axis by axis, last (contiguous) first
    for (size_t dd = ndim; dd-- > 0;) {
        const int64_t L = (int64_t)nn[dd];
        int64_t inner = 1;
        for (size_t e = dd + 1; e < ndim; ++e) inner *= (int64_t)nn[e];
        const int64_t outer = total / (L * inner);
        if (inner == 1) {
            if (L > cap) return fail(NRFFT_E_NOMEM, "device budget below one row of the last axis");
            if ((rc = fs_rows(c, host, L, outer, std::max<int64_t>(1, cap / L), isign))) return rc;
        } else {
            if (L > (int64_t(1) << C2C_MAX_LOGL_))
                return fail(NRFFT_E_UNSUPPORTED, "out-of-core strided axes up to 8192 points");
            if ((rc = fs_strided(c, host, L, inner, outer, half, isign, 0))) return rc;
        }
    }
    HIPCHK(hipStreamSynchronize(c.s));
    return NRFFT_OK;
}

}  // namespace nrfft

using namespace nrfft;

extern "C" {

int nrfft_fourfs_f64(double* data, double* scratch, const size_t* nn, size_t ndim, int isign, size_t device_bytes) {
    g_err.clear();
    return fourfs_host(NRFFT_F64, data, scratch, nn, ndim, isign, device_bytes);
}
int nrfft_fourfs_f32(float* data, float* scratch, const size_t* nn, size_t ndim, int isign, size_t device_bytes) {
    g_err.clear();
    return fourfs_host(NRFFT_F32, data, scratch, nn, ndim, isign, device_bytes);
}

}  // extern "C"
