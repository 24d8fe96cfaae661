// fourfs_api.hip -- out-of-core n-D FFT (numrs Four_FS.rs, NR fourfs): an
// in-place transform of a host array of any size (an mmap-ed file works)
// streamed through HBM with at most `device_bytes` of device memory.
//
//   fits the budget              one upload, the in-memory fourn plan, one
//                                download
//   last axis (contiguous rows)  chunks of whole rows -> batched C2C plan
//   other axes (stride inner)    column bands of w columns gathered by the
//                                CPU into pinned staging, one DMA to [L][w],
//                                an axis-0 pass (column mapping, element
//                                stride w), scattered back the same way;
//                                whole slabs are batched when they fit
//   1-D beyond the budget        four-step N = N1 N2 over the host array:
//                                column bands of [N1][N2] + the W_N^{k1 n2}
//                                twiddle into `scratch`, then chunks of its
//                                rows transformed and transposed on the
//                                device, stored at natural positions
// Every phase streams its chunks through the two-slot pipeline of
// host_pipe.hip, so host copies, both PCIe directions and the transforms
// overlap.  Results equal fourn's (C order, unnormalised, kernel
// exp(+isign 2 pi i ..)).
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

static Plan fs_plan(int dtype) {
    Plan p;
    p.kind = PK_C2C;
    p.dtype = dtype;
    p.tune = Tune{};
    return p;
}

constexpr size_t FS_SLOT_MAX = size_t(64) << 20;  // pinned / device bytes per pipeline slot

struct FsCtx {
    int dtype;
    size_t es;       // bytes per complex value
    size_t budget;   // device bytes the call may hold
};

// Slots for chunks of whole units of `unit` bytes, with a workspace as large
// as a slot when `ws`: two slots of a quarter of the budget (at most
// FS_SLOT_MAX) when a unit fits, else one slot as large as the budget allows.
static int fs_slots(const FsCtx& c, size_t unit, bool ws, const char* what, PipeJob* job) {
    const size_t w = ws ? 1 : 0;
    size_t s = std::max(unit, std::min(c.budget / 4, FS_SLOT_MAX));
    if ((2 + w) * s <= c.budget) {
        job->slots = 2;
    } else {
        s = std::max(unit, std::min(c.budget / 2, FS_SLOT_MAX));
        if ((1 + w) * s > c.budget) return fail(NRFFT_E_NOMEM, what);
        job->slots = 1;
    }
    job->slot_bytes = s;
    job->ws_bytes = w * s;
    return NRFFT_OK;
}

template <typename T>
static void launch_fs_twiddle(hipStream_t s, DeviceCtx* ctx, void* x, int64_t L, int64_t w, int64_t c0, int logn,
                              int isign) {
    const unsigned g = (unsigned)std::min<int64_t>(65536, (L * w + 255) / 256);
    const Twiddles<T> tw = ctx_twiddles<T>(ctx);
    if (isign > 0)
        hipLaunchKernelGGL((k_fs_twiddle<T, 1>), dim3(g), dim3(256), 0, s, (cpx<T>*)x, L, w, c0, logn, tw);
    else
        hipLaunchKernelGGL((k_fs_twiddle<T, -1>), dim3(g), dim3(256), 0, s, (cpx<T>*)x, L, w, c0, logn, tw);
}

// Batched contiguous transforms of length L over `rows` rows of host data.
static int fs_rows(const FsCtx& c, char* host, int64_t L, int64_t rows, int isign) {
    const size_t rb = size_t(L) * c.es;
    PipeJob job;
    if (int rc = fs_slots(c, rb, L > (int64_t(1) << C2C_MAX_LOGL_), "device budget below one row of the last axis", &job))
        return rc;
    const int64_t per = std::min<int64_t>(rows, int64_t(job.slot_bytes / rb));
    job.chunks = size_t((rows + per - 1) / per);
    const int64_t tail = rows - int64_t(job.chunks - 1) * per;
    Plan pf = fs_plan(c.dtype), pt = fs_plan(c.dtype);
    size_t tf = 0, tt = 0;
    int rc = plan_c2c_1d(pf, ilog2((size_t)L), per, B_IN, B_OUT, &tf);
    if (!rc && tail != per) rc = plan_c2c_1d(pt, ilog2((size_t)L), tail, B_IN, B_OUT, &tt);
    if (rc) return rc;
    if (std::max(tf, tt) * c.es > job.ws_bytes) return fail(NRFFT_E_NOMEM, "device budget below the row plan's workspace");
    auto nb = [&](size_t i) { return size_t(i + 1 < job.chunks ? per : tail) * rb; };
    job.gather = [&](size_t i, char* pin) {
        par_copy2d(pin, 0, host + i * size_t(per) * rb, 0, nb(i), 1);
        return nb(i);
    };
    job.compute = [&](size_t i, char* d, void* ws, hipStream_t s, const char** o, size_t* ob) -> int {
        *o = d;
        *ob = nb(i);
        return exec_plan(i + 1 < job.chunks || tail == per ? pf : pt, d, d, ws, nullptr, isign, s);
    };
    job.scatter = [&](size_t i, const char* pin, size_t bytes) { par_copy2d(host + i * size_t(per) * rb, 0, pin, 0, bytes, 1); };
    return run_pipe(job);
}

// Transforms along an axis of length L <= 8192 with element stride `inner`,
// for `outer` slabs of L * inner values read from `src` and stored at the
// same positions in `dst` (src == dst: in place); optional four-step twiddle
// (1-D split, logn > 0) on the band [L][w] after the transform.  Whole slabs
// go in contiguous chunks when they fit a slot, else bands of w columns
// gathered by rows of w values.
static int fs_strided(const FsCtx& c, const char* src, char* dst, int64_t L, int64_t inner, int64_t outer, int isign,
                      int tw_logn) {
    const int64_t slab = L * inner;
    const size_t sb = size_t(slab) * c.es;
    DeviceCtx* ctx = nullptr;
    int rc = get_ctx(&ctx);
    if (rc) return rc;
    PipeJob job;
    if (tw_logn == 0 && !fs_slots(c, sb, false, "", &job) && job.slots == 2) {
        const int64_t k = std::min<int64_t>(outer, int64_t(job.slot_bytes / sb));
        job.chunks = size_t((outer + k - 1) / k);
        const int64_t tail = outer - int64_t(job.chunks - 1) * k;
        Plan pf = fs_plan(c.dtype), pt = fs_plan(c.dtype);
        if ((rc = plan_axis0(pf, L, inner, k)) || (tail != k && (rc = plan_axis0(pt, L, inner, tail)))) return rc;
        auto nb = [&](size_t i) { return size_t(i + 1 < job.chunks ? k : tail) * sb; };
        job.gather = [&](size_t i, char* pin) {
            par_copy2d(pin, 0, src + i * size_t(k) * sb, 0, nb(i), 1);
            return nb(i);
        };
        job.compute = [&](size_t i, char* d, void*, hipStream_t s, const char** o, size_t* ob) -> int {
            *o = d;
            *ob = nb(i);
            return exec_plan(i + 1 < job.chunks || tail == k ? pf : pt, d, d, nullptr, nullptr, isign, s);
        };
        job.scatter = [&](size_t i, const char* pin, size_t bytes) { par_copy2d(dst + i * size_t(k) * sb, 0, pin, 0, bytes, 1); };
        return run_pipe(job);
    }
    g_err.clear();
    // column bands of w (a power of two, so bands never straddle a slab) columns
    if ((rc = fs_slots(c, size_t(L) * c.es, false, "device budget below one column of the transform", &job))) return rc;
    int64_t w = 1;
    while (w * 2 <= inner && size_t(L * w * 2) * c.es <= job.slot_bytes) w *= 2;
    const int64_t bands = inner / w;
    job.chunks = size_t(outer * bands);
    const size_t rb = size_t(w) * c.es, pitch = size_t(inner) * c.es;
    Plan p = fs_plan(c.dtype);
    if ((rc = plan_axis0(p, L, w, 1))) return rc;
    auto at = [&](size_t i) { return size_t(int64_t(i) / bands) * sb + size_t(int64_t(i) % bands) * rb; };
    job.gather = [&](size_t i, char* pin) {
        par_copy2d(pin, rb, src + at(i), pitch, rb, size_t(L));
        return size_t(L) * rb;
    };
    job.compute = [&](size_t i, char* d, void*, hipStream_t s, const char** o, size_t* ob) -> int {
        *o = d;
        *ob = size_t(L) * rb;
        int rc2 = exec_plan(p, d, d, nullptr, nullptr, isign, s);
        if (rc2 || !tw_logn) return rc2;
        const int64_t c0 = (int64_t(i) % bands) * w;
        if (c.dtype == NRFFT_F32)
            launch_fs_twiddle<float>(s, ctx, d, L, w, c0, tw_logn, isign);
        else
            launch_fs_twiddle<double>(s, ctx, d, L, w, c0, tw_logn, isign);
        const hipError_t e = hipGetLastError();
        return e == hipSuccess ? NRFFT_OK : fail(NRFFT_E_HIP, std::string("fs twiddle: ") + hipGetErrorString(e));
    };
    job.scatter = [&](size_t i, const char* pin, size_t) { par_copy2d(dst + at(i), pitch, pin, rb, rb, size_t(L)); };
    return run_pipe(job);
}

// The whole array on the device at once: one upload, the in-memory fourn
// plan, one download.  Returns 1 when it does not fit the budget.
static int fs_whole(const FsCtx& c, char* host, const std::vector<size_t>& nn, int64_t total, int isign) {
    Plan p = fs_plan(c.dtype);
    p.nn = nn;
    if (plan_c2c_nd(p, nn, 1)) {
        g_err.clear();
        return 1;
    }
    const size_t bytes = size_t(total) * c.es, tb = p.tmp_elems * c.es;
    if (bytes + tb > c.budget) return 1;
    const hipStream_t s = hipStreamPerThread;
    DevBuf buf, tmp;
    int rc = dev_alloc(buf, bytes, s);
    if (!rc && tb) rc = dev_alloc(tmp, tb, s);
    if (rc) return rc;
    HIPCHK(hipMemcpyAsync(buf.p, host, bytes, hipMemcpyHostToDevice, s));
    if ((rc = exec_plan(p, buf.p, buf.p, tmp.p, nullptr, isign, s))) return rc;
    HIPCHK(hipMemcpyAsync(host, buf.p, bytes, hipMemcpyDeviceToHost, s));
    HIPCHK(hipStreamSynchronize(s));
    return NRFFT_OK;
}

// 1-D four-step N = N1 N2 over the host array: columns of [N1][N2] with the
// W_N^{isign k1 n2} twiddle, from the data into `scratch`, then chunks of
// rows of `scratch` transformed and transposed on the device and stored at
// their natural positions in the data.
static int fs_1d(const FsCtx& c, char* host, char* sc, int64_t total, int isign) {
    const int logn = ilog2((size_t)total);
    if (logn > 26) return fail(NRFFT_E_UNSUPPORTED, "1-D out-of-core transforms up to 2^26 points");
    if (!sc) return fail(NRFFT_E_INVALID, "a 1-D transform larger than the device budget needs a scratch buffer");
    const int l1 = std::min(C2C_MAX_LOGL_, logn / 2);
    const int64_t N1 = int64_t(1) << l1, N2 = total >> l1;  // N2 <= 8192: one-pass rows
    const size_t rb = size_t(N2) * c.es;
    // a row chunk and its transposed copy share a slot
    PipeJob job;
    int rc = fs_slots(c, 2 * rb, false, "device budget below two rows of the four-step split", &job);
    if (rc) return rc;
    if ((rc = fs_strided(c, host, sc, N1, N2, 1, isign, logn))) return rc;
    const int64_t per = std::min<int64_t>(N1, int64_t(job.slot_bytes / (2 * rb)));
    job.chunks = size_t((N1 + per - 1) / per);
    const int64_t tail = N1 - int64_t(job.chunks - 1) * per;
    Plan pf = fs_plan(c.dtype), pt = fs_plan(c.dtype);
    size_t tf = 0;
    if ((rc = plan_c2c_1d(pf, ilog2((size_t)N2), per, B_IN, B_OUT, &tf)) ||
        (tail != per && (rc = plan_c2c_1d(pt, ilog2((size_t)N2), tail, B_IN, B_OUT, &tf))))
        return rc;
    if (tf) return fail(NRFFT_E_UNSUPPORTED, "four-step rows need a workspace");
    auto rows = [&](size_t i) { return i + 1 < job.chunks ? per : tail; };
    job.gather = [&](size_t i, char* pin) {
        par_copy2d(pin, 0, sc + i * size_t(per) * rb, 0, size_t(rows(i)) * rb, 1);
        return size_t(rows(i)) * rb;
    };
    job.compute = [&](size_t i, char* d, void*, hipStream_t s, const char** o, size_t* ob) -> int {
        const int64_t b = rows(i);
        const size_t bytes = size_t(b) * rb;
        int rc2 = exec_plan(i + 1 < job.chunks || tail == per ? pf : pt, d, d, nullptr, nullptr, isign, s);
        if (rc2) return rc2;
        char* t = d + bytes;  // [N2][b] after the chunk
        const dim3 grid((unsigned)((N2 + 31) / 32), (unsigned)((b + 31) / 32));
        if (c.dtype == NRFFT_F32)
            hipLaunchKernelGGL((k_fs_transpose<float>), grid, dim3(32, 8), 0, s, (const cpx<float>*)d, (cpx<float>*)t, b, N2);
        else
            hipLaunchKernelGGL((k_fs_transpose<double>), grid, dim3(32, 8), 0, s, (const cpx<double>*)d, (cpx<double>*)t, b,
                               N2);
        const hipError_t e = hipGetLastError();
        if (e != hipSuccess) return fail(NRFFT_E_HIP, std::string("fs transpose: ") + hipGetErrorString(e));
        *o = t;
        *ob = bytes;
        return NRFFT_OK;
    };
    // [N2][b]: row k2 goes to data[k2 * N1 + r0 .. + b]
    job.scatter = [&](size_t i, const char* pin, size_t) {
        const size_t w = size_t(rows(i)) * c.es;
        par_copy2d(host + i * size_t(per) * c.es, size_t(N1) * c.es, pin, w, w, size_t(N2));
    };
    return run_pipe(job);
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
    c.budget = device_bytes ? device_bytes : (size_t(1) << 30);
    if (c.budget < 128 * c.es) return fail(NRFFT_E_NOMEM, "device budget too small");
    char* host = static_cast<char*>(data);
    int64_t total = 1;
    for (size_t i = 0; i < ndim; ++i) total *= (int64_t)nn[i];
    int rc = fs_whole(c, host, std::vector<size_t>(nn, nn + ndim), total, isign);
    if (rc <= 0) return rc;
    if (ndim == 1) return fs_1d(c, host, static_cast<char*>(scratch), total, isign);
    // axis by axis, last (contiguous) first
    for (size_t dd = ndim; dd-- > 0;) {
        const int64_t L = (int64_t)nn[dd];
        int64_t inner = 1;
        for (size_t e = dd + 1; e < ndim; ++e) inner *= (int64_t)nn[e];
        const int64_t outer = total / (L * inner);
        if (inner == 1) {
            if ((rc = fs_rows(c, host, L, outer, isign))) return rc;
        } else {
            if (L > (int64_t(1) << C2C_MAX_LOGL_))
                return fail(NRFFT_E_UNSUPPORTED, "out-of-core strided axes up to 8192 points");
            if ((rc = fs_strided(c, host, host, L, inner, outer, isign, 0))) return rc;
        }
    }
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
