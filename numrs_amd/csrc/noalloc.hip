// noalloc.hip -- batched 1-D C2C on device pointers that allocates nothing:
// the kernels and default tunings of nrfft_plan_c2c + nrfft_execute_c2c,
// with the four-step intermediate in a file-scope device array instead of a
// plan-owned workspace.
//
// `scratch` is a static __device__ array: the code object's loader places
// one copy on every GPU that loads the library, held until the process
// ends.  Host code never uses its name as an address (on the host it is a
// stand-in whose address means nothing to the GPU): nrfft_init calls
// noalloc_setup with the device current, which asks hipGetSymbolAddress for
// that device's copy and keeps the address per device; the entry point
// passes that address to the kernels as the four-step workspace (B_TMP).
//
// Nothing in the scratch has to start at zero: pass A writes every element
// pass B of the same chunk reads.  Calls on one GPU share its scratch, so
// two calls that overlap in time on that GPU (different streams, or host
// threads without ordering) race; calls ordered on one stream do not.
#include <mutex>
#include <string>

#include "../../include/nrfft.h"
#include "internal.hpp"

#ifndef NRFFT_NOALLOC_SCRATCH_BYTES
#define NRFFT_NOALLOC_SCRATCH_BYTES (size_t(64) << 20)
#endif

static __device__ __align__(256) unsigned char scratch[NRFFT_NOALLOC_SCRATCH_BYTES];

namespace nrfft {

constexpr int NOALLOC_MAX_DEVICES = 64;
static std::mutex g_scratch_mu;
static void* g_scratch_ptr[NOALLOC_MAX_DEVICES] = {};

int noalloc_setup(int device) {
    if (device < 0 || device >= NOALLOC_MAX_DEVICES) return fail(NRFFT_E_INVALID, "device index out of range");
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (g_scratch_ptr[device]) return NRFFT_OK;
    void* ptr = nullptr;
    const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !ptr)
        return fail(NRFFT_E_HIP, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
    g_scratch_ptr[device] = ptr;
    return NRFFT_OK;
}

static int scratch_of(int device, void** out) {
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (device < 0 || device >= NOALLOC_MAX_DEVICES || !g_scratch_ptr[device])
        return fail(NRFFT_E_INVALID, "nrfft_init(" + std::to_string(device) + ") has not run on this device");
    *out = g_scratch_ptr[device];
    return NRFFT_OK;
}

static int noalloc_c2c(int dtype, const void* d_in, void* d_out, size_t nn, size_t batch, int isign,
                       hipStream_t stream) {
    if (dtype != NRFFT_F32 && dtype != NRFFT_F64) return fail(NRFFT_E_INVALID, "dtype");
    if (isign != 1 && isign != -1) return fail(NRFFT_E_ISIGN, "isign must be 1 or -1");
    if (nn == 0 || batch == 0) return fail(NRFFT_E_EMPTY, "nn == 0 or batch == 0");
    if (!is_pow2(nn)) return fail(NRFFT_E_NOT_POW2, "nn must be a power of two");
    if (!d_in || !d_out) return fail(NRFFT_E_INVALID, "null buffer");
    const int logn = ilog2(nn);
    if (logn > 2 * C2C_MAX_LOGL_) return fail(NRFFT_E_UNSUPPORTED, "transform longer than 2^26");
    const size_t cs = csize(dtype);
    if (logn > C2C_MAX_LOGL_ && nn * cs > NRFFT_NOALLOC_SCRATCH_BYTES)
        return fail(NRFFT_E_UNSUPPORTED, "a transform of " + std::to_string(nn) + " points needs " +
                                             std::to_string(nn * cs >> 20) + " MiB of scratch; the allocation-free "
                                             "entry point holds " +
                                             std::to_string(NRFFT_NOALLOC_SCRATCH_BYTES >> 20) + " MiB");
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess) return fail(NRFFT_E_HIP, "hipGetDevice");
    DeviceCtx* ctx = nullptr;
    if (int rc = find_ctx(&ctx)) return rc;
    void* ws = nullptr;
    if (int rc = scratch_of(dev, &ws)) return rc;
    // A host-side plan description on the stack (its vectors stay empty).
    Plan p;
    p.kind = PK_C2C;
    p.dtype = dtype;
    p.batch = batch;
    p.device = dev;
    p.tune = default_tune_1d(dtype, nn, batch, 0);
    void* bufs[B_COUNT] = {const_cast<void*>(d_in), d_out, ws, nullptr, nullptr, nullptr};
    const int64_t N = int64_t(nn);
    if (logn == 0) {
        if (d_in != d_out && hipMemcpyAsync(d_out, d_in, batch * cs, hipMemcpyDeviceToDevice, stream) != hipSuccess)
            return fail(NRFFT_E_HIP, "hipMemcpyAsync");
        return NRFFT_OK;
    }
    if (logn <= C2C_MAX_LOGL_) {
        // One pass (no scratch), batch in slices whose in-tile offsets fit 32 bits.
        PassDesc d{};
        d.nb0 = (int64_t)batch;
        d.nb1 = 1;
        d.nb2 = 1;
        d.si0 = (int32_t)N;
        d.so0 = (int32_t)N;
        d.esi = 1;
        d.eso = 1;
        if (N * std::min<int64_t>((int64_t)batch, 256) > 0x7fffffffLL) return fail(NRFFT_E_UNSUPPORTED, "too large");
        Launch l;
        if (int rc = c2c_pass_launch(p, MODE_JJ, logn, 0, d, B_IN, B_OUT, &l)) return rc;
        return exec_launch(p, l, bufs, isign, stream, ctx);
    }
    FourStep fs;
    if (int rc = fourstep_setup(p, logn, &fs)) return rc;
    const int64_t cap = int64_t(NRFFT_NOALLOC_SCRATCH_BYTES / (nn * cs));
    int64_t chunk = p.tune.chunk > 0 ? std::min<int64_t>(p.tune.chunk, (int64_t)batch) : (int64_t)batch;
    chunk = std::min(chunk, cap);
    Launch la, lb;
    int64_t built = 0;  // chunk size the launches were built for
    for (int64_t c0 = 0; c0 < (int64_t)batch; c0 += chunk) {
        const int64_t cb = std::min<int64_t>(chunk, (int64_t)batch - c0);
        if (cb != built) {
            if (int rc = fourstep_launches(p, fs, cb, B_IN, B_OUT, &la, &lb)) return rc;
            built = cb;
        }
        la.off_src = c0 * N;
        lb.off_dst = c0 * N;
        if (int rc = exec_launch(p, la, bufs, isign, stream, ctx)) return rc;
        if (int rc = exec_launch(p, lb, bufs, isign, stream, ctx)) return rc;
    }
    return NRFFT_OK;
}

}  // namespace nrfft

using namespace nrfft;

extern "C" int nrfft_execute_four1_batch_noalloc(int dtype, const void* d_in, void* d_out, size_t nn, size_t batch,
                                                 int isign, void* stream) {
    g_err.clear();
    return noalloc_c2c(dtype, d_in, d_out, nn, batch, isign, (hipStream_t)stream);
}
