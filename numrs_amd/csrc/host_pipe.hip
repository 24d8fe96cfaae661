// host_pipe.hip -- pipelined host-pointer batches: the drop-in calls on host
// arrays (four1 batch, fourn batch) stream a large batch through HBM in
// chunks of whole transforms, so the PCIe transfers in both directions, the
// transforms and the host-side copies overlap.
//
// Host arrays are pageable memory (the reference hands over Rust slices);
// DMA needs pinned memory, so each chunk is copied by several host threads
// into a pinned staging slot, uploaded, transformed, downloaded into a second
// pinned slot and copied back by the host threads.  Two slots per direction;
// the upload, compute and download of a chunk run on three streams ordered
// by events, so chunk i's upload overlaps chunk i-1's transform and chunk
// i-2's download.  The pinned slots are allocated once per process and
// device and kept (pinning costs milliseconds).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <condition_variable>
#include <cstring>
#include <mutex>
#include <thread>
#include <vector>

#include "../../include/nrfft.h"
#include "internal.hpp"

#define HIPCHK(x)                                                                                      \
    do {                                                                                               \
        hipError_t e_ = (x);                                                                           \
        if (e_ != hipSuccess) return fail(NRFFT_E_HIP, std::string(#x) + ": " + hipGetErrorString(e_)); \
    } while (0)

namespace nrfft {

namespace {

constexpr size_t CHUNK_TARGET = size_t(32) << 20;  // bytes of data per chunk
constexpr int COPY_THREADS = 8;

// Host copies on a small persistent pool (one host thread moves ~10 GB/s;
// starting threads per chunk would cost more than the copy saves).  The
// workers are detached and wait for jobs until the process ends.
class CopyPool {
public:
    static CopyPool& get() {
        static CopyPool* p = new CopyPool();  // never destroyed: workers outlive static teardown
        return *p;
    }
    // Copy `bytes` in COPY_THREADS parts, this thread taking part 0.
    void copy(char* dst, const char* src, size_t bytes) {
        const size_t min_part = size_t(2) << 20;
        const int np = (int)std::min<size_t>(COPY_THREADS, std::max<size_t>(1, bytes / min_part));
        if (np <= 1) {
            std::memcpy(dst, src, bytes);
            return;
        }
        const size_t part = (bytes + np - 1) / np;
        std::unique_lock<std::mutex> lk(m_);
        job_dst_ = dst;
        job_src_ = src;
        job_bytes_ = bytes;
        job_part_ = part;
        next_ = 1;
        parts_ = np;
        pending_ = np - 1;
        ++gen_;
        cv_.notify_all();
        lk.unlock();
        std::memcpy(dst, src, std::min(part, bytes));
        lk.lock();
        done_.wait(lk, [&] { return pending_ == 0; });
    }

private:
    CopyPool() {
        for (int i = 1; i < COPY_THREADS; ++i) std::thread([this] { worker(); }).detach();
    }
    void worker() {
        uint64_t seen = 0;
        std::unique_lock<std::mutex> lk(m_);
        for (;;) {
            cv_.wait(lk, [&] { return gen_ != seen && next_ < parts_; });
            seen = gen_;
            while (next_ < parts_) {
                const int i = next_++;
                char* d = job_dst_;
                const char* s = job_src_;
                const size_t o = size_t(i) * job_part_, n = o < job_bytes_ ? std::min(job_part_, job_bytes_ - o) : 0;
                lk.unlock();
                if (n) std::memcpy(d + o, s + o, n);
                lk.lock();
                if (--pending_ == 0) done_.notify_one();
            }
        }
    }
    std::mutex m_;
    std::condition_variable cv_, done_;
    char* job_dst_ = nullptr;
    const char* job_src_ = nullptr;
    size_t job_bytes_ = 0, job_part_ = 0;
    int next_ = 0, parts_ = 0, pending_ = 0;
    uint64_t gen_ = 0;
};

void par_memcpy(char* dst, const char* src, size_t bytes) { CopyPool::get().copy(dst, src, bytes); }

// Pinned staging of one device: in[2], out[2], each `slot` bytes.
struct Staging {
    size_t slot = 0;
    char* in[2] = {nullptr, nullptr};
    char* out[2] = {nullptr, nullptr};
};
std::mutex g_stage_mu;
Staging g_stage[64];

int get_staging(int dev, size_t slot, Staging** out) {
    if (dev < 0 || dev >= 64) return fail(NRFFT_E_INVALID, "device index");
    Staging& s = g_stage[dev];
    if (s.slot < slot) {
        for (int k = 0; k < 2; ++k) {
            if (s.in[k]) (void)hipHostFree(s.in[k]);
            if (s.out[k]) (void)hipHostFree(s.out[k]);
            s.in[k] = s.out[k] = nullptr;
        }
        s.slot = 0;
        for (int k = 0; k < 2; ++k) {
            if (hipHostMalloc(reinterpret_cast<void**>(&s.in[k]), slot, hipHostMallocDefault) != hipSuccess ||
                hipHostMalloc(reinterpret_cast<void**>(&s.out[k]), slot, hipHostMallocDefault) != hipSuccess)
                return fail(NRFFT_E_NOMEM, "pinned staging");
        }
        s.slot = slot;
    }
    *out = &s;
    return NRFFT_OK;
}

// Streams and events of one call, released on every return path.
struct PipeRes {
    hipStream_t st[3] = {nullptr, nullptr, nullptr};  // upload, compute, download
    hipEvent_t up[2] = {nullptr, nullptr}, comp[2] = {nullptr, nullptr}, down[2] = {nullptr, nullptr};
    ~PipeRes() {
        for (auto s : st)
            if (s) {
                (void)hipStreamSynchronize(s);
                (void)hipStreamDestroy(s);
            }
        for (int k = 0; k < 2; ++k) {
            if (up[k]) (void)hipEventDestroy(up[k]);
            if (comp[k]) (void)hipEventDestroy(comp[k]);
            if (down[k]) (void)hipEventDestroy(down[k]);
        }
    }
};

}  // namespace

// In place on `batch` contiguous transforms of shape nn at host address data,
// in chunks of `chunk` transforms (the last may be shorter).
int host_c2c_pipelined(int dtype, void* data, const std::vector<size_t>& nn, size_t batch, size_t chunk,
                       int isign) {
    size_t per = csize(dtype);
    for (size_t v : nn) per *= v;
    const size_t nch = (batch + chunk - 1) / chunk;
    std::shared_ptr<Plan> pf, pt;
    int rc = cached_plan(PK_C2C, dtype, nn, chunk, 0, &pf);
    if (rc) return rc;
    const size_t tail = batch - (nch - 1) * chunk;
    if (tail != chunk && (rc = cached_plan(PK_C2C, dtype, nn, tail, 0, &pt))) return rc;
    int dev = 0;
    HIPCHK(hipGetDevice(&dev));
    std::lock_guard<std::mutex> lk(g_stage_mu);  // one pipelined call per process at a time
    Staging* sg = nullptr;
    if ((rc = get_staging(dev, chunk * per, &sg))) return rc;
    PipeRes r;
    for (auto& s : r.st) HIPCHK(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
    for (int k = 0; k < 2; ++k) {
        HIPCHK(hipEventCreateWithFlags(&r.up[k], hipEventDisableTiming));
        HIPCHK(hipEventCreateWithFlags(&r.comp[k], hipEventDisableTiming));
        HIPCHK(hipEventCreateWithFlags(&r.down[k], hipEventDisableTiming));
    }
    DevBuf buf[2], tmp;
    for (int k = 0; k < 2; ++k)
        if ((rc = dev_alloc(buf[k], chunk * per, r.st[1]))) return rc;
    const size_t tmp_elems = std::max(pf->tmp_elems, pt ? pt->tmp_elems : size_t(0));
    if ((rc = dev_alloc(tmp, tmp_elems * csize(dtype), r.st[1]))) return rc;
    HIPCHK(hipStreamSynchronize(r.st[1]));  // allocations visible to the other streams
    // On every return, the three streams drain before the buffers above are
    // released (declared after them, so destroyed first).
    struct Drain {
        PipeRes& r;
        ~Drain() {
            for (auto s : r.st)
                if (s) (void)hipStreamSynchronize(s);
        }
    } drain{r};

    char* host = static_cast<char*>(data);
    auto chunk_bytes = [&](size_t i) { return (i + 1 < nch ? chunk : tail) * per; };
    auto copy_out = [&](size_t i) -> int {
        const int k = int(i % 2);
        HIPCHK(hipEventSynchronize(r.down[k]));
        par_memcpy(host + i * chunk * per, sg->out[k], chunk_bytes(i));
        return NRFFT_OK;
    };
    for (size_t i = 0; i < nch; ++i) {
        const int k = int(i % 2);
        const size_t b = chunk_bytes(i);
        // Slot k was last used by chunk i - 2: its download has to land (and be
        // copied back) before the staging and device buffers are reused.
        if (i >= 2 && (rc = copy_out(i - 2))) return rc;
        if (i >= 2) HIPCHK(hipEventSynchronize(r.up[k]));
        par_memcpy(sg->in[k], host + i * chunk * per, b);
        if (i >= 2) HIPCHK(hipStreamWaitEvent(r.st[0], r.down[k], 0));  // device slot free
        HIPCHK(hipMemcpyAsync(buf[k].p, sg->in[k], b, hipMemcpyHostToDevice, r.st[0]));
        HIPCHK(hipEventRecord(r.up[k], r.st[0]));
        HIPCHK(hipStreamWaitEvent(r.st[1], r.up[k], 0));
        const Plan& p = (i + 1 < nch || !pt) ? *pf : *pt;
        if ((rc = exec_plan(p, buf[k].p, buf[k].p, tmp.p, nullptr, isign, r.st[1]))) return rc;
        HIPCHK(hipEventRecord(r.comp[k], r.st[1]));
        HIPCHK(hipStreamWaitEvent(r.st[2], r.comp[k], 0));
        HIPCHK(hipMemcpyAsync(sg->out[k], buf[k].p, b, hipMemcpyDeviceToHost, r.st[2]));
        HIPCHK(hipEventRecord(r.down[k], r.st[2]));
    }
    for (size_t i = nch >= 2 ? nch - 2 : 0; i < nch; ++i)
        if ((rc = copy_out(i))) return rc;
    return NRFFT_OK;
}

// Chunk size (transforms) for a host batch, or 0 for the one-shot path: the
// pipeline pays off once a batch spans several chunks of whole transforms.
size_t host_pipe_chunk(int dtype, const std::vector<size_t>& nn, size_t batch) {
    size_t per = csize(dtype);
    for (size_t v : nn) per *= v;
    if (batch < 4 || per > CHUNK_TARGET) return 0;
    const size_t chunk = std::max<size_t>(1, CHUNK_TARGET / per);
    return batch >= 3 * chunk ? chunk : 0;
}

}  // namespace nrfft
