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
#include <functional>
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
constexpr size_t COPY_MIN_PART = size_t(2) << 20;

// Host copies on a small persistent pool (one host thread moves ~10 GB/s;
// starting threads per chunk would cost more than the copy saves).  The
// workers are detached and wait for jobs until the process ends; one job
// runs at a time.
class CopyPool {
public:
    static CopyPool& get() {
        static CopyPool* p = new CopyPool();  // never destroyed: workers outlive static teardown
        return *p;
    }
    // fn(0) .. fn(np - 1), this thread taking part in the work.
    void run(int np, const std::function<void(int)>& fn) {
        if (np <= 1) {
            fn(0);
            return;
        }
        std::lock_guard<std::mutex> job(job_mu_);
        std::unique_lock<std::mutex> lk(m_);
        fn_ = &fn;
        next_ = 0;
        parts_ = np;
        pending_ = np;
        ++gen_;
        cv_.notify_all();
        take(lk);
        done_.wait(lk, [&] { return pending_ == 0; });
        fn_ = nullptr;
    }

private:
    CopyPool() {
        for (int i = 1; i < COPY_THREADS; ++i) std::thread([this] { worker(); }).detach();
    }
    // Runs parts until none is left; called and returns with m_ held.
    void take(std::unique_lock<std::mutex>& lk) {
        while (next_ < parts_) {
            const int i = next_++;
            const std::function<void(int)>* fn = fn_;
            lk.unlock();
            (*fn)(i);
            lk.lock();
            if (--pending_ == 0) done_.notify_one();
        }
    }
    void worker() {
        uint64_t seen = 0;
        std::unique_lock<std::mutex> lk(m_);
        for (;;) {
            cv_.wait(lk, [&] { return gen_ != seen && next_ < parts_; });
            seen = gen_;
            take(lk);
        }
    }
    std::mutex job_mu_, m_;
    std::condition_variable cv_, done_;
    const std::function<void(int)>* fn_ = nullptr;
    int next_ = 0, parts_ = 0, pending_ = 0;
    uint64_t gen_ = 0;
};

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

void par_copy2d(char* dst, size_t dpitch, const char* src, size_t spitch, size_t width, size_t rows) {
    if (!width || !rows) return;
    if (rows == 1 || (dpitch == width && spitch == width)) {  // one contiguous run, split by bytes
        const size_t bytes = width * rows;
        const int np = (int)std::min<size_t>(COPY_THREADS, std::max<size_t>(1, bytes / COPY_MIN_PART));
        const size_t part = ((bytes + np - 1) / np + 63) & ~size_t(63);
        CopyPool::get().run(np, [&](int i) {
            const size_t o = size_t(i) * part;
            if (o < bytes) std::memcpy(dst + o, src + o, std::min(part, bytes - o));
        });
        return;
    }
    const int np = (int)std::min<size_t>({size_t(COPY_THREADS), rows, std::max<size_t>(1, width * rows / COPY_MIN_PART)});
    const size_t per = (rows + np - 1) / np;
    CopyPool::get().run(np, [&](int i) {
        const size_t r1 = std::min(rows, size_t(i + 1) * per);
        for (size_t r = size_t(i) * per; r < r1; ++r) std::memcpy(dst + r * dpitch, src + r * spitch, width);
    });
}

// Chunk i goes through slot i % ns: gathered on the host threads into the
// pinned input slot, uploaded (stream 0), transformed (stream 1), downloaded
// into the pinned output slot (stream 2) and scattered.  With two slots the
// upload of chunk i overlaps the transform of chunk i-1 and the download of
// chunk i-2; a slot is reused once its previous chunk has been scattered.
int run_pipe(const PipeJob& job) {
    if (job.chunks == 0) return NRFFT_OK;
    const int ns = job.slots == 1 ? 1 : 2;
    int dev = 0;
    HIPCHK(hipGetDevice(&dev));
    std::lock_guard<std::mutex> lk(g_stage_mu);  // one pipelined call per process at a time
    Staging* sg = nullptr;
    int rc = get_staging(dev, job.slot_bytes, &sg);
    if (rc) return rc;
    PipeRes r;
    for (auto& s : r.st) HIPCHK(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
    for (int k = 0; k < 2; ++k) {
        HIPCHK(hipEventCreateWithFlags(&r.up[k], hipEventDisableTiming));
        HIPCHK(hipEventCreateWithFlags(&r.comp[k], hipEventDisableTiming));
        HIPCHK(hipEventCreateWithFlags(&r.down[k], hipEventDisableTiming));
    }
    DevBuf buf[2], ws;
    for (int k = 0; k < ns; ++k)
        if ((rc = dev_alloc(buf[k], job.slot_bytes, r.st[1]))) return rc;
    if (job.ws_bytes && (rc = dev_alloc(ws, job.ws_bytes, r.st[1]))) return rc;
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

    size_t out_bytes[2] = {0, 0};
    auto copy_out = [&](size_t i) -> int {
        const int k = int(i % ns);
        HIPCHK(hipEventSynchronize(r.down[k]));
        job.scatter(i, sg->out[k], out_bytes[k]);
        return NRFFT_OK;
    };
    for (size_t i = 0; i < job.chunks; ++i) {
        const int k = int(i % ns);
        // Slot k was last used by chunk i - ns: its download has to land (and
        // be scattered) before the staging and device buffers are reused.
        if (i >= size_t(ns)) {
            if ((rc = copy_out(i - ns))) return rc;
            HIPCHK(hipEventSynchronize(r.up[k]));
        }
        const size_t b = job.gather(i, sg->in[k]);
        if (b > job.slot_bytes) return fail(NRFFT_E_INVALID, "pipeline chunk larger than its slot");
        if (i >= size_t(ns)) HIPCHK(hipStreamWaitEvent(r.st[0], r.down[k], 0));  // device slot free
        HIPCHK(hipMemcpyAsync(buf[k].p, sg->in[k], b, hipMemcpyHostToDevice, r.st[0]));
        HIPCHK(hipEventRecord(r.up[k], r.st[0]));
        HIPCHK(hipStreamWaitEvent(r.st[1], r.up[k], 0));
        const char* o = nullptr;
        size_t ob = 0;
        if ((rc = job.compute(i, static_cast<char*>(buf[k].p), ws.p, r.st[1], &o, &ob))) return rc;
        if (ob > job.slot_bytes) return fail(NRFFT_E_INVALID, "pipeline result larger than its slot");
        out_bytes[k] = ob;
        HIPCHK(hipEventRecord(r.comp[k], r.st[1]));
        HIPCHK(hipStreamWaitEvent(r.st[2], r.comp[k], 0));
        HIPCHK(hipMemcpyAsync(sg->out[k], o, ob, hipMemcpyDeviceToHost, r.st[2]));
        HIPCHK(hipEventRecord(r.down[k], r.st[2]));
    }
    for (size_t i = job.chunks >= size_t(ns) ? job.chunks - ns : 0; i < job.chunks; ++i)
        if ((rc = copy_out(i))) return rc;
    return NRFFT_OK;
}

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
    char* host = static_cast<char*>(data);
    auto chunk_bytes = [&](size_t i) { return (i + 1 < nch ? chunk : tail) * per; };
    PipeJob job;
    job.chunks = nch;
    job.slot_bytes = chunk * per;
    job.ws_bytes = std::max(pf->tmp_elems, pt ? pt->tmp_elems : size_t(0)) * csize(dtype);
    job.gather = [&](size_t i, char* pin) {
        par_copy2d(pin, 0, host + i * chunk * per, 0, chunk_bytes(i), 1);
        return chunk_bytes(i);
    };
    job.compute = [&](size_t i, char* d, void* ws, hipStream_t s, const char** o, size_t* ob) -> int {
        *o = d;
        *ob = chunk_bytes(i);
        const Plan& p = (i + 1 < nch || !pt) ? *pf : *pt;
        return exec_plan(p, d, d, ws, nullptr, isign, s);
    };
    job.scatter = [&](size_t i, const char* pin, size_t bytes) { par_copy2d(host + i * chunk * per, 0, pin, 0, bytes, 1); };
    return run_pipe(job);
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
