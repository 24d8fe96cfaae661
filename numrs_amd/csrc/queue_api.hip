// queue_api.hip -- uneven-batch plans: task lists, the persistent queue
// kernel, hipGraph capture and the extern "C" entry points.
#include <algorithm>
#include <cstring>
#include <memory>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/nrfft.h"
#include "internal.hpp"
#include "queue.hpp"

namespace nrfft {

#define HIPCHK(x)                                                                                      \
    do {                                                                                               \
        hipError_t e_ = (x);                                                                           \
        if (e_ != hipSuccess) return fail(NRFFT_E_HIP, std::string(#x) + ": " + hipGetErrorString(e_)); \
    } while (0)

struct UnevenItem {
    int ndim;
    size_t n0, n1;  // ndim == 1: n0 = N, n1 = 1
    size_t off;     // complex elements
};

struct UnevenState {
    int dtype = 0;
    std::vector<UnevenItem> items;
    std::vector<QTask> tasks;  // phase 1 then phase 2
    int n1 = 0, n2 = 0;
    size_t ws_elems = 0;
    size_t total_elems = 0;
    QTask* d_tasks = nullptr;
    QState* d_state = nullptr;
    void* d_ws = nullptr;
    unsigned grid[2] = {0, 0};  // per direction (different kernels)
    // graph cache
    hipGraphExec_t gexec = nullptr;
    hipGraph_t graph = nullptr;
    hipStream_t cap = nullptr;
    const void* g_in = nullptr;
    void* g_out = nullptr;
    int g_isign = 0;
    std::mutex mu;
    ~UnevenState() {
        if (gexec) (void)hipGraphExecDestroy(gexec);
        if (graph) (void)hipGraphDestroy(graph);
        if (cap) (void)hipStreamDestroy(cap);
        if (d_tasks) (void)hipFree(d_tasks);
        if (d_state) (void)hipFree(d_state);
        if (d_ws) (void)hipFree(d_ws);
    }
};

template <typename T>
static const void* queue_kernel(int isign) {
    return isign > 0 ? reinterpret_cast<const void*>(&k_queue<T, 1>) : reinterpret_cast<const void*>(&k_queue<T, -1>);
}

static int q_cls(int logl, int mode) { return (logl - Q_MIN_LOGL) * Q_MODES + mode; }

// Tasks of one pass over `count` transforms of length L = 2^logl: tiles of TI
// transforms along b (b-th transform at base + b * bstride), element stride es.
static void add_pass(std::vector<QTask>& out, int dtype, int logl, int mode, int64_t count, int src, int dst,
                     int64_t ibase, int64_t ibstride, int32_t iesi, int64_t obase, int64_t obstride, int32_t oeso,
                     int tw_logn) {
    const int qp = dtype == NRFFT_F32 ? qpts<float>() : qpts<double>();
    const int64_t TI = qp >> logl;
    for (int64_t b0 = 0; b0 < count; b0 += TI) {
        QTask q{};
        q.cls = q_cls(logl, mode);
        q.nvalid = (int32_t)std::min<int64_t>(TI, count - b0);
        q.src = src;
        q.dst = dst;
        q.ioff = ibase + b0 * ibstride;
        q.ooff = obase + b0 * obstride;
        q.si0 = (int32_t)ibstride;
        q.esi = iesi;
        q.so0 = (int32_t)obstride;
        q.eso = oeso;
        q.tw_logn = tw_logn;
        q.tw_n2base = (int32_t)b0;
        out.push_back(q);
    }
}

// Four-step split of a 1-D item of 2^logn points: the autotuned choice for
// its size class when the tune cache holds one, else N1 = 2^floor(logn / 2).
static int uneven_split(int dtype, int logn, int lo, int hi, const std::vector<int>* force) {
    if (force && (size_t)logn < force->size() && (*force)[logn] > 0) return (*force)[logn];
    Tune t;
    if (tune_find(PK_UNEVEN, dtype, {size_t(1) << logn}, 0, &t) && t.split_log1 >= lo && t.split_log1 <= hi)
        return t.split_log1;
    return std::min(hi, std::max(lo, logn / 2));
}

static int build_uneven(UnevenState& u, const std::vector<int>* force_split = nullptr) {
    const bool f32 = u.dtype == NRFFT_F32;
    const int qlog = ilog2(f32 ? qpts<float>() : qpts<double>());
    const int maxl = f32 ? q_max_logl<float>() : q_max_logl<double>();
    // a column tile (TT / JT) holds at least 4 transforms
    const int maxl_col = std::min(maxl, qlog - 2);
    std::vector<std::pair<int64_t, std::vector<QTask>>> p1, p2;  // (work, tasks) per item
    size_t ws = 0;
    for (const UnevenItem& it : u.items) {
        std::vector<QTask> a, b;
        if (it.ndim == 1) {
            const int logn = ilog2(it.n0);
            if (logn <= maxl) {
                if (logn < Q_MIN_LOGL) return fail(NRFFT_E_UNSUPPORTED, "uneven items need N >= 64");
                add_pass(a, u.dtype, logn, 0, 1, B_IN, B_OUT, it.off, it.n0, 1, it.off, it.n0, 1, 0);
            } else {
                const int lo = std::max(Q_MIN_LOGL, logn - maxl_col), hi = std::min(maxl_col, logn - Q_MIN_LOGL);
                if (lo > hi) return fail(NRFFT_E_UNSUPPORTED, "uneven 1-D item longer than the queue supports");
                const int l1 = uneven_split(u.dtype, logn, lo, hi, force_split);
                const int l2 = logn - l1;
                const int64_t N1 = int64_t(1) << l1, N2 = int64_t(1) << l2;
                // pass A: columns n2 (stride N2), length N1, twiddle -> workspace [k1][n2]
                add_pass(a, u.dtype, l1, 1, N2, B_IN, B_TMP, (int64_t)it.off, 1, (int32_t)N2, (int64_t)ws, 1,
                         (int32_t)N2, logn);
                // pass B: rows k1 of the workspace, length N2, stored at X[k2 N1 + k1]
                add_pass(b, u.dtype, l2, 2, N1, B_TMP, B_OUT, (int64_t)ws, N2, 1, (int64_t)it.off, 1, (int32_t)N1, 0);
                ws += size_t(N1 * N2);
            }
        } else {
            const int l0 = ilog2(it.n0), l1 = ilog2(it.n1);
            if (l0 < Q_MIN_LOGL || l1 < Q_MIN_LOGL || l1 > maxl || l0 > std::min(maxl, qlog - 1))
                return fail(NRFFT_E_UNSUPPORTED, "uneven 2-D sides out of the supported range");
            // rows along n1 (contiguous), then columns along n0 (stride n1) in place in out
            add_pass(a, u.dtype, l1, 0, (int64_t)it.n0, B_IN, B_OUT, it.off, it.n1, 1, it.off, it.n1, 1, 0);
            add_pass(b, u.dtype, l0, 1, (int64_t)it.n1, B_OUT, B_OUT, it.off, 1, (int32_t)it.n1, it.off, 1,
                     (int32_t)it.n1, 0);
        }
        const int64_t work = int64_t(it.n0 * it.n1);
        p1.emplace_back(work, std::move(a));
        if (!b.empty()) p2.emplace_back(work, std::move(b));
    }
    // Largest items first in each phase (longest-processing-time order).
    auto by_work = [](const auto& x, const auto& y) { return x.first > y.first; };
    std::stable_sort(p1.begin(), p1.end(), by_work);
    std::stable_sort(p2.begin(), p2.end(), by_work);
    u.tasks.clear();
    for (auto& e : p1) u.tasks.insert(u.tasks.end(), e.second.begin(), e.second.end());
    u.n1 = (int)u.tasks.size();
    for (auto& e : p2) u.tasks.insert(u.tasks.end(), e.second.begin(), e.second.end());
    u.n2 = (int)u.tasks.size() - u.n1;
    u.ws_elems = ws;
    return NRFFT_OK;
}

static int uneven_setup_device(UnevenState& u) {
    DeviceCtx* ctx = nullptr;
    int rc = get_ctx(&ctx);
    if (rc) return rc;
    HIPCHK(hipMalloc(&u.d_tasks, std::max<size_t>(1, u.tasks.size()) * sizeof(QTask)));
    if (!u.tasks.empty())
        HIPCHK(hipMemcpy(u.d_tasks, u.tasks.data(), u.tasks.size() * sizeof(QTask), hipMemcpyHostToDevice));
    HIPCHK(hipMalloc(&u.d_state, sizeof(QState)));
    if (u.ws_elems) {
        const size_t bytes = u.ws_elems * csize(u.dtype);
        if (hipMalloc(&u.d_ws, bytes) != hipSuccess) return fail(NRFFT_E_NOMEM, "uneven workspace");
    }
    for (int d = 0; d < 2; ++d) {
        const void* fn = u.dtype == NRFFT_F32 ? queue_kernel<float>(d == 0 ? 1 : -1)
                                              : queue_kernel<double>(d == 0 ? 1 : -1);
        const int nt = u.dtype == NRFFT_F32 ? qpts<float>() / 16 : qpts<double>() / 16;
        int per_cu = 0;
        HIPCHK(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, fn, nt, 0));
        if (per_cu < 1) return fail(NRFFT_E_UNSUPPORTED, "queue kernel does not fit a CU");
        u.grid[d] = (unsigned)(per_cu * ctx->ncu);
    }
    HIPCHK(hipStreamCreateWithFlags(&u.cap, hipStreamNonBlocking));
    return NRFFT_OK;
}

template <typename T>
static int uneven_enqueue(UnevenState& u, const void* in, void* out, int isign, hipStream_t s) {
    DeviceCtx* ctx = nullptr;
    int rc = get_ctx(&ctx);
    if (rc) return rc;
    Twiddles<T> tw = ctx_twiddles<T>(ctx);
    cpx<T>* pin = const_cast<cpx<T>*>(static_cast<const cpx<T>*>(in));
    cpx<T>* pout = static_cast<cpx<T>*>(out);
    cpx<T>* pws = static_cast<cpx<T>*>(u.d_ws);
    const QTask* tk = u.d_tasks;
    QState* st = u.d_state;
    const int d = isign > 0 ? 0 : 1;
    hipLaunchKernelGGL(k_queue_reset, dim3(1), dim3(64), 0, s, st);
    HIPCHK(hipGetLastError());
    const int nt = (int)(qpts<T>() / 16);
    for (int phase = 0; phase < 2; ++phase) {
        int base = phase ? u.n1 : 0, count = phase ? u.n2 : u.n1, ph = phase;
        if (count == 0) continue;
        const unsigned grid = std::min<unsigned>(u.grid[d], (unsigned)count);
        void* args[] = {&pin, &pout, &pws, &tk, &base, &count, &ph, &st, &tw};
        HIPCHK(hipLaunchKernel(queue_kernel<T>(isign), dim3(grid), dim3(nt), args, 0, s));
    }
    return NRFFT_OK;
}

static int uneven_execute(UnevenState& u, const void* in, void* out, int isign, hipStream_t stream, bool use_graph) {
    std::lock_guard<std::mutex> lk(u.mu);
    if (!use_graph) {
        return u.dtype == NRFFT_F32 ? uneven_enqueue<float>(u, in, out, isign, stream)
                                    : uneven_enqueue<double>(u, in, out, isign, stream);
    }
    if (!u.gexec || u.g_in != in || u.g_out != out || u.g_isign != isign) {
        if (u.gexec) {
            (void)hipGraphExecDestroy(u.gexec);
            u.gexec = nullptr;
        }
        if (u.graph) {
            (void)hipGraphDestroy(u.graph);
            u.graph = nullptr;
        }
        HIPCHK(hipStreamBeginCapture(u.cap, hipStreamCaptureModeThreadLocal));
        int rc = u.dtype == NRFFT_F32 ? uneven_enqueue<float>(u, in, out, isign, u.cap)
                                      : uneven_enqueue<double>(u, in, out, isign, u.cap);
        hipGraph_t g = nullptr;
        hipError_t e = hipStreamEndCapture(u.cap, &g);
        if (rc) {
            if (g) (void)hipGraphDestroy(g);
            return rc;
        }
        if (e != hipSuccess) return fail(NRFFT_E_HIP, std::string("hipStreamEndCapture: ") + hipGetErrorString(e));
        u.graph = g;
        HIPCHK(hipGraphInstantiate(&u.gexec, u.graph, nullptr, nullptr, 0));
        u.g_in = in;
        u.g_out = out;
        u.g_isign = isign;
    }
    HIPCHK(hipGraphLaunch(u.gexec, stream));
    return NRFFT_OK;
}

// Average ms of `iters` executions of an uneven plan (plain launches).
template <typename T>
static int time_uneven(UnevenState& u, void* in, void* out, int iters, hipStream_t s, hipEvent_t e0, hipEvent_t e1,
                       float* ms) {
    int rc = uneven_enqueue<T>(u, in, out, 1, s);
    if (rc) return rc;
    HIPCHK(hipEventRecord(e0, s));
    for (int i = 0; i < iters; ++i)
        if ((rc = uneven_enqueue<T>(u, in, out, 1, s))) return rc;
    HIPCHK(hipEventRecord(e1, s));
    HIPCHK(hipEventSynchronize(e1));
    HIPCHK(hipEventElapsedTime(ms, e0, e1));
    *ms /= (float)iters;
    return NRFFT_OK;
}

// Every four-step split of the 1-D size classes 2^lmin .. 2^lmax, timed on a
// batch of that class (2^25 FP32 / 2^24 FP64 points); the fastest goes to the
// tune cache under (PK_UNEVEN, dtype, {N}).
static int autotune_uneven(int dtype, int lmin, int lmax, int iters, int* best_l1, float* best_ms) {
    const bool f32 = dtype == NRFFT_F32;
    const int qlog = ilog2(f32 ? qpts<float>() : qpts<double>());
    const int maxl = f32 ? q_max_logl<float>() : q_max_logl<double>();
    const int maxl_col = std::min(maxl, qlog - 2);
    const size_t pts = size_t(1) << (f32 ? 25 : 24);
    struct Res {
        hipStream_t s = nullptr;
        hipEvent_t e0 = nullptr, e1 = nullptr;
        void* in = nullptr;
        void* out = nullptr;
        ~Res() {
            if (e0) (void)hipEventDestroy(e0);
            if (e1) (void)hipEventDestroy(e1);
            if (in) (void)hipFree(in);
            if (out) (void)hipFree(out);
            if (s) (void)hipStreamDestroy(s);
        }
    } r;
    HIPCHK(hipStreamCreateWithFlags(&r.s, hipStreamNonBlocking));
    HIPCHK(hipEventCreate(&r.e0));
    HIPCHK(hipEventCreate(&r.e1));
    if (hipMalloc(&r.in, pts * csize(dtype)) != hipSuccess || hipMalloc(&r.out, pts * csize(dtype)) != hipSuccess)
        return fail(NRFFT_E_NOMEM, "autotune buffers");
    HIPCHK(hipMemsetAsync(r.in, 0, pts * csize(dtype), r.s));
    for (int logn = lmin; logn <= lmax; ++logn) {
        const int slot = logn - lmin;
        if (best_l1) best_l1[slot] = 0;
        if (best_ms) best_ms[slot] = 0.f;
        const int lo = std::max(Q_MIN_LOGL, logn - maxl_col), hi = std::min(maxl_col, logn - Q_MIN_LOGL);
        if (logn <= maxl || lo > hi) continue;  // single pass: nothing to choose
        const size_t n = size_t(1) << logn;
        const size_t count = std::max<size_t>(1, pts / n);
        float best = 0.f;
        int arg = 0;
        for (int l1 = lo; l1 <= hi; ++l1) {
            UnevenState u;
            u.dtype = dtype;
            for (size_t k = 0; k < count; ++k) u.items.push_back(UnevenItem{1, n, 1, k * n});
            u.total_elems = count * n;
            std::vector<int> force(size_t(logn) + 1, 0);
            force[logn] = l1;
            int rc = build_uneven(u, &force);
            if (!rc) rc = uneven_setup_device(u);
            float ms = 0.f;
            if (!rc) rc = f32 ? time_uneven<float>(u, r.in, r.out, iters, r.s, r.e0, r.e1, &ms)
                              : time_uneven<double>(u, r.in, r.out, iters, r.s, r.e0, r.e1, &ms);
            if (rc) return rc;
            if (arg == 0 || ms < best) {
                best = ms;
                arg = l1;
            }
        }
        Tune t;
        t.split_log1 = arg;
        tune_store(PK_UNEVEN, dtype, {n}, 0, t);
        if (best_l1) best_l1[slot] = arg;
        if (best_ms) best_ms[slot] = best;
    }
    HIPCHK(hipStreamSynchronize(r.s));
    return NRFFT_OK;
}

}  // namespace nrfft

using namespace nrfft;

struct nrfft_uneven_s {
    UnevenState u;
};

// FFTProcessor::fft_batch (FFT_1.rs:185-189) over host buffers of mixed
// lengths: packed into one HBM buffer, run as one uneven queue launch.
template <typename T>
static int host_fft_batch(T* const* data, const size_t* nn, size_t count, int isign, int dtype) {
    if (isign != 1 && isign != -1) return fail(NRFFT_E_ISIGN, "isign must be 1 or -1");
    if (count == 0) return NRFFT_OK;
    if (!data || !nn) return fail(NRFFT_E_INVALID, "null array");
    std::vector<int> nd(count, 1);
    std::vector<size_t> dims(2 * count, 1), offs(count);
    size_t total = 0;
    for (size_t i = 0; i < count; ++i) {
        if (!data[i]) return fail(NRFFT_E_INVALID, "null buffer");
        if (nn[i] == 0) return fail(NRFFT_E_EMPTY, "nn == 0");
        if (!is_pow2(nn[i])) return fail(NRFFT_E_NOT_POW2, "nn must be a power of two");
        dims[2 * i] = nn[i];
        offs[i] = total;
        total += nn[i];
    }
    // Items shorter than the queue's minimum go through the batched C2C call.
    std::vector<size_t> big, small;
    for (size_t i = 0; i < count; ++i) (nn[i] >= 64 ? big : small).push_back(i);
    for (size_t i : small) {
        int rc = dtype == NRFFT_F32 ? nrfft_four1_f32(reinterpret_cast<float*>(data[i]), nn[i], isign)
                                    : nrfft_four1_f64(reinterpret_cast<double*>(data[i]), nn[i], isign);
        if (rc) return rc;
    }
    if (big.empty()) return NRFFT_OK;
    std::vector<int> bnd(big.size(), 1);
    std::vector<size_t> bdims(2 * big.size(), 1), boffs(big.size());
    size_t btotal = 0;
    for (size_t k = 0; k < big.size(); ++k) {
        bdims[2 * k] = nn[big[k]];
        boffs[k] = btotal;
        btotal += nn[big[k]];
    }
    nrfft_uneven plan = nullptr;
    int rc = nrfft_plan_uneven(&plan, dtype, big.size(), bnd.data(), bdims.data(), boffs.data());
    if (rc) return rc;
    std::unique_ptr<nrfft_uneven_s> guard(plan);
    hipStream_t s = hipStreamPerThread;
    DevBuf buf;
    const size_t bytes = btotal * 2 * sizeof(T);
    if ((rc = dev_alloc(buf, bytes, s))) return rc;
    for (size_t k = 0; k < big.size(); ++k)
        HIPCHK(hipMemcpyAsync(static_cast<T*>(buf.p) + 2 * boffs[k], data[big[k]], nn[big[k]] * 2 * sizeof(T),
                              hipMemcpyHostToDevice, s));
    if ((rc = uneven_execute(plan->u, buf.p, buf.p, isign, s, false))) return rc;
    for (size_t k = 0; k < big.size(); ++k)
        HIPCHK(hipMemcpyAsync(data[big[k]], static_cast<T*>(buf.p) + 2 * boffs[k], nn[big[k]] * 2 * sizeof(T),
                              hipMemcpyDeviceToHost, s));
    HIPCHK(hipStreamSynchronize(s));
    if ((rc = nrfft_uneven_check(plan, s))) return rc;
    return NRFFT_OK;
}

extern "C" {

int nrfft_plan_uneven(nrfft_uneven* plan, int dtype, size_t nitems, const int* ndim, const size_t* dims,
                      const size_t* offs) {
    g_err.clear();
    if (!plan) return fail(NRFFT_E_INVALID, "null plan pointer");
    *plan = nullptr;
    if (dtype != NRFFT_F32 && dtype != NRFFT_F64) return fail(NRFFT_E_INVALID, "dtype");
    if (nitems == 0) return fail(NRFFT_E_EMPTY, "no items");
    if (!ndim || !dims || !offs) return fail(NRFFT_E_INVALID, "null array");
    auto p = std::make_unique<nrfft_uneven_s>();
    p->u.dtype = dtype;
    size_t end = 0;
    for (size_t i = 0; i < nitems; ++i) {
        UnevenItem it;
        it.ndim = ndim[i];
        if (it.ndim != 1 && it.ndim != 2) return fail(NRFFT_E_DIMS, "uneven items are 1-D or 2-D");
        it.n0 = dims[2 * i];
        it.n1 = it.ndim == 2 ? dims[2 * i + 1] : 1;
        if (it.n0 <= 1 || (it.ndim == 2 && it.n1 <= 1)) return fail(NRFFT_E_DIMS, "Invalid dimension size");
        if (!is_pow2(it.n0) || !is_pow2(it.n1)) return fail(NRFFT_E_NOT_POW2, "dimension must be a power of two");
        it.off = offs[i];
        end = std::max(end, it.off + it.n0 * it.n1);
        p->u.items.push_back(it);
    }
    p->u.total_elems = end;
    int rc = build_uneven(p->u);
    if (rc) return rc;
    rc = uneven_setup_device(p->u);
    if (rc) return rc;
    *plan = p.release();
    return NRFFT_OK;
}

int nrfft_execute_uneven(nrfft_uneven plan, const void* d_in, void* d_out, int isign, void* stream) {
    g_err.clear();
    if (!plan) return fail(NRFFT_E_INVALID, "null plan");
    if (isign != 1 && isign != -1) return fail(NRFFT_E_ISIGN, "isign must be 1 or -1");
    if (!d_in || !d_out) return fail(NRFFT_E_INVALID, "null buffer");
    return uneven_execute(plan->u, d_in, d_out, isign, (hipStream_t)stream, true);
}

int nrfft_execute_uneven_nograph(nrfft_uneven plan, const void* d_in, void* d_out, int isign, void* stream) {
    g_err.clear();
    if (!plan) return fail(NRFFT_E_INVALID, "null plan");
    if (isign != 1 && isign != -1) return fail(NRFFT_E_ISIGN, "isign must be 1 or -1");
    if (!d_in || !d_out) return fail(NRFFT_E_INVALID, "null buffer");
    return uneven_execute(plan->u, d_in, d_out, isign, (hipStream_t)stream, false);
}

int nrfft_uneven_check(nrfft_uneven plan, void* stream) {
    g_err.clear();
    if (!plan) return fail(NRFFT_E_INVALID, "null plan");
    QState st{};
    HIPCHK(hipMemcpyAsync(&st, plan->u.d_state, sizeof st, hipMemcpyDeviceToHost, (hipStream_t)stream));
    HIPCHK(hipStreamSynchronize((hipStream_t)stream));
    // every group leaves after drawing one index past the end of its phase
    if (st.next[0] < (uint32_t)plan->u.n1 || st.next[1] < (uint32_t)plan->u.n2)
        return fail(NRFFT_E_HIP, "uneven queue: a phase did not drain its task list");
    return NRFFT_OK;
}

int nrfft_uneven_info(nrfft_uneven plan, size_t* tasks_phase1, size_t* tasks_phase2, size_t* workspace_bytes,
                      unsigned* groups) {
    g_err.clear();
    if (!plan) return fail(NRFFT_E_INVALID, "null plan");
    if (tasks_phase1) *tasks_phase1 = (size_t)plan->u.n1;
    if (tasks_phase2) *tasks_phase2 = (size_t)plan->u.n2;
    if (workspace_bytes) *workspace_bytes = plan->u.ws_elems * csize(plan->u.dtype);
    if (groups) *groups = plan->u.grid[0];
    return NRFFT_OK;
}

int nrfft_uneven_destroy(nrfft_uneven plan) {
    delete plan;
    return NRFFT_OK;
}

int nrfft_autotune_uneven(int dtype, int log2n_min, int log2n_max, int iters, int* best_l1, float* best_ms) {
    g_err.clear();
    if (dtype != NRFFT_F32 && dtype != NRFFT_F64) return fail(NRFFT_E_INVALID, "dtype");
    if (log2n_min < Q_MIN_LOGL || log2n_max < log2n_min || log2n_max > 26)
        return fail(NRFFT_E_INVALID, "size-class range");
    if (iters < 1) iters = 1;
    return autotune_uneven(dtype, log2n_min, log2n_max, iters, best_l1, best_ms);
}

int nrfft_fft_batch_f64(double* const* data, const size_t* nn, size_t count, int isign) {
    g_err.clear();
    return host_fft_batch<double>(data, nn, count, isign, NRFFT_F64);
}
int nrfft_fft_batch_f32(float* const* data, const size_t* nn, size_t count, int isign) {
    g_err.clear();
    return host_fft_batch<float>(data, nn, count, isign, NRFFT_F32);
}

}  // extern "C"
