// tune.hip -- autotuner and tuning cache for C2C plans.
//
// A plan's tunables (four-step split point, tile widths of the column and
// transpose passes -- which set threads per group, LDS per group and so
// groups (waves) per CU -- and the batch chunk that keeps a chunk's
// intermediate in the last-level cache) are chosen per (shape, batch, dtype).
// nrfft_autotune_c2c times every candidate on the GPU with HIP events and
// stores the fastest in an in-memory cache, which plan creation consults.
// The cache can be saved to / loaded from a text file, and is loaded at the
// first lookup from $NRFFT_TUNE_CACHE when that variable is set.
#include <algorithm>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <mutex>
#include <string>
#include <tuple>
#include <vector>

#include "../../include/nrfft.h"
#include "internal.hpp"

namespace nrfft {

using TuneKey = std::tuple<int, int, std::vector<size_t>, size_t>;  // kind, dtype, nn, batch

static std::mutex g_tune_mu;
static std::map<TuneKey, Tune> g_tune;
static bool g_tune_env_loaded = false;

static int load_file_locked(const char* path) {
    FILE* f = std::fopen(path, "r");
    if (!f) return fail(NRFFT_E_INVALID, std::string("cannot open ") + path);
    // One entry per line: kind dtype batch ndim nn... : split tilog_a tilog_b chunk wg_per_cu xcd_remap
    char line[512];
    while (std::fgets(line, sizeof line, f)) {
        std::vector<long long> v;
        char* p = line;
        char* colon = std::strchr(line, ':');
        if (!colon || line[0] == '#') continue;
        *colon = ' ';
        while (*p) {
            char* e = nullptr;
            long long x = std::strtoll(p, &e, 10);
            if (e == p) break;
            v.push_back(x);
            p = e;
        }
        if (v.size() < 4) continue;
        const size_t ndim = (size_t)v[3];
        if (v.size() != 4 + ndim + 6) continue;
        std::vector<size_t> nn(v.begin() + 4, v.begin() + 4 + ndim);
        Tune t;
        t.split_log1 = (int)v[4 + ndim];
        t.tilog_a = (int)v[5 + ndim];
        t.tilog_b = (int)v[6 + ndim];
        t.chunk = (int)v[7 + ndim];
        t.wg_per_cu = (int)v[8 + ndim];
        t.xcd_remap = (int)v[9 + ndim];
        g_tune[TuneKey((int)v[0], (int)v[1], nn, (size_t)v[2])] = t;
    }
    std::fclose(f);
    return NRFFT_OK;
}

// Defaults for shapes the tuner has not seen, from sweeps on MI355X
// (64 x 2^20 FP32: split 2^9 x 2^11, 16-wide tiles, 16-transform chunks).
Tune default_tune(const Plan& p) {
    if (p.kind == PK_C2C && p.nn.size() == 1) return default_tune_1d(p.dtype, p.nn[0], p.batch, p.extra);
    Tune t;
    if (p.kind == PK_C2C) t.tilog_a = 4;
    return t;
}

Tune default_tune_1d(int dtype, size_t n, size_t batch, size_t extra) {
    Tune t;
    {
        const int logn = ilog2(n);
        if (logn > C2C_MAX_LOGL_) {
            // Four-step split and batch chunking per length, from a sweep on
            // MI355X (256-512 MiB batches, pass B with non-temporal result
            // stores).  From 2^19, 2048-point pass-B rows (the column pass at
            // most 2048 points); below, the best split measured per length.
            // Chunks of 32 MiB (2^14), 64 MiB (short transforms) or 128 MiB
            // whenever the batch is larger, so a chunk's intermediate stays in
            // the MALL.  Against the earlier rule (split near sqrt N, chunks
            // only above 256 MiB): FP32 2^15 0.215 -> 0.203 ms, 2^17 0.237 ->
            // 0.210, 2^19 0.260 -> 0.213 (256 MiB), 2^21 0.913 -> 0.495, 2^22
            // 0.938 -> 0.626 (512 MiB); FP64 (256 MiB) 2^14 0.218 -> 0.206,
            // 2^20 0.259 -> 0.211, 2^21 0.529 -> 0.229, 2^22 0.642 -> 0.313.
            static const int short_f32[5] = {7, 8, 7, 8, 9}, short_f64[5] = {7, 8, 8, 9, 9};  // 2^14..2^18
            const bool f32 = dtype == NRFFT_F32;
            t.split_log1 = logn >= 19 ? std::max(logn - C2C_MAX_LOGL_, std::min(logn - 11, 11))
                                      : (f32 ? short_f32 : short_f64)[logn - 14];
            // A split-precision plan needs a pass of 2^8..2^10 points for its
            // MFMA stages; if neither pass has one, the split near sqrt N.
            auto mfma_len = [](int l) { return l >= 8 && l <= 10; };
            if (extra != 0 && !mfma_len(t.split_log1) && !mfma_len(logn - t.split_log1))
                t.split_log1 = std::max(logn - C2C_MAX_LOGL_, logn / 2 - 1);
            // Tile widths (log2 transforms per tile; capped by LDS per length):
            // 16 for FP32; FP64 8-wide pass-B tiles and 8-wide pass-A tiles from
            // 512-point columns (FP64 2^14 0.222 -> 0.205 ms, 2^17 0.226 ->
            // 0.215, 2^20 0.224 -> 0.209 against 16-wide).
            t.tilog_a = f32 || t.split_log1 < 9 ? 4 : 3;
            t.tilog_b = f32 ? 4 : 3;
            // A small problem (a few transforms) is launch-latency bound: 4-wide
            // tiles spread each pass over more CUs (one FP64 2^16 transform:
            // 10.6 -> 9.8 us).
            if (batch * n <= (size_t(1) << 18)) t.tilog_a = t.tilog_b = 2;
            const size_t per = n * (f32 ? 8 : 16);
            const size_t mib = logn == 14 ? 32 : (f32 ? logn <= 18 : logn <= 17) ? 64 : 128;
            const size_t chunk = std::max<size_t>(1, (mib << 20) / per);
            if (chunk < batch) t.chunk = (int)chunk;
        }
    }
    return t;
}

// The inner C2C of a multi-pass plan runs with the tuning of a C2C plan of
// its length and batch (four-step split, tile widths, batch chunking).
void use_c2c_tune(Plan& p, int logn) {
    Plan q;
    q.kind = PK_C2C;
    q.dtype = p.dtype;
    q.nn = {size_t(1) << logn};
    q.batch = p.batch;
    tune_lookup(q);
    p.tune = q.tune;
}

// The cache holds plain-precision C2C tunings (and uneven size classes);
// split-precision C2C plans (extra != 0) need a split with an MFMA-capable
// pass and always take the defaults.
void tune_lookup(Plan& p) {
    if (p.kind == PK_C2C && p.extra != 0) {
        p.tune = default_tune(p);
        return;
    }
    std::lock_guard<std::mutex> lk(g_tune_mu);
    if (!g_tune_env_loaded) {
        g_tune_env_loaded = true;
        if (const char* path = std::getenv("NRFFT_TUNE_CACHE")) (void)load_file_locked(path);
    }
    auto it = g_tune.find(TuneKey(p.kind, p.dtype, p.nn, p.batch));
    p.tune = it != g_tune.end() ? it->second : default_tune(p);
}


bool tune_find(int kind, int dtype, const std::vector<size_t>& nn, size_t batch, Tune* out) {
    std::lock_guard<std::mutex> lk(g_tune_mu);
    if (!g_tune_env_loaded) {
        g_tune_env_loaded = true;
        if (const char* path = std::getenv("NRFFT_TUNE_CACHE")) (void)load_file_locked(path);
    }
    auto it = g_tune.find(TuneKey(kind, dtype, nn, batch));
    if (it == g_tune.end()) return false;
    *out = it->second;
    return true;
}

void tune_store(int kind, int dtype, const std::vector<size_t>& nn, size_t batch, const Tune& t) {
    std::lock_guard<std::mutex> lk(g_tune_mu);
    g_tune[TuneKey(kind, dtype, nn, batch)] = t;
}

static std::vector<Tune> candidates(int dtype, const std::vector<size_t>& nn, size_t batch) {
    std::vector<Tune> out;
    const int maxt = dtype == NRFFT_F32 ? 5 : 4;
    if (nn.size() == 1) {
        const int logn = ilog2(nn[0]);
        if (logn <= C2C_MAX_LOGL_) return {Tune{}};
        std::vector<int> chunks = {0};
        for (int c = 1; c < (int)batch; c *= 2) chunks.push_back(c);
        for (int l1 = std::max(1, logn - C2C_MAX_LOGL_); l1 <= std::min(C2C_MAX_LOGL_, logn - 1); ++l1) {
            if (std::abs(l1 - logn / 2) > 2) continue;
            for (int ta = 1; ta <= maxt; ++ta)
                for (int tb = 1; tb <= maxt; ++tb)
                    for (int c : chunks)
                        for (int xr = 0; xr <= 1; ++xr) {
                            Tune t;
                            t.split_log1 = l1;
                            t.tilog_a = ta;
                            t.tilog_b = tb;
                            t.chunk = c;
                            t.xcd_remap = xr;
                            out.push_back(t);
                        }
        }
    } else {
        for (int ta = 0; ta <= maxt + 1; ++ta) {
            Tune t;
            t.tilog_a = ta;
            out.push_back(t);
        }
    }
    return out;
}

static int time_plan(const Plan& p, void* in, void* out, int iters, hipStream_t s, hipEvent_t e0, hipEvent_t e1,
                     float* ms) {
    void* tmp = nullptr;
    if (p.tmp_elems) {
        if (hipMalloc(&tmp, p.tmp_elems * csize(p.dtype)) != hipSuccess) return NRFFT_E_NOMEM;
    }
    int rc = NRFFT_OK;
    for (int w = 0; w < 2 && !rc; ++w) rc = exec_plan(p, in, out, tmp, nullptr, 1, s);
    if (!rc) {
        (void)hipEventRecord(e0, s);
        for (int i = 0; i < iters && !rc; ++i) rc = exec_plan(p, in, out, tmp, nullptr, 1, s);
        (void)hipEventRecord(e1, s);
        (void)hipEventSynchronize(e1);
        float t = 0;
        (void)hipEventElapsedTime(&t, e0, e1);
        *ms = t / iters;
    }
    (void)hipStreamSynchronize(s);
    if (tmp) (void)hipFree(tmp);
    return rc;
}

}  // namespace nrfft

using namespace nrfft;

extern "C" {

int nrfft_plan_c2c_tuned(nrfft_plan* plan, int dtype, const size_t* nn, size_t ndim, size_t batch, const int* tune,
                         size_t ntune) {
    g_err.clear();
    if (!tune && ntune) return fail(NRFFT_E_INVALID, "null tune");
    if (ntune > 7) return fail(NRFFT_E_INVALID, "ntune > 7");
    Tune t;
    if (ntune > 0) t.split_log1 = tune[0];
    if (ntune > 1) t.tilog_a = tune[1];
    if (ntune > 2) t.tilog_b = tune[2];
    if (ntune > 3) t.chunk = tune[3];
    if (ntune > 4) t.wg_per_cu = tune[4];
    if (ntune > 5) t.xcd_remap = tune[5];
    const int prec = ntune > 6 ? tune[6] : 0;
    if (prec != 0 && prec != NRFFT_SPLIT_F16 && prec != NRFFT_SPLIT_BF16) return fail(NRFFT_E_INVALID, "precision");
    if (prec != 0 && dtype != NRFFT_F32) return fail(NRFFT_E_INVALID, "split precision needs dtype NRFFT_F32");
    if (!nn || ndim == 0) return fail(NRFFT_E_DIMS, "Invalid dimensions");
    for (size_t i = 0; i < ndim; ++i)
        if (nn[i] <= 1) return fail(NRFFT_E_DIMS, "Invalid dimension size");
    for (size_t i = 0; i < ndim; ++i)
        if (!is_pow2(nn[i])) return fail(NRFFT_E_NOT_POW2, "dimension must be a power of two");
    // The tune applies to this plan only: the cache is left as it is.
    return make_plan(plan, PK_C2C, dtype, std::vector<size_t>(nn, nn + ndim), batch, (size_t)prec, &t);
}

int nrfft_autotune_c2c(int dtype, const size_t* nn, size_t ndim, size_t batch, int iters, int* best_tune,
                       double* best_ms) {
    g_err.clear();
    if (!nn || ndim == 0 || batch == 0 || iters <= 0) return fail(NRFFT_E_INVALID, "bad argument");
    if (dtype != NRFFT_F32 && dtype != NRFFT_F64) return fail(NRFFT_E_INVALID, "dtype");
    std::vector<size_t> d(nn, nn + ndim);
    size_t elems = batch;
    for (size_t v : d) {
        if (v < 2 || !is_pow2(v)) return fail(NRFFT_E_NOT_POW2, "dimension must be a power of two");
        elems *= v;
    }
    void *in = nullptr, *out = nullptr;
    hipStream_t s = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    int rc = NRFFT_OK;
    if (hipMalloc(&in, elems * csize(dtype)) != hipSuccess || hipMalloc(&out, elems * csize(dtype)) != hipSuccess) {
        rc = fail(NRFFT_E_NOMEM, "autotune buffers");
    }
    if (!rc && (hipStreamCreate(&s) != hipSuccess || hipEventCreate(&e0) != hipSuccess ||
                hipEventCreate(&e1) != hipSuccess))
        rc = fail(NRFFT_E_HIP, "autotune stream/events");
    if (!rc && hipMemsetAsync(in, 0, elems * csize(dtype), s) != hipSuccess) rc = fail(NRFFT_E_HIP, "memset");
    Tune best;
    float best_t = 1e30f;
    bool found = false;
    if (!rc) {
        for (const Tune& t : candidates(dtype, d, batch)) {
            Plan p;
            p.kind = PK_C2C;
            p.dtype = dtype;
            p.nn = d;
            p.batch = batch;
            p.tune = t;
            if (build_plan(p) != NRFFT_OK) continue;
            float ms = 0;
            if (time_plan(p, in, out, iters, s, e0, e1, &ms) != NRFFT_OK) continue;
            if (ms < best_t) {
                best_t = ms;
                best = t;
                found = true;
            }
        }
        g_err.clear();
        if (!found) rc = fail(NRFFT_E_UNSUPPORTED, "no candidate plan ran");
    }
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    if (s) (void)hipStreamDestroy(s);
    if (in) (void)hipFree(in);
    if (out) (void)hipFree(out);
    if (rc) return rc;
    {
        std::lock_guard<std::mutex> lk(g_tune_mu);
        g_tune_env_loaded = true;
        g_tune[TuneKey(PK_C2C, dtype, d, batch)] = best;
    }
    if (best_tune) {
        best_tune[0] = best.split_log1;
        best_tune[1] = best.tilog_a;
        best_tune[2] = best.tilog_b;
        best_tune[3] = best.chunk;
        best_tune[4] = best.wg_per_cu;
        best_tune[5] = best.xcd_remap;
    }
    if (best_ms) *best_ms = best_t;
    return NRFFT_OK;
}

int nrfft_plan_tune(nrfft_plan plan, int* tune, size_t ntune) {
    g_err.clear();
    if (!plan || (!tune && ntune)) return fail(NRFFT_E_INVALID, "bad argument");
    const Tune& t = plan->plan.tune;
    const int v[6] = {t.split_log1, t.tilog_a, t.tilog_b, t.chunk, t.wg_per_cu, t.xcd_remap};
    for (size_t i = 0; i < ntune && i < 6; ++i) tune[i] = v[i];
    return NRFFT_OK;
}

int nrfft_tune_cache_save(const char* path) {
    g_err.clear();
    if (!path) return fail(NRFFT_E_INVALID, "null path");
    std::lock_guard<std::mutex> lk(g_tune_mu);
    FILE* f = std::fopen(path, "w");
    if (!f) return fail(NRFFT_E_INVALID, std::string("cannot write ") + path);
    std::fprintf(f, "# kind dtype batch ndim nn... : split_log1 tilog_a tilog_b chunk wg_per_cu xcd_remap\n");
    for (const auto& kv : g_tune) {
        const auto& k = kv.first;
        std::fprintf(f, "%d %d %zu %zu", std::get<0>(k), std::get<1>(k), std::get<3>(k), std::get<2>(k).size());
        for (size_t v : std::get<2>(k)) std::fprintf(f, " %zu", v);
        const Tune& t = kv.second;
        std::fprintf(f, " : %d %d %d %d %d %d\n", t.split_log1, t.tilog_a, t.tilog_b, t.chunk, t.wg_per_cu,
                     t.xcd_remap);
    }
    std::fclose(f);
    return NRFFT_OK;
}

int nrfft_tune_cache_load(const char* path) {
    g_err.clear();
    if (!path) return fail(NRFFT_E_INVALID, "null path");
    std::lock_guard<std::mutex> lk(g_tune_mu);
    g_tune_env_loaded = true;
    return load_file_locked(path);
}

}  // extern "C"
