// api.hip -- the extern "C" boundary (include/nrfft.h): device contexts,
// planning, execution and the host-pointer drop-in calls.
#include <algorithm>
#include <cstring>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <tuple>
#include <vector>

#include "../../include/nrfft.h"
#include "internal.hpp"

namespace nrfft {

thread_local std::string g_err;

int fail(int code, const std::string& msg) {
    g_err = msg;
    return code;
}

#define HIPCHK(x)                                                                                      \
    do {                                                                                               \
        hipError_t e_ = (x);                                                                           \
        if (e_ != hipSuccess) return fail(NRFFT_E_HIP, std::string(#x) + ": " + hipGetErrorString(e_)); \
    } while (0)

bool is_pow2(size_t n) { return n != 0 && (n & (n - 1)) == 0; }
int ilog2(size_t n) {
    int l = 0;
    while ((size_t(1) << l) < n) ++l;
    return l;
}

// ---------------------------------------------------------------------------
// Device context: twiddle tables generated in FP64 on the GPU.
// ---------------------------------------------------------------------------
__global__ void k_twiddles(c64* c64c, c64* c64f, c32* c32c, c32* c32f) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= TW_COARSE) return;
    double s, c;
    sincospi((double)i / (double)(TW_COARSE / 2), &s, &c);  // exp(2 pi i m / 2^13)
    c64c[i] = c64{c, s};
    c32c[i] = c32{(float)c, (float)s};
    sincospi((double)i / (double)(1 << (TW_LOG_VIRT - 1)), &s, &c);  // exp(2 pi i m / 2^26)
    c64f[i] = c64{c, s};
    c32f[i] = c32{(float)c, (float)s};
}

// Per-stage tables: entry e in [2L, 4L) belongs to length L = 2^(floor(log2 e) - 1);
// within L's block, stage s holds NS_s rows of R_s values exp(2 pi i r k / (NS_s R_s)).
__global__ void k_stage_twiddles(c64* t64, c32* t32) {
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= TW_STAGE_ELEMS) return;
    double s = 0.0, c = 1.0;
    if (e >= 4) {
        int lg = 0;
        while ((2 << lg) <= e) ++lg;  // 2^lg <= e < 2^(lg+1)
        const int logl = lg - 1, L = 1 << logl;
        int o = e - 2 * L;
        int ns = 1;
        const int nst = L < 16 ? 1 : (logl + 3) / 4;
        for (int st = 0; st < nst; ++st) {
            const int R = plan_radix(L, logl, nst, st);
            if (o < ns * R) {
                const int k = o / R, r = o % R;
                sincospi(2.0 * (double)(r * k) / (double)(ns * R), &s, &c);
                break;
            }
            o -= ns * R;
            ns *= R;
        }
    }
    t64[e] = c64{c, s};
    t32[e] = c32{(float)c, (float)s};
}

static std::mutex g_ctx_mu;
static std::map<int, DeviceCtx*> g_ctx;

int get_ctx(DeviceCtx** out) {
    int dev = 0;
    HIPCHK(hipGetDevice(&dev));
    std::lock_guard<std::mutex> lk(g_ctx_mu);
    auto it = g_ctx.find(dev);
    if (it != g_ctx.end()) {
        *out = it->second;
        return NRFFT_OK;
    }
    auto* c = new DeviceCtx();
    c->device = dev;
    HIPCHK(hipMalloc(&c->c64c, TW_COARSE * sizeof(c64)));
    HIPCHK(hipMalloc(&c->c64f, TW_COARSE * sizeof(c64)));
    HIPCHK(hipMalloc(&c->c32c, TW_COARSE * sizeof(c32)));
    HIPCHK(hipMalloc(&c->c32f, TW_COARSE * sizeof(c32)));
    HIPCHK(hipMalloc(&c->c64s, TW_STAGE_ELEMS * sizeof(c64)));
    HIPCHK(hipMalloc(&c->c32s, TW_STAGE_ELEMS * sizeof(c32)));
    HIPCHK(hipDeviceGetAttribute(&c->ncu, hipDeviceAttributeMultiprocessorCount, dev));
    hipLaunchKernelGGL(k_twiddles, dim3(TW_COARSE / 256), dim3(256), 0, 0, c->c64c, c->c64f, c->c32c, c->c32f);
    HIPCHK(hipGetLastError());
    hipLaunchKernelGGL(k_stage_twiddles, dim3(TW_STAGE_ELEMS / 256), dim3(256), 0, 0, c->c64s, c->c32s);
    HIPCHK(hipGetLastError());
    if (int rc = build_mfma_tables(c)) return rc;
    HIPCHK(hipDeviceSynchronize());
    g_ctx[dev] = c;
    *out = c;
    return NRFFT_OK;
}

// The current device's context if nrfft_init (or any earlier call) built
// it; never allocates.
int find_ctx(DeviceCtx** out) {
    int dev = 0;
    HIPCHK(hipGetDevice(&dev));
    std::lock_guard<std::mutex> lk(g_ctx_mu);
    auto it = g_ctx.find(dev);
    if (it == g_ctx.end()) return fail(NRFFT_E_INVALID, "device " + std::to_string(dev) + " not initialised (nrfft_init)");
    *out = it->second;
    return NRFFT_OK;
}

// ---------------------------------------------------------------------------
// Kernel lookup.
// ---------------------------------------------------------------------------
// NRFFT_C2C_16K = 1: FP32 C2C of 16384 points runs as one row pass (one
// transform per group, one group per CU) instead of a four-step.
#ifndef NRFFT_C2C_16K
#define NRFFT_C2C_16K 1
#endif
constexpr int64_t C2C_16K_MIN_BATCH = 128;

const KernelRef& c2c_lookup(int dtype, int sign, int mode, int logl, int tilog) {
    static const KernelRef none{nullptr, 0, 0, 0};
    if (NRFFT_C2C_16K && dtype == NRFFT_F32 && logl == C2C_MAX_LOGL_ + 1 && mode == MODE_JJ && tilog == 0)
        return sign > 0 ? c2c16k_lookup_f32p() : c2c16k_lookup_f32m();
    if (logl < 1 || logl > C2C_MAX_LOGL_ || tilog < 0 || tilog >= C2C_NTI_ || mode < 0 || mode > 3) return none;
    if (dtype == NRFFT_F32) return sign > 0 ? c2c_lookup_f32p(mode, logl, tilog) : c2c_lookup_f32m(mode, logl, tilog);
    return sign > 0 ? c2c_lookup_f64p(mode, logl, tilog) : c2c_lookup_f64m(mode, logl, tilog);
}

// Largest valid tile width <= 2^want_log (and <= nb0 rounded up to pow2).
int pick_tilog(int dtype, int mode, int logl, int64_t nb0, int want_log) {
    int cap = 0;
    while ((int64_t(1) << cap) < nb0 && cap < want_log) ++cap;
    for (int t = std::min(cap, want_log); t >= 0; --t)
        if (c2c_lookup(dtype, +1, mode, logl, t).fn) return t;
    return -1;
}

// ---------------------------------------------------------------------------
// Plans.
// ---------------------------------------------------------------------------
int plan_add_c2c_pass(Plan& p, int mode, int logl, int tilog, const PassDesc& d, int src, int dst) {
    Launch l;
    if (int rc = c2c_pass_launch(p, mode, logl, tilog, d, src, dst, &l)) return rc;
    p.launches.push_back(l);
    return NRFFT_OK;
}

// The launch of one C2C pass (kernel, persistent grid, descriptor); reads
// p.kind, p.dtype, p.extra and p.tune only.
int c2c_pass_launch(const Plan& p, int mode, int logl, int tilog, const PassDesc& d, int src, int dst, Launch* out) {
    const int prec = (p.kind == PK_C2C && p.dtype == NRFFT_F32) ? (int)p.extra : 0;
    const bool use_mfma = prec != 0 && mfma_lookup(prec, +1, mode, logl).fn != nullptr;
    const KernelRef& kp = use_mfma ? mfma_lookup(prec, +1, mode, logl) : c2c_lookup(p.dtype, +1, mode, logl, tilog);
    const KernelRef& km = use_mfma ? mfma_lookup(prec, -1, mode, logl) : c2c_lookup(p.dtype, -1, mode, logl, tilog);
    if (!kp.fn || !km.fn) return fail(NRFFT_E_UNSUPPORTED, "no kernel for pass (mode " + std::to_string(mode) +
                                                               ", log2 L " + std::to_string(logl) + ")");
    Launch l;
    l.kind = LK_C2C;
    l.fn[0] = kp.fn;
    l.fn[1] = km.fn;
    l.nt = kp.nt;
    l.prec = use_mfma ? prec : 0;
    const int64_t tiles0 = (d.nb0 + kp.ti - 1) / kp.ti;
    const int64_t ntiles = tiles0 * d.nb1 * d.nb2;
    if (ntiles <= 0) return fail(NRFFT_E_UNSUPPORTED, "empty pass");
    // Persistent grid: as many groups as are resident at once (or the tuned
    // groups per CU), never more than there are tiles.
    DeviceCtx* ctx = nullptr;
    int rc = get_ctx(&ctx);
    if (rc) return rc;
    int per_cu = 0;
    HIPCHK(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kp.fn, kp.nt, 0));
    if (per_cu < 1) per_cu = 1;
    if (p.tune.wg_per_cu > 0 && p.tune.wg_per_cu < per_cu) per_cu = p.tune.wg_per_cu;
    const int64_t grid = std::min<int64_t>(ntiles, int64_t(per_cu) * ctx->ncu);
    // In-tile offsets must fit 32 bits.
    const int64_t span_in = int64_t(kp.ti - 1) * d.si0 + int64_t((1 << logl) - 1) * d.esi;
    const int64_t span_out = int64_t(kp.ti - 1) * d.so0 + int64_t((1 << logl) - 1) * d.eso;
    if (span_in > 0x7fffffffLL || span_out > 0x7fffffffLL) return fail(NRFFT_E_UNSUPPORTED, "tile span too large");
    l.grid = (unsigned)grid;
    l.desc = d;
    l.desc.xcd_remap = p.tune.xcd_remap > 0 ? 1 : 0;
    // DMA-staged column pass: full 16-wide tiles of 128-byte rows whose
    // 16-byte pieces stay 16-byte aligned (checked for the pointer at launch).
    if (!use_mfma && p.dtype == NRFFT_F32 && mode == MODE_TT && kp.ti == DMA_TI && d.si0 == 1 &&
        d.nb0 % DMA_TI == 0 && d.esi % 2 == 0 && d.si1 % 2 == 0 && d.si2 % 2 == 0 && span_out < (1LL << 28) &&
        NRFFT_PASS_DMA) {
        const DmaRef& dp = dma_lookup_f32p(logl);
        const DmaRef& dm = dma_lookup_f32m(logl);
        if (dp.fn && dm.fn) {
            int dcu = 0;
            HIPCHK(hipOccupancyMaxActiveBlocksPerMultiprocessor(&dcu, dp.fn, dp.nt, 0));
            if (dcu >= 1) {
                l.fn_dma[0] = dp.fn;
                l.fn_dma[1] = dm.fn;
                l.nt_dma = dp.nt;
                l.grid_dma = (unsigned)std::min<int64_t>(ntiles, int64_t(dcu) * ctx->ncu);
            }
        }
    }
    l.src = src;
    l.dst = dst;
    l.lds = kp.lds_bytes;
    l.logl = logl;
    l.ti = kp.ti;
    l.mode = mode;
    *out = l;
    return NRFFT_OK;
}

// Element size of one complex value.
size_t csize(int dtype) { return dtype == NRFFT_F32 ? 8 : 16; }

// Preferred tile width (log2) for column-mapped passes.
int col_tilog(int dtype) { return dtype == NRFFT_F32 ? 4 : 3; }

// NRFFT_C2C_OVERLAP = 1: batched 1-D C2C plans of several chunks run each
// chunk's second pass on a second stream, beside the next chunk's first
// (exec_overlap).  Results are bit-identical, but on MI355X it is slower:
// FP32 64 x 2^20 0.569 ms against 0.441, FP64 32 x 2^20 0.532 against
// 0.409, FP32 256 x 2^18 0.525 against 0.434, FP32 32 x 2^22 1.570 against
// 1.208.  Both passes are bound by the same HBM stream, so running them
// side by side adds no bandwidth, and the two scratch halves push the
// intermediate out of the MALL.  Off.
#ifndef NRFFT_C2C_OVERLAP
#define NRFFT_C2C_OVERLAP 0
#endif

static int use_map_kernel(const Plan& p, const KernelRef& km, const PassDesc& d, Launch* l);

// Three-pass C2C of N = N1 * Na * Mb (M = Na * Mb), chunk by chunk:
//   pass 1: N1-point FFTs down the columns m of x[m + M n1] (stride M), times
//           W_N^{k1 m} -> T[k1][m];
//   pass 2: in place, the four-step's first pass of each row k1 as an
//           M-point transform: Na-point FFTs down its columns mb (stride Mb),
//           times W_M^{ka mb} -> T[k1][ka][mb];
//   pass 3: Mb-point FFTs along the rows (k1, ka), tiled over k1, stored to
//           X[k1 + N1 (ka + Na kb)].
// X[k1 + N1 kM] is the M-point transform of row k1 (the outer four-step) and
// kM = ka + Na kb is the inner four-step's output order.  Every pass is a
// column pass or a row pass with 16-wide (FP32) / 8-wide (FP64) tiles.
// lmap != LMAP_PLAIN: pass 1 reads src through that load map (sig_reals
// reals per signal, argument map_arg); the launches run with isign `force`
// and only for calls of direction `only` when those are non-zero.
int plan_c2c_3pass(Plan& p, int logn, int64_t batch, int src, int dst, size_t* tmp_elems, int lmap, int64_t sig_reals,
                   int64_t map_arg, int force, int only) {
    int l1, la, lb;
    three_pass_split(logn, &l1, &la, &lb);
    const int64_t N = int64_t(1) << logn, N1 = int64_t(1) << l1, Na = int64_t(1) << la, Mb = int64_t(1) << lb;
    const int64_t M = Na * Mb;
    const int want = col_tilog(p.dtype);
    const int t1 = pick_tilog(p.dtype, MODE_TT, l1, M, want), t2 = pick_tilog(p.dtype, MODE_TT, la, Mb, want);
    const int t3 = pick_tilog(p.dtype, MODE_JT, lb, N1, want);
    if (t1 < 0 || t2 < 0 || t3 < 0) return fail(NRFFT_E_UNSUPPORTED, "no three-pass kernels");
    const int64_t chunk = (p.tune.chunk > 0 && p.tune.chunk < batch) ? p.tune.chunk : batch;
    for (int64_t c0 = 0; c0 < batch; c0 += chunk) {
        const int64_t cb = std::min<int64_t>(chunk, batch - c0);
        PassDesc a{};
        a.nb0 = M;
        a.si0 = 1;
        a.so0 = 1;
        a.esi = (int32_t)M;
        a.eso = (int32_t)M;
        a.nb1 = cb;
        a.si1 = lmap != LMAP_PLAIN ? sig_reals : N;
        a.so1 = N;
        a.nb2 = 1;
        a.tw_logn = logn;
        Launch l1a;
        if (int rc = c2c_pass_launch(p, MODE_TT, l1, t1, a, src, B_TMP, &l1a)) return rc;
        if (lmap != LMAP_PLAIN) {
            if (int rc = use_map_kernel(p, map_lookup(p.dtype, lmap, l1, t1), a, &l1a)) return rc;
            l1a.desc.map_arg = map_arg;
            l1a.off_src_real = c0 * sig_reals;
        } else {
            l1a.off_src = c0 * N;
        }
        PassDesc b{};
        b.nb0 = Mb;
        b.si0 = 1;
        b.so0 = 1;
        b.esi = (int32_t)Mb;
        b.eso = (int32_t)Mb;
        b.nb1 = N1;
        b.si1 = M;
        b.so1 = M;
        b.nb2 = cb;
        b.si2 = N;
        b.so2 = N;
        b.tw_logn = la + lb;
        Launch l2;
        if (int rc = c2c_pass_launch(p, MODE_TT, la, t2, b, B_TMP, B_TMP, &l2)) return rc;
        PassDesc c{};
        c.nb0 = N1;
        c.si0 = (int32_t)M;
        c.esi = 1;
        c.so0 = 1;
        c.eso = (int32_t)(N1 * Na);
        c.nb1 = Na;
        c.si1 = Mb;
        c.so1 = N1;
        c.nb2 = cb;
        c.si2 = N;
        c.so2 = N;
        c.nt_out = 1;  // the plan's result, not re-read
        Launch l3;
        if (int rc = c2c_pass_launch(p, MODE_JT, lb, t3, c, B_TMP, dst, &l3)) return rc;
        l3.off_dst = c0 * N;
        for (Launch* l : {&l1a, &l2, &l3}) {
            if (force) l->args[6] = force;
            if (only) l->args[7] = only;
            p.launches.push_back(*l);
        }
    }
    *tmp_elems = std::max<size_t>(*tmp_elems, size_t(N) * size_t(chunk));
    return NRFFT_OK;
}

// 1-D batched C2C of length N = 2^p over `batch` contiguous transforms, from
// buffer src to buffer dst (B_TMP used as the intermediate of two-pass sizes).
// Output element k of transform b lands at dst[b*N*ostride_b... ] -- contiguous.
int plan_c2c_1d(Plan& p, int logn, int64_t batch, int src, int dst, size_t* tmp_elems) {
    const int64_t N = int64_t(1) << logn;
    if (logn == 0) {
        // nn == 1: the identity; a copy when out of place.
        Launch l;
        l.kind = LK_COPY;
        l.src = src;
        l.dst = dst;
        l.bytes = size_t(batch) * csize(p.dtype);
        p.launches.push_back(l);
        return NRFFT_OK;
    }
    // (FP32 16384 points: one pass as well for batches of C2C_16K_MIN_BATCH
    // and more -- one transform per CU, so a small batch keeps the
    // four-step, which spreads each transform over many CUs -- except for
    // split-precision plans.)
    if (logn <= C2C_MAX_LOGL_ || (logn == C2C_MAX_LOGL_ + 1 && batch >= C2C_16K_MIN_BATCH &&
                                  !(p.kind == PK_C2C && p.extra != 0) && c2c_lookup(p.dtype, +1, MODE_JJ, logn, 0).fn)) {
        PassDesc d{};
        d.nb0 = batch;
        d.nb1 = 1;
        d.nb2 = 1;
        d.si0 = (int32_t)N;
        d.so0 = (int32_t)N;
        d.esi = 1;
        d.eso = 1;
        if (N * std::min<int64_t>(batch, 256) > 0x7fffffffLL) return fail(NRFFT_E_UNSUPPORTED, "too large");
        return plan_add_c2c_pass(p, MODE_JJ, logn, 0, d, src, dst);
    }
    if (logn > 2 * C2C_MAX_LOGL_) return fail(NRFFT_E_UNSUPPORTED, "transform longer than 2^26");
    // (Split-precision plans keep the two-pass form and its MFMA pass lengths.)
    if (logn >= C2C_3PASS_MIN_LOGN && !(p.kind == PK_C2C && p.extra != 0))
        return plan_c2c_3pass(p, logn, batch, src, dst, tmp_elems);
    FourStep fs;
    if (int rc = fourstep_setup(p, logn, &fs)) return rc;
    // Batch chunking: pass A and pass B of one chunk run back to back so the
    // chunk's intermediate can stay in the last-level cache (256 MB MALL).
    const int64_t chunk = (p.tune.chunk > 0 && p.tune.chunk < batch) ? p.tune.chunk : batch;
    // A C2C plan of several chunks alone in its plan may overlap chunk k's
    // pass B with chunk k+1's pass A (two scratch halves).
    const bool overlap = NRFFT_C2C_OVERLAP && p.kind == PK_C2C && p.launches.empty() && chunk < batch;
    *tmp_elems = std::max<size_t>(*tmp_elems, size_t(N) * size_t(chunk) * (overlap ? 2 : 1));
    for (int64_t c0 = 0, k = 0; c0 < batch; c0 += chunk, ++k) {
        const int64_t cb = std::min<int64_t>(chunk, batch - c0);
        Launch la, lb;
        if (int rc = fourstep_launches(p, fs, cb, src, dst, &la, &lb)) return rc;
        la.off_src = c0 * N;
        lb.off_dst = c0 * N;
        if (overlap) la.off_dst = lb.off_src = (k & 1) * N * chunk;
        p.launches.push_back(la);
        p.launches.push_back(lb);
    }
    p.overlap = overlap;
    return NRFFT_OK;
}

// Four-step split N = N1 * N2 (x viewed as [n1][n2]) and the two passes'
// tile widths.
int fourstep_setup(const Plan& p, int logn, FourStep* fs) {
    int l1 = logn / 2, l2 = logn - l1;
    if (p.tune.split_log1 > 0 && p.tune.split_log1 < logn && logn - p.tune.split_log1 <= C2C_MAX_LOGL_ &&
        p.tune.split_log1 <= C2C_MAX_LOGL_) {
        l1 = p.tune.split_log1;
        l2 = logn - l1;
    }
    if (p.kind == PK_C2C && p.extra != 0 && p.dtype == NRFFT_F32 && logn <= 20 && p.tune.split_log1 <= 0) {
        // split-precision plans: both passes on MFMA kernels (L <= 1024)
        l1 = logn / 2;
        l2 = logn - l1;
    }
    if (l2 > C2C_MAX_LOGL_ || l1 > C2C_MAX_LOGL_) return fail(NRFFT_E_UNSUPPORTED, "no four-step split");
    const int want_a = p.tune.tilog_a >= 0 ? p.tune.tilog_a : col_tilog(p.dtype);
    const int want_b = p.tune.tilog_b >= 0 ? p.tune.tilog_b : col_tilog(p.dtype);
    fs->logn = logn;
    fs->l1 = l1;
    fs->l2 = l2;
    fs->ta = pick_tilog(p.dtype, MODE_TT, l1, int64_t(1) << l2, want_a);
    fs->tb = pick_tilog(p.dtype, MODE_JT, l2, int64_t(1) << l1, want_b);
    if (fs->ta < 0 || fs->tb < 0) return fail(NRFFT_E_UNSUPPORTED, "no four-step kernel");
    return NRFFT_OK;
}

// Swaps a load-mapped kernel of the same tile shape into a pass launch that
// c2c_pass_launch sized for the plain kernel (the plan forces the map's
// isign).
static int use_map_kernel(const Plan& p, const KernelRef& km, const PassDesc& d, Launch* l) {
    if (!km.fn || km.ti != l->ti || km.nt != l->nt) return fail(NRFFT_E_UNSUPPORTED, "no load-mapped first pass");
    DeviceCtx* ctx = nullptr;
    if (int rc = get_ctx(&ctx)) return rc;
    int per_cu = 0;
    HIPCHK(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, km.fn, km.nt, 0));
    if (per_cu < 1) per_cu = 1;
    if (p.tune.wg_per_cu > 0 && p.tune.wg_per_cu < per_cu) per_cu = p.tune.wg_per_cu;
    const int64_t ntiles = ((d.nb0 + km.ti - 1) / km.ti) * d.nb1 * d.nb2;
    l->fn[0] = l->fn[1] = km.fn;
    l->fn_dma[0] = l->fn_dma[1] = nullptr;
    l->grid = (unsigned)std::min<int64_t>(ntiles, int64_t(per_cu) * ctx->ncu);
    l->lds = km.lds_bytes;
    return NRFFT_OK;
}

// Pass A and pass B of one chunk of cb transforms (offsets left at 0).
int fourstep_launches(const Plan& p, const FourStep& fs, int64_t cb, int src, int dst, Launch* la, Launch* lb,
                      int lmap, int64_t sig_reals) {
    const int64_t N = int64_t(1) << fs.logn, N1 = int64_t(1) << fs.l1, N2 = int64_t(1) << fs.l2;
    // Pass A: FFT over n1 (stride N2), tile over n2, twiddle W_N^{k1 n2}; -> tmp[k1][n2].
    PassDesc a{};
    a.nb0 = N2;
    a.si0 = 1;
    a.so0 = 1;
    a.esi = (int32_t)N2;
    a.eso = (int32_t)N2;
    a.nb1 = cb;
    a.si1 = N;
    a.so1 = N;
    a.nb2 = 1;
    a.tw_logn = fs.logn;
    a.tw_b0_shift = 0;
    if (lmap != LMAP_PLAIN) a.si1 = sig_reals;
    if (int rc = c2c_pass_launch(p, MODE_TT, fs.l1, fs.ta, a, src, B_TMP, la)) return rc;
    if (lmap != LMAP_PLAIN) {
        if (int rc = use_map_kernel(p, map_lookup(p.dtype, lmap, fs.l1, fs.ta), a, la)) return rc;
    }
    // Pass B: FFT over n2 (contiguous rows of tmp), rows k1 tiled; X[k2 N1 + k1].
    PassDesc b{};
    b.nb0 = N1;
    b.si0 = (int32_t)N2;
    b.esi = 1;
    b.so0 = 1;
    b.eso = (int32_t)N1;
    b.nb1 = cb;
    b.si1 = N;
    b.so1 = N;
    b.nb2 = 1;
    // The transform's result: stored non-temporally, so the writes do not
    // evict the next chunk's input and intermediate from L2 / MALL
    // (64 x 2^20 FP32: 0.487 -> 0.454 ms on MI355X).  Whatever reads it
    // next (the caller, or a finishing kernel of a real or spectral plan)
    // runs after the last chunk.
    b.nt_out = 1;
    return c2c_pass_launch(p, MODE_JT, fs.l2, fs.tb, b, B_TMP, dst, lb);
}

// The C2C of length N = 2^logn over `batch` signals whose first pass reads
// src through load map lmap (sig_reals reals per signal, map argument
// map_arg), chunk by chunk, into dst (N complex values per signal); the
// launches run with isign `force` and only for calls of direction `only`
// (0 = both).  In place (src == dst) is safe: a chunk's first pass has read
// all of its signals before its second pass writes them.  The caller checks
// map_fits first.
int plan_c2c_mapped(Plan& p, int logn, int64_t batch, int src, int dst, int lmap, int64_t sig_reals, int64_t map_arg,
                    int force, int only, size_t* tmp_elems) {
    if (logn >= FUSED_3PASS_MIN_LOGN)
        return plan_c2c_3pass(p, logn, batch, src, dst, tmp_elems, lmap, sig_reals, map_arg, force, only);
    FourStep fs;
    if (int rc = fourstep_setup(p, logn, &fs)) return rc;
    const int64_t N = int64_t(1) << logn;
    const int64_t chunk = (p.tune.chunk > 0 && p.tune.chunk < batch) ? p.tune.chunk : batch;
    for (int64_t c0 = 0; c0 < batch; c0 += chunk) {
        const int64_t cb = std::min<int64_t>(chunk, batch - c0);
        Launch la, lb;
        if (int rc = fourstep_launches(p, fs, cb, src, dst, &la, &lb, lmap, sig_reals)) return rc;
        la.desc.map_arg = map_arg;
        la.off_src_real = c0 * sig_reals;
        lb.off_dst = c0 * N;
        for (Launch* l : {&la, &lb}) {
            l->args[6] = force;
            l->args[7] = only;
            p.launches.push_back(*l);
        }
    }
    *tmp_elems = std::max<size_t>(*tmp_elems, size_t(N) * size_t(chunk));
    return NRFFT_OK;
}

// Strided axes of >= 4096 points take the two-pass four-step along the axis
// when at least this many columns sit side by side (inner), so pass B gets
// 16-wide tiles; narrower axes run as one column pass.
#ifndef NRFFT_STRIDED_4STEP_MIN_INNER
#define NRFFT_STRIDED_4STEP_MIN_INNER 16
#endif
constexpr int64_t STRIDED_4STEP_MIN_INNER = NRFFT_STRIDED_4STEP_MIN_INNER;
constexpr int64_t STRIDED_4STEP_CHUNK_BYTES = int64_t(128) << 20;

// n-D (C order) batched C2C; axes processed fastest first.
int plan_c2c_nd(Plan& p, const std::vector<size_t>& nn, int64_t batch) {
    size_t tmp = 0;
    const size_t ndim = nn.size();
    if (ndim == 1) {
        int rc = plan_c2c_1d(p, ilog2(nn[0]), batch, B_IN, B_OUT, &tmp);
        p.tmp_elems = tmp;
        return rc;
    }
    int64_t total = 1;
    for (size_t v : nn) total *= (int64_t)v;
    bool first = true;
    for (size_t dd = ndim; dd-- > 0;) {
        const int64_t L = (int64_t)nn[dd];
        const int logl = ilog2(nn[dd]);
        int64_t inner = 1;
        for (size_t e = dd + 1; e < ndim; ++e) inner *= (int64_t)nn[e];
        const int64_t outer = (total / (L * inner)) * batch;
        const int src = first ? B_IN : B_OUT;
        first = false;
        if (logl > C2C_MAX_LOGL_) {
            if (inner != 1) return fail(NRFFT_E_UNSUPPORTED, "strided axis longer than 8192");
            int rc = plan_c2c_1d(p, logl, outer, src, B_OUT, &tmp);
            if (rc) return rc;
            continue;
        }
        PassDesc d{};
        if (inner == 1) {
            d.nb0 = outer;
            d.si0 = (int32_t)L;
            d.so0 = (int32_t)L;
            d.esi = 1;
            d.eso = 1;
            d.nb1 = 1;
            d.nb2 = 1;
            int rc = plan_add_c2c_pass(p, MODE_JJ, logl, 0, d, src, B_OUT);
            if (rc) return rc;
        } else if (logl >= 12 && inner >= STRIDED_4STEP_MIN_INNER && !p.strided_one_pass) {
            // A long strided axis as a four-step along the axis, L = L1 L2:
            // pass A transforms over l1 (element stride L2 inner) and applies
            // W_L^{k1 l2}, pass B over l2 (stride inner) into natural order
            // k1 + L1 k2.  Both tile contiguous columns 16 wide, where one pass
            // over L >= 4096 points gets 2- or 4-wide tiles (16-32-byte rows:
            // FP32 4096 x 4096 column pass 1.9 TB/s, 8192 x 2048 0.9 TB/s).
            // The pass-A tile index is b0 = l2 inner + column, so the kernel's
            // four-step twiddle W^{k1 (b0 >> log2 inner)} is W_L^{k1 l2}.
            // The outer slabs run in chunks (pass A then pass B per chunk) so
            // the workspace stays near STRIDED_4STEP_CHUNK_BYTES (one slab at
            // least) and a chunk's intermediate can stay in the MALL.
            if (L * inner > 0x7fffffffLL) return fail(NRFFT_E_UNSUPPORTED, "axis span too large");
            const int l1 = logl / 2, l2 = logl - l1;
            const int64_t L1 = int64_t(1) << l1, L2 = int64_t(1) << l2;
            const int want = p.tune.tilog_a >= 0 ? p.tune.tilog_a : col_tilog(p.dtype);
            const int64_t slab = L * inner;
            const int64_t chunk =
                std::max<int64_t>(1, std::min<int64_t>(outer, STRIDED_4STEP_CHUNK_BYTES / (slab * (int64_t)csize(p.dtype))));
            for (int64_t c0 = 0; c0 < outer; c0 += chunk) {
                const int64_t cb = std::min<int64_t>(chunk, outer - c0);
                PassDesc a{};
                a.nb0 = L2 * inner;
                a.si0 = 1;
                a.so0 = 1;
                a.esi = (int32_t)(L2 * inner);
                a.eso = (int32_t)(L2 * inner);
                a.nb1 = cb;
                a.si1 = slab;
                a.so1 = slab;
                a.nb2 = 1;
                a.tw_logn = logl;
                a.tw_b0_shift = ilog2((size_t)inner);
                const int ta = pick_tilog(p.dtype, MODE_TT, l1, a.nb0, want);
                if (ta < 0) return fail(NRFFT_E_UNSUPPORTED, "no column kernel");
                int rc = plan_add_c2c_pass(p, MODE_TT, l1, ta, a, src, B_TMP);
                if (rc) return rc;
                p.launches.back().off_src = c0 * slab;
                PassDesc b{};
                b.nb0 = inner;
                b.si0 = 1;
                b.so0 = 1;
                b.esi = (int32_t)inner;
                b.eso = (int32_t)(L1 * inner);
                b.nb1 = L1;
                b.si1 = L2 * inner;
                b.so1 = inner;
                b.nb2 = cb;
                b.si2 = slab;
                b.so2 = slab;
                const int tb = pick_tilog(p.dtype, MODE_TT, l2, inner, want);
                if (tb < 0) return fail(NRFFT_E_UNSUPPORTED, "no column kernel");
                if ((rc = plan_add_c2c_pass(p, MODE_TT, l2, tb, b, B_TMP, B_OUT))) return rc;
                p.launches.back().off_dst = c0 * slab;
            }
            tmp = std::max<size_t>(tmp, size_t(chunk) * size_t(slab));
        } else {
            if (L * inner > 0x7fffffffLL) return fail(NRFFT_E_UNSUPPORTED, "axis span too large");
            d.nb0 = inner;
            d.si0 = 1;
            d.so0 = 1;
            d.esi = (int32_t)inner;
            d.eso = (int32_t)inner;
            d.nb1 = outer;
            d.si1 = L * inner;
            d.so1 = L * inner;
            d.nb2 = 1;
            int t = pick_tilog(p.dtype, MODE_TT, logl, inner, p.tune.tilog_a >= 0 ? p.tune.tilog_a : col_tilog(p.dtype));
            if (t < 0) return fail(NRFFT_E_UNSUPPORTED, "no column kernel");
            int rc = plan_add_c2c_pass(p, MODE_TT, logl, t, d, src, B_OUT);
            if (rc) return rc;
        }
    }
    // Later axis passes re-read what earlier ones wrote: plain stores only.
    for (auto& l : p.launches) l.desc.nt_out = 0;
    p.tmp_elems = tmp;
    return NRFFT_OK;
}

// ---------------------------------------------------------------------------
// Execution.
// ---------------------------------------------------------------------------
// Overlapped execution of a plan with p.overlap: launches (A_k, B_k) per
// chunk k, A_k writing scratch half k mod 2 and B_k reading it.  A_k runs on
// the caller's stream s0, B_k on the device's auxiliary stream s1 (default
// priority, non-blocking), ordered by events: B_k after A_k, A_k after
// B_{k-2} (the scratch half it overwrites), and s0 after the last B, so
// whatever the caller enqueues next on s0 sees the whole result.  Events
// and the stream are per device, reused under one lock while enqueuing.
struct OverlapCtx {
    hipStream_t s1 = nullptr;
    std::vector<hipEvent_t> ev;  // [2k] after A_k, [2k+1] after B_k
};
static std::mutex g_ov_mu;
static std::map<int, OverlapCtx> g_ov;

static int exec_overlap(const Plan& p, void* const* bufs, int isign, hipStream_t s0, DeviceCtx* ctx) {
    std::lock_guard<std::mutex> lk(g_ov_mu);
    OverlapCtx& o = g_ov[ctx->device];
    if (!o.s1) HIPCHK(hipStreamCreateWithFlags(&o.s1, hipStreamNonBlocking));
    const size_t nl = p.launches.size();
    while (o.ev.size() < nl) {
        hipEvent_t e = nullptr;
        HIPCHK(hipEventCreateWithFlags(&e, hipEventDisableTiming));
        o.ev.push_back(e);
    }
    for (size_t k = 0; 2 * k + 1 < nl; ++k) {
        if (k >= 2) HIPCHK(hipStreamWaitEvent(s0, o.ev[2 * (k - 2) + 1], 0));
        if (int rc = exec_launch(p, p.launches[2 * k], bufs, isign, s0, ctx)) return rc;
        HIPCHK(hipEventRecord(o.ev[2 * k], s0));
        HIPCHK(hipStreamWaitEvent(o.s1, o.ev[2 * k], 0));
        if (int rc = exec_launch(p, p.launches[2 * k + 1], bufs, isign, o.s1, ctx)) return rc;
        HIPCHK(hipEventRecord(o.ev[2 * k + 1], o.s1));
    }
    HIPCHK(hipStreamWaitEvent(s0, o.ev[nl - 1], 0));
    return NRFFT_OK;
}

int exec_plan(const Plan& p, const void* in, void* out, void* tmp, const void* aux, int isign, hipStream_t stream,
              int first, int count, void* aux2) {
    DeviceCtx* ctx = nullptr;
    int rc = get_ctx(&ctx);
    if (rc) return rc;
    int idx = -1;
    void* tmp2 = tmp ? static_cast<char*>(tmp) + p.tmp2_offset * csize(p.dtype) : nullptr;
    void* bufs[B_COUNT] = {const_cast<void*>(in), out, tmp, const_cast<void*>(aux), tmp2, aux2};
    if (p.overlap && first == 0 && count < 0) return exec_overlap(p, bufs, isign, stream, ctx);
    for (const Launch& l : p.launches) {
        ++idx;
        if (idx < first || (count >= 0 && idx >= first + count)) continue;
        if ((rc = exec_launch(p, l, bufs, isign, stream, ctx))) return rc;
    }
    return NRFFT_OK;
}

// Enqueue one launch of plan p on the buffers bufs (indexed by BufId).
int exec_launch(const Plan& p, const Launch& l, void* const* bufs, int isign, hipStream_t stream, DeviceCtx* ctx) {
    if (l.args[7] != 0 && ((l.args[7] > 0) != (isign > 0))) return NRFFT_OK;
    const int eff = l.args[6] != 0 ? (int)l.args[6] : isign;
    const int si = eff > 0 ? 0 : 1;
    void* a = static_cast<char*>(bufs[l.src]) + l.off_src * (int64_t)csize(p.dtype) +
              l.off_src_real * (int64_t)(csize(p.dtype) / 2);
    void* b = static_cast<char*>(bufs[l.dst]) + l.off_dst * (int64_t)csize(p.dtype);
    if (l.kind == LK_COPY) {
        if (a != b) HIPCHK(hipMemcpyAsync(b, a, l.bytes, hipMemcpyDeviceToDevice, stream));
        return NRFFT_OK;
    }
    if (l.kind == LK_C2C) {
        PassDesc d = l.desc;
        if (p.dtype == NRFFT_F32 && l.prec) {
            Twiddles<float> tw = ctx_twiddles<float>(ctx);
            const void* amat = l.prec == NRFFT_SPLIT_F16 ? ctx->mfma_f16 : ctx->mfma_bf16;
            void* args[] = {&a, &b, &d, &tw, &amat};
            HIPCHK(hipLaunchKernel(l.fn[si], dim3(l.grid), dim3(l.nt), args, 0, stream));
        } else if (p.dtype == NRFFT_F32) {
            Twiddles<float> tw = ctx_twiddles<float>(ctx);
            void* args[] = {&a, &b, &d, &tw};
            if (l.fn_dma[si] && (reinterpret_cast<uintptr_t>(a) & 15) == 0)
                HIPCHK(hipLaunchKernel(l.fn_dma[si], dim3(l.grid_dma), dim3(l.nt_dma), args, 0, stream));
            else
                HIPCHK(hipLaunchKernel(l.fn[si], dim3(l.grid), dim3(l.nt), args, 0, stream));
        } else {
            Twiddles<double> tw = ctx_twiddles<double>(ctx);
            void* args[] = {&a, &b, &d, &tw};
            HIPCHK(hipLaunchKernel(l.fn[si], dim3(l.grid), dim3(l.nt), args, 0, stream));
        }
        return NRFFT_OK;
    }
    return l.kind == LK_SPEC ? exec_spectral(p, l, bufs, eff, stream, ctx) : exec_special(p, l, bufs, eff, stream, ctx);
}

// Plan cache for the host-pointer calls (plans without workspace).
static std::mutex g_plan_mu;
static std::map<std::tuple<int, int, std::vector<size_t>, size_t, size_t>, std::shared_ptr<Plan>> g_plans;

int cached_plan(int kind, int dtype, const std::vector<size_t>& nn, size_t batch, size_t extra,
                std::shared_ptr<Plan>* out) {
    int dev = 0;
    HIPCHK(hipGetDevice(&dev));
    auto key = std::make_tuple(kind * 64 + dev, dtype, nn, batch, extra);
    {
        std::lock_guard<std::mutex> lk(g_plan_mu);
        auto it = g_plans.find(key);
        if (it != g_plans.end()) {
            *out = it->second;
            return NRFFT_OK;
        }
    }
    auto p = std::make_shared<Plan>();
    p->kind = kind;
    p->dtype = dtype;
    p->nn = nn;
    p->batch = batch;
    p->extra = extra;
    p->device = dev;
    tune_lookup(*p);
    int rc = build_plan(*p);
    if (rc) return rc;
    std::lock_guard<std::mutex> lk(g_plan_mu);
    g_plans[key] = p;
    *out = p;
    return NRFFT_OK;
}

int build_plan(Plan& p) {
    switch (p.kind) {
        case PK_C2C:
            return plan_c2c_nd(p, p.nn, (int64_t)p.batch);
        case PK_TWOFFT:
        case PK_COSFT1:
        case PK_COSFT2:
        case PK_RLFT3:
            return build_plan_spectral(p);
        default:
            return build_plan_real(p);
    }
}

int dev_alloc(DevBuf& b, size_t bytes, hipStream_t s) {
    b.s = s;
    if (bytes == 0) return NRFFT_OK;
    hipError_t e = hipMallocAsync(&b.p, bytes, s);
    if (e != hipSuccess) {
        b.p = nullptr;
        return fail(NRFFT_E_NOMEM, std::string("hipMallocAsync: ") + hipGetErrorString(e));
    }
    return NRFFT_OK;
}

// Host-pointer C2C (four1 / fourn / batch): in place on host data.
int host_c2c(int dtype, void* data, const std::vector<size_t>& nn, size_t batch, int isign) {
    if (const size_t chunk = host_pipe_chunk(dtype, nn, batch)) return host_c2c_pipelined(dtype, data, nn, batch, chunk, isign);
    std::shared_ptr<Plan> p;
    int rc = cached_plan(PK_C2C, dtype, nn, batch, 0, &p);
    if (rc) return rc;
    size_t elems = batch;
    for (size_t v : nn) elems *= v;
    const size_t bytes = elems * csize(dtype);
    hipStream_t s = hipStreamPerThread;
    DevBuf buf, tmp;
    if ((rc = dev_alloc(buf, bytes, s))) return rc;
    if ((rc = dev_alloc(tmp, p->tmp_elems * csize(dtype), s))) return rc;
    HIPCHK(hipMemcpyAsync(buf.p, data, bytes, hipMemcpyHostToDevice, s));
    if ((rc = exec_plan(*p, buf.p, buf.p, tmp.p, nullptr, isign, s))) return rc;
    HIPCHK(hipMemcpyAsync(data, buf.p, bytes, hipMemcpyDeviceToHost, s));
    HIPCHK(hipStreamSynchronize(s));
    return NRFFT_OK;
}

int make_plan(nrfft_plan* plan, int kind, int dtype, std::vector<size_t> nn, size_t batch, size_t extra,
              const Tune* explicit_tune) {
    if (!plan) return fail(NRFFT_E_INVALID, "null plan pointer");
    *plan = nullptr;
    if (dtype != NRFFT_F32 && dtype != NRFFT_F64) return fail(NRFFT_E_INVALID, "dtype");
    if (batch == 0) return fail(NRFFT_E_EMPTY, "batch == 0");
    auto* p = new (std::nothrow) nrfft_plan_s();
    if (!p) return fail(NRFFT_E_NOMEM, "plan");
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) {
        delete p;
        return fail(NRFFT_E_HIP, hipGetErrorString(e));
    }
    p->plan.kind = kind;
    p->plan.dtype = dtype;
    p->plan.nn = std::move(nn);
    p->plan.batch = batch;
    p->plan.extra = extra;
    p->plan.device = dev;
    if (explicit_tune) p->plan.tune = *explicit_tune;
    else tune_lookup(p->plan);
    int rc = build_plan(p->plan);
    if (!rc && p->plan.tmp_elems) {
        size_t bytes = p->plan.tmp_elems * csize(dtype);
        e = hipMalloc(&p->tmp, bytes);
        if (e != hipSuccess && kind == PK_C2C && p->plan.nn.size() > 1 && !p->plan.strided_one_pass) {
            // The four-step of a long strided axis needs a workspace; without
            // one, such axes run as single column passes.
            (void)hipGetLastError();
            Plan q = p->plan;
            q.launches.clear();
            q.tmp_elems = 0;
            q.strided_one_pass = true;
            rc = build_plan(q);
            if (!rc) {
                p->plan = std::move(q);
                bytes = p->plan.tmp_elems * csize(dtype);
                e = bytes ? hipMalloc(&p->tmp, bytes) : hipSuccess;
            }
        }
        if (rc) {
        } else if (e != hipSuccess) {
            (void)hipGetLastError();
            rc = fail(NRFFT_E_NOMEM, std::string("workspace: ") + hipGetErrorString(e));
        } else {
            p->tmp_bytes = bytes;
        }
    }
    if (rc) {
        delete p;
        return rc;
    }
    *plan = p;
    return NRFFT_OK;
}
}  // namespace nrfft

using namespace nrfft;

// ---------------------------------------------------------------------------
// extern "C"
// ---------------------------------------------------------------------------
extern "C" {

const char* nrfft_version(void) { return "nrfft-mi355x 0.1 (gfx950)"; }

const char* nrfft_status_string(int s) {
    switch (s) {
        case NRFFT_OK: return "ok";
        case NRFFT_E_ISIGN: return "isign must be 1 or -1";
        case NRFFT_E_DIMS: return "Invalid dimensions";
        case NRFFT_E_NOT_POW2: return "length must be a power of two";
        case NRFFT_E_LEN: return "Input arrays must have the same length";
        case NRFFT_E_EMPTY: return "Input arrays cannot be empty";
        case NRFFT_E_RESP_TOO_LONG: return "Response function longer than data";
        case NRFFT_E_HIP: return "HIP runtime error";
        case NRFFT_E_NOMEM: return "out of memory";
        case NRFFT_E_UNSUPPORTED: return "size not supported";
        case NRFFT_E_INVALID: return "invalid argument";
        default: return "unknown status";
    }
}

const char* nrfft_last_error(void) { return g_err.c_str(); }

int nrfft_init(int device) {
    g_err.clear();
    HIPCHK(hipSetDevice(device));
    DeviceCtx* c = nullptr;
    if (int rc = get_ctx(&c)) return rc;
    return noalloc_setup(device);
}

static int check_four1(const void* data, size_t nn, int isign) {
    if (isign != 1 && isign != -1) return fail(NRFFT_E_ISIGN, "isign must be 1 or -1");
    if (nn == 0) return fail(NRFFT_E_EMPTY, "nn == 0");
    if (!data) return fail(NRFFT_E_INVALID, "null data");
    if (!is_pow2(nn)) return fail(NRFFT_E_NOT_POW2, "nn must be a power of two");
    return NRFFT_OK;
}

int nrfft_four1_f64(double* data, size_t nn, int isign) {
    g_err.clear();
    int rc = check_four1(data, nn, isign);
    return rc ? rc : host_c2c(NRFFT_F64, data, {nn}, 1, isign);
}
int nrfft_four1_f32(float* data, size_t nn, int isign) {
    g_err.clear();
    int rc = check_four1(data, nn, isign);
    return rc ? rc : host_c2c(NRFFT_F32, data, {nn}, 1, isign);
}
int nrfft_four1_batch_f64(double* data, size_t nn, size_t batch, int isign) {
    g_err.clear();
    int rc = check_four1(data, nn, isign);
    if (rc) return rc;
    if (batch == 0) return NRFFT_OK;
    return host_c2c(NRFFT_F64, data, {nn}, batch, isign);
}
int nrfft_four1_batch_f32(float* data, size_t nn, size_t batch, int isign) {
    g_err.clear();
    int rc = check_four1(data, nn, isign);
    if (rc) return rc;
    if (batch == 0) return NRFFT_OK;
    return host_c2c(NRFFT_F32, data, {nn}, batch, isign);
}

// Fourn.rs:367-378 order of checks: dims, isign, sizes.
static int check_fourn(const void* data, const size_t* nn, size_t ndim, int isign, std::vector<size_t>* dims) {
    if (ndim == 0 || !nn) return fail(NRFFT_E_DIMS, "Invalid dimensions");
    if (isign != 1 && isign != -1) return fail(NRFFT_E_ISIGN, "isign must be 1 or -1");
    for (size_t i = 0; i < ndim; ++i)
        if (nn[i] <= 1) return fail(NRFFT_E_DIMS, "Invalid dimension size");
    for (size_t i = 0; i < ndim; ++i)
        if (!is_pow2(nn[i])) return fail(NRFFT_E_NOT_POW2, "dimension must be a power of two");
    if (!data) return fail(NRFFT_E_INVALID, "null data");
    dims->assign(nn, nn + ndim);
    return NRFFT_OK;
}

int nrfft_fourn_f64(double* data, const size_t* nn, size_t ndim, int isign) {
    g_err.clear();
    std::vector<size_t> d;
    int rc = check_fourn(data, nn, ndim, isign, &d);
    return rc ? rc : host_c2c(NRFFT_F64, data, d, 1, isign);
}
int nrfft_fourn_f32(float* data, const size_t* nn, size_t ndim, int isign) {
    g_err.clear();
    std::vector<size_t> d;
    int rc = check_fourn(data, nn, ndim, isign, &d);
    return rc ? rc : host_c2c(NRFFT_F32, data, d, 1, isign);
}

// ---- device plans ---------------------------------------------------------

int nrfft_plan_c2c(nrfft_plan* plan, int dtype, const size_t* nn, size_t ndim, size_t batch) {
    g_err.clear();
    std::vector<size_t> d;
    int rc = check_fourn(nn, nn, ndim, 1, &d);
    if (rc) return rc;
    return make_plan(plan, PK_C2C, dtype, d, batch, 0);
}

int nrfft_plan_c2c_split(nrfft_plan* plan, const size_t* nn, size_t ndim, size_t batch, int prec) {
    g_err.clear();
    if (prec != NRFFT_SPLIT_F16 && prec != NRFFT_SPLIT_BF16) return fail(NRFFT_E_INVALID, "precision");
    std::vector<size_t> d;
    int rc = check_fourn(nn, nn, ndim, 1, &d);
    if (rc) return rc;
    return make_plan(plan, PK_C2C, NRFFT_F32, d, batch, (size_t)prec);
}

int nrfft_execute_c2c(nrfft_plan plan, const void* d_in, void* d_out, int isign, void* stream) {
    g_err.clear();
    if (!plan || plan->plan.kind != PK_C2C) return fail(NRFFT_E_INVALID, "not a c2c plan");
    if (isign != 1 && isign != -1) return fail(NRFFT_E_ISIGN, "isign must be 1 or -1");
    if (!d_in || !d_out) return fail(NRFFT_E_INVALID, "null buffer");
    return exec_plan(plan->plan, d_in, d_out, plan->tmp, nullptr, isign, (hipStream_t)stream);
}

int nrfft_execute_c2c_launches(nrfft_plan plan, const void* d_in, void* d_out, int isign, void* stream, int first,
                               int count) {
    g_err.clear();
    if (!plan || plan->plan.kind != PK_C2C) return fail(NRFFT_E_INVALID, "not a c2c plan");
    if (isign != 1 && isign != -1) return fail(NRFFT_E_ISIGN, "isign must be 1 or -1");
    if (!d_in || !d_out || first < 0 || count < 0) return fail(NRFFT_E_INVALID, "bad argument");
    return exec_plan(plan->plan, d_in, d_out, plan->tmp, nullptr, isign, (hipStream_t)stream, first, count);
}

int nrfft_plan_destroy(nrfft_plan plan) {
    g_err.clear();
    if (!plan) return NRFFT_OK;
    if (plan->tmp) (void)hipFree(plan->tmp);
    delete plan;
    return NRFFT_OK;
}

size_t nrfft_plan_workspace_bytes(nrfft_plan plan) { return plan ? plan->tmp_bytes : 0; }

int nrfft_plan_num_launches(nrfft_plan plan) { return plan ? (int)plan->plan.launches.size() : 0; }

}  // extern "C"
