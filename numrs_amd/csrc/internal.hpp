// internal.hpp -- declarations shared by the library's translation units.
#pragma once
#include <hip/hip_runtime.h>

#include <memory>
#include <string>
#include <functional>
#include <vector>

#include "pass_c2c.hpp"

namespace nrfft {

const KernelRef& c2c_lookup_f32p(int mode, int logl, int tilog);
const KernelRef& c2c_lookup_f32m(int mode, int logl, int tilog);
const KernelRef& c2c_lookup_f64p(int mode, int logl, int tilog);
const KernelRef& c2c_lookup_f64m(int mode, int logl, int tilog);
const KernelRef& c2c_lookup(int dtype, int sign, int mode, int logl, int tilog);
const KernelRef& c2c16k_lookup_f32p();
const KernelRef& c2c16k_lookup_f32m();
const DmaRef& dma_lookup_f32p(int logl);
const DmaRef& dma_lookup_f32m(int logl);
int pick_tilog(int dtype, int mode, int logl, int64_t nb0, int want_log);

// Per-device state: twiddle tables (FP64-generated; FP32 copies rounded).
struct DeviceCtx {
    int device = 0;
    c64* c64c = nullptr;
    c64* c64f = nullptr;
    c32* c32c = nullptr;
    c32* c32f = nullptr;
    c64* c64s = nullptr;  // per-stage tables (TW_STAGE_ELEMS entries)
    c32* c32s = nullptr;
    void* mfma_f16 = nullptr;   // split-precision DFT-matrix operands
    void* mfma_bf16 = nullptr;
    int ncu = 256;        // compute units
};
int build_mfma_tables(DeviceCtx* c);
const KernelRef& mfma_lookup(int prec, int sign, int mode, int logl);
template <typename T>
inline Twiddles<T> ctx_twiddles(const DeviceCtx* c) {
    if constexpr (sizeof(T) == 4) return Twiddles<T>{c->c32c, c->c32f, c->c32s};
    else return Twiddles<T>{c->c64c, c->c64f, c->c64s};
}
int get_ctx(DeviceCtx** out);
int find_ctx(DeviceCtx** out);
// Per-device set-up of the allocation-free entry point's scratch (its
// device address); called by nrfft_init with the device current.
int noalloc_setup(int device);

// Buffers a launch reads or writes.
// B_AUX2: a second caller buffer (twofft's second spectrum, rlft3's speq).
enum BufId { B_IN = 0, B_OUT = 1, B_TMP = 2, B_AUX = 3, B_TMP2 = 4, B_AUX2 = 5, B_COUNT = 6 };
// Plan kinds.
enum PlanKind {
    PK_C2C = 0,
    PK_REALFT = 1,
    PK_CONVLV = 2,
    PK_CORREL = 3,
    PK_TWOFFT = 4,
    PK_COSFT1 = 5,
    PK_COSFT2 = 6,
    PK_RLFT3 = 7,
    PK_UNEVEN = 8  // tune-cache key of an uneven-batch size class (nn = {N}, batch 0)
};
// Launch kinds.
enum LaunchKind { LK_C2C = 0, LK_COPY = 1, LK_REAL = 2, LK_SPEC = 3 };

// Tunables chosen per (kind, N, batch, dtype); -1 = default.
struct Tune {
    int split_log1 = -1;  // four-step split: log2 N1 (column pass length)
    int tilog_a = -1;     // tile width (log2) of column passes
    int tilog_b = -1;     // tile width (log2) of transpose passes
    int chunk = -1;       // transforms per four-step chunk (0/-1 = whole batch)
    int wg_per_cu = -1;   // resident groups per CU of persistent passes (-1 = occupancy limit)
    int xcd_remap = -1;   // 1 = contiguous tiles per XCD (-1 = default)
    bool operator==(const Tune& o) const {
        return split_log1 == o.split_log1 && tilog_a == o.tilog_a && tilog_b == o.tilog_b && chunk == o.chunk &&
               wg_per_cu == o.wg_per_cu && xcd_remap == o.xcd_remap;
    }
};

struct Launch {
    int kind = LK_C2C;
    const void* fn[2] = {nullptr, nullptr};  // [0] isign = +1, [1] isign = -1
    unsigned grid = 0;
    int nt = 0;
    int lds = 0;
    int logl = 0;
    int ti = 0;
    int mode = 0;
    PassDesc desc{};
    int src = B_IN;
    int dst = B_OUT;
    int prec = 0;         // 0 = plain kernel, else split-precision MFMA kernel (MfmaPrec)
    // DMA-staged variant of an FP32 MODE_TT pass (k_c2c_dma), used when the
    // input pointer is 16-byte aligned; nullptr when the pass does not qualify.
    const void* fn_dma[2] = {nullptr, nullptr};
    unsigned grid_dma = 0;
    int nt_dma = 0;
    int64_t off_src = 0;  // element offsets (complex units) into src / dst
    int64_t off_dst = 0;
    int64_t off_src_real = 0;  // further offset into src in reals (load-mapped passes)
    size_t bytes = 0;   // LK_COPY
    int special = 0;    // LK_REAL: which real-FFT kernel
    // args[0..5]: kernel arguments of LK_REAL launches.
    // args[6]: if non-zero, the isign this launch always runs with.
    // args[7]: if non-zero, the launch runs only for calls with that isign sign.
    int64_t args[8] = {0, 0, 0, 0, 0, 0, 0, 0};
};

struct Plan {
    int kind = PK_C2C;
    int dtype = 0;
    std::vector<size_t> nn;
    size_t batch = 1;
    size_t extra = 0;  // convlv: response length m; C2C: split precision (0 = plain FP32/FP64)
    int device = 0;
    Tune tune;
    std::vector<Launch> launches;
    size_t tmp_elems = 0;    // complex elements of workspace
    size_t tmp2_offset = 0;  // B_TMP2 = workspace + tmp2_offset complex elements
    // Strided n-D axes of >= 4096 points run as one column pass (no
    // workspace) instead of the two-pass four-step along the axis: set when
    // the four-step's workspace could not be allocated.
    bool strided_one_pass = false;
    // 1-D C2C four-step over several chunks whose launches alternate pass A,
    // pass B per chunk with the scratch double-buffered: executed with each
    // chunk's pass B on a second stream (exec_plan).
    bool overlap = false;
};

extern thread_local std::string g_err;
int fail(int code, const std::string& msg);
bool is_pow2(size_t n);
int ilog2(size_t n);
size_t csize(int dtype);
int build_plan(Plan& p);
int build_plan_real(Plan& p);
int exec_special(const Plan& p, const Launch& l, void* const* bufs, int isign, hipStream_t stream, DeviceCtx* ctx);
int build_plan_spectral(Plan& p);
int exec_spectral(const Plan& p, const Launch& l, void* const* bufs, int isign, hipStream_t stream, DeviceCtx* ctx);
int plan_c2c_nd(Plan& p, const std::vector<size_t>& nn, int64_t batch);
void tune_lookup(Plan& p);
// Give a plan whose C2C runs inside a larger transform the tuning of a plain
// C2C plan of length 2^logn and the plan's batch.
void use_c2c_tune(Plan& p, int logn);
bool tune_find(int kind, int dtype, const std::vector<size_t>& nn, size_t batch, Tune* out);
void tune_store(int kind, int dtype, const std::vector<size_t>& nn, size_t batch, const Tune& t);
Tune default_tune(const Plan& p);
// Defaults of a 1-D C2C plan of length n (kind PK_C2C; extra = split precision).
Tune default_tune_1d(int dtype, size_t n, size_t batch, size_t extra);
int plan_c2c_1d(Plan& p, int logn, int64_t batch, int src, int dst, size_t* tmp_elems);
int plan_add_c2c_pass(Plan& p, int mode, int logl, int tilog, const PassDesc& d, int src, int dst);
int c2c_pass_launch(const Plan& p, int mode, int logl, int tilog, const PassDesc& d, int src, int dst, Launch* out);
struct FourStep {
    int logn = 0, l1 = 0, l2 = 0;  // N = 2^l1 * 2^l2
    int ta = 0, tb = 0;            // log2 tile widths of pass A / pass B
};
int fourstep_setup(const Plan& p, int logn, FourStep* fs);
// lmap != LMAP_PLAIN: pass A reads a real array of sig_reals reals per
// signal through that load map (isign +1, or -1 for LMAP_CF2INV).
int fourstep_launches(const Plan& p, const FourStep& fs, int64_t cb, int src, int dst, Launch* la, Launch* lb,
                      int lmap = LMAP_PLAIN, int64_t sig_reals = 0);
// The load-mapped first pass of (dtype, load map, log2 L, log2 tile width);
// fn == nullptr if none.
const KernelRef& map_lookup(int dtype, int lmap, int logl, int tilog);
// Long 1-D C2C: from 2^C2C_3PASS_MIN_LOGN points plain C2C plans run three
// passes of N1 x Na x Mb points instead of the two-pass four-step (whose
// passes beyond 2^22 points get tiles of 8 or fewer transforms; 16 x 2^22
// FP32: 0.577 ms against 0.595, 8 x 2^23: 0.580 against 1.376).  Paths that
// fold other steps into the four-step's passes (load maps, realft's
// per-chunk untangle) keep the two-pass form below 2^FUSED_3PASS_MIN_LOGN,
// where it is within a few percent and their fused steps gain more.
#ifndef NRFFT_C2C_3PASS_MIN_LOGN
#define NRFFT_C2C_3PASS_MIN_LOGN 22
#endif
constexpr int C2C_3PASS_MIN_LOGN = NRFFT_C2C_3PASS_MIN_LOGN;
constexpr int FUSED_3PASS_MIN_LOGN = C2C_3PASS_MIN_LOGN > 23 ? C2C_3PASS_MIN_LOGN : 23;
inline void three_pass_split(int logn, int* l1, int* la, int* lb) {
    *l1 = (logn + 2) / 3;
    *la = (logn - *l1) / 2;
    *lb = logn - *l1 - *la;
}
int plan_c2c_3pass(Plan& p, int logn, int64_t batch, int src, int dst, size_t* tmp_elems, int lmap = LMAP_PLAIN,
                   int64_t sig_reals = 0, int64_t map_arg = 0, int force = 0, int only = 0);
// The chunked C2C whose first pass reads src through load map lmap (api.hip).
int plan_c2c_mapped(Plan& p, int logn, int64_t batch, int src, int dst, int lmap, int64_t sig_reals, int64_t map_arg,
                    int force, int only, size_t* tmp_elems);
// True when the C2C of length 2^logn can read its input through load map
// lmap: a four-step length whose first pass has a load-mapped kernel of the
// plan's tile width.
bool map_fits(const Plan& p, int logn, int lmap);
int exec_launch(const Plan& p, const Launch& l, void* const* bufs, int isign, hipStream_t stream, DeviceCtx* ctx);
int col_tilog(int dtype);
int exec_plan(const Plan& p, const void* in, void* out, void* tmp, const void* aux, int isign, hipStream_t stream,
              int first = 0, int count = -1, void* aux2 = nullptr);
int cached_plan(int kind, int dtype, const std::vector<size_t>& nn, size_t batch, size_t extra,
                std::shared_ptr<Plan>* out);

// Stream-ordered device scratch of the host-pointer calls.
struct DevBuf {
    void* p = nullptr;
    hipStream_t s = nullptr;
    ~DevBuf() {
        if (p) (void)hipFreeAsync(p, s);
    }
};
int dev_alloc(DevBuf& b, size_t bytes, hipStream_t s);
// A batch of chunks streamed host -> HBM -> host through pinned staging:
// gather(i, pinned) fills the input slot and returns its bytes; compute(i,
// dev, ws, stream, &out, &out_bytes) transforms the uploaded chunk in the
// device slot `dev` (slot_bytes) with the shared workspace `ws` (ws_bytes)
// and names the device bytes to download; scatter(i, pinned, bytes) stores
// them.  Two slots overlap transfers, transforms and host copies; one slot
// runs the chunks in turn.  Chunks must cover disjoint host ranges.
struct PipeJob {
    size_t chunks = 0;
    size_t slot_bytes = 0;
    size_t ws_bytes = 0;
    int slots = 2;
    std::function<size_t(size_t, char*)> gather;
    std::function<int(size_t, char*, void*, hipStream_t, const char**, size_t*)> compute;
    std::function<void(size_t, const char*, size_t)> scatter;
};
int run_pipe(const PipeJob& job);
// dst[r * dpitch ..+ width] = src[r * spitch ..+ width] for r < rows, on the host copy threads.
void par_copy2d(char* dst, size_t dpitch, const char* src, size_t spitch, size_t width, size_t rows);
int host_c2c_pipelined(int dtype, void* data, const std::vector<size_t>& nn, size_t batch, size_t chunk, int isign);
size_t host_pipe_chunk(int dtype, const std::vector<size_t>& nn, size_t batch);

}  // namespace nrfft

struct nrfft_plan_s {
    nrfft::Plan plan;
    void* tmp = nullptr;
    size_t tmp_bytes = 0;
};

namespace nrfft {
// explicit_tune: tunables for this plan only (the tune cache is neither read
// nor written); nullptr = the cached tuning for the key, else the defaults.
int make_plan(nrfft_plan_s** plan, int kind, int dtype, std::vector<size_t> nn, size_t batch, size_t extra,
              const Tune* explicit_tune = nullptr);
}
