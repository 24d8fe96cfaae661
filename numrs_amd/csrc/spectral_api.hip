// spectral_api.hip -- twofft / cosft1 / cosft2 / rlft3: planning, execution
// and the extern "C" entry points of include/nrfft.h.
#include <atomic>
#include <string>
#include <vector>

#include "../../include/nrfft.h"
#include "internal.hpp"
#include "spectral_kern.hpp"

namespace nrfft {

#define HIPCHK(x)                                                                                      \
    do {                                                                                               \
        hipError_t e_ = (x);                                                                           \
        if (e_ != hipSuccess) return fail(NRFFT_E_HIP, std::string(#x) + ": " + hipGetErrorString(e_)); \
    } while (0)

constexpr int COSFT_MAX_LOGN = 24;   // cosft2's exp(i pi k / 2n) comes from the 2^26-point table
constexpr int TWOFFT_MAX_LOGN = 26;  // the C2C plans' limit

enum SpecSpecial {
    SP_TW_FUSED = 1,  // args: nsig, n
    SP_TW_REAL,       // args: nsig, n
    SP_CF1_FUSED,     // args: nsig, n
    SP_CF1_PACK,      // args: nsig, n
    SP_CF1_POST,      // args: nsig, n
    SP_CF2_FUSED,     // args: nsig, L
    SP_CF2_N2,        // args: nsig
    SP_CF2_FPACK,     // args: nsig, L
    SP_CF2_FPOST,     // args: nsig, L
    SP_CF2_IPRE,      // args: nsig, L
    SP_CF2_IUNPACK,   // args: nsig, L
    SP_R3_COPY,       // args: nsig, plane, h
    SP_R3_UNTANGLE,   // args: nsig, nn1, nn2, h, log2 nn3
    SP_R3_ROWS,       // args: row pairs of the batch, nn1, nn2, h
};

// NRFFT_R3_FUSED = 0: the rlft3 forward always runs fourn, the speq copy and
// the untangle as separate passes; 1: where the shape allows, the last axis
// (rows of h = nn3/2 points) runs last and carries the untangle and speq.
#ifndef NRFFT_R3_FUSED
#define NRFFT_R3_FUSED 1
#endif

static Launch spec_launch(int special, int src, int dst, std::initializer_list<int64_t> args) {
    Launch l;
    l.kind = LK_SPEC;
    l.special = special;
    l.src = src;
    l.dst = dst;
    int i = 0;
    for (int64_t a : args) l.args[i++] = a;
    return l;
}

static int spec_fused_max(int dtype) { return dtype == NRFFT_F32 ? spec_max_logl<float>() : spec_max_logl<double>(); }

static const SpecRef& spec_lookup(int dtype, int logl) {
    return dtype == NRFFT_F32 ? spec_lookup_f32(logl) : spec_lookup_f64(logl);
}

// C2C launches appended from index `from` run with isign `force` and only for
// calls of direction `only` (0 = both).
static void tag_launches(Plan& p, size_t from, int force, int only) {
    for (size_t i = from; i < p.launches.size(); ++i) {
        if (force) p.launches[i].args[6] = force;
        if (only) p.launches[i].args[7] = only;
    }
}

// The C2C of length 2^logn of a spectral plan, chunk by chunk: each chunk's
// packing launch right before its first pass, its finishing launch right
// after its second pass (which stores plainly), all through a chunk-sized
// scratch TMP2 (complex values [0, chunk * 2^logn)), so the chunk's data
// passes between the steps through the caches (MALL) rather than HBM.
// pre(c0, cb) / post(c0, cb) return the launches for signals [c0, c0 + cb)
// with their byte offsets set; the C2C launches run with isign `force` and
// every launch only for calls of direction `only` (0 = both).
template <class PRE, class POST>
static int plan_spec_chunked(Plan& p, int logn, int64_t batch, int force, int only, PRE pre, POST post) {
    FourStep fs;
    if (int rc = fourstep_setup(p, logn, &fs)) return rc;
    const int64_t N = int64_t(1) << logn;
    const int64_t chunk = (p.tune.chunk > 0 && p.tune.chunk < batch) ? p.tune.chunk : batch;
    for (int64_t c0 = 0; c0 < batch; c0 += chunk) {
        const int64_t cb = std::min<int64_t>(chunk, batch - c0);
        Launch a = pre(c0, cb);
        a.args[7] = only;
        p.launches.push_back(a);
        Launch la, lb;
        if (int rc = fourstep_launches(p, fs, cb, B_TMP2, B_TMP2, &la, &lb)) return rc;
        lb.desc.nt_out = 0;
        for (Launch* l : {&la, &lb}) {
            l->args[6] = force;
            l->args[7] = only;
            p.launches.push_back(*l);
        }
        Launch b = post(c0, cb);
        b.args[7] = only;
        p.launches.push_back(b);
    }
    p.tmp2_offset = std::max<size_t>(p.tmp2_offset, size_t(N) * size_t(chunk));  // the C2C scratch (B_TMP) comes first
    p.tmp_elems = std::max<size_t>(p.tmp_elems, p.tmp2_offset + size_t(N) * size_t(chunk));
    return NRFFT_OK;
}

const KernelRef& map_lookup(int dtype, int lmap, int logl, int tilog) {
    return dtype == NRFFT_F32 ? map_lookup_f32(lmap, logl, tilog) : map_lookup_f64(lmap, logl, tilog);
}

// NRFFT_SPEC_LOAD_MAP = 0: the DCTs beyond one group pack their sequence
// into the workspace before the C2C instead of reading it in the first pass.
#ifndef NRFFT_SPEC_LOAD_MAP
#define NRFFT_SPEC_LOAD_MAP 1
#endif

// True when the C2C of length 2^logn can read its input through load map
// lmap: a four-step length whose first pass has a load-mapped kernel of the
// plan's tile width.
bool map_fits(const Plan& p, int logn, int lmap) {
    if (!NRFFT_SPEC_LOAD_MAP || logn <= C2C_MAX_LOGL_ || logn > 2 * C2C_MAX_LOGL_) return false;
    if (logn >= FUSED_3PASS_MIN_LOGN) {  // the three-pass plan's first pass
        int l1, la, lb;
        three_pass_split(logn, &l1, &la, &lb);
        const int t1 = pick_tilog(p.dtype, MODE_TT, l1, int64_t(1) << (la + lb), col_tilog(p.dtype));
        return t1 >= 0 && map_lookup(p.dtype, lmap, l1, t1).fn != nullptr;
    }
    FourStep fs;
    if (fourstep_setup(p, logn, &fs)) return false;
    return map_lookup(p.dtype, lmap, fs.l1, fs.ta).fn != nullptr;
}

// A DCT's C2C of length N = 2^logn whose first pass reads the caller's reals
// (sig_reals per signal) through load map lmap, chunk by chunk, with isign
// `force`, for calls of direction `only` (0 = both), into a whole-batch TMP2;
// then the finishing launch post(0, batch) once after the last chunk.
// (Finishing each chunk right after its second pass, through a chunk-sized
// TMP2, measured slower at 64 x 2^20 FP32: cosft1 0.783 ms against 0.652,
// cosft2 forward 0.451 against 0.379; a chunk's intermediate and its second
// pass's output together fill the MALL.)  Call only when map_fits.
template <class POST>
static int plan_spec_mapped(Plan& p, int logn, int64_t batch, int lmap, int64_t sig_reals, int force, int only,
                            POST post) {
    size_t tmp = 0;
    if (int rc = plan_c2c_mapped(p, logn, batch, B_IN, B_TMP2, lmap, sig_reals, 0, force, only, &tmp)) return rc;
    Launch b = post(int64_t(0), batch);
    b.args[7] = only;
    p.launches.push_back(b);
    p.tmp2_offset = std::max<size_t>(p.tmp2_offset, tmp);  // the C2C scratch (B_TMP) comes first
    p.tmp_elems = std::max<size_t>(p.tmp_elems, p.tmp2_offset + (size_t(batch) << logn));
    return NRFFT_OK;
}

int build_plan_spectral(Plan& p) {
    const int64_t batch = (int64_t)p.batch;
    size_t tmp = 0;
    int rc = NRFFT_OK;
    switch (p.kind) {
        case PK_TWOFFT: {
            const int64_t n = (int64_t)p.nn.at(0);
            const int logn = ilog2((size_t)n);
            if (logn <= spec_fused_max(p.dtype)) {
                p.launches.push_back(spec_launch(SP_TW_FUSED, B_IN, B_OUT, {batch, n}));
                return NRFFT_OK;
            }
            // data1 and data2 read as n/2 complex values each: two C2C plans
            // into the workspace (Z1 then Z2), then the realft untangle of
            // both into the full n-bin spectra -- no packing pass.  (Writing
            // Z1 / Z2 into the outputs with stride n and untangling in place
            // measured 5% slower; so did untangling chunk by chunk right after
            // both arrays' passes, 64 x 2^20 FP32 0.81-0.86 ms against 0.78.)
            const int64_t N = n / 2;
            use_c2c_tune(p, logn - 1);
            const size_t from = p.launches.size();
            if ((rc = plan_c2c_1d(p, logn - 1, batch, B_IN, B_TMP2, &tmp))) return rc;
            const size_t from2 = p.launches.size();
            if ((rc = plan_c2c_1d(p, logn - 1, batch, B_AUX, B_TMP2, &tmp))) return rc;
            for (size_t i = from2; i < p.launches.size(); ++i)
                if (p.launches[i].dst == B_TMP2) p.launches[i].off_dst += batch * N;
            tag_launches(p, from, 1, 0);
            p.launches.push_back(spec_launch(SP_TW_REAL, B_TMP2, B_OUT, {batch, n}));
            p.tmp2_offset = tmp;
            p.tmp_elems = tmp + 2 * size_t(batch) * size_t(N);
            return NRFFT_OK;
        }
        case PK_COSFT1: {
            const int64_t n = (int64_t)p.nn.at(0);
            const int logn = ilog2((size_t)n);
            if (logn <= spec_fused_max(p.dtype)) {
                p.launches.push_back(spec_launch(SP_CF1_FUSED, B_IN, B_OUT, {batch, n}));
                return NRFFT_OK;
            }
            use_c2c_tune(p, logn);
            if (map_fits(p, logn, LMAP_EVENEXT)) {
                const int64_t row = (n + 1) * (int64_t(csize(p.dtype)) / 2);  // bytes of one signal
                return plan_spec_mapped(p, logn, batch, LMAP_EVENEXT, n + 1, 1, 0, [&](int64_t c0, int64_t cb) {
                    Launch l = spec_launch(SP_CF1_POST, B_TMP2, B_OUT, {cb, n});
                    l.off_dst = c0 * row;
                    return l;
                });
            }
            // Whole batch per step: the chunked form (pack, passes and finish
            // per chunk through a chunk-sized scratch) measured 0.833 ms
            // against 0.818 at 64 x 2^20 FP32.
            p.launches.push_back(spec_launch(SP_CF1_PACK, B_IN, B_TMP2, {batch, n}));
            const size_t from = p.launches.size();
            if ((rc = plan_c2c_1d(p, logn, batch, B_TMP2, B_TMP2, &tmp))) return rc;
            tag_launches(p, from, 1, 0);
            p.launches.push_back(spec_launch(SP_CF1_POST, B_TMP2, B_OUT, {batch, n}));
            p.tmp2_offset = tmp;
            p.tmp_elems = tmp + size_t(batch) * size_t(n);
            return NRFFT_OK;
        }
        case PK_COSFT2: {
            const int64_t n = (int64_t)p.nn.at(0), L = n / 2;
            if (n == 2) {
                p.launches.push_back(spec_launch(SP_CF2_N2, B_IN, B_OUT, {batch}));
                return NRFFT_OK;
            }
            const int logl = ilog2((size_t)L);
            if (logl <= spec_fused_max(p.dtype)) {
                p.launches.push_back(spec_launch(SP_CF2_FUSED, B_IN, B_OUT, {batch, L}));
                return NRFFT_OK;
            }
            use_c2c_tune(p, logl);
            const int64_t sig = n * (int64_t(csize(p.dtype)) / 2);  // bytes of one signal
            // forward: permute-pack -> C2C(+1) -> untangle and rotate
            if (map_fits(p, logl, LMAP_EVENODD)) {
                rc = plan_spec_mapped(p, logl, batch, LMAP_EVENODD, n, 1, 1, [&](int64_t c0, int64_t cb) {
                                          Launch l = spec_launch(SP_CF2_FPOST, B_TMP2, B_OUT, {cb, L});
                                          l.off_dst = c0 * sig;
                                          return l;
                                      });
            } else rc = plan_spec_chunked(
                p, logl, batch, 1, 1,
                [&](int64_t c0, int64_t cb) {
                    Launch l = spec_launch(SP_CF2_FPACK, B_IN, B_TMP2, {cb, L});
                    l.off_src = c0 * sig;
                    return l;
                },
                [&](int64_t c0, int64_t cb) {
                    Launch l = spec_launch(SP_CF2_FPOST, B_TMP2, B_OUT, {cb, L});
                    l.off_dst = c0 * sig;
                    return l;
                });
            if (rc) return rc;
            // inverse: rotate and untangle -> C2C(-1) -> unpermute.  With a
            // load-mapped first pass the rotation and untangle happen in its
            // loads (64 x 2^20 FP32: 0.362 ms against 0.444 for the separate
            // pre-step).  Otherwise whole batch per step (chunked: 0.455 ms
            // against 0.444; the C2C's second pass storing through the
            // permutation straight into y instead of the unpermute step:
            // 0.448 against 0.447, its 4-byte stores filling half of each
            // line, so it was not kept).
            if (map_fits(p, logl, LMAP_CF2INV)) {
                // The rotation and untangle as the C2C's first-pass load map.
                return plan_spec_mapped(p, logl, batch, LMAP_CF2INV, n, -1, -1, [&](int64_t, int64_t cb) {
                    Launch d = spec_launch(SP_CF2_IUNPACK, B_TMP2, B_OUT, {cb, L});
                    return d;
                });
            }
            Launch c = spec_launch(SP_CF2_IPRE, B_IN, B_TMP2, {batch, L});
            c.args[7] = -1;
            p.launches.push_back(c);
            const size_t from = p.launches.size();
            if ((rc = plan_c2c_1d(p, logl, batch, B_TMP2, B_TMP2, &tmp))) return rc;
            tag_launches(p, from, -1, -1);
            Launch d = spec_launch(SP_CF2_IUNPACK, B_TMP2, B_OUT, {batch, L});
            d.args[7] = -1;
            p.launches.push_back(d);
            // one TMP2 for both directions: after the C2C scratch, the whole batch
            p.tmp2_offset = std::max<size_t>(p.tmp2_offset, tmp);
            p.tmp_elems = std::max<size_t>(p.tmp_elems, p.tmp2_offset + size_t(batch) * size_t(L));
            return NRFFT_OK;
        }
        case PK_RLFT3: {
            const int64_t nn1 = (int64_t)p.nn.at(0), nn2 = (int64_t)p.nn.at(1), nn3 = (int64_t)p.nn.at(2);
            const int64_t h = nn3 / 2;
            std::vector<size_t> dims;
            for (int64_t v : {nn1, nn2, h})
                if (v > 1) dims.push_back((size_t)v);
            const int64_t vol = nn1 * nn2 * h;
            const bool fused = NRFFT_R3_FUSED && nn2 >= 2 && nn1 < 4096 && nn2 < 4096 && h >= 16 && h <= 2048 &&
                               vol <= 0x7fffffffLL && spec_lookup(p.dtype, ilog2((size_t)h)).r3rows;
            size_t from = p.launches.size();
            if (fused) {
                // forward: axis 1 (in -> data, in place for a device plan;
                // in -> TMP2 when nn1 = 1), axis 0 (data -> TMP2), then the
                // rows of h points with the untangle and speq (TMP2 -> data).
                const int want = col_tilog(p.dtype);
                PassDesc d1{};
                d1.nb0 = h;
                d1.si0 = 1;
                d1.so0 = 1;
                d1.esi = (int32_t)h;
                d1.eso = (int32_t)h;
                d1.nb1 = nn1 * batch;
                d1.si1 = nn2 * h;
                d1.so1 = nn2 * h;
                d1.nb2 = 1;
                const int t1 = pick_tilog(p.dtype, MODE_TT, ilog2((size_t)nn2), h, want);
                if (t1 < 0) return fail(NRFFT_E_UNSUPPORTED, "no column kernel");
                if ((rc = plan_add_c2c_pass(p, MODE_TT, ilog2((size_t)nn2), t1, d1, B_IN, nn1 >= 2 ? B_OUT : B_TMP2)))
                    return rc;
            }
            if (fused && nn1 >= 2) {
                const int want = col_tilog(p.dtype);
                PassDesc d0{};
                d0.nb0 = nn2 * h;
                d0.si0 = 1;
                d0.so0 = 1;
                d0.esi = (int32_t)(nn2 * h);
                d0.eso = (int32_t)(nn2 * h);
                d0.nb1 = batch;
                d0.si1 = vol;
                d0.so1 = vol;
                d0.nb2 = 1;
                const int t0 = pick_tilog(p.dtype, MODE_TT, ilog2((size_t)nn1), nn2 * h, want);
                if (t0 < 0) return fail(NRFFT_E_UNSUPPORTED, "no column kernel");
                if ((rc = plan_add_c2c_pass(p, MODE_TT, ilog2((size_t)nn1), t0, d0, B_OUT, B_TMP2))) return rc;
            }
            if (fused) {
                tag_launches(p, from, 1, 1);
                const int64_t per = nn1 >= 2 ? nn1 * nn2 / 2 + 2 : nn2 / 2 + 1;  // row pairs of a volume
                Launch f = spec_launch(SP_R3_ROWS, B_TMP2, B_OUT, {batch * per, nn1, nn2, h});
                f.args[7] = 1;
                p.launches.push_back(f);
            } else {
                // forward: fourn(+1) on the data (in place) -> speq <- k3 = 0 column -> untangle
                if (!dims.empty() && (rc = plan_c2c_nd(p, dims, batch))) return rc;
                tag_launches(p, from, 1, 1);
                Launch a = spec_launch(SP_R3_COPY, B_OUT, B_AUX2, {batch, nn1 * nn2, h});
                a.args[7] = 1;
                p.launches.push_back(a);
                Launch b = spec_launch(SP_R3_UNTANGLE, B_OUT, B_OUT, {batch, nn1, nn2, h, ilog2((size_t)nn3)});
                b.args[7] = 1;
                p.launches.push_back(b);
            }
            // inverse: untangle -> fourn(-1)
            Launch c = spec_launch(SP_R3_UNTANGLE, B_OUT, B_OUT, {batch, nn1, nn2, h, ilog2((size_t)nn3)});
            c.args[7] = -1;
            p.launches.push_back(c);
            from = p.launches.size();
            if (!dims.empty() && (rc = plan_c2c_nd(p, dims, batch))) return rc;
            tag_launches(p, from, -1, -1);
            if (fused) {  // TMP2 (the forward's axis-0 output) after the inverse plan's own scratch
                p.tmp2_offset = std::max<size_t>(p.tmp2_offset, p.tmp_elems);
                p.tmp_elems = p.tmp2_offset + size_t(batch) * size_t(vol);
            }
            return NRFFT_OK;
        }
        default:
            return fail(NRFFT_E_INVALID, "not a spectral plan");
    }
}

static unsigned grid_of(int64_t work, int nt) {
    int64_t g = (work + nt - 1) / nt;
    if (g > 65536) g = 65536;
    if (g < 1) g = 1;
    return (unsigned)g;
}

// 2-D grid of the multi-pass helpers: y over signals, x over the `per`
// elements of one signal, 256 threads a block, at most ~64K blocks.
static dim3 grid2(int64_t per, int64_t nsig, int64_t max_blocks = 65536) {
    const int64_t gy = std::max<int64_t>(1, std::min<int64_t>(nsig, 65535));
    const int64_t cap = std::max<int64_t>(1, max_blocks / gy);
    const int64_t gx = std::max<int64_t>(1, std::min<int64_t>((per + 255) / 256, cap));
    return dim3((unsigned)gx, (unsigned)gy);
}

template <typename T>
static int exec_spec_t(const Plan& p, const Launch& l, void* const* bufs, int isign, hipStream_t s, DeviceCtx* ctx) {
    Twiddles<T> tw = ctx_twiddles<T>(ctx);
    // Spectral launches carry byte offsets into their buffers (chunks).
    void* src = static_cast<char*>(bufs[l.src]) + l.off_src;
    void* dst = static_cast<char*>(bufs[l.dst]) + l.off_dst;
    const int64_t* a = l.args;
    const int64_t nsig = a[0];
    switch (l.special) {
        case SP_TW_FUSED: {
            const SpecRef& r = spec_lookup(p.dtype, ilog2((size_t)a[1]));
            void* d2 = bufs[B_AUX];
            void* f2 = bufs[B_AUX2];
            int64_t ns = nsig;
            void* args[] = {&src, &d2, &dst, &f2, &ns, &tw};
            HIPCHK(hipLaunchKernel(r.twofft, dim3((unsigned)((nsig + r.ti - 1) / r.ti)), dim3(r.nt), args, 0, s));
            return NRFFT_OK;
        }
        case SP_TW_REAL: {
            const cpx<T>* z1 = (const cpx<T>*)src;
            hipLaunchKernelGGL((k_twofft_real<T>), grid_pairs(a[1] / 4, 2 * nsig), dim3(PAIR_NT), 0, s, z1,
                               z1 + nsig * (a[1] / 2), (cpx<T>*)dst,
                               (cpx<T>*)(static_cast<char*>(bufs[B_AUX2]) + l.off_dst), a[1], nsig,
                               ilog2((size_t)a[1]), tw);
            break;
        }
        case SP_CF1_FUSED: {
            const int logl = ilog2((size_t)a[1]);
            const SpecRef& r = spec_lookup(p.dtype, logl);
            int64_t ns = nsig;
            void* args[] = {&src, &dst, &ns, &tw};
            int64_t grid = (nsig + r.ti - 1) / r.ti;
            if (r.cf1_persist) {
                // Persistent: as many groups as the GPU holds at once.
                static std::atomic<int> per_cu[2][SPEC_MAX_LOGL + 1];
                std::atomic<int>& pc = per_cu[p.dtype == NRFFT_F32 ? 0 : 1][logl];
                int g = pc.load();
                if (g == 0) {
                    HIPCHK(hipOccupancyMaxActiveBlocksPerMultiprocessor(&g, r.cosft1, r.nt, 0));
                    g = std::max(g, 1);
                    pc.store(g);
                }
                grid = std::min<int64_t>(grid, int64_t(g) * ctx->ncu);
            }
            HIPCHK(hipLaunchKernel(r.cosft1, dim3((unsigned)grid), dim3(r.nt), args, 0, s));
            return NRFFT_OK;
        }
        case SP_CF1_PACK:
            hipLaunchKernelGGL((k_cosft1_pack<T>), grid2(a[1], nsig), dim3(256), 0, s, (const T*)src,
                               (cpx<T>*)dst, a[1], nsig);
            break;
        case SP_CF1_POST:
            hipLaunchKernelGGL((k_cosft1_post<T>), grid_pairs(a[1] / 2, nsig), dim3(256), 0, s,
                               (const cpx<T>*)src, (T*)dst, a[1], nsig, ilog2((size_t)a[1]) + 1, tw);
            break;
        case SP_CF2_FUSED: {
            const SpecRef& r = spec_lookup(p.dtype, ilog2((size_t)a[1]));
            int64_t ns = nsig;
            void* args[] = {&src, &dst, &ns, &tw};
            HIPCHK(hipLaunchKernel(r.cosft2[isign > 0 ? 0 : 1], dim3((unsigned)((nsig + r.ti - 1) / r.ti)),
                                   dim3(r.nt), args, 0, s));
            return NRFFT_OK;
        }
        case SP_CF2_N2:
            if (isign > 0)
                hipLaunchKernelGGL((k_cosft2_n2<T, 1>), dim3(grid_of(nsig, 256)), dim3(256), 0, s, (const T*)src,
                                   (T*)dst, nsig);
            else
                hipLaunchKernelGGL((k_cosft2_n2<T, -1>), dim3(grid_of(nsig, 256)), dim3(256), 0, s, (const T*)src,
                                   (T*)dst, nsig);
            break;
        case SP_CF2_FPACK:
            hipLaunchKernelGGL((k_cosft2_fpack<T>), grid_pairs(a[1] / 2, nsig), dim3(256), 0, s, (const T*)src,
                               (cpx<T>*)dst, a[1], nsig);
            break;
        case SP_CF2_FPOST:
            hipLaunchKernelGGL((k_cosft2_fpost<T>), grid_pairs(a[1] / 2, nsig), dim3(256), 0, s,
                               (const cpx<T>*)src, (T*)dst, a[1], nsig, ilog2((size_t)a[1]) + 1, tw);
            break;
        case SP_CF2_IPRE:
            hipLaunchKernelGGL((k_cosft2_ipre<T>), grid_pairs(a[1] / 2, nsig), dim3(256), 0, s,
                               (const T*)src, (cpx<T>*)dst, a[1], nsig, ilog2((size_t)a[1]) + 1, tw);
            break;
        case SP_CF2_IUNPACK:
            hipLaunchKernelGGL((k_cosft2_iunpack<T>), grid_pairs(a[1] / 2, nsig), dim3(256), 0, s,
                               (const cpx<T>*)src, (T*)dst, a[1], nsig);
            break;
        case SP_R3_COPY:
            hipLaunchKernelGGL((k_rlft3_copycol<T>), dim3(grid_of(nsig * a[1], 256)), dim3(256), 0, s,
                               (const cpx<T>*)src, (cpx<T>*)dst, a[1], a[2], nsig);
            break;
        case SP_R3_ROWS: {
            const SpecRef& r = spec_lookup(p.dtype, ilog2((size_t)a[3]));
            const cpx<T>* zp = (const cpx<T>*)src;
            cpx<T>* dp = (cpx<T>*)dst;
            cpx<T>* sp = (cpx<T>*)bufs[B_AUX2];
            int64_t npairs = a[0], n1v = a[1], n2v = a[2];
            void* args[] = {&zp, &dp, &sp, &npairs, &n1v, &n2v, &tw};
            const int64_t half = r.r3_ti / 2;
            HIPCHK(hipLaunchKernel(r.r3rows, dim3((unsigned)((npairs + half - 1) / half)), dim3(r.r3_nt), args, 0, s));
            return NRFFT_OK;
        }
        case SP_R3_UNTANGLE: {
            const int64_t nk = a[3] > 1 ? a[3] / 2 : 1;
            const unsigned g = grid_of(nsig * a[1] * a[2] * nk, 256);
            cpx<T>* speq = (cpx<T>*)bufs[B_AUX2];
            if (isign > 0)
                hipLaunchKernelGGL((k_rlft3_untangle<T, 1>), dim3(g), dim3(256), 0, s, (cpx<T>*)dst, speq, a[1], a[2],
                                   a[3], nsig, (int)a[4], tw);
            else
                hipLaunchKernelGGL((k_rlft3_untangle<T, -1>), dim3(g), dim3(256), 0, s, (cpx<T>*)dst, speq, a[1],
                                   a[2], a[3], nsig, (int)a[4], tw);
            break;
        }
        default:
            return fail(NRFFT_E_INVALID, "unknown spectral launch");
    }
    HIPCHK(hipGetLastError());
    return NRFFT_OK;
}

int exec_spectral(const Plan& p, const Launch& l, void* const* bufs, int isign, hipStream_t stream, DeviceCtx* ctx) {
    if (p.dtype == NRFFT_F32) return exec_spec_t<float>(p, l, bufs, isign, stream, ctx);
    return exec_spec_t<double>(p, l, bufs, isign, stream, ctx);
}

// ---------------------------------------------------------------------------
// Validation (reference order first, then this build's power-of-two limits).
// ---------------------------------------------------------------------------
static int check_twofft(size_t n) {
    if (n == 0) return fail(NRFFT_E_EMPTY, "Input arrays cannot be empty");
    if (!is_pow2(n) || n < 2) return fail(NRFFT_E_NOT_POW2, "n must be a power of two >= 2");
    if (ilog2(n) > TWOFFT_MAX_LOGN) return fail(NRFFT_E_UNSUPPORTED, "n > 2^26");
    return NRFFT_OK;
}

static int check_cosft(size_t n) {
    if (n == 0) return fail(NRFFT_E_EMPTY, "n == 0");
    if (!is_pow2(n) || n < 2) return fail(NRFFT_E_NOT_POW2, "n must be a power of two >= 2");
    if (ilog2(n) > COSFT_MAX_LOGN) return fail(NRFFT_E_UNSUPPORTED, "n > 2^24");
    return NRFFT_OK;
}

// Real_FT3.rs:16-19 (isign, then the shapes); dims powers of two, nn3 >= 2.
static int check_rlft3(size_t nn1, size_t nn2, size_t nn3, int isign) {
    if (isign != 1 && isign != -1) return fail(NRFFT_E_ISIGN, "isign must be 1 or -1");
    if (nn1 == 0 || nn2 == 0 || nn3 == 0) return fail(NRFFT_E_EMPTY, "empty dimension");
    if (!is_pow2(nn1) || !is_pow2(nn2) || !is_pow2(nn3) || nn3 < 2)
        return fail(NRFFT_E_NOT_POW2, "dimensions must be powers of two, nn3 >= 2");
    return NRFFT_OK;
}

// ---------------------------------------------------------------------------
// Host-pointer calls: stage through HBM on the per-thread stream.
// ---------------------------------------------------------------------------
struct HostBuf {
    const void* src;  // copied in (nullptr: not an input)
    void* dst;        // copied out (nullptr: not an output)
    size_t bytes;
};

static int host_spectral(int kind, int dtype, std::vector<size_t> nn, HostBuf in, HostBuf aux, HostBuf out,
                         HostBuf aux2, int isign) {
    std::shared_ptr<Plan> p;
    int rc = cached_plan(kind, dtype, nn, 1, 0, &p);
    if (rc) return rc;
    hipStream_t s = hipStreamPerThread;
    DevBuf dbin, daux, dout, daux2, dtmp;
    // `out` may be the same host buffer as `in` (in-place calls): then one
    // device buffer serves both.
    const bool same = out.dst != nullptr && out.dst == in.src;
    if ((rc = dev_alloc(dbin, in.bytes, s))) return rc;
    if (!same && out.bytes && (rc = dev_alloc(dout, out.bytes, s))) return rc;
    if (aux.bytes && (rc = dev_alloc(daux, aux.bytes, s))) return rc;
    if (aux2.bytes && (rc = dev_alloc(daux2, aux2.bytes, s))) return rc;
    if ((rc = dev_alloc(dtmp, p->tmp_elems * csize(dtype), s))) return rc;
    void* dout_p = same ? dbin.p : dout.p;
    HIPCHK(hipMemcpyAsync(dbin.p, in.src, in.bytes, hipMemcpyHostToDevice, s));
    if (aux.src) HIPCHK(hipMemcpyAsync(daux.p, aux.src, aux.bytes, hipMemcpyHostToDevice, s));
    if (aux2.src) HIPCHK(hipMemcpyAsync(daux2.p, aux2.src, aux2.bytes, hipMemcpyHostToDevice, s));
    if ((rc = exec_plan(*p, dbin.p, dout_p, dtmp.p, daux.p, isign, s, 0, -1, daux2.p))) return rc;
    if (out.dst) HIPCHK(hipMemcpyAsync(out.dst, dout_p, out.bytes, hipMemcpyDeviceToHost, s));
    if (aux2.dst) HIPCHK(hipMemcpyAsync(aux2.dst, daux2.p, aux2.bytes, hipMemcpyDeviceToHost, s));
    HIPCHK(hipStreamSynchronize(s));
    return NRFFT_OK;
}

template <typename T>
static int host_twofft(const T* d1, const T* d2, size_t n, T* fft1, T* fft2, int dtype) {
    if (!d1 || !d2 || !fft1 || !fft2) return fail(NRFFT_E_INVALID, "null pointer");
    int rc = check_twofft(n);
    if (rc) return rc;
    const size_t ib = n * sizeof(T), ob = 2 * n * sizeof(T);
    rc = host_spectral(PK_TWOFFT, dtype, {n}, HostBuf{d1, nullptr, ib}, HostBuf{d2, nullptr, ib},
                       HostBuf{nullptr, fft1, ob}, HostBuf{nullptr, fft2, ob}, 1);
    if (rc) return rc;
    // FFT_2.rs:6-7 buffers hold 2n + 2 values; bins 0..n-1 fill the first 2n.
    fft1[2 * n] = fft1[2 * n + 1] = (T)0;
    fft2[2 * n] = fft2[2 * n + 1] = (T)0;
    return NRFFT_OK;
}

template <typename T>
static int host_cosft1(T* y, size_t n, int dtype) {
    if (!y) return fail(NRFFT_E_INVALID, "null pointer");
    int rc = check_cosft(n);
    if (rc) return rc;
    const size_t b = (n + 1) * sizeof(T);
    return host_spectral(PK_COSFT1, dtype, {n}, HostBuf{y, nullptr, b}, HostBuf{nullptr, nullptr, 0},
                         HostBuf{nullptr, y, b}, HostBuf{nullptr, nullptr, 0}, 1);
}

template <typename T>
static int host_cosft2(T* y, size_t n, int isign, int dtype) {
    if (isign != 1 && isign != -1) return fail(NRFFT_E_ISIGN, "Invalid isign value: must be 1 or -1");
    if (!y) return fail(NRFFT_E_INVALID, "null pointer");
    int rc = check_cosft(n);
    if (rc) return rc;
    const size_t b = n * sizeof(T);
    return host_spectral(PK_COSFT2, dtype, {n}, HostBuf{y, nullptr, b}, HostBuf{nullptr, nullptr, 0},
                         HostBuf{nullptr, y, b}, HostBuf{nullptr, nullptr, 0}, isign);
}

template <typename T>
static int host_rlft3(T* data, T* speq, size_t nn1, size_t nn2, size_t nn3, int isign, int dtype) {
    int rc = check_rlft3(nn1, nn2, nn3, isign);
    if (rc) return rc;
    if (!data || !speq) return fail(NRFFT_E_INVALID, "null pointer");
    const size_t db = nn1 * nn2 * nn3 * sizeof(T), sb = nn1 * 2 * nn2 * sizeof(T);
    return host_spectral(PK_RLFT3, dtype, {nn1, nn2, nn3}, HostBuf{data, nullptr, db}, HostBuf{nullptr, nullptr, 0},
                         HostBuf{nullptr, data, db}, HostBuf{isign < 0 ? speq : nullptr, speq, sb}, isign);
}

}  // namespace nrfft

using namespace nrfft;

extern "C" {

int nrfft_twofft_f64(const double* data1, const double* data2, size_t n, double* fft1, double* fft2) {
    g_err.clear();
    return host_twofft<double>(data1, data2, n, fft1, fft2, NRFFT_F64);
}
int nrfft_twofft_f32(const float* data1, const float* data2, size_t n, float* fft1, float* fft2) {
    g_err.clear();
    return host_twofft<float>(data1, data2, n, fft1, fft2, NRFFT_F32);
}
int nrfft_cosft1_f64(double* y, size_t n) {
    g_err.clear();
    return host_cosft1<double>(y, n, NRFFT_F64);
}
int nrfft_cosft1_f32(float* y, size_t n) {
    g_err.clear();
    return host_cosft1<float>(y, n, NRFFT_F32);
}
int nrfft_cosft2_f64(double* y, size_t n, int isign) {
    g_err.clear();
    return host_cosft2<double>(y, n, isign, NRFFT_F64);
}
int nrfft_cosft2_f32(float* y, size_t n, int isign) {
    g_err.clear();
    return host_cosft2<float>(y, n, isign, NRFFT_F32);
}
int nrfft_rlft3_f64(double* data, double* speq, size_t nn1, size_t nn2, size_t nn3, int isign) {
    g_err.clear();
    return host_rlft3<double>(data, speq, nn1, nn2, nn3, isign, NRFFT_F64);
}
int nrfft_rlft3_f32(float* data, float* speq, size_t nn1, size_t nn2, size_t nn3, int isign) {
    g_err.clear();
    return host_rlft3<float>(data, speq, nn1, nn2, nn3, isign, NRFFT_F32);
}

int nrfft_plan_twofft(nrfft_plan* plan, int dtype, size_t n, size_t batch) {
    g_err.clear();
    int rc = check_twofft(n);
    return rc ? rc : make_plan(plan, PK_TWOFFT, dtype, {n}, batch, 0);
}
int nrfft_execute_twofft(nrfft_plan plan, const void* d_data1, const void* d_data2, void* d_fft1, void* d_fft2,
                         void* stream) {
    g_err.clear();
    if (!plan || plan->plan.kind != PK_TWOFFT) return fail(NRFFT_E_INVALID, "not a twofft plan");
    if (!d_data1 || !d_data2 || !d_fft1 || !d_fft2) return fail(NRFFT_E_INVALID, "null buffer");
    return exec_plan(plan->plan, d_data1, d_fft1, plan->tmp, d_data2, 1, (hipStream_t)stream, 0, -1, d_fft2);
}

int nrfft_plan_cosft1(nrfft_plan* plan, int dtype, size_t n, size_t batch) {
    g_err.clear();
    int rc = check_cosft(n);
    return rc ? rc : make_plan(plan, PK_COSFT1, dtype, {n}, batch, 0);
}
int nrfft_execute_cosft1(nrfft_plan plan, const void* d_in, void* d_out, void* stream) {
    g_err.clear();
    if (!plan || plan->plan.kind != PK_COSFT1) return fail(NRFFT_E_INVALID, "not a cosft1 plan");
    if (!d_in || !d_out) return fail(NRFFT_E_INVALID, "null buffer");
    return exec_plan(plan->plan, d_in, d_out, plan->tmp, nullptr, 1, (hipStream_t)stream);
}

int nrfft_plan_cosft2(nrfft_plan* plan, int dtype, size_t n, size_t batch) {
    g_err.clear();
    int rc = check_cosft(n);
    return rc ? rc : make_plan(plan, PK_COSFT2, dtype, {n}, batch, 0);
}
int nrfft_execute_cosft2(nrfft_plan plan, const void* d_in, void* d_out, int isign, void* stream) {
    g_err.clear();
    if (!plan || plan->plan.kind != PK_COSFT2) return fail(NRFFT_E_INVALID, "not a cosft2 plan");
    if (isign != 1 && isign != -1) return fail(NRFFT_E_ISIGN, "Invalid isign value: must be 1 or -1");
    if (!d_in || !d_out) return fail(NRFFT_E_INVALID, "null buffer");
    return exec_plan(plan->plan, d_in, d_out, plan->tmp, nullptr, isign, (hipStream_t)stream);
}

int nrfft_plan_rlft3(nrfft_plan* plan, int dtype, size_t nn1, size_t nn2, size_t nn3, size_t batch) {
    g_err.clear();
    int rc = check_rlft3(nn1, nn2, nn3, 1);
    return rc ? rc : make_plan(plan, PK_RLFT3, dtype, {nn1, nn2, nn3}, batch, 0);
}
int nrfft_execute_rlft3(nrfft_plan plan, void* d_data, void* d_speq, int isign, void* stream) {
    g_err.clear();
    if (!plan || plan->plan.kind != PK_RLFT3) return fail(NRFFT_E_INVALID, "not an rlft3 plan");
    if (isign != 1 && isign != -1) return fail(NRFFT_E_ISIGN, "isign must be 1 or -1");
    if (!d_data || !d_speq) return fail(NRFFT_E_INVALID, "null buffer");
    return exec_plan(plan->plan, d_data, d_data, plan->tmp, nullptr, isign, (hipStream_t)stream, 0, -1, d_speq);
}

}  // extern "C"
