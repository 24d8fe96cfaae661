// queue.hpp -- uneven batches: one persistent work-queue kernel for many
// transforms of different lengths and dimensionality.
//
// The host turns every item into tiles (tasks) of one Stockham pass each:
//   1-D, N <= QPTS            one row tile (the whole transform)
//   1-D, N  > QPTS            four-step: N1 column tiles with the W_N twiddle
//                             (in -> workspace), then N2 row tiles stored
//                             transposed (workspace -> out)
//   2-D, n0 x n1              row tiles along n1 (in -> out), then column
//                             tiles along n0 (in place in out)
// Phase-1 tasks (first passes and single-pass items) are ordered largest
// first; phase-2 tasks (second passes) follow in the same item order.  Each
// phase is one dispatch of the persistent kernel; its groups take tasks from
// an atomic counter (dynamic load balance).  The two dispatches are the only
// synchronisation: the kernel boundary makes phase-1 results written on one
// XCD's L2 visible to phase-2 readers on another.  (Per-task release/acquire
// fences inside one dispatch write back / invalidate a whole L2 each on this
// chip -- hundreds of thousands of them cost far more than a kernel
// boundary.)
#pragma once
#include "pass_c2c.hpp"

namespace nrfft {

// Tile points (TI x L) of the queue kernel: EPT = 16 values per thread, so a
// group has qpts / 16 threads.  Build switches: NRFFT_QPTS_F32 / _F64.
#ifndef NRFFT_QPTS_F32
#define NRFFT_QPTS_F32 16384
#endif
#ifndef NRFFT_QPTS_F64
#define NRFFT_QPTS_F64 8192
#endif
template <typename T>
constexpr int qpts() {
    return sizeof(T) == 4 ? NRFFT_QPTS_F32 : NRFFT_QPTS_F64;
}
constexpr int Q_MIN_LOGL = 6;
// Longest pass of a tile: at most the tile itself and 8192 points (a single
// 16384-point FP32 tile -- four stages, one transform -- measured slower on
// the 10,000-item batch than the two-pass four-step it replaces).
constexpr int Q_MAX_LOGL = 13;
template <typename T>
constexpr int q_max_logl() {
    return ilog2c(qpts<T>()) < Q_MAX_LOGL ? ilog2c(qpts<T>()) : Q_MAX_LOGL;
}
constexpr int Q_MODES = 3;  // 0: row (JJ), 1: column (TT), 2: row load / column store (JT)

template <typename T>
constexpr int q_nclasses() {
    return (q_max_logl<T>() - Q_MIN_LOGL + 1) * Q_MODES;
}

template <typename T>
constexpr int q_lds_elems() {
    return qpts<T>() + qpts<T>() / NRFFT_LDS_PAD + 1;
}

struct QTask {
    int32_t cls;        // (log2 L - Q_MIN_LOGL) * Q_MODES + mode
    int32_t nvalid;     // transforms of the tile that exist
    int32_t src, dst;   // 0 = in, 1 = out, 2 = workspace
    int64_t ioff, ooff; // element offsets of the tile base
    int32_t si0, esi, so0, eso;
    int32_t tw_logn;    // four-step twiddle W_N^{k (tw_n2base + t)}, N = 2^tw_logn (0 = none)
    int32_t tw_n2base;
};

struct QState {
    uint32_t next[2];  // task counters of the two phases
};

// Stages 1 .. NST-2 of a tile class read their twiddles from an LDS copy
// (at most the 8192-point tables: 612 entries); a group refreshes it when
// its next task has another transform length.
template <int L>
constexpr int q_ltm() {
    return RadixPlan<L>::NST >= 3 ? RadixPlan<L>::NST - 2 : 0;
}
constexpr int Q_LTW_ELEMS = 612;

template <typename T, int S, int C>
__device__ __forceinline__ void q_run_class(const QTask& q, cpx<T>* const* bufs, cpx<T>* lds, cpx<T>* ltw,
                                            bool fresh_tw, const Twiddles<T>& tw, int tid) {
    constexpr int LOGL = C / Q_MODES + Q_MIN_LOGL;
    constexpr int MODE = C % Q_MODES;
    constexpr int L = 1 << LOGL;
    if constexpr (L > qpts<T>()) {
        return;
    } else {
        constexpr int TI = qpts<T>() / L;
        constexpr int NT = qpts<T>() / 16;
        constexpr int ML = MODE == 1 ? MAP_T : MAP_J;
        constexpr int MS = MODE == 0 ? MAP_J : MAP_T;
        using E = Engine<T, S, L, TI, NT, ML, MS>;
        static_assert(E::LDS_ELEMS <= q_lds_elems<T>(), "queue LDS too small for a tile class");
        static_assert(TI * L == qpts<T>() && E::EPT == 16, "queue tile shape");
        constexpr int LTM = q_ltm<L>();
        static_assert(LTM == 0 || E::tw_elems(LTM) <= Q_LTW_ELEMS, "queue twiddle copy too small");
        // Published by the first exchange barrier; the loop-top barrier of
        // k_queue ordered the previous task's reads before this overwrite.
        if constexpr (LTM > 0)
            if (fresh_tw) E::template stage_tw_to_lds<LTM>(ltw, tw.stage, tid);
        const cpx<T>* ib = bufs[q.src] + q.ioff;
        cpx<T>* ob = bufs[q.dst] + q.ooff;
        const int nvalid = q.nvalid, si0 = q.si0, esi = q.esi, so0 = q.so0, eso = q.eso;
        cpx<T> v[E::EPT];
        E::load0(v, tid, [&](int t, int i) -> cpx<T> {
            if (t < nvalid) return ib[t * si0 + i * esi];
            return cpx<T>{(T)0, (T)0};
        });
        E::template compute<0>(v, tid, tw.stage);
        E::template middle<1, LTM>(v, tid, lds, tw.stage, false, ltw);
        if (q.tw_logn) {
            const int logn = q.tw_logn;
            const uint32_t mask = (1u << logn) - 1u;
            E::store_last(v, tid, [&](int t, int j, const cpx<T>* w) {
                if (t < nvalid) {
                    const uint32_t n2 = uint32_t(q.tw_n2base + t);
                    cpx<T> f[E::RLAST];
                    twiddle_run<E::RLAST, S>(tw, (uint32_t(j) * n2) & mask, (uint32_t(E::NBLAST) * n2) & mask, logn, f);
#pragma unroll
                    for (int r = 0; r < E::RLAST; ++r) ob[t * so0 + (j + r * E::NBLAST) * eso] = cmul(w[r], f[r]);
                }
            });
        } else {
            E::store_last(v, tid, [&](int t, int j, const cpx<T>* w) {
                if (t < nvalid) {
#pragma unroll
                    for (int r = 0; r < E::RLAST; ++r) ob[t * so0 + (j + r * E::NBLAST) * eso] = w[r];
                }
            });
        }
    }
}

template <typename T, int S, int... Cs>
__device__ __forceinline__ void q_dispatch(int cls, const QTask& q, cpx<T>* const* bufs, cpx<T>* lds, cpx<T>* ltw,
                                           bool fresh_tw, const Twiddles<T>& tw, int tid,
                                           std::integer_sequence<int, Cs...>) {
    ((cls == Cs ? q_run_class<T, S, Cs>(q, bufs, lds, ltw, fresh_tw, tw, tid) : void()), ...);
}

// Reset of the queue counters (a fill kernel, so the sequence can live in a
// captured graph).
__global__ void k_queue_reset(QState* st) {
    if (threadIdx.x < 2) st->next[threadIdx.x] = 0;
}

// One phase: tasks [base, base + count) of the task list.
template <typename T, int S>
__global__ __launch_bounds__(qpts<T>() / 16) void k_queue(cpx<T>* in, cpx<T>* out, cpx<T>* ws, const QTask* tasks,
                                                            int base, int count, int phase, QState* st,
                                                            Twiddles<T> tw) {
    __shared__ cpx<T> lds[q_lds_elems<T>()];
    __shared__ __align__(16) cpx<T> ltw[Q_LTW_ELEMS];
    __shared__ int sh_task;
    cpx<T>* bufs[3] = {in, out, ws};
    const int tid = threadIdx.x;
    int tw_logl = -1;  // transform length whose twiddles ltw holds
    for (;;) {
        lds_barrier();  // previous task's LDS readers are done; sh_task free
        if (tid == 0) sh_task = (int)atomicAdd(&st->next[phase], 1u);
        lds_barrier();
        const int k = sh_task;
        if (k >= count) break;
        const QTask q = tasks[base + k];
        const int logl = q.cls / Q_MODES + Q_MIN_LOGL;
        q_dispatch<T, S>(q.cls, q, bufs, lds, ltw, logl != tw_logl, tw, fresh_tid(tid),
                         std::make_integer_sequence<int, q_nclasses<T>()>{});
        tw_logl = logl;
    }
}

}  // namespace nrfft
