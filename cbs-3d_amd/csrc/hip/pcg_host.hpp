// Host side of the device PCG loops shared by the solver forms
// (k_pressure.hip: FP64 and mixed; k_amg.hip: multigrid-preconditioned):
// the device-resident scalars, chunked iteration launches and their HIP
// graph capture, and the result record.
#pragma once

#include <algorithm>
#include <climits>
#include <cstddef>

#include "ctx.hpp"
#include "device_common.hpp"

namespace cbsdev
{
    // check codes of the PCG error keys (ctx.hip maps them to the
    // reference's messages)
    enum CgCheck : uint32_t
    {
        CK_BAD_DIAG = 1, CK_BAD_R = 2, CK_RZ_INIT = 3, CK_BREAKDOWN = 4, CK_RZ_UPD = 5, CK_NOT_FREE = 6
    };

    // iterations per captured graph block
    constexpr int kGraphIters = 8;
    // without graphs, every kTimedStride-th iteration carries the timers
    constexpr int kTimedStride = 8;

    // Host copies of the CG scalars of a new solve (only the part from rz_old
    // on is uploaded).
    inline void reset_cg_scalars(cbs_hip_ctx* c)
    {
        for (int k = 0; k < c->nseg; ++k)
        {
            DevScalars& h = c->h_scal[k];
            h = DevScalars{};
            h.err = ~0ull;
            h.pending_err = ~0ull;
            h.max_iter = c->cfg.cg_max_iter;
            h.conv_test = c->cfg.cg_conv_test;
            h.rel_tol = c->cfg.relToler;
            h.abs_tol = c->cfg.absToler;
        }
    }

    // Pushes only the CG part of every segment's scalars (from rz_old on;
    // the time-step and norm scalars before it stay device-owned).
    inline void upload_cg_scalars(cbs_hip_ctx* c)
    {
        const size_t off = offsetof(DevScalars, rz_old);
        const size_t len = sizeof(DevScalars) - off;
        hip_check(hipMemcpy2DAsync(reinterpret_cast<char*>(c->scal) + off, sizeof(DevScalars),
                                   reinterpret_cast<char*>(c->h_scal) + off, sizeof(DevScalars), len, size_t(c->nseg),
                                   hipMemcpyHostToDevice, c->stream),
                  "upload CG scalars");
    }

    inline bool all_done(const cbs_hip_ctx* c, int DevScalars::*flag)
    {
        for (int k = 0; k < c->nseg; ++k)
            if (!(c->h_scal[k].*flag)) return false;
        return true;
    }

    inline int max_iter_of(const cbs_hip_ctx* c)
    {
        int m = 0;
        for (int k = 0; k < c->nseg; ++k) m = std::max(m, c->h_scal[k].iter);
        return m;
    }

    // iter_before >= 0: the scalars close a chunk of iterations that began at
    // that iteration count (single segment); only its applies that ran are
    // kept by the timer.
    inline void download_scalars(cbs_hip_ctx* c, int iter_before = -1)
    {
        hip_check(hipMemcpyAsync(c->h_scal, c->scal, sizeof(DevScalars) * size_t(c->nseg), hipMemcpyDeviceToHost,
                                 c->stream),
                  "download scalars");
        hip_check(hipStreamSynchronize(c->stream), "stream sync");
        const int keep = iter_before >= 0 ? max_iter_of(c) - iter_before : INT_MAX;
        timer_collect(c, keep);
        if (c->profile) phase_collect(c, keep);
    }

    // With profiling on, `launch` timed as one group of phase `ph` (for the
    // launches of a solve that run once or a few times per solve).
    template <typename F>
    inline void timed_group(cbs_hip_ctx* c, int ph, F launch)
    {
        if (c->profile) phase_mark(c, -1);
        launch();
        if (c->profile) phase_mark(c, ph);
    }

    // Copies the per-segment outcome of a solve into c->seg_result and the
    // call's result: one segment reports itself; a batch reports the member
    // with the most iterations, converged only when every member converged.
    inline void collect_results(cbs_hip_ctx* c, cbs_cg_result* res, double secs, double bytes_per_iter)
    {
        c->seg_result.assign(size_t(c->nseg), cbs_cg_result{});
        int worst = 0;
        bool all_conv = true;
        for (int k = 0; k < c->nseg; ++k)
        {
            const DevScalars& h = c->h_scal[k];
            cbs_cg_result& r = c->seg_result[size_t(k)];
            r.iterations = h.iter;
            r.converged = h.converged;
            r.initial_l2 = h.initial_l2;
            r.final_l2 = h.final_l2;
            r.final_relative_l2 = h.rel_l2;
            r.final_max_abs = h.max_abs;
            r.matvec_calls = 1 + h.iter;
            r.dot_calls = 2 + 3 * h.iter;
            r.preconditioner_calls = 1 + h.iter;
            r.total_seconds = secs;
            r.algorithmic_bytes = bytes_per_iter * h.iter;
            all_conv = all_conv && h.converged;
            if (h.iter > c->h_scal[worst].iter) worst = k;
        }
        *res = c->seg_result[size_t(worst)];
        res->converged = all_conv ? 1 : 0;
        if (c->nseg > 1)
        {
            double bytes = 0.0;
            for (const auto& r : c->seg_result) bytes += r.algorithmic_bytes;
            res->algorithmic_bytes = bytes;
        }
    }

    // Captures the block of form `form` (0 FP64, 1 mixed, 2 multigrid);
    // apply() and rest() launch one iteration.
    template <typename Apply, typename Rest>
    inline bool capture_block(cbs_hip_ctx* c, int form, Apply apply, Rest rest)
    {
        cbs_hip_ctx::PcgGraph& g = c->pcg_graph[form];
        if (hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal) != hipSuccess)
        {
            (void)hipGetLastError();
            return false;
        }
        for (int k = 0; k < kGraphIters; ++k)
        {
            apply();
            rest(PhaseMark{});
        }
        bool ok = hipGetLastError() == hipSuccess;
        hipGraph_t graph = nullptr;
        ok = (hipStreamEndCapture(c->stream, &graph) == hipSuccess) && ok && graph;
        if (ok) ok = hipGraphInstantiate(&g.exec, graph, nullptr, nullptr, 0) == hipSuccess;
        if (graph) (void)hipGraphDestroy(graph);
        (void)hipGetLastError();
        if (!ok)
        {
            g.exec = nullptr;
            return false;
        }
        return true;
    }

    // Launches at least n iterations of form `form` (precision code prec for
    // the timer) and returns how many were launched: replays of the captured
    // block when graphs are on, else one launch sequence per iteration.
    // rest(mark) launches the iteration's kernels after the apply and calls
    // mark(phase) after each group of them; with profiling on, the
    // iterations that carry the timing record events there (one eager
    // iteration in the middle of a chunk of graph replays, away from the
    // host synchronisation before it; every kTimedStride-th otherwise).
    // allow_timer: the apply timer of `prec` also times that iteration.
    template <typename Apply, typename Rest>
    inline int launch_iterations(cbs_hip_ctx* c, int form, int prec, int n, bool allow_timer, int64_t& seq,
                                 Apply apply, Rest rest)
    {
        // the apply timer's event pair brackets the apply alone; the phase
        // marks sit outside it
        auto timed_iteration = [&](int idx, int64_t sq) {
            phase_mark(c, -1, idx);
            const bool t = allow_timer && timer_begin(c, prec, idx, sq);
            apply();
            timer_end(c, t);
            phase_mark(c, PH_MATVEC);
            rest(PhaseMark{c, true});
        };
        if (c->use_graphs)
        {
            cbs_hip_ctx::PcgGraph& g = c->pcg_graph[form];
            if (!g.exec && !capture_block(c, form, apply, rest)) c->use_graphs = false;
            if (g.exec)
            {
                const bool timed = c->profile;
                const int reps = std::max(timed ? 0 : 1, (n - (timed ? 1 : 0) + kGraphIters - 1) / kGraphIters);
                const int before = timed ? reps / 2 : reps;
                for (int r = 0; r < before; ++r) hip_check(hipGraphLaunch(g.exec, c->stream), "hipGraphLaunch");
                if (timed) timed_iteration(before * kGraphIters, 0);
                for (int r = before; r < reps; ++r) hip_check(hipGraphLaunch(g.exec, c->stream), "hipGraphLaunch");
                return reps * kGraphIters + (timed ? 1 : 0);
            }
        }
        for (int k = 0; k < n; ++k)
        {
            const int64_t sq = seq++;
            if (c->profile && sq % kTimedStride == 0)
            {
                timed_iteration(k, sq);
                continue;
            }
            apply();
            rest(PhaseMark{});
        }
        return n;
    }
}
