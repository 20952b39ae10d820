// Batched configuration: every q of a list prepared up front (each with its
// own chi cutoff and empty-band window), the whole launch sequence (pair
// densities, weights, contraction, Dyson step of every q, every frequency)
// captured once into a HIP graph and replayed as one submission; and the
// autotuner's record.  No counterpart in the reference, which runs one q per
// loop iteration of one process (chiqw.F:662-687).
//
// The graph's kernels take their buffers as arguments, so everything a
// captured node touches must stay put: each q keeps its inputs and its eps^-1
// in its own slot, and the stages' work buffers belong to the batch (sized to
// the largest q before capture), never to the handle's eager calls.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <string>

#include "rp_internal.h"

namespace rp {

void swap_q(rp_ctx* h, QSlot& s) {
  std::swap(h->qready, s.qready);
  std::swap(h->q0, s.q0);
  std::swap(h->have_A, s.have_A);
  std::swap(h->have_w, s.have_w);
  std::swap(h->have_chi, s.have_chi);
  for (int d = 0; d < 3; ++d) std::swap(h->q[d], s.q[d]);
  std::swap(h->ng, s.ng);
  std::swap(h->g0, s.g0);
  std::swap(h->npair, s.npair);
  h->lg.swap(s.lg);
  h->ikq.swap(s.ikq);
  h->shift.swap(s.shift);
  h->pocc.swap(s.pocc);
  h->pvir.swap(s.pvir);
  h->intra.swap(s.intra);
  h->pdplan.swap(s.pdplan);
  h->qn.swap(s.qn);
  std::swap(h->lda, s.lda);
  std::swap(h->kpad, s.kpad);
  std::swap(h->ecut_eps_win, s.ecut_eps_win);
  std::swap(h->n_vir_win, s.n_vir_win);
  std::swap(h->q_span, s.q_span);
  h->d_lg.swap(s.d_lg);
  h->d_ikq.swap(s.d_ikq);
  h->d_shift.swap(s.d_shift);
  h->d_pocc.swap(s.d_pocc);
  h->d_pvir.swap(s.d_pvir);
  h->d_intra.swap(s.d_intra);
  h->d_pdplan.swap(s.d_pdplan);
  h->d_qn.swap(s.d_qn);
  h->d_eps.swap(s.d_eps);
}

void swap_scratch(rp_ctx* h, Scratch& s) {
  h->d_A.swap(s.d_A);
  h->d_A3.swap(s.d_A3);
  h->d_wf.swap(s.d_wf);
  h->d_Ao.swap(s.d_Ao);
  h->d_sync.swap(s.d_sync);
  h->d_w.swap(s.d_w);
  h->d_part.swap(s.d_part);
  h->d_chi.swap(s.d_chi);
  h->d_drude.swap(s.d_drude);
  h->d_wd.swap(s.d_wd);
  h->d_status.swap(s.d_status);
  h->d_ipiv.swap(s.d_ipiv);
  h->d_gjwork.swap(s.d_gjwork);
}

namespace {

// Restores the handle's own per-q state and work buffers however the plan
// ends, and ends a capture left open by an error.
struct BatchScope {
  rp_ctx* h;
  Batch* b;
  QSlot saved;
  int current = -1;  // slot swapped into the handle, if any
  BatchScope(rp_ctx* h_, Batch* b_) : h(h_), b(b_) {
    swap_q(h, saved);
    swap_scratch(h, b->scratch);
  }
  void enter(int i) {
    swap_q(h, b->slots[i]);
    current = i;
  }
  void leave() {
    if (current >= 0) swap_q(h, b->slots[current]);
    current = -1;
  }
  ~BatchScope() {
    leave();
    if (h->capturing) {
      hipGraph_t g = nullptr;
      (void)hipStreamEndCapture(h->stream, &g);
      if (g) (void)hipGraphDestroy(g);
      h->capturing = false;
    }
    swap_scratch(h, b->scratch);
    swap_q(h, saved);
  }
};

}  // namespace

}  // namespace rp

using namespace rp;

extern "C" {

int rp_chiqw_autotune(rp_handle h, int32_t enable) {
  return guarded([&] {
    RP_CHECK(h, "rp_chiqw_autotune: null handle");
    h->autotune = enable ? 1 : 0;
  });
}

int rp_chiqw_tune_info(rp_handle h, int32_t idx, int32_t* iout, double* dout) {
  return guarded([&] {
    RP_CHECK(h, "rp_chiqw_tune_info: null handle");
    RP_CHECK(idx >= 0 && idx < (int)h->tune.size(), "rp_chiqw_tune_info: no such entry");
    const TuneEntry& t = h->tune[idx];
    const int32_t v[10] = {t.prec,     t.ng,           t.nen,         t.ndir,          (int32_t)t.kpad,
                           t.ncand,    t.best.nsplit,  t.best.nwg,    t.heuristic.nsplit, t.heuristic.nwg};
    std::memcpy(iout, v, sizeof v);
    dout[0] = t.best_ms;
    dout[1] = t.heuristic_ms;
  });
}

int rp_chiqw_tune_count(rp_handle h, int32_t* n) {
  return guarded([&] {
    RP_CHECK(h, "rp_chiqw_tune_count: null handle");
    *n = (int32_t)h->tune.size();
  });
}

int rp_chiqw_batch_plan(rp_handle h, int32_t nq, const double* qs, const double* ecut_for_eps,
                        const int32_t* n_vir, int32_t precision) {
  return guarded([&] {
    RP_CHECK(h && h->open, "rp_chiqw_batch_plan: system not open");
    RP_CHECK(nq > 0 && qs, "rp_chiqw_batch_plan: empty q list");
    RP_CHECK(precision >= RP_PREC_FP64 && precision <= RP_PREC_BF16X3, "unknown precision");
    RP_HIP(hipSetDevice(h->device));
    h->batch.reset();
    auto batch = std::make_unique<Batch>();
    Batch* b = batch.get();
    b->prec = precision;
    b->slots.resize(nq);
    {
      BatchScope scope(h, b);
      h->d_status.reserve(1);  // the batch's own status word
      // 1. prepare every q into its own slot
      for (int i = 0; i < nq; ++i) {
        h->ecut_eps_win = ecut_for_eps ? ecut_for_eps[i] : 0.0;
        h->n_vir_win = n_vir ? n_vir[i] : 0;
        RP_CHECK(h->ecut_eps_win >= 0.0 && h->ecut_eps_win <= h->ecut_psi && h->n_vir_win >= 0,
                 "rp_chiqw_batch_plan: bad window");
        if (rp_chiqw_prepare_q(h, qs + 3 * i, nullptr, nullptr, 0)) throw Error(g_err);
        swap_q(h, b->slots[i]);  // the slot takes q i; the handle is empty again
      }
      // 2. one eager pass over every q: checks each stage, tunes each shape
      //    (when autotuning is on) and grows the work buffers to the largest q
      for (int i = 0; i < nq; ++i) {
        scope.enter(i);
        issue_pairdens(h);
        h->have_A = true;
        issue_weights(h);
        check_status(h, "batch: tetrahedron weights");
        h->have_w = true;
        // (the launch is not part of the swapped state: it stays with the slot)
        b->slots[i].launch = chi_launch(h, precision);
        issue_accumulate(h, precision, b->slots[i].launch, false);
        h->have_chi = true;
        issue_dyson(h);
        check_status(h, "batch: Dyson step");
        // the span this q's eps^-1 calls for (split-BF16): the graph uses it
        if (const int sp = adapt_span(h, precision, b->slots[i].launch.span)) {
          b->slots[i].launch.span = sp;
          h->q_span = sp;
          issue_accumulate(h, precision, b->slots[i].launch, false);
          issue_dyson(h);
          check_status(h, "batch: Dyson step");
        }
        scope.leave();
      }
      // 3. capture every q's launches into one graph.  The pair densities,
      //    weights and contraction operands of q go on the second stream, after
      //    q-1's contraction (the last reader of their buffers), so that they
      //    run beside q-1's Dyson step, which leaves most compute units idle;
      //    the contraction of q waits for them.
      h->capturing = true;
      RP_HIP(hipStreamBeginCapture(h->stream, hipStreamCaptureModeRelaxed));
      launch_zero(h->d_status.p, sizeof(int), h->stream);
      RP_HIP(hipEventRecord(h->bev[0], h->stream));
      RP_HIP(hipStreamWaitEvent(h->stream2, h->bev[0], 0));
      for (int i = 0; i < nq; ++i) {
        scope.enter(i);
        {
          OnStream2 on2(h);  // the stages issue on h->stream
          issue_pairdens(h);
          issue_weights(h);
          issue_operands(h, precision);
          RP_HIP(hipEventRecord(h->bev[1], h->stream));
        }
        RP_HIP(hipStreamWaitEvent(h->stream, h->bev[1], 0));
        issue_accumulate(h, precision, b->slots[i].launch, false, true);
        RP_HIP(hipEventRecord(h->bev[2], h->stream));
        RP_HIP(hipStreamWaitEvent(h->stream2, h->bev[2], 0));
        issue_dyson(h);
        scope.leave();
      }
      // the second stream joins back before the capture ends
      RP_HIP(hipEventRecord(h->bev[1], h->stream2));
      RP_HIP(hipStreamWaitEvent(h->stream, h->bev[1], 0));
      hipGraph_t g = nullptr;
      h->capturing = false;
      RP_HIP(hipStreamEndCapture(h->stream, &g));
      b->graph = g;
      RP_HIP(hipGraphInstantiate(&b->exec, b->graph, nullptr, nullptr, 0));
    }
    h->batch = std::move(batch);
  });
}

int rp_chiqw_batch_run(rp_handle h, double* seconds) {
  return guarded([&] {
    RP_CHECK(h && h->batch && h->batch->exec, "rp_chiqw_batch_run: no batch planned");
    RP_HIP(hipSetDevice(h->device));
    Batch* b = h->batch.get();
    RP_HIP(hipEventRecord(h->ev[4], h->stream));
    RP_HIP(hipGraphLaunch(b->exec, h->stream));
    RP_HIP(hipEventRecord(h->ev[5], h->stream));
    RP_HIP(hipEventSynchronize(h->ev[5]));
    // the graph's own status word
    swap_scratch(h, b->scratch);
    try {
      check_status(h, "batched run");
    } catch (...) {
      swap_scratch(h, b->scratch);
      throw;
    }
    swap_scratch(h, b->scratch);
    if (seconds) {
      float ms = 0;
      RP_HIP(hipEventElapsedTime(&ms, h->ev[4], h->ev[5]));
      *seconds = ms * 1e-3;
    }
  });
}

int rp_chiqw_batch_result(rp_handle h, int32_t iq, int32_t* ng, int32_t* ndir, int32_t* lg,
                          int32_t lg_cap, double* epsinv_out) {
  return guarded([&] {
    RP_CHECK(h && h->batch, "rp_chiqw_batch_result: no batch planned");
    Batch* b = h->batch.get();
    RP_CHECK(iq >= 0 && iq < (int)b->slots.size(), "rp_chiqw_batch_result: no such q");
    const QSlot& s = b->slots[iq];
    if (ng) *ng = s.ng;
    if (ndir) *ndir = s.q0 ? 3 : 1;
    if (lg) std::memcpy(lg, s.lg.data(), sizeof(int) * std::min<size_t>(s.lg.size(), 3 * (size_t)lg_cap));
    if (epsinv_out) {
      RP_HIP(hipSetDevice(h->device));
      const size_t n = (size_t)(s.q0 ? 3 : 1) * h->nen * s.ng * s.ng;
      RP_HIP(hipMemcpyAsync(epsinv_out, s.d_eps.p, sizeof(double2) * n, hipMemcpyDeviceToHost, h->stream));
      RP_HIP(hipStreamSynchronize(h->stream));
    }
  });
}

int rp_chiqw_batch_free(rp_handle h) {
  return guarded([&] {
    RP_CHECK(h, "rp_chiqw_batch_free: null handle");
    h->batch.reset();
  });
}

}  // extern "C"
