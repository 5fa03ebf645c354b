// mint_sched.hip — step scheduling for the three forms, hipGraph capture of
// the batched launch window, and the launch-shape autotuner.
//
// Reference: the host loop src/mercurius.cu:179-198 launches the fused kernel
// once per 1e6-step batch; the kernel's own loop (src/sim.cuh:679-742) runs
// the steps with six cluster barriers each.  Here the massive subsystem runs
// ahead into an ephemeris buffer and the TPs read it without any barrier.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <climits>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <sstream>
#include <string>
#include <vector>

#include "mint_api_internal.h"

namespace mint {

std::string g_create_error;

namespace {

constexpr int DEF_BLOCK = 256, DEF_P = 1, DEF_C = 64;

struct Shape {
  int block, P, C;
};

// The launch shape of an nsteps call: the configuration's held values, else
// the tuned ones.  A call of one launch window has its own tuned block and
// TPs per lane (where neither is held): with no window after it, its queue
// work only has to fit beside one main kernel, which favours smaller blocks.
Shape batched_shape(const Ctx *c, int64_t nsteps) {
  Shape s;
  s.block = c->cfg.block_size > 0 ? c->cfg.block_size : (c->tuned_block > 0 ? c->tuned_block : DEF_BLOCK);
  s.P = c->cfg.tps_per_lane > 0 ? c->cfg.tps_per_lane : (c->tuned_P > 0 ? c->tuned_P : DEF_P);
  s.C = c->cfg.chunk_steps > 0 ? c->cfg.chunk_steps : (c->tuned_C > 0 ? c->tuned_C : DEF_C);
  if (nsteps > 0 && nsteps <= s.C && c->cfg.block_size <= 0 && c->cfg.tps_per_lane <= 0 && c->tuned_block1 > 0) {
    s.block = c->tuned_block1;
    s.P = c->tuned_P1;
  }
  return s;
}

// encounter kernel: one-wave blocks, enough to put a wave on every SIMD twice
int enc_grid(const Ctx *c) {
  long long g = (c->Ntp + 63) / 64;
  return (int)std::max(1LL, std::min(g, 2048LL));
}



// Ephemeris of one launch window into buffer (xs0, rec) on stream st, from the
// massive state (in, in + nb) (positions, then velocities) to the snapshot
// `next`, the window's end state.
int window_eph(Ctx *c, hipStream_t st, int C, const double *in, const double *in_v, double *xs0, double *rec,
               double *next, KTimer *kt) {
  const StepConst sc = c->step_const();
  if (kt) kt->start();
  unsigned long long *stats = (c->cfg.collect_stats & 1) ? c->d_stats : nullptr;
  const size_t nb = (size_t)c->S * c->NM * 3;
  int rc = c->hip(c->L->ephemeris(st, 0, c->S, C, c->d_prm, in, in_v, nullptr, nullptr, xs0, rec, next, next + nb,
                                  c->d_status, stats, sc, c->cfg.central_mass),
                  "ephemeris launch");
  if (rc) return rc;
  if (kt) kt->stop(c->timing.ms_ephemeris, c->timing.n_ephemeris);
  return MINT_OK;
}

// Event i of a per-window event list (created on first use).
int window_event(Ctx *c, std::vector<hipEvent_t> &v, size_t i, hipEvent_t *out) {
  while (v.size() <= i) {
    hipEvent_t e = nullptr;
    int rc = c->hip(hipEventCreateWithFlags(&e, hipEventDisableTiming), "event create");
    if (rc) return rc;
    v.push_back(e);
  }
  *out = v[i];
  return MINT_OK;
}

// Enqueue nsteps of the batched form as launch windows of up to sh.C steps.
//
// Window w, on the context stream: reset of the queue of parity w%2 (for
// w > 0), the queue-head pass (TPs carried from window w-1, plus the scorer's
// picks), then the main kernel, whose mid-window deferrals are queued for
// their tail and carried into window w+1.  On the queue stream, forked after
// the head pass: the tails of window w-1 (perturber cache, encounter kernel),
// then the head of window w (cache, encounter kernel); joined back before
// window w+1.  A tail thus runs beside the next main kernel instead of after
// its own.  The ephemeris stream computes window w into buffer w%3 while
// window w-1 runs; that buffer was last read by window w-3 and by its tails,
// served in window w-2.  The last window's tails run on the context stream at
// the end.  Without pipelining (or when timing each kernel) the same launches
// run in the same order on the context stream.
//
// Nothing writes the resident massive state (d_mx, d_mv) until the call's
// last kernel: window 0's ephemeris and queue head read it as the call's start
// state, each window's ephemeris writes its end state to the next window's
// snapshot, and the last kernel copies the last snapshot back and clears the
// queue of parity 0 for the next call (upload leaves it clear).  So the first
// ephemeris and the first queue head depend on nothing and start at once.
int enqueue_batched(Ctx *c, const Shape &sh, int64_t nsteps, bool pipelined, KTimer *kt, bool capturing) {
  std::vector<int> win;
  for (int64_t done = 0; done < nsteps;) {
    const int C = (int)std::min<int64_t>(sh.C, nsteps - done);
    win.push_back(C);
    done += C;
  }
  const int W = (int)win.size();
  const StepConst sc = c->step_const();
  const TPArrays tp = c->tp_arrays();
  unsigned long long *stats = (c->cfg.collect_stats & 1) ? c->d_stats : nullptr;
  const float *nnw = c->d_nnw_kep;
  const float *scw = c->cfg.use_scorer ? c->d_nnw_sc : nullptr;
  const int guess = c->cfg.kepler_guess;
  const int tpx = (c->ntiles + 7) / 8;
  const int nblocks = 8 * tpx;
  const int S = c->S;
  double *xs[3] = {c->d_xs0, c->d_xs0b, c->d_xs0c};
  double *rec[3] = {c->d_rec, c->d_recb, c->d_recc};
  QEntry *q[2] = {c->d_q, c->d_q2};
  // per parity: [0] queue-head entries, [1] systems listed by the head, [3]
  // systems listed by the tails, [4] mid-window deferrals (the tails, queued
  // after the head)
  unsigned int *ctl[2] = {c->d_qctl, c->d_qctl + 8};
  unsigned int *need[2] = {c->d_need, c->d_need + S};
  int *slist_head[2] = {c->d_slist, c->d_slist + 2 * S};
  int *slist_tail[2] = {c->d_slist + S, c->d_slist + 3 * S};
  const bool pipe = pipelined && !kt;
  hipStream_t sm = c->stream;
  hipStream_t sp = pipe ? c->stream2 : sm;  // ephemeris
  hipStream_t sq = pipe ? c->stream3 : sm;  // queue work
  int rc;
  auto record = [&](std::vector<hipEvent_t> &v, size_t i, hipStream_t st) -> int {
    if (!pipe) return MINT_OK;
    hipEvent_t e;
    int r = window_event(c, v, i, &e);
    return r ? r : c->hip(hipEventRecord(e, st), "event record");
  };
  // collect_stats bit 2: the main kernels stamp their wavefronts' entry and
  // exit (buffer sized by batched_setup), read by mint_step after the call
  const bool ktime = (c->cfg.collect_stats & 4) && !kt && c->d_kstamp != nullptr;
  auto wait = [&](hipStream_t st, std::vector<hipEvent_t> &v, size_t i) -> int {
    if (!pipe) return MINT_OK;
    hipEvent_t e;
    int r = window_event(c, v, i, &e);
    return r ? r : c->hip(hipStreamWaitEvent(st, e, 0), "event wait");
  };
  // encounter work of window u on stream st: the queue head (entries
  // [0, head count), systems flagged 1) or the tails (entries after the head,
  // systems flagged 2; those flagged 1 are cached already).  Beside a main
  // kernel long enough to hide it (a large batch) the entries are packed one
  // per lane densely into few waves; otherwise the chains are on the critical
  // path and each entry gets a group of lanes that splits its BS work.
  const bool hidden = c->Ntp >= (1LL << 20);
  auto serve = [&](int u, bool tail, hipStream_t st) -> int {
    const int dense = (st != sm && hidden) ? 1 : 0;
    const int pu = u % 2, bu = u % 3;
    if (kt) kt->start();
    int r = c->hip(c->L->pcache(st, tail ? slist_tail[pu] : slist_head[pu], tail ? ctl[pu] + 3 : ctl[pu] + 1,
                                win[u], rec[bu], c->d_pc, sc),
                   "perturber cache launch");
    if (r) return r;
    const int grid = dense ? std::min(enc_grid(c), 256) : enc_grid(c);
    r = c->hip(c->L->tp_encounter(st, guess, grid, 64, q[pu], tail ? ctl[pu] + 4 : ctl[pu],
                                  tail ? ctl[pu] : nullptr, dense, !dense, tp, c->d_prm, xs[bu], rec[bu], c->d_pc,
                                  win[u], sc, nnw, c->d_status, stats),
               "encounter launch");
    if (r) return r;
    if (kt) kt->stop(c->timing.ms_encounter, c->timing.n_encounter);
    return MINT_OK;
  };
  const long long nb = (long long)S * c->NM * 3;
  if ((rc = record(c->ev_side, 0, sm)) || (rc = wait(sp, c->ev_side, 0))) return rc;
  for (int w = 0; w < W; ++w) {
    const int p = w % 2, B = w % 3, C = win[w];
    const double *wx = w == 0 ? c->d_mx : c->d_snap[B];  // window start state
    const double *wv = w == 0 ? c->d_mv : c->d_snap[B] + nb;
    // ephemeris of window w (its end state is window w+1's snapshot, buffer
    // (w+1) mod 3, last read by the queue head of window w-2); enqueued first
    // so that its few waves are placed before the queue head fills the CUs,
    // except in a call of one window: there the queue head, which the main
    // kernel waits for, goes first (the graph starts its second root node
    // 25-37 us after the first; 1.442 -> 1.434 and 1.459 -> 1.442 ms per
    // 20-step call in same-process A/Bs, while calls of several windows were
    // 0.2-0.5 % slower that way)
    const bool head_first = W == 1;
    auto eph = [&]() -> int {
      int r2;
      if (w >= 2 && (r2 = wait(sp, c->ev_tp, w - 2))) return r2;
      if ((r2 = window_eph(c, sp, C, wx, wv, xs[B], rec[B], c->d_snap[(w + 1) % 3], kt))) return r2;
      return record(c->ev_eph, w, sp);
    };
    if (!head_first && (rc = eph())) return rc;
    // queue head of window w, from the window-start state, beside the
    // window's ephemeris (window 0's queue was cleared by the previous call or
    // the upload)
    if (w > 0) {
      window_start_kernel<<<(unsigned)((S + 255) / 256), 256, 0, sm>>>(ctl[p], need[p], S, nullptr, nullptr,
                                                                       nullptr, nullptr, 0);
      if ((rc = c->hip(hipGetLastError(), "queue reset launch"))) return rc;
    }
    if ((rc = c->hip(c->L->tp_prebucket(sm, c->Ntp, c->d_tp_sys, tp, c->d_prm, wx, wv, C, sc, scw,
                                        c->score_thresh, c->scorer_skip_on() ? c->sc_cut : INFINITY, c->d_carry, c->d_pre, q[p], ctl[p], need[p],
                                        slist_head[p], stats),
                     "queue head launch")))
      return rc;
    if (head_first && (rc = eph())) return rc;
    if ((rc = wait(sm, c->ev_eph, w))) return rc;
    // queue stream: tails of window w-1, then the head of window w
    if ((rc = record(c->ev_fork, w, sm)) || (rc = wait(sq, c->ev_fork, w))) return rc;
    if (w >= 1 && (rc = serve(w - 1, true, sq))) return rc;
    if ((rc = serve(w, false, sq))) return rc;
    if ((rc = record(c->ev_join, w, sq))) return rc;
    // main kernel of window w; its deferrals are carried unless it is the last
    if (c->ntiles > 0) {
      if (kt) kt->start();
      if ((rc = c->hip(c->L->tp_main(sm, guess, sh.P, sh.block, nblocks, c->d_tiles, c->ntiles, tpx, tp,
                                     c->d_prm, xs[B], rec[B], C, sc, nnw, c->d_pre, q[p], ctl[p], need[p],
                                     slist_tail[p], w + 1 < W ? c->d_carry : nullptr, c->d_status, stats,
                                     ktime ? c->d_kstamp + (size_t)w * 2 * c->kstamp_wpl : nullptr),
                       "main launch")))
        return rc;
      if (kt) kt->stop(c->timing.ms_tp_main, c->timing.n_tp_main);
    }
    if ((rc = wait(sm, c->ev_join, w)) || (rc = record(c->ev_tp, w, sm))) return rc;
  }
  if (ktime && c->ntiles > 0) c->kwin_pending = W;
  // the last window's tails
  if ((rc = serve(W - 1, true, sm))) return rc;
  // the end state back into the resident buffers; parity 0's queue cleared
  double *last = c->d_snap[W % 3];
  window_start_kernel<<<(unsigned)((nb + 255) / 256), 256, 0, sm>>>(ctl[0], need[0], S, c->d_mx, c->d_mv, last,
                                                                    last + nb, nb);
  return c->hip(hipGetLastError(), "call end launch");
}

// The cached graph of an nsteps call at shape sh, captured now if there is
// none for the context's current constants (GraphSig).  *nwin receives the
// number of launch windows in it.
int graph_for(Ctx *c, const Shape &sh, int64_t nsteps, hipGraphExec_t *out, int *nwin) {
  const GraphSig sig = c->graph_sig();
  for (auto &g : c->graphs)
    if (g.nsteps == nsteps && g.block == sh.block && g.P == sh.P && g.C == sh.C && g.sig.same(sig)) {
      *out = g.exec;
      *nwin = g.nwin;
      return MINT_OK;
    }
  int rc = c->hip(hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal), "begin capture");
  if (rc) return rc;
  rc = enqueue_batched(c, sh, nsteps, true, nullptr, true);
  const int W = c->kwin_pending;
  c->kwin_pending = 0;
  hipGraph_t g = nullptr;
  hipError_t e = hipStreamEndCapture(c->stream, &g);
  if (rc) {
    if (g) hipGraphDestroy(g);
    return rc;
  }
  if ((rc = c->hip(e, "end capture"))) return rc;
  hipGraphExec_t ex = nullptr;
  e = hipGraphInstantiate(&ex, g, nullptr, nullptr, 0);
  hipGraphDestroy(g);
  if ((rc = c->hip(e, "graph instantiate"))) return rc;
  // make the graph resident on the device now, so its first replay costs no
  // more than any other
  if ((rc = c->hip(hipGraphUpload(ex, c->stream), "graph upload")) ||
      (rc = c->hip(hipStreamSynchronize(c->stream), "graph upload"))) {
    hipGraphExecDestroy(ex);
    return rc;
  }
  if (c->graphs.size() >= 8) {
    hipGraphExecDestroy(c->graphs.front().exec);
    c->graphs.erase(c->graphs.begin());
  }
  GraphEntry ge;
  ge.nsteps = nsteps;
  ge.block = sh.block;
  ge.P = sh.P;
  ge.C = sh.C;
  ge.nwin = W;
  ge.sig = sig;
  ge.exec = ex;
  c->graphs.push_back(ge);
  *out = ex;
  *nwin = W;
  return MINT_OK;
}

// Shape checks and buffers of an nsteps batched call at the context's launch
// shape (all allocation happens here, before any capture).
int batched_setup(Ctx *c, int64_t nsteps, Shape *out) {
  const Shape sh = batched_shape(c, nsteps);
  if (!(sh.P == 1 || sh.P == 2 || sh.P == 4)) return c->fail(MINT_EINVAL, "tps_per_lane must be 1, 2 or 4");
  if (sh.block < 64 || sh.block > 512 || (sh.block & 63)) return c->fail(MINT_EINVAL, "block_size must be a multiple of 64 in [64,512]");
  if (sh.C < 1 || sh.C > 4096) return c->fail(MINT_EINVAL, "chunk_steps must be in [1,4096]");
  int rc = c->ensure_rec(sh.C);
  if (rc) return rc;
  if ((rc = c->ensure_rec2(sh.C))) return rc;
  if ((rc = c->ensure_tiles(sh.block, sh.P))) return rc;
  if (c->cfg.collect_stats & 4) {
    // main-kernel launch stamps: two per wavefront per launch window
    const int wpl = 8 * ((c->ntiles + 7) / 8) * (sh.block / 64);
    const size_t need = (size_t)((nsteps + sh.C - 1) / sh.C) * 2 * (size_t)std::max(wpl, 1);
    if (need > c->kstamp_cap || wpl != c->kstamp_wpl) {
      if (c->d_kstamp) {  // a captured graph holds the old pointer
        c->drop_graphs();
        hipFree(c->d_kstamp);
        c->allocs.erase(std::find(c->allocs.begin(), c->allocs.end(), (void *)c->d_kstamp));
        c->d_kstamp = nullptr;
        c->kstamp_cap = 0;
      }
      const size_t cap = std::max(need, c->kstamp_cap);
      if ((rc = c->dalloc((void **)&c->d_kstamp, cap * sizeof(long long), "launch stamps"))) return rc;
      c->kstamp_cap = cap;
      c->kstamp_wpl = wpl;
    }
    if (c->wall_khz <= 0.0) {
      int khz = 0;
      if ((rc = c->hip(hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, c->cfg.device), "wall clock rate")))
        return rc;
      c->wall_khz = (double)khz;
    }
  }
  *out = sh;
  return MINT_OK;
}

int run_batched(Ctx *c, int64_t nsteps) {
  Shape sh;
  int rc = batched_setup(c, nsteps, &sh);
  if (rc) return rc;
  c->kwin_pending = 0;
  if (timing_on(c)) {
    KTimer kt(c);
    return enqueue_batched(c, sh, nsteps, false, &kt, false);
  }
  if (!c->cfg.use_graph) return enqueue_batched(c, sh, nsteps, true, nullptr, false);
  hipGraphExec_t ex;
  int W = 0;
  if ((rc = graph_for(c, sh, nsteps, &ex, &W))) return rc;
  if ((rc = c->hip(hipGraphLaunch(ex, c->stream), "graph launch"))) return rc;
  if (c->cfg.collect_stats & 4) c->kwin_pending = W;
  return MINT_OK;
}

// Plain (groups of systems_per_launch systems) and flat (all systems): one
// ephemeris step and one TP step kernel per step, like the reference's
// per-sweep kernel but relaunched every step.
int run_per_step(Ctx *c, int64_t nsteps, bool flat) {
  int rc = c->ensure_rec(1);
  if (rc) return rc;
  const StepConst sc = c->step_const();
  const TPArrays tp = c->tp_arrays();
  unsigned long long *stats = (c->cfg.collect_stats & 1) ? c->d_stats : nullptr;
  const int block = c->cfg.block_size > 0 ? std::min(c->cfg.block_size, 256) : 256;
  const int spl = flat ? c->S : std::max(1, c->cfg.systems_per_launch);
  KTimer kt(c);
  for (int64_t st = 0; st < nsteps; ++st) {
    for (int s0 = 0; s0 < c->S; s0 += spl) {
      const int s1 = std::min(c->S, s0 + spl);
      kt.start();
      rc = c->hip(c->L->ephemeris(c->stream, s0, s1, 1, c->d_prm, c->d_mx, c->d_mv, c->d_mx, c->d_mv, c->d_xs0, c->d_rec,
                                  nullptr, nullptr, c->d_status, stats, sc, c->cfg.central_mass),
                  "ephemeris launch");
      if (rc) return rc;
      kt.stop(c->timing.ms_ephemeris, c->timing.n_ephemeris);
      kt.start();
      rc = c->hip(c->L->tp_step(c->stream, c->cfg.kepler_guess, block, c->tp_offset[s0], c->tp_offset[s1],
                                c->d_tp_sys, tp, c->d_prm, c->d_xs0, c->d_rec, sc, c->d_nnw_kep,
                                c->d_status, stats),
                  "step launch");
      if (rc) return rc;
      kt.stop(c->timing.ms_tp_main, c->timing.n_tp_main);
    }
  }
  return MINT_OK;
}

}  // namespace

int step_impl(Ctx *c, int64_t nsteps) {
  if (c->many) return run_many(c, nsteps);
  switch (c->cfg.form) {
    case MINT_FORM_PLAIN: return run_per_step(c, nsteps, false);
    case MINT_FORM_FLAT: return run_per_step(c, nsteps, true);
    case MINT_FORM_BATCHED: return run_batched(c, nsteps);
    default: return c->fail(MINT_EINVAL, "unknown form");
  }
}

}  // namespace mint

using namespace mint;

extern "C" {

int mint_step(void *ctx, int64_t nsteps) {
  Ctx *c = (Ctx *)ctx;
  if (!c) return MINT_EINVAL;
  if (!c->uploaded) return c->fail(MINT_ESTATE, "mint_step: nothing uploaded");
  if (nsteps < 0) return c->fail(MINT_EINVAL, "mint_step: negative step count");
  if (nsteps == 0) return MINT_OK;
  if (c->cfg.kepler_guess == MINT_GUESS_LEARNED && c->w_kep.empty())
    return c->fail(MINT_ESTATE, "mint_step: learned guess selected but the Kepler MLP has no weights");
  if (c->cfg.form == MINT_FORM_BATCHED && c->cfg.use_scorer && c->w_sc.empty())
    return c->fail(MINT_ESTATE, "mint_step: scorer selected but it has no weights");
  hipSetDevice(c->cfg.device);
  int rc = c->hip(hipEventRecord(c->ev0, c->stream), "event");
  if (rc) return rc;
  rc = step_impl(c, nsteps);
  if (rc) {
    c->kwin_pending = 0;
    hipStreamSynchronize(c->stream);
    return rc;
  }
  if ((rc = c->hip(hipEventRecord(c->ev1, c->stream), "event"))) return rc;
  if ((rc = c->hip(hipEventSynchronize(c->ev1), "step sync"))) return rc;
  float ms = 0.f;
  hipEventElapsedTime(&ms, c->ev0, c->ev1);
  c->timing.ms_total += ms;
  if (c->kwin_pending > 0) {
    // each launch: its last wavefront's exit minus its first wavefront's entry
    const size_t per = 2 * (size_t)c->kstamp_wpl;
    std::vector<long long> h((size_t)c->kwin_pending * per);
    rc = c->hip(hipMemcpy(h.data(), c->d_kstamp, h.size() * sizeof(long long), hipMemcpyDeviceToHost),
                "D2H launch stamps");
    for (int w = 0; w < c->kwin_pending && !rc; ++w) {
      long long t0 = LLONG_MAX, t1 = LLONG_MIN;
      for (size_t i = 0; i < per; i += 2) {
        t0 = std::min(t0, h[w * per + i]);
        t1 = std::max(t1, h[w * per + i + 1]);
      }
      c->timing.ms_tp_main += (double)(t1 - t0) / c->wall_khz;
      c->timing.n_tp_main += 1;
    }
    c->kwin_pending = 0;
    if (rc) return rc;
  }
  return c->hip(hipGetLastError(), "step");
}

int mint_prepare(void *ctx, int64_t nsteps) {
  Ctx *c = (Ctx *)ctx;
  if (!c) return MINT_EINVAL;
  if (!c->uploaded) return c->fail(MINT_ESTATE, "mint_prepare: nothing uploaded");
  if (nsteps < 0) return c->fail(MINT_EINVAL, "mint_prepare: negative step count");
  if (nsteps == 0 || c->many || c->cfg.form != MINT_FORM_BATCHED || !c->cfg.use_graph || timing_on(c))
    return MINT_OK;
  hipSetDevice(c->cfg.device);
  Shape sh;
  int rc = batched_setup(c, nsteps, &sh);
  if (rc) return rc;
  hipGraphExec_t ex;
  int W = 0;
  return graph_for(c, sh, nsteps, &ex, &W);
}

// ---------------------------------------------------------------------------
int mint_launch_shape(void *ctx, int64_t nsteps, int32_t *shape) {
  Ctx *c = (Ctx *)ctx;
  if (!c || !shape || nsteps < 0) return MINT_EINVAL;
  const Shape sh = batched_shape(c, nsteps);
  shape[0] = sh.block;
  shape[1] = sh.P;
  shape[2] = sh.C;
  return MINT_OK;
}

// Autotuner: time each candidate (block, TPs/lane, window) on the uploaded
// batch for a probe of 256 steps, then block and TPs/lane for calls of one
// window, restoring the device state after each run.
int mint_autotune(void *ctx, const char *cache_path, int32_t *best) {
  Ctx *c = (Ctx *)ctx;
  if (!c) return MINT_EINVAL;
  if (!c->uploaded) return c->fail(MINT_ESTATE, "mint_autotune: nothing uploaded");
  if (c->many) {  // the many-body engine has a fixed launch shape
    if (best) best[0] = best[1] = best[2] = 0;
    return MINT_OK;
  }
  hipSetDevice(c->cfg.device);
  hipDeviceProp_t prop;
  hipGetDeviceProperties(&prop, c->cfg.device);
  char key[512];
  std::snprintf(key, sizeof(key), "%s|%s|nm=%d|S=%d|N=%lld|dt=%.17g|guess=%d|scorer=%d|fixed=%d,%d,%d",
                prop.gcnArchName, prop.name, c->NM, c->S, c->Ntp, c->cfg.dt, c->cfg.kepler_guess,
                c->cfg.use_scorer, std::max(0, c->cfg.block_size), std::max(0, c->cfg.tps_per_lane),
                std::max(0, c->cfg.chunk_steps));
  for (char *p = key; *p; ++p)
    if (*p == ' ') *p = '_';
  if (cache_path) {
    std::ifstream in(cache_path);
    std::string line;
    while (std::getline(in, line)) {
      std::istringstream is(line);
      std::string k;
      int b, p, cc, b1 = 0, p1 = 0;
      if (is >> k >> b >> p >> cc && k == key) {
        if (!(is >> b1 >> p1)) b1 = p1 = 0;  // no one-window shape on the line
        c->tuned_block = b;
        c->tuned_P = p;
        c->tuned_C = cc;
        c->tuned_block1 = b1;
        c->tuned_P1 = p1;
        if (best) {
          best[0] = b;
          best[1] = p;
          best[2] = cc;
        }
        return MINT_OK;
      }
    }
  }
  // save state
  const size_t mb = (size_t)c->S * c->NM * 3 * sizeof(double);
  const size_t tb = (size_t)c->Ntp * sizeof(double);
  std::vector<void *> save;
  auto bk = [&](const void *src, size_t bytes) -> void * {
    void *p = nullptr;
    if (hipMalloc(&p, std::max<size_t>(bytes, 8)) != hipSuccess) return nullptr;
    hipMemcpy(p, src, bytes, hipMemcpyDeviceToDevice);
    save.push_back(p);
    return p;
  };
  void *smx = bk(c->d_mx, mb), *smv = bk(c->d_mv, mb), *sst = bk(c->d_status, (size_t)c->S * 4);
  void *stp[6];
  bool okmem = smx && smv && sst;
  for (int d = 0; d < 6; ++d) {
    stp[d] = bk(c->d_tp[d], tb);
    okmem = okmem && stp[d];
  }
  if (!okmem) {
    for (void *p : save) hipFree(p);
    return c->fail(MINT_ENOMEM, "mint_autotune: cannot save state");
  }
  auto restore = [&]() {
    hipMemcpy(c->d_mx, smx, mb, hipMemcpyDeviceToDevice);
    hipMemcpy(c->d_mv, smv, mb, hipMemcpyDeviceToDevice);
    hipMemcpy(c->d_status, sst, (size_t)c->S * 4, hipMemcpyDeviceToDevice);
    for (int d = 0; d < 6; ++d) hipMemcpy(c->d_tp[d], stp[d], tb, hipMemcpyDeviceToDevice);
  };
  const mint_config saved = c->cfg;
  // candidates; a parameter set in the configuration is held
  std::vector<int> blocks = {64, 128, 256, 512}, Ps = {1, 2, 4}, Cs = {32, 64, 128};
  if (saved.block_size > 0) blocks = {saved.block_size};
  if (saved.tps_per_lane > 0) Ps = {saved.tps_per_lane};
  if (saved.chunk_steps > 0) Cs = {saved.chunk_steps};
  // Three passes over the candidates, each candidate ranked by the median of
  // its three times (rank = 1) or by the slowest (rank = 2): a single pass in
  // a fixed order favoured the candidates timed last, once the clock had
  // settled under the load, and the best of two passes let through shapes
  // whose queue work only sometimes finds room beside the main kernel (small
  // workgroups at calls of one window: 1.47 ms once, 1.8-2.0 ms mostly),
  // which the slowest of three rules out.
  int rc = MINT_OK;
  c->cfg.form = MINT_FORM_BATCHED;
  c->cfg.collect_stats = 0;
  constexpr int NPASS = 3;
  auto tune = [&](const std::vector<Shape> &cand, int probe, int rank, Shape *out) -> int {
    std::vector<double> times(cand.size() * NPASS, 0.0);
    for (int pass = 0; pass < NPASS; ++pass)
      for (size_t k = 0; k < cand.size(); ++k) {
        c->cfg.block_size = cand[k].block;
        c->cfg.tps_per_lane = cand[k].P;
        c->cfg.chunk_steps = cand[k].C;
        int r;
        // first pass: an untimed run first (buffers, tiles and the graph for this shape)
        if (pass == 0) {
          restore();
          if ((r = step_impl(c, probe))) return r;
        }
        restore();
        hipDeviceSynchronize();
        hipEventRecord(c->ev0, c->stream);
        r = step_impl(c, probe);
        hipEventRecord(c->ev1, c->stream);
        hipEventSynchronize(c->ev1);
        if (r) return r;
        float ms = 0.f;
        hipEventElapsedTime(&ms, c->ev0, c->ev1);
        times[k * NPASS + pass] = ms;
      }
    double best_ms = 1e300;
    for (size_t k = 0; k < cand.size(); ++k) {
      double *t = times.data() + k * NPASS;
      std::sort(t, t + NPASS);
      const double v = t[rank == 2 ? NPASS - 1 : NPASS / 2];
      if (v < best_ms) {
        best_ms = v;
        *out = cand[k];
      }
    }
    return MINT_OK;
  };
  // calls of more than one window: a 256-step probe (two or more windows, so
  // it includes the pipeline's steady state)
  std::vector<Shape> cand;
  for (int b : blocks)
    for (int p : Ps)
      for (int C : Cs) cand.push_back(Shape{b, p, C});
  Shape bs = cand[0];
  rc = tune(cand, 256, 1, &bs);
  // calls of one window (where block and TPs per lane are both free): a
  // probe of one window of at most 32 steps at the chosen window length,
  // ranked by the slowest pass (there the queue work has to fit beside the
  // one main kernel)
  Shape b1{0, 0, bs.C};
  if (rc == MINT_OK && saved.block_size <= 0 && saved.tps_per_lane <= 0) {
    std::vector<Shape> cand1;
    for (int b : blocks)
      for (int p : Ps) cand1.push_back(Shape{b, p, bs.C});
    rc = tune(cand1, std::min(bs.C, 32), 2, &b1);
  }
  c->cfg = saved;
  restore();
  hipDeviceSynchronize();
  for (void *p : save) hipFree(p);
  c->drop_graphs();
  if (rc) return rc;
  c->tuned_block = bs.block;
  c->tuned_P = bs.P;
  c->tuned_C = bs.C;
  c->tuned_block1 = b1.block;
  c->tuned_P1 = b1.P;
  if (best) {
    best[0] = bs.block;
    best[1] = bs.P;
    best[2] = bs.C;
  }
  if (cache_path) {
    std::ofstream out(cache_path, std::ios::app);
    if (out) out << key << " " << bs.block << " " << bs.P << " " << bs.C << " " << b1.block << " " << b1.P << "\n";
  }
  hipMemset(c->d_stats, 0, S_COUNT * sizeof(unsigned long long));
  return MINT_OK;
}

}  // extern "C"
