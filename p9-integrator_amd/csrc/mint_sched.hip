// mint_sched.hip — step scheduling for the three forms, hipGraph capture of
// the batched launch window, and the launch-shape autotuner.
//
// Reference: the host loop src/mercurius.cu:179-198 launches the fused kernel
// once per 1e6-step batch; the kernel's own loop (src/sim.cuh:679-742) runs
// the steps with six cluster barriers each.  Here the massive subsystem runs
// ahead into an ephemeris buffer and the TPs read it without any barrier.
#include <hip/hip_runtime.h>

#include <algorithm>
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

Shape batched_shape(const Ctx *c) {
  Shape s;
  s.block = c->cfg.block_size > 0 ? c->cfg.block_size : (c->tuned_block > 0 ? c->tuned_block : DEF_BLOCK);
  s.P = c->cfg.tps_per_lane > 0 ? c->cfg.tps_per_lane : (c->tuned_P > 0 ? c->tuned_P : DEF_P);
  s.C = c->cfg.chunk_steps > 0 ? c->cfg.chunk_steps : (c->tuned_C > 0 ? c->tuned_C : DEF_C);
  return s;
}

// encounter kernel: one-wave blocks, enough to put a wave on every SIMD twice
int enc_grid(const Ctx *c) {
  long long g = (c->Ntp + 63) / 64;
  return (int)std::max(1LL, std::min(g, 2048LL));
}



// Ephemeris of one launch window into buffer (xs0, rec) on stream st.
int window_eph(Ctx *c, hipStream_t st, int C, double *xs0, double *rec, KTimer *kt) {
  const StepConst sc = c->step_const();
  if (kt) kt->start();
  unsigned long long *stats = (c->cfg.collect_stats & 1) ? c->d_stats : nullptr;
  int rc = c->hip(c->L->ephemeris(st, 0, c->S, C, c->d_prm, c->d_mx, c->d_mv, xs0, rec, c->d_status, stats,
                                  sc, c->cfg.central_mass),
                  "ephemeris launch");
  if (rc) return rc;
  if (kt) kt->stop(c->timing.ms_ephemeris, c->timing.n_ephemeris);
  return MINT_OK;
}

hipEvent_t window_event(std::vector<hipEvent_t> &v, size_t i);

// TP side of one launch window of C steps on the context stream: queue reset,
// then (with the scorer) the pre-bucketing pass whose entries are carried by
// an encounter kernel on the side stream while the main kernel runs, then the
// main kernel, the perturber cache of systems that deferred mid-window, and
// the encounter kernel over those deferrals.
int window_tp(Ctx *c, const Shape &sh, int C, const double *xs0, const double *rec, KTimer *kt) {
  const StepConst sc = c->step_const();
  const TPArrays tp = c->tp_arrays();
  unsigned long long *stats = (c->cfg.collect_stats & 1) ? c->d_stats : nullptr;
  const float *nnw = c->d_nnw_kep;
  const float *scw = c->cfg.use_scorer ? c->d_nnw_sc : nullptr;
  const int tpx = (c->ntiles + 7) / 8;
  const int nblocks = 8 * tpx;
  const int guess = c->cfg.kepler_guess;
  int rc;
  zero_u32_kernel<<<1, 64, 0, c->stream>>>(c->d_qctl, 4);
  zero_u32_kernel<<<(c->S + 255) / 256, 256, 0, c->stream>>>(c->d_need, c->S);
  if ((rc = c->hip(hipGetLastError(), "queue reset launch"))) return rc;
  const bool side = scw && c->Ntp > 0;
  if (side) {
    rc = c->hip(c->L->tp_prebucket(c->stream, c->Ntp, c->d_tp_sys, tp, c->d_prm, xs0, rec, C, sc, scw,
                                   c->score_thresh, c->d_pre, c->d_q, c->d_qctl, c->d_need, stats),
                "prebucket launch");
    if (rc) return rc;
    copy_u32_kernel<<<1, 1, 0, c->stream>>>(c->d_qctl + 2, c->d_qctl);
    rc = c->hip(c->L->pcache(c->stream, 0, c->S, C, rec, c->d_need, 1u, c->d_pc, sc), "perturber cache launch");
    if (rc) return rc;
    if (kt) {
      kt->start();
      rc = c->hip(c->L->tp_encounter(c->stream, guess, enc_grid(c), 64, c->d_q, c->d_qctl + 2, nullptr, tp,
                                     c->d_prm, xs0, rec, c->d_pc, C, sc, nnw, c->d_status, stats),
                  "encounter launch");
      if (rc) return rc;
      kt->stop(c->timing.ms_encounter, c->timing.n_encounter);
    } else {
      hipEvent_t f = window_event(c->ev_side, 0);
      if ((rc = c->hip(hipEventRecord(f, c->stream), "side fork"))) return rc;
      if ((rc = c->hip(hipStreamWaitEvent(c->stream3, f, 0), "side wait"))) return rc;
      rc = c->hip(c->L->tp_encounter(c->stream3, guess, enc_grid(c), 64, c->d_q, c->d_qctl + 2, nullptr, tp,
                                     c->d_prm, xs0, rec, c->d_pc, C, sc, nnw, c->d_status, stats),
                  "encounter launch");
      if (rc) return rc;
    }
  }
  if (c->ntiles > 0) {
    if (kt) kt->start();
    rc = c->hip(c->L->tp_main(c->stream, guess, sh.P, sh.block, nblocks, c->d_tiles, c->ntiles, tpx, tp, c->d_prm,
                              xs0, rec, C, sc, nnw, scw ? c->d_pre : nullptr, c->d_q, c->d_qctl, c->d_need,
                              c->d_status, stats),
                "main launch");
    if (rc) return rc;
    if (kt) kt->stop(c->timing.ms_tp_main, c->timing.n_tp_main);
    if (kt) kt->start();
    rc = c->hip(c->L->pcache(c->stream, 0, c->S, C, rec, c->d_need, 2u, c->d_pc, sc), "perturber cache launch");
    if (rc) return rc;
    rc = c->hip(c->L->tp_encounter(c->stream, guess, enc_grid(c), 64, c->d_q, c->d_qctl,
                                   side ? c->d_qctl + 2 : nullptr, tp, c->d_prm, xs0, rec, c->d_pc, C, sc, nnw,
                                   c->d_status, stats),
                "encounter launch");
    if (rc) return rc;
    if (kt) kt->stop(c->timing.ms_encounter, c->timing.n_encounter);
  }
  if (side && !kt) {
    hipEvent_t j = window_event(c->ev_side, 1);
    if ((rc = c->hip(hipEventRecord(j, c->stream3), "side join"))) return rc;
    if ((rc = c->hip(hipStreamWaitEvent(c->stream, j, 0), "side join wait"))) return rc;
  }
  return MINT_OK;
}

hipEvent_t window_event(std::vector<hipEvent_t> &v, size_t i) {
  while (v.size() <= i) {
    hipEvent_t e = nullptr;
    hipEventCreateWithFlags(&e, hipEventDisableTiming);
    v.push_back(e);
  }
  return v[i];
}

// Enqueue nsteps of the batched form.  Pipelined: the ephemeris of window
// w+1 runs on the second stream while the TP kernels of window w run on the
// context stream, with the two ephemeris buffers alternating; window w's TP
// kernels wait for its ephemeris, and the ephemeris of window w+2 waits for
// window w's TP kernels (it overwrites their buffer).
int enqueue_batched(Ctx *c, const Shape &sh, int64_t nsteps, bool pipelined, KTimer *kt) {
  std::vector<int> win;
  for (int64_t done = 0; done < nsteps;) {
    const int C = (int)std::min<int64_t>(sh.C, nsteps - done);
    win.push_back(C);
    done += C;
  }
  double *xs[2] = {c->d_xs0, c->d_xs0b};
  double *rec[2] = {c->d_rec, c->d_recb};
  int rc;
  if (!pipelined || win.size() < 2) {
    for (int C : win) {
      if ((rc = window_eph(c, c->stream, C, xs[0], rec[0], kt))) return rc;
      if ((rc = window_tp(c, sh, C, xs[0], rec[0], kt))) return rc;
    }
    return MINT_OK;
  }
  hipEvent_t fork = window_event(c->ev_eph, 0);
  if ((rc = c->hip(hipEventRecord(fork, c->stream), "fork event"))) return rc;
  if ((rc = c->hip(hipStreamWaitEvent(c->stream2, fork, 0), "fork wait"))) return rc;
  for (size_t w = 0; w < win.size(); ++w) {
    const int b = (int)(w & 1);
    if (w >= 2 && (rc = c->hip(hipStreamWaitEvent(c->stream2, window_event(c->ev_tp, w - 2), 0), "buffer wait")))
      return rc;
    if ((rc = window_eph(c, c->stream2, win[w], xs[b], rec[b], nullptr))) return rc;
    hipEvent_t e = window_event(c->ev_eph, w + 1);
    if ((rc = c->hip(hipEventRecord(e, c->stream2), "ephemeris event"))) return rc;
    if ((rc = c->hip(hipStreamWaitEvent(c->stream, e, 0), "ephemeris wait"))) return rc;
    if ((rc = window_tp(c, sh, win[w], xs[b], rec[b], nullptr))) return rc;
    if ((rc = c->hip(hipEventRecord(window_event(c->ev_tp, w), c->stream), "window event"))) return rc;
  }
  return MINT_OK;
}

int graph_for(Ctx *c, const Shape &sh, int64_t nsteps, hipGraphExec_t *out) {
  for (auto &g : c->graphs)
    if (g.nsteps == nsteps && g.block == sh.block && g.P == sh.P && g.C == sh.C) {
      *out = g.exec;
      return MINT_OK;
    }
  int rc = c->hip(hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal), "begin capture");
  if (rc) return rc;
  rc = enqueue_batched(c, sh, nsteps, true, nullptr);
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
  if (c->graphs.size() >= 8) {
    hipGraphExecDestroy(c->graphs.front().exec);
    c->graphs.erase(c->graphs.begin());
  }
  GraphEntry ge;
  ge.nsteps = nsteps;
  ge.block = sh.block;
  ge.P = sh.P;
  ge.C = sh.C;
  ge.exec = ex;
  c->graphs.push_back(ge);
  *out = ex;
  return MINT_OK;
}

int run_batched(Ctx *c, int64_t nsteps) {
  const Shape sh = batched_shape(c);
  if (!(sh.P == 1 || sh.P == 2 || sh.P == 4)) return c->fail(MINT_EINVAL, "tps_per_lane must be 1, 2 or 4");
  if (sh.block < 64 || sh.block > 512 || (sh.block & 63)) return c->fail(MINT_EINVAL, "block_size must be a multiple of 64 in [64,512]");
  if (sh.C < 1 || sh.C > 4096) return c->fail(MINT_EINVAL, "chunk_steps must be in [1,4096]");
  int rc = c->ensure_rec(sh.C);
  if (rc) return rc;
  if ((rc = c->ensure_rec2(sh.C))) return rc;
  if ((rc = c->ensure_tiles(sh.block, sh.P))) return rc;
  if (timing_on(c)) {
    KTimer kt(c);
    return enqueue_batched(c, sh, nsteps, false, &kt);
  }
  if (!c->cfg.use_graph) return enqueue_batched(c, sh, nsteps, true, nullptr);
  hipGraphExec_t ex;
  if ((rc = graph_for(c, sh, nsteps, &ex))) return rc;
  return c->hip(hipGraphLaunch(ex, c->stream), "graph launch");
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
      rc = c->hip(c->L->ephemeris(c->stream, s0, s1, 1, c->d_prm, c->d_mx, c->d_mv, c->d_xs0, c->d_rec,
                                  c->d_status, stats, sc, c->cfg.central_mass),
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
    hipStreamSynchronize(c->stream);
    return rc;
  }
  if ((rc = c->hip(hipEventRecord(c->ev1, c->stream), "event"))) return rc;
  if ((rc = c->hip(hipEventSynchronize(c->ev1), "step sync"))) return rc;
  float ms = 0.f;
  hipEventElapsedTime(&ms, c->ev0, c->ev1);
  c->timing.ms_total += ms;
  return c->hip(hipGetLastError(), "step");
}

// ---------------------------------------------------------------------------
// Autotuner: time each candidate (block, TPs/lane, window) on the uploaded
// batch for a probe of 128 steps, restoring the device state after each.
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
  std::snprintf(key, sizeof(key), "%s|%s|nm=%d|S=%d|N=%lld|dt=%.17g|guess=%d|scorer=%d", prop.gcnArchName,
                prop.name, c->NM, c->S, c->Ntp, c->cfg.dt, c->cfg.kepler_guess, c->cfg.use_scorer);
  for (char *p = key; *p; ++p)
    if (*p == ' ') *p = '_';
  if (cache_path) {
    std::ifstream in(cache_path);
    std::string line;
    while (std::getline(in, line)) {
      std::istringstream is(line);
      std::string k;
      int b, p, cc;
      if (is >> k >> b >> p >> cc && k == key) {
        c->tuned_block = b;
        c->tuned_P = p;
        c->tuned_C = cc;
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
  const int blocks[] = {64, 128, 256, 512};
  const int Ps[] = {1, 2, 4};
  const int Cs[] = {32, 64, 128};
  const int probe = 128;
  double best_ms = 1e300;
  Shape bs{DEF_BLOCK, DEF_P, DEF_C};
  int rc = MINT_OK;
  c->cfg.form = MINT_FORM_BATCHED;
  c->cfg.collect_stats = 0;
  for (int b : blocks)
    for (int p : Ps)
      for (int C : Cs) {
        c->cfg.block_size = b;
        c->cfg.tps_per_lane = p;
        c->cfg.chunk_steps = C;
        restore();
        hipDeviceSynchronize();
        hipEventRecord(c->ev0, c->stream);
        rc = step_impl(c, probe);
        hipEventRecord(c->ev1, c->stream);
        hipEventSynchronize(c->ev1);
        if (rc) break;
        float ms = 0.f;
        hipEventElapsedTime(&ms, c->ev0, c->ev1);
        if (ms < best_ms) {
          best_ms = ms;
          bs = Shape{b, p, C};
        }
      }
  c->cfg = saved;
  restore();
  hipDeviceSynchronize();
  for (void *p : save) hipFree(p);
  for (auto &g : c->graphs)
    if (g.exec) hipGraphExecDestroy(g.exec);
  c->graphs.clear();
  if (rc) return rc;
  c->tuned_block = bs.block;
  c->tuned_P = bs.P;
  c->tuned_C = bs.C;
  if (best) {
    best[0] = bs.block;
    best[1] = bs.P;
    best[2] = bs.C;
  }
  if (cache_path) {
    std::ofstream out(cache_path, std::ios::app);
    if (out) out << key << " " << bs.block << " " << bs.P << " " << bs.C << "\n";
  }
  hipMemset(c->d_stats, 0, S_COUNT * sizeof(unsigned long long));
  return MINT_OK;
}

}  // extern "C"
