// mint_api_internal.h — the context behind the opaque mint handle.
#pragma once
#include <hip/hip_runtime.h>

#include <string>
#include <vector>

#include "../../include/mint.h"
#include "mint_launch.h"

namespace mint {

// What a captured graph holds by value besides device pointers (the step
// constants, the central mass, the scorer threshold, the Kepler guess, the
// scorer and counter switches): a cached graph is replayed only while all of
// these are unchanged, so a change of dt or tolerance (mint_load_checkpoint)
// or of the threshold can never replay old constants.
struct GraphSig {
  double h = 0, mu = 0, eps2 = 0, tol = 0, gate_ir0 = 0, M = 0;
  float thresh = 0.f, cut = 0.f;
  int guess = -1, scorer = -1, stats = -1, skip = -1;
  bool same(const GraphSig &o) const {
    return h == o.h && mu == o.mu && eps2 == o.eps2 && tol == o.tol && gate_ir0 == o.gate_ir0 && M == o.M &&
           thresh == o.thresh && cut == o.cut && guess == o.guess && scorer == o.scorer && stats == o.stats && skip == o.skip;
  }
};

// The captured launch windows of one mint_step(nsteps) call of the batched form.
struct GraphEntry {
  long long nsteps = 0;
  int block = 0, P = 0, C = 0;  // launch shape the graph was captured with
  int nwin = 0;                 // launch windows in the graph
  GraphSig sig;
  hipGraphExec_t exec = nullptr;
};

struct Ctx {
  mint_config cfg;
  std::string err;
  hipStream_t stream = nullptr;
  hipStream_t stream2 = nullptr;   // ephemeris runs one window ahead on this stream
  hipStream_t stream3 = nullptr;   // queue work (window heads, carried tails) beside the main kernel
  std::vector<hipEvent_t> ev_side;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  // per-window events of the pipeline: ephemeris done, window done, queue
  // work forked from / joined back into the context stream
  std::vector<hipEvent_t> ev_eph, ev_tp, ev_fork, ev_join;
  // collect_stats bit 2: every main-kernel launch of a batched call stamps
  // the wall clock at each wavefront's entry and exit, window w into
  // d_kstamp[w * 2 * kstamp_wpl ...]; kwin_pending windows of the call just
  // enqueued await reading
  long long *d_kstamp = nullptr;
  size_t kstamp_cap = 0;  // long longs
  int kstamp_wpl = 0;     // wavefronts per main-kernel launch
  int kwin_pending = 0;
  double wall_khz = 0.0;  // wall-clock rate

  // batch shape
  int S = 0, NM = 0;
  long long Ntp = 0;
  std::vector<long long> tp_offset;
  std::vector<double> m, rc;
  const Launchers *L = nullptr;
  bool uploaded = false;

  // device buffers
  std::vector<void *> allocs;
  double *d_prm = nullptr, *d_mx = nullptr, *d_mv = nullptr, *d_xs0 = nullptr, *d_rec = nullptr;
  // ephemeris buffers 2 and 3 of the batched form (three rotate: a window's
  // buffer is in use until its carried tails finish in the next window)
  double *d_xs0b = nullptr, *d_recb = nullptr, *d_xs0c = nullptr, *d_recc = nullptr;
  // massive positions and velocities at the start of launch window w, in
  // buffer w mod 3 (positions [S][NM][3], then velocities): what the queue
  // head of window w reads, written by the ephemeris of window w-1 (by a copy
  // for the first window of a call), so that the queue head does not wait for
  // its own window's ephemeris
  double *d_snap[3] = {nullptr, nullptr, nullptr};
  double *d_pc = nullptr;          // perturber cache of the window being served
  unsigned int *d_need = nullptr;  // [2][S] per window parity: 1 queue head, 2 mid-window deferral
  int *d_slist = nullptr;          // [2][2][S] per parity: systems flagged 1, then flagged 2
  double *d_tp[6] = {nullptr, nullptr, nullptr, nullptr, nullptr, nullptr};
  int *d_tp_sys = nullptr;
  unsigned char *d_pre = nullptr;    // per-TP: starts the current window in the queue
  unsigned char *d_carry = nullptr;  // per-TP: deferred this window, queued for the next
  Tile *d_tiles = nullptr;
  int ntiles = 0;
  long long tiles_key = -1;
  // tile lists of the two most recent (block, TPs per lane) shapes, so that
  // calls alternating between two shapes keep both lists (and the graphs that
  // hold them); d_tiles, ntiles and tiles_key are the current one's
  struct TileSet {
    long long key = -1;
    Tile *d = nullptr;
    int n = 0;
  };
  TileSet tilesets[2];
  int tiles_next = 0;  // the set a new shape replaces
  QEntry *d_q = nullptr, *d_q2 = nullptr;  // BS queues of even / odd windows
  // [2][8] per window parity: [0] entries of the queue head, [1] systems
  // flagged 1, [3] systems flagged 2, [4] mid-window deferrals (queued after
  // the head); [2], [5..7] unused
  unsigned int *d_qctl = nullptr;
  unsigned int *d_status = nullptr;
  unsigned long long *d_stats = nullptr;
  float *d_nnw_kep = nullptr, *d_nnw_sc = nullptr;
  int rec_steps_alloc = 0;
  int rec2_steps_alloc = 0;

  // learned parts (host copies; empty = not trained / not set)
  std::vector<float> w_kep, w_sc;
  float score_thresh = 0.0f;
  // scorer pairs with both distance features at or above sc_cut (feature
  // units) are left out of the network's list when mode 1: sc_cut is set with
  // the weights and the threshold, from an interval bound of the logit over
  // such pairs (sc_sat_bound) that lies below the threshold
  int scorer_skip_mode = 1;
  float sc_sat_bound = INFINITY, sc_cut = INFINITY;
  bool scorer_skip_on() const { return scorer_skip_mode == 1 && !w_sc.empty() && sc_cut <= 8.0f; }

  // launch shape chosen for the batched form (calls of more than one launch
  // window), and the block and TPs per lane chosen for calls of one window
  // (0: the same as for longer calls)
  int tuned_block = 0, tuned_P = 0, tuned_C = 0, tuned_block1 = 0, tuned_P1 = 0;

  std::vector<GraphEntry> graphs;
  mint_timing timing{};

  // many-body engine state (n_massive > MAX_NM, no TPs)
  bool many = false;
  std::vector<double> rad;          // physical radii [S*NM] (zeros: none)
  double *d_bp = nullptr;           // [S][5][NM]: m, 0.1 r_c, 1/(0.9 r_c), r_c^2, R
  double *d_mxb = nullptr, *d_mvb = nullptr;  // post-drift x, v  [S][NM][3]
  double *d_mbmax = nullptr;        // [S][2]: largest r_c^2 and physical radius (pair-check bounds)
  int *d_mflag = nullptr;           // [S][NM] body is in an encounter pair this step
  int *d_medge = nullptr;           // [S][MB_EMAX][2] encounter pairs
  unsigned int *d_mctl = nullptr;   // [S][4]: [0], [1] pairs in the list of even, odd steps
  std::vector<hipEvent_t> tev;       // event pool of the per-kernel timing mode
  hipGraphExec_t mb_graph = nullptr;
  long long mb_graph_steps = 0;
  GraphSig mb_sig;                  // what mb_graph was captured with

  int fail(int code, const std::string &msg);
  int hip(hipError_t e, const char *what);
  void free_device();
  int dalloc(void **p, size_t bytes, const char *what);
  int ensure_rec(int C);
  int ensure_rec2(int C);
  int ensure_tiles(int block, int P);
  StepConst step_const() const;
  GraphSig graph_sig() const;
  void drop_graphs();
  TPArrays tp_arrays() const;
};

inline bool timing_on(const Ctx *c) { return (c->cfg.collect_stats & 2) != 0; }

// HIP-event timer around single launches on the context stream (profiling
// runs only).  Event pairs are recorded without waiting and resolved when the
// timer is flushed (at the latest on destruction), so the launches still run
// back to back and each pair brackets one kernel.
struct KTimer {
  Ctx *c;
  bool on;
  std::vector<std::pair<double *, int64_t *>> dst;
  explicit KTimer(Ctx *cc) : c(cc), on(timing_on(cc)) {}
  ~KTimer() { flush(); }
  hipEvent_t ev(size_t i) {
    while (c->tev.size() <= i) {
      hipEvent_t e = nullptr;
      hipEventCreate(&e);
      c->tev.push_back(e);
    }
    return c->tev[i];
  }
  void start() {
    if (on) hipEventRecord(ev(2 * dst.size()), c->stream);
  }
  void stop(double &ms, int64_t &n) {
    if (!on) return;
    hipEventRecord(ev(2 * dst.size() + 1), c->stream);
    dst.emplace_back(&ms, &n);
    if (dst.size() >= 1024) flush();  // bound the event pool
  }
  void flush() {
    if (!on || dst.empty()) return;
    hipEventSynchronize(c->tev[2 * dst.size() - 1]);
    for (size_t i = 0; i < dst.size(); ++i) {
      float t = 0.f;
      hipEventElapsedTime(&t, c->tev[2 * i], c->tev[2 * i + 1]);
      *dst[i].first += t;
      *dst[i].second += 1;
    }
    dst.clear();
  }
};

// many-body engine (mint_mb.hip): systems of more than MAX_NM massive bodies
constexpr int MB_MAXN = 2048;   // massive bodies per system
constexpr int MB_EMAX = 1024;   // encounter pairs per system and step
constexpr int MB_GMAX = 64;     // bodies per encounter group (one wavefront)
int mb_alloc(Ctx *c);       // engine buffers for the uploaded shape
int mb_reset(Ctx *c);       // body parameters from m, rc, rad; clear flags and counters
int mb_set_radii(Ctx *c);   // body parameters again after a change of rad
int run_many(Ctx *c, int64_t nsteps);

extern std::string g_create_error;
void noalloc_init(int device);  // mint_noalloc.hip
float scorer_cutoff(const std::vector<float> &w, float t, float *bound);
const char *form_name(int f);

}  // namespace mint
