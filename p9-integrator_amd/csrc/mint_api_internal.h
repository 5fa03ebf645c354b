// mint_api_internal.h — the context behind the opaque mint handle.
#pragma once
#include <hip/hip_runtime.h>

#include <string>
#include <vector>

#include "../../include/mint.h"
#include "mint_launch.h"

namespace mint {

// A captured launch window (ephemeris + queue reset + main + encounter) for
// one window length C; replayed once per window.
struct GraphEntry {
  long long nsteps = 0;
  int block = 0, P = 0, C = 0;  // launch shape the graph was captured with
  hipGraphExec_t exec = nullptr;
};

struct Ctx {
  mint_config cfg;
  std::string err;
  hipStream_t stream = nullptr;
  hipStream_t stream2 = nullptr;   // ephemeris runs one window ahead on this stream
  hipStream_t stream3 = nullptr;   // pre-bucketed encounter work beside the main kernel
  std::vector<hipEvent_t> ev_side;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  std::vector<hipEvent_t> ev_eph, ev_tp;  // per-window completion events of the pipeline

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
  double *d_xs0b = nullptr, *d_recb = nullptr;  // second buffer of the pipelined batched form
  double *d_pc = nullptr;          // perturber cache of the current window
  unsigned int *d_need = nullptr;  // per-system: a TP entered the encounter queue this window
  double *d_tp[6] = {nullptr, nullptr, nullptr, nullptr, nullptr, nullptr};
  int *d_tp_sys = nullptr;
  unsigned char *d_pre = nullptr;  // per-TP pre-bucket flag of the current window
  Tile *d_tiles = nullptr;
  int ntiles = 0;
  long long tiles_key = -1;
  QEntry *d_q = nullptr;
  unsigned int *d_qctl = nullptr;  // [0] entries so far, [2] entries from the pre-bucketing pass
  unsigned int *d_status = nullptr;
  unsigned long long *d_stats = nullptr;
  float *d_nnw_kep = nullptr, *d_nnw_sc = nullptr;
  int rec_steps_alloc = 0;
  int rec2_steps_alloc = 0;

  // learned parts (host copies; empty = not trained / not set)
  std::vector<float> w_kep, w_sc;
  float score_thresh = 0.0f;

  // launch shape chosen for the batched form
  int tuned_block = 0, tuned_P = 0, tuned_C = 0;

  std::vector<GraphEntry> graphs;
  mint_timing timing{};

  int fail(int code, const std::string &msg);
  int hip(hipError_t e, const char *what);
  void free_device();
  int dalloc(void **p, size_t bytes, const char *what);
  int ensure_rec(int C);
  int ensure_rec2(int C);
  int ensure_tiles(int block, int P);
  StepConst step_const() const;
  TPArrays tp_arrays() const;
};

extern std::string g_create_error;
const char *form_name(int f);

}  // namespace mint
