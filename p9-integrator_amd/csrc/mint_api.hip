// mint_api.hip — host side of libmint: context, uploads, step scheduling,
// downloads, energy and the element conversions of the configuration edge.
//
// The reference drives its single kernel from main (src/mercurius.cu:17-263):
// it allocates the element arrays, replicates bodies per sweep, launches one
// cluster per sweep and relaunches every 1e6 steps.  Here the context owns a
// CSR batch of systems, and mint_step schedules the ephemeris kernel and one
// of three TP forms on the context's stream.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <new>
#include <sstream>
#include <string>
#include <vector>

#include "../../include/mint.h"
#include "../../include/mint_flops.h"
#include "mint_api_internal.h"

using namespace mint;

namespace mint {

const char *form_name(int f) {
  switch (f) {
    case MINT_FORM_PLAIN: return "plain";
    case MINT_FORM_FLAT: return "flat";
    case MINT_FORM_BATCHED: return "batched";
    default: return "?";
  }
}

int Ctx::fail(int code, const std::string &msg) {
  err = msg;
  return code;
}

int Ctx::hip(hipError_t e, const char *what) {
  if (e == hipSuccess) return MINT_OK;
  err = std::string(what) + ": " + hipGetErrorString(e);
  return MINT_EHIP;
}

void Ctx::drop_graphs() {
  for (auto &g : graphs)
    if (g.exec) hipGraphExecDestroy(g.exec);
  graphs.clear();
  if (mb_graph) hipGraphExecDestroy(mb_graph);
  mb_graph = nullptr;
  mb_graph_steps = 0;
}

void Ctx::free_device() {
  for (void *p : allocs) hipFree(p);
  allocs.clear();
  drop_graphs();
  d_prm = d_mx = d_mv = d_xs0 = d_rec = nullptr;
  d_xs0b = d_recb = d_xs0c = d_recc = nullptr;
  for (auto &p : d_snap) p = nullptr;
  d_kstamp = nullptr;
  kstamp_cap = 0;
  d_pc = nullptr;
  d_carry = nullptr;
  d_slist = nullptr;
  d_q2 = nullptr;
  d_need = nullptr;
  rec2_steps_alloc = 0;
  for (auto &p : d_tp) p = nullptr;
  d_tp_sys = nullptr;
  d_pre = nullptr;
  d_tiles = nullptr;
  d_q = nullptr;
  d_qctl = nullptr;
  d_status = nullptr;
  d_stats = nullptr;
  d_nnw_kep = d_nnw_sc = nullptr;
  d_bp = d_mxb = d_mvb = d_mbmax = nullptr;
  d_mflag = d_medge = nullptr;
  d_mctl = nullptr;
  many = false;
  rec_steps_alloc = 0;
  tiles_key = -1;
  ntiles = 0;
  for (auto &ts : tilesets) ts = TileSet{};
  tiles_next = 0;
  uploaded = false;
}

int Ctx::dalloc(void **p, size_t bytes, const char *what) {
  *p = nullptr;
  if (bytes == 0) bytes = 8;
  hipError_t e = hipMalloc(p, bytes);
  if (e != hipSuccess) return fail(MINT_ENOMEM, std::string("hipMalloc ") + what + ": " + hipGetErrorString(e));
  allocs.push_back(*p);
  return MINT_OK;
}

int Ctx::ensure_rec(int C) {
  if (C <= rec_steps_alloc) return MINT_OK;
  if (d_rec) {
    hipFree(d_rec);
    allocs.erase(std::find(allocs.begin(), allocs.end(), (void *)d_rec));
    d_rec = nullptr;
    drop_graphs();
  }
  if (d_pc) {
    hipFree(d_pc);
    allocs.erase(std::find(allocs.begin(), allocs.end(), (void *)d_pc));
    d_pc = nullptr;
  }
  const size_t bytes = (size_t)S * (size_t)C * (size_t)rec_size(NM) * sizeof(double);
  int rc = dalloc((void **)&d_rec, bytes, "ephemeris");
  if (rc) return rc;
  if ((rc = dalloc((void **)&d_pc, (size_t)S * C * pc_size(NM) * sizeof(double), "perturber cache"))) return rc;
  rec_steps_alloc = C;
  return MINT_OK;
}

// the batched form's other buffers: ephemeris buffers 2 and 3 (with their
// chunk-start positions) and the odd windows' queue
int Ctx::ensure_rec2(int C) {
  int rc;
  if (!d_q2 && (rc = dalloc((void **)&d_q2, (size_t)std::max<long long>(Ntp, 1) * sizeof(QEntry), "BS queue (odd windows)")))
    return rc;
  for (auto &p : d_snap)
    if (!p && (rc = dalloc((void **)&p, (size_t)S * NM * 6 * sizeof(double), "window-start snapshot"))) return rc;
  if (C <= rec2_steps_alloc) return MINT_OK;
  for (double **p : {&d_recb, &d_xs0b, &d_recc, &d_xs0c}) {
    if (*p) {
      hipFree(*p);
      allocs.erase(std::find(allocs.begin(), allocs.end(), (void *)*p));
      *p = nullptr;
    }
  }
  drop_graphs();
  const size_t bytes = (size_t)S * (size_t)C * (size_t)rec_size(NM) * sizeof(double);
  const size_t xb = (size_t)S * NM * 3 * sizeof(double);
  if ((rc = dalloc((void **)&d_recb, bytes, "ephemeris (second buffer)"))) return rc;
  if ((rc = dalloc((void **)&d_recc, bytes, "ephemeris (third buffer)"))) return rc;
  if ((rc = dalloc((void **)&d_xs0b, xb, "chunk start positions (second buffer)"))) return rc;
  if ((rc = dalloc((void **)&d_xs0c, xb, "chunk start positions (third buffer)"))) return rc;
  rec2_steps_alloc = C;
  return MINT_OK;
}

GraphSig Ctx::graph_sig() const {
  const StepConst sc = step_const();
  GraphSig g;
  g.h = sc.h;
  g.mu = sc.mu;
  g.eps2 = sc.eps2;
  g.tol = sc.tol;
  g.gate_ir0 = sc.gate_ir0;
  g.M = cfg.central_mass;
  g.thresh = score_thresh;
  g.guess = cfg.kepler_guess;
  g.scorer = cfg.use_scorer;
  g.stats = cfg.collect_stats;
  g.skip = scorer_skip_on() ? 1 : 0;
  g.cut = sc_cut;
  return g;
}

StepConst Ctx::step_const() const {
  StepConst sc;
  sc.h = cfg.dt;
  sc.hh = 0.5 * cfg.dt;
  sc.mu = cfg.central_mass;
  // Where 1/r0 <= gate_ir0 the learned guess's gate (dt^2 mu / r0^3 >
  // KM_TAU_GATE^2) cannot open: the bound is 1e-6 below the exact threshold,
  // far more than the five roundings of the gate's product can move it.
  const double gd = sc.mu * (sc.h * sc.h);
  sc.gate_ir0 = gd > 0.0 ? std::cbrt((KM_TAU_GATE * KM_TAU_GATE) / gd) * (1.0 - 1e-6) : INFINITY;
  sc.eps2 = cfg.softening2;
  sc.tol = cfg.bs_tol;
  return sc;
}

TPArrays Ctx::tp_arrays() const {
  TPArrays t;
  t.x = d_tp[0];
  t.y = d_tp[1];
  t.z = d_tp[2];
  t.vx = d_tp[3];
  t.vy = d_tp[4];
  t.vz = d_tp[5];
  return t;
}

// Tiles: each system's TPs cut into runs of block*P, in system order.
int Ctx::ensure_tiles(int block, int P) {
  const long long key = (long long)block * 16 + P;
  if (key == tiles_key) return MINT_OK;
  for (int k = 0; k < 2; ++k)
    if (tilesets[k].key == key) {
      d_tiles = tilesets[k].d;
      ntiles = tilesets[k].n;
      tiles_key = key;
      tiles_next = 1 - k;
      return MINT_OK;
    }
  std::vector<Tile> tl;
  const long long per = (long long)block * P;
  for (int s = 0; s < S; ++s) {
    for (long long a = tp_offset[s]; a < tp_offset[s + 1]; a += per) {
      Tile t;
      t.sys = s;
      t.start = a;
      t.count = (int)std::min(per, tp_offset[s + 1] - a);
      tl.push_back(t);
    }
  }
  // replace the set used least recently; graphs captured with it hold its
  // pointer, so they go
  TileSet &ts = tilesets[tiles_next];
  if (ts.d) {
    drop_graphs();
    hipFree(ts.d);
    allocs.erase(std::find(allocs.begin(), allocs.end(), (void *)ts.d));
    ts = TileSet{};
  }
  d_tiles = nullptr;
  tiles_key = -1;
  int rc = dalloc((void **)&ts.d, tl.size() * sizeof(Tile), "tiles");
  if (rc) return rc;
  ts.n = (int)tl.size();
  if (!tl.empty()) {
    rc = hip(hipMemcpy(ts.d, tl.data(), tl.size() * sizeof(Tile), hipMemcpyHostToDevice), "tiles H2D");
    if (rc) return rc;
  }
  ts.key = key;
  d_tiles = ts.d;
  ntiles = ts.n;
  tiles_key = key;
  tiles_next = 1 - tiles_next;
  return MINT_OK;
}

}  // namespace mint

// ---------------------------------------------------------------------------
// Host numerics shared by upload/download/energy (velocity frame changes).
// Democratic-heliocentric velocity: v_dh = v_h - (sum m_j v_h,j)/(M + sum m_j)
// (democratic_heliocentric_conversion, src/sim.cuh:555-587, forward), and the
// reverse v_h = v_dh + (sum m_j v_dh,j)/M.
static void dh_shift_forward(int nm, const double *m, const double *mv_h, double M, double out[3]) {
  double px = 0.0, py = 0.0, pz = 0.0, msum = 0.0;
  for (int j = 0; j < nm; ++j) {
    px = std::fma(m[j], mv_h[3 * j + 0], px);
    py = std::fma(m[j], mv_h[3 * j + 1], py);
    pz = std::fma(m[j], mv_h[3 * j + 2], pz);
    msum = msum + m[j];
  }
  const double den = M + msum;
  out[0] = px / den;
  out[1] = py / den;
  out[2] = pz / den;
}

static void dh_shift_reverse(int nm, const double *m, const double *mv_dh, double M, double out[3]) {
  double px = 0.0, py = 0.0, pz = 0.0;
  for (int j = 0; j < nm; ++j) {
    px = std::fma(m[j], mv_dh[3 * j + 0], px);
    py = std::fma(m[j], mv_dh[3 * j + 1], py);
    pz = std::fma(m[j], mv_dh[3 * j + 2], pz);
  }
  out[0] = px / M;
  out[1] = py / M;
  out[2] = pz / M;
}

static double dot3h(const double *a, const double *b) {
  return std::fma(a[0], b[0], std::fma(a[1], b[1], a[2] * b[2]));
}

// device arithmetic self-check (mint_selftest_arith)
namespace {
__global__ void selftest_arith_kernel(int op, long long n, const double *__restrict__ a,
                                      const double *__restrict__ b, double *__restrict__ out) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  if (op == 0)
    out[i] = mint::sqrt_dist2(a[i]);
  else if (op == 1)
    out[i] = mint::div_normal(a[i], b[i]);
  else
    out[i] = mint::sqrt_dist2_pos(a[i]);
}
}  // namespace

extern "C" {

int32_t mint_abi_version(void) { return MINT_ABI_VERSION; }

void mint_config_default(mint_config *c) {
  if (!c) return;
  std::memset(c, 0, sizeof(*c));
  c->dt = 0.8219;
  c->central_mass = 1.0;
  c->n_hill = 3.0;
  c->softening2 = 1e-8;
  c->bs_tol = 1e-12;
  c->form = MINT_FORM_BATCHED;
  c->kepler_guess = MINT_GUESS_SERIES;
  c->systems_per_launch = 4;
  c->block_size = 0;
  c->tps_per_lane = 0;
  c->chunk_steps = 0;
  c->use_scorer = 0;
  c->use_graph = 1;
  c->device = 0;
  c->collect_stats = 0;
}

int mint_create(const mint_config *cfg, void **out) {
  if (!out) return MINT_EINVAL;
  *out = nullptr;
  Ctx *c = new Ctx();
  if (cfg)
    c->cfg = *cfg;
  else
    mint_config_default(&c->cfg);
  const mint_config &k = c->cfg;
  if (!(k.dt > 0.0) || !(k.central_mass > 0.0) || !(k.n_hill > 0.0) || !(k.softening2 >= 0.0) ||
      !(k.bs_tol > 0.0) || k.form < 0 || k.form > 2 || k.kepler_guess < 0 || k.kepler_guess > 2) {
    g_create_error = "mint_create: invalid configuration";
    delete c;
    return MINT_EINVAL;
  }
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0) {
    g_create_error = "mint_create: no HIP device";
    delete c;
    return MINT_EHIP;
  }
  if (k.device < 0 || k.device >= ndev) {
    g_create_error = "mint_create: device ordinal out of range";
    delete c;
    return MINT_EINVAL;
  }
  if (hipSetDevice(k.device) != hipSuccess || hipStreamCreate(&c->stream) != hipSuccess ||
      hipStreamCreate(&c->stream2) != hipSuccess || hipStreamCreate(&c->stream3) != hipSuccess) {
    g_create_error = "mint_create: cannot initialise device";
    delete c;
    return MINT_EHIP;
  }
  hipEventCreate(&c->ev0);
  hipEventCreate(&c->ev1);
  noalloc_init(k.device);  // mint_step_noalloc's scratch address on this device, once per device
  *out = c;
  return MINT_OK;
}

void mint_destroy(void *ctx) {
  Ctx *c = (Ctx *)ctx;
  if (!c) return;
  hipSetDevice(c->cfg.device);
  hipStreamSynchronize(c->stream);
  c->free_device();
  hipStreamSynchronize(c->stream2);
  hipStreamSynchronize(c->stream3);
  for (hipEvent_t e : c->ev_side) hipEventDestroy(e);
  hipEventDestroy(c->ev0);
  hipEventDestroy(c->ev1);
  for (hipEvent_t e : c->tev) hipEventDestroy(e);
  for (hipEvent_t e : c->ev_eph) hipEventDestroy(e);
  for (hipEvent_t e : c->ev_tp) hipEventDestroy(e);
  for (hipEvent_t e : c->ev_fork) hipEventDestroy(e);
  for (hipEvent_t e : c->ev_join) hipEventDestroy(e);
  hipStreamDestroy(c->stream);
  hipStreamDestroy(c->stream2);
  hipStreamDestroy(c->stream3);
  delete c;
}

const char *mint_last_error(void *ctx) {
  Ctx *c = (Ctx *)ctx;
  if (!c) return g_create_error.c_str();
  return c->err.c_str();
}

int mint_selftest_arith(int32_t device, int32_t op, int64_t n, const double *a, const double *b,
                        double *out) {
  if (n < 0 || !a || !out || op < 0 || op > 2 || (op == 1 && !b)) return MINT_EINVAL;
  if (n == 0) return MINT_OK;
  if (hipSetDevice(device) != hipSuccess) return MINT_EHIP;
  const size_t bytes = (size_t)n * sizeof(double);
  double *da = nullptr, *db = nullptr, *dout = nullptr;
  int rc = MINT_OK;
  if (hipMalloc(&da, bytes) != hipSuccess || hipMalloc(&dout, bytes) != hipSuccess ||
      (op == 1 && hipMalloc(&db, bytes) != hipSuccess)) {
    rc = MINT_ENOMEM;
  } else if (hipMemcpy(da, a, bytes, hipMemcpyHostToDevice) != hipSuccess ||
             (op == 1 && hipMemcpy(db, b, bytes, hipMemcpyHostToDevice) != hipSuccess)) {
    rc = MINT_EHIP;
  } else {
    selftest_arith_kernel<<<(unsigned)((n + 255) / 256), 256>>>(op, n, da, db, dout);
    if (hipGetLastError() != hipSuccess || hipMemcpy(out, dout, bytes, hipMemcpyDeviceToHost) != hipSuccess)
      rc = MINT_EHIP;
  }
  hipFree(da);
  hipFree(db);
  hipFree(dout);
  return rc;
}

int mint_hill_radii(int64_t n, double M, double n_hill, const double *m, const double *x,
                    const double *v, double *rc) {
  if (n < 0 || !m || !x || !v || !rc || !(M > 0.0)) return MINT_EINVAL;
  for (int64_t i = 0; i < n; ++i) {
    const double r = std::sqrt(dot3h(x + 3 * i, x + 3 * i));
    const double v2 = dot3h(v + 3 * i, v + 3 * i);
    const double inva = 2.0 / r - v2 / M;
    const double a = inva > 0.0 ? 1.0 / inva : r;
    rc[i] = n_hill * (a * std::cbrt(m[i] / (3.0 * M)));
  }
  return MINT_OK;
}

namespace {

// mint_upload proper.  Every argument is checked before the context is
// touched; a failure after that point leaves the context with nothing
// uploaded (the wrapper clears the flag), never with mixed old and new state.
int upload_impl(Ctx *c, const mint_state *st) {
  if (!st) return c->fail(MINT_EINVAL, "mint_upload: null state");
  const int S = st->n_systems, NM = st->n_massive;
  if (S <= 0 || NM < 1 || NM > MB_MAXN)
    return c->fail(MINT_EINVAL, "mint_upload: need n_systems > 0 and 1 <= n_massive <= 2048");
  if (!st->tp_offset || !st->m || !st->mx || !st->mv)
    return c->fail(MINT_EINVAL, "mint_upload: missing tp_offset/m/mx/mv");
  if (st->tp_offset[0] != 0) return c->fail(MINT_EINVAL, "mint_upload: tp_offset[0] must be 0");
  for (int s = 0; s < S; ++s)
    if (st->tp_offset[s + 1] < st->tp_offset[s])
      return c->fail(MINT_EINVAL, "mint_upload: tp_offset must be non-decreasing");
  const int64_t N = st->tp_offset[S];
  if (N >= (int64_t)0x7fffffff) return c->fail(MINT_EINVAL, "mint_upload: at most 2^31-1 TPs");
  if (N > 0 && (!st->tx || !st->ty || !st->tz || !st->tvx || !st->tvy || !st->tvz))
    return c->fail(MINT_EINVAL, "mint_upload: missing TP arrays");
  const bool many = NM > MAX_NM || (c->cfg.many_body && N == 0);
  if (many && N > 0)
    return c->fail(MINT_EINVAL, "mint_upload: test particles need at most 8 massive bodies per system");
  for (int i = 0; i < S * NM; ++i)
    if (!(st->m[i] > 1e-12))
      return c->fail(MINT_EINVAL, "mint_upload: massive masses must exceed 1e-12 (src/mercurius.cu:9)");
  if (st->rc)
    for (int i = 0; i < S * NM; ++i)
      if (!(st->rc[i] > 0.0) || !std::isfinite(st->rc[i]))
        return c->fail(MINT_EINVAL, "mint_upload: rc must be positive and finite");
  hipSetDevice(c->cfg.device);
  hipStreamSynchronize(c->stream);
  // same shape as the resident batch: keep the buffers (and captured graphs)
  const bool reuse = c->uploaded && c->S == S && c->NM == NM && c->Ntp == N &&
                     std::equal(c->tp_offset.begin(), c->tp_offset.end(), st->tp_offset);
  if (!reuse) c->free_device();
  c->uploaded = false;
  c->S = S;
  c->NM = NM;
  c->Ntp = N;
  c->many = many;
  c->L = many ? nullptr : launchers_for(NM);
  c->rad.assign((size_t)S * NM, 0.0);
  c->tp_offset.assign(st->tp_offset, st->tp_offset + S + 1);
  c->m.assign(st->m, st->m + (size_t)S * NM);
  c->rc.resize((size_t)S * NM);
  if (st->rc) {
    for (int i = 0; i < S * NM; ++i) c->rc[i] = st->rc[i];
  } else {
    mint_hill_radii((int64_t)S * NM, c->cfg.central_mass, c->cfg.n_hill, st->m, st->mx, st->mv,
                    c->rc.data());
  }
  // parameter blocks (the many-body engine keeps per-body parameters instead)
  const int PS = many ? 0 : prm_size(NM);
  std::vector<double> prm((size_t)S * PS, 0.0);
  for (int s = 0; s < S && !many; ++s) {
    double *p = prm.data() + (size_t)s * PS;
    const double *rcs = c->rc.data() + (size_t)s * NM;
    for (int j = 0; j < NM; ++j) {
      p[P_M(NM) + j] = c->m[(size_t)s * NM + j];
      p[P_RCLO(NM) + j] = 0.1 * rcs[j];
      p[P_RCWINV(NM) + j] = 1.0 / (0.9 * rcs[j]);
      p[P_RC2(NM) + j] = rcs[j] * rcs[j];
      p[P_FAR2(NM) + j] = (rcs[j] * rcs[j]) * (1.0 + 1e-3);
      for (int k = 0; k < NM; ++k) {
        const double r = std::max(rcs[j], rcs[k]);
        p[P_PRCLO(NM) + j * NM + k] = 0.1 * r;
        p[P_PRCWINV(NM) + j * NM + k] = 1.0 / (0.9 * r);
        p[P_PRC2(NM) + j * NM + k] = r * r;
      }
    }
  }
  // democratic-heliocentric velocities
  const double M = c->cfg.central_mass;
  std::vector<double> mv((size_t)S * NM * 3);
  std::vector<double> tv[3];
  for (int d = 0; d < 3; ++d) tv[d].resize((size_t)N);
  const double *tvin[3] = {st->tvx, st->tvy, st->tvz};
  for (int s = 0; s < S; ++s) {
    double sh[3];
    dh_shift_forward(NM, st->m + (size_t)s * NM, st->mv + (size_t)s * NM * 3, M, sh);
    for (int j = 0; j < NM; ++j)
      for (int d = 0; d < 3; ++d)
        mv[((size_t)s * NM + j) * 3 + d] = st->mv[((size_t)s * NM + j) * 3 + d] - sh[d];
    for (int64_t i = st->tp_offset[s]; i < st->tp_offset[s + 1]; ++i)
      for (int d = 0; d < 3; ++d) tv[d][i] = tvin[d][i] - sh[d];
  }
  std::vector<int> tp_sys((size_t)N);
  for (int s = 0; s < S; ++s)
    for (int64_t i = st->tp_offset[s]; i < st->tp_offset[s + 1]; ++i) tp_sys[i] = s;

  int rc;
  const size_t mb = (size_t)S * NM * 3 * sizeof(double);
  if (!reuse) {
  if ((rc = c->dalloc((void **)&c->d_mx, mb, "massive x"))) return rc;
  if ((rc = c->dalloc((void **)&c->d_mv, mb, "massive v"))) return rc;
  if ((rc = c->dalloc((void **)&c->d_status, (size_t)S * sizeof(unsigned int), "status"))) return rc;
  if ((rc = c->dalloc((void **)&c->d_stats, S_COUNT * sizeof(unsigned long long), "stats"))) return rc;
  if (many && (rc = mb_alloc(c))) return rc;
  }
  if (!reuse && !many) {
  if ((rc = c->dalloc((void **)&c->d_prm, prm.size() * sizeof(double), "params"))) return rc;
  if ((rc = c->dalloc((void **)&c->d_xs0, mb, "chunk start positions"))) return rc;
  for (int d = 0; d < 6; ++d)
    if ((rc = c->dalloc((void **)&c->d_tp[d], (size_t)N * sizeof(double), "TP state"))) return rc;
  if ((rc = c->dalloc((void **)&c->d_tp_sys, (size_t)N * sizeof(int), "TP system index"))) return rc;
  if ((rc = c->dalloc((void **)&c->d_pre, (size_t)N, "queue-head flags"))) return rc;
  if ((rc = c->dalloc((void **)&c->d_carry, (size_t)N, "carry flags"))) return rc;
  if ((rc = c->dalloc((void **)&c->d_need, (size_t)2 * S * sizeof(unsigned int), "cache flags"))) return rc;
  if ((rc = c->dalloc((void **)&c->d_slist, (size_t)4 * S * sizeof(int), "cache system lists"))) return rc;
  if ((rc = c->dalloc((void **)&c->d_q, (size_t)std::max<int64_t>(N, 1) * sizeof(QEntry), "BS queue"))) return rc;
  if ((rc = c->dalloc((void **)&c->d_qctl, 16 * sizeof(unsigned int), "queue counters"))) return rc;
  if ((rc = c->dalloc((void **)&c->d_nnw_kep, KM_NP * sizeof(float), "kepler mlp"))) return rc;
  if ((rc = c->dalloc((void **)&c->d_nnw_sc, SC_NP * sizeof(float), "scorer"))) return rc;
  }

  if (!many && (rc = c->hip(hipMemcpy(c->d_prm, prm.data(), prm.size() * sizeof(double), hipMemcpyHostToDevice), "H2D params"))) return rc;
  if (many && (rc = mb_reset(c))) return rc;
  if ((rc = c->hip(hipMemcpy(c->d_mx, st->mx, mb, hipMemcpyHostToDevice), "H2D massive x"))) return rc;
  if ((rc = c->hip(hipMemcpy(c->d_mv, mv.data(), mb, hipMemcpyHostToDevice), "H2D massive v"))) return rc;
  const double *tpin[3] = {st->tx, st->ty, st->tz};
  if (N > 0) {
    for (int d = 0; d < 3; ++d) {
      if ((rc = c->hip(hipMemcpy(c->d_tp[d], tpin[d], (size_t)N * sizeof(double), hipMemcpyHostToDevice), "H2D TP x"))) return rc;
      if ((rc = c->hip(hipMemcpy(c->d_tp[3 + d], tv[d].data(), (size_t)N * sizeof(double), hipMemcpyHostToDevice), "H2D TP v"))) return rc;
    }
    if ((rc = c->hip(hipMemcpy(c->d_tp_sys, tp_sys.data(), (size_t)N * sizeof(int), hipMemcpyHostToDevice), "H2D TP sys"))) return rc;
  }
  if ((rc = c->hip(hipMemset(c->d_status, 0, (size_t)S * sizeof(unsigned int)), "status clear"))) return rc;
  if ((rc = c->hip(hipMemset(c->d_stats, 0, S_COUNT * sizeof(unsigned long long)), "stats clear"))) return rc;
  if (many) {
    if ((rc = c->hip(hipDeviceSynchronize(), "upload"))) return rc;
    c->uploaded = true;
    c->timing = mint_timing{};
    return MINT_OK;
  }
  // the queues start clear (a batched call clears parity 0's again at its end)
  if ((rc = c->hip(hipMemset(c->d_qctl, 0, 16 * sizeof(unsigned int)), "queue clear"))) return rc;
  if ((rc = c->hip(hipMemset(c->d_need, 0, (size_t)2 * S * sizeof(unsigned int)), "cache flag clear"))) return rc;
  if (N > 0 && (rc = c->hip(hipMemset(c->d_carry, 0, (size_t)N), "carry clear"))) return rc;
  // networks: keep any weights set earlier on this context
  if (!c->w_kep.empty())
    hipMemcpy(c->d_nnw_kep, c->w_kep.data(), KM_NP * sizeof(float), hipMemcpyHostToDevice);
  else
    hipMemset(c->d_nnw_kep, 0, KM_NP * sizeof(float));
  if (!c->w_sc.empty())
    hipMemcpy(c->d_nnw_sc, c->w_sc.data(), SC_NP * sizeof(float), hipMemcpyHostToDevice);
  else
    hipMemset(c->d_nnw_sc, 0, SC_NP * sizeof(float));
  if ((rc = c->hip(hipDeviceSynchronize(), "upload"))) return rc;
  c->uploaded = true;
  c->timing = mint_timing{};
  return MINT_OK;
}

}  // namespace

int mint_upload(void *ctx, const mint_state *st) {
  Ctx *c = (Ctx *)ctx;
  if (!c) return MINT_EINVAL;
  int rc;
  try {
    rc = upload_impl(c, st);
  } catch (const std::bad_alloc &) {
    rc = c->fail(MINT_ENOMEM, "mint_upload: host allocation failed");
  }
  if (rc) c->uploaded = false;
  return rc;
}

int mint_download(void *ctx, mint_state *st) {
  Ctx *c = (Ctx *)ctx;
  if (!c) return MINT_EINVAL;
  if (!c->uploaded) return c->fail(MINT_ESTATE, "mint_download: nothing uploaded");
  if (!st || st->n_systems != c->S || st->n_massive != c->NM)
    return c->fail(MINT_EINVAL, "mint_download: shape differs from the upload");
  hipSetDevice(c->cfg.device);
  int rc = c->hip(hipStreamSynchronize(c->stream), "sync");
  if (rc) return rc;
  const int S = c->S, NM = c->NM;
  const int64_t N = c->Ntp;
  const size_t mb = (size_t)S * NM * 3 * sizeof(double);
  std::vector<double> mx((size_t)S * NM * 3), mv((size_t)S * NM * 3);
  if ((rc = c->hip(hipMemcpy(mx.data(), c->d_mx, mb, hipMemcpyDeviceToHost), "D2H massive x"))) return rc;
  if ((rc = c->hip(hipMemcpy(mv.data(), c->d_mv, mb, hipMemcpyDeviceToHost), "D2H massive v"))) return rc;
  std::vector<double> shift((size_t)S * 3);
  for (int s = 0; s < S; ++s)
    dh_shift_reverse(NM, c->m.data() + (size_t)s * NM, mv.data() + (size_t)s * NM * 3,
                     c->cfg.central_mass, shift.data() + 3 * s);
  if (st->mx) std::memcpy(st->mx, mx.data(), mb);
  if (st->mv) {
    for (int s = 0; s < S; ++s)
      for (int j = 0; j < NM; ++j)
        for (int d = 0; d < 3; ++d)
          st->mv[((size_t)s * NM + j) * 3 + d] = mv[((size_t)s * NM + j) * 3 + d] + shift[3 * s + d];
  }
  if (st->m) std::memcpy(st->m, c->m.data(), c->m.size() * sizeof(double));
  if (st->rc) std::memcpy(st->rc, c->rc.data(), c->rc.size() * sizeof(double));
  double *px[3] = {st->tx, st->ty, st->tz};
  double *pv[3] = {st->tvx, st->tvy, st->tvz};
  std::vector<double> buf((size_t)N);
  for (int d = 0; d < 3; ++d) {
    if (px[d] && N > 0)
      if ((rc = c->hip(hipMemcpy(px[d], c->d_tp[d], (size_t)N * sizeof(double), hipMemcpyDeviceToHost), "D2H TP x"))) return rc;
    if (pv[d] && N > 0) {
      if ((rc = c->hip(hipMemcpy(buf.data(), c->d_tp[3 + d], (size_t)N * sizeof(double), hipMemcpyDeviceToHost), "D2H TP v"))) return rc;
      for (int s = 0; s < S; ++s)
        for (int64_t i = c->tp_offset[s]; i < c->tp_offset[s + 1]; ++i) pv[d][i] = buf[i] + shift[3 * s + d];
    }
  }
  return MINT_OK;
}

// ---------------------------------------------------------------------------
// Checkpoint file: magic, S, NM, format version (2), Ntp, step index, dt, M,
// n_hill, eps^2, tol, then tp_offset[S+1], m, rc, physical radii, mx, mv (DH),
// and the six TP arrays (DH velocities), all little-endian raw.
static const char MINT_CKPT_MAGIC[8] = {'M', 'I', 'N', 'T', 'C', 'K', 'P', '1'};

int mint_save_checkpoint(void *ctx, const char *path, int64_t step_index) {
  Ctx *c = (Ctx *)ctx;
  if (!c) return MINT_EINVAL;
  if (!c->uploaded) return c->fail(MINT_ESTATE, "mint_save_checkpoint: nothing uploaded");
  if (!path) return c->fail(MINT_EINVAL, "mint_save_checkpoint: null path");
  hipSetDevice(c->cfg.device);
  int rc = c->hip(hipStreamSynchronize(c->stream), "sync");
  if (rc) return rc;
  FILE *f = std::fopen(path, "wb");
  if (!f) return c->fail(MINT_EIO, std::string("mint_save_checkpoint: cannot open ") + path);
  const int32_t hdr_i[3] = {c->S, c->NM, 2};
  const int64_t hdr_l[2] = {c->Ntp, step_index};
  const double hdr_d[5] = {c->cfg.dt, c->cfg.central_mass, c->cfg.n_hill, c->cfg.softening2, c->cfg.bs_tol};
  bool ok = std::fwrite(MINT_CKPT_MAGIC, 1, 8, f) == 8 && std::fwrite(hdr_i, sizeof(hdr_i), 1, f) == 1 &&
            std::fwrite(hdr_l, sizeof(hdr_l), 1, f) == 1 && std::fwrite(hdr_d, sizeof(hdr_d), 1, f) == 1;
  ok = ok && std::fwrite(c->tp_offset.data(), sizeof(long long), c->tp_offset.size(), f) == c->tp_offset.size();
  ok = ok && std::fwrite(c->m.data(), sizeof(double), c->m.size(), f) == c->m.size();
  ok = ok && std::fwrite(c->rc.data(), sizeof(double), c->rc.size(), f) == c->rc.size();
  ok = ok && std::fwrite(c->rad.data(), sizeof(double), c->rad.size(), f) == c->rad.size();
  const size_t nb = (size_t)c->S * c->NM * 3;
  std::vector<double> buf(std::max<size_t>(nb, (size_t)c->Ntp));
  for (double *d : {c->d_mx, c->d_mv}) {
    if (ok && (rc = c->hip(hipMemcpy(buf.data(), d, nb * sizeof(double), hipMemcpyDeviceToHost), "D2H"))) {
      std::fclose(f);
      return rc;
    }
    ok = ok && std::fwrite(buf.data(), sizeof(double), nb, f) == nb;
  }
  for (int k = 0; k < 6 && ok; ++k) {
    if (c->Ntp > 0 && (rc = c->hip(hipMemcpy(buf.data(), c->d_tp[k], (size_t)c->Ntp * sizeof(double),
                                             hipMemcpyDeviceToHost), "D2H"))) {
      std::fclose(f);
      return rc;
    }
    ok = ok && std::fwrite(buf.data(), sizeof(double), (size_t)c->Ntp, f) == (size_t)c->Ntp;
  }
  ok = (std::fclose(f) == 0) && ok;
  return ok ? MINT_OK : c->fail(MINT_EIO, "mint_save_checkpoint: write failed");
}

namespace {

// mint_load_checkpoint proper.  The header is checked against the file's
// size before anything is allocated, and the context's step parameters are
// only replaced once the whole file has been read; if the upload then fails,
// the previous parameters are restored and nothing is uploaded.
int load_checkpoint_impl(Ctx *c, const char *path, int64_t *step_index) {
  if (!path) return c->fail(MINT_EINVAL, "mint_load_checkpoint: null path");
  FILE *f = std::fopen(path, "rb");
  if (!f) return c->fail(MINT_EIO, std::string("mint_load_checkpoint: cannot open ") + path);
  long fsize = -1;
  if (std::fseek(f, 0, SEEK_END) == 0) fsize = std::ftell(f);
  std::rewind(f);
  char magic[8];
  int32_t hdr_i[3];
  int64_t hdr_l[2];
  double hdr_d[5];
  bool ok = fsize > 0 && std::fread(magic, 1, 8, f) == 8 && std::memcmp(magic, MINT_CKPT_MAGIC, 8) == 0 &&
            std::fread(hdr_i, sizeof(hdr_i), 1, f) == 1 && std::fread(hdr_l, sizeof(hdr_l), 1, f) == 1 &&
            std::fread(hdr_d, sizeof(hdr_d), 1, f) == 1;
  const int S = ok ? hdr_i[0] : 0, NM = ok ? hdr_i[1] : 0;
  const int64_t N = ok ? hdr_l[0] : 0;
  ok = ok && S > 0 && NM >= 1 && NM <= MB_MAXN && N >= 0 && N < 0x7fffffff && hdr_i[2] == 2;
  // the body: tp_offset[S+1], then m, rc, radii [S*NM], mx, mv [S*NM*3], six TP arrays [N]
  const long long head = 8 + (long long)sizeof(hdr_i) + (long long)sizeof(hdr_l) + (long long)sizeof(hdr_d);
  const long long expect = ok ? head + 8LL * ((long long)S + 1) + 8LL * 9 * (long long)S * NM + 8LL * 6 * N : -1;
  if (!ok || expect != (long long)fsize) {
    std::fclose(f);
    return c->fail(MINT_EIO, "mint_load_checkpoint: not a mint checkpoint, or its size disagrees with its header");
  }
  std::vector<int64_t> off((size_t)S + 1);
  std::vector<double> m((size_t)S * NM), rcv((size_t)S * NM), rad((size_t)S * NM), mx((size_t)S * NM * 3),
      mv((size_t)S * NM * 3);
  std::vector<double> t[6];
  for (auto &a : t) a.resize((size_t)N);
  ok = std::fread(off.data(), sizeof(int64_t), off.size(), f) == off.size() &&
       std::fread(m.data(), sizeof(double), m.size(), f) == m.size() &&
       std::fread(rcv.data(), sizeof(double), rcv.size(), f) == rcv.size() &&
       std::fread(rad.data(), sizeof(double), rad.size(), f) == rad.size() &&
       std::fread(mx.data(), sizeof(double), mx.size(), f) == mx.size() &&
       std::fread(mv.data(), sizeof(double), mv.size(), f) == mv.size();
  for (int k = 0; k < 6 && ok; ++k) ok = std::fread(t[k].data(), sizeof(double), (size_t)N, f) == (size_t)N;
  std::fclose(f);
  if (!ok || off[0] != 0 || off[S] != N) return c->fail(MINT_EIO, "mint_load_checkpoint: truncated or inconsistent file");
  const mint_config saved = c->cfg;
  c->cfg.dt = hdr_d[0];
  c->cfg.central_mass = hdr_d[1];
  c->cfg.n_hill = hdr_d[2];
  c->cfg.softening2 = hdr_d[3];
  c->cfg.bs_tol = hdr_d[4];
  // upload the geometry with zero massive velocities (no DH shift), then
  // overwrite the velocities with the saved internal (DH) ones, bit for bit
  std::vector<double> zeros_v(mv.size(), 0.0);
  mint_state st{S, NM, off.data(), m.data(), mx.data(), zeros_v.data(), t[0].data(), t[1].data(), t[2].data(),
                t[3].data(), t[4].data(), t[5].data(), rcv.data()};
  int rc = mint_upload(c, &st);
  if (!rc) {
    hipSetDevice(c->cfg.device);
    rc = c->hip(hipMemcpy(c->d_mv, mv.data(), mv.size() * sizeof(double), hipMemcpyHostToDevice), "H2D");
    for (int k = 3; k < 6 && !rc; ++k)
      if (N > 0)
        rc = c->hip(hipMemcpy(c->d_tp[k], t[k].data(), (size_t)N * sizeof(double), hipMemcpyHostToDevice), "H2D");
    if (!rc) rc = mint_set_radii(c, rad.data());
  }
  if (rc) {
    c->cfg = saved;
    c->uploaded = false;
    return rc;
  }
  if (step_index) *step_index = hdr_l[1];
  return MINT_OK;
}

}  // namespace

int mint_load_checkpoint(void *ctx, const char *path, int64_t *step_index) {
  Ctx *c = (Ctx *)ctx;
  if (!c) return MINT_EINVAL;
  try {
    return load_checkpoint_impl(c, path, step_index);
  } catch (const std::bad_alloc &) {
    c->uploaded = false;
    return c->fail(MINT_ENOMEM, "mint_load_checkpoint: host allocation failed");
  }
}

int mint_set_radii(void *ctx, const double *radius) {
  Ctx *c = (Ctx *)ctx;
  if (!c) return MINT_EINVAL;
  if (!c->uploaded) return c->fail(MINT_ESTATE, "mint_set_radii: nothing uploaded");
  if (!radius) return c->fail(MINT_EINVAL, "mint_set_radii: null radii");
  const size_t n = (size_t)c->S * c->NM;
  for (size_t i = 0; i < n; ++i)
    if (!(radius[i] >= 0.0) || !std::isfinite(radius[i]))
      return c->fail(MINT_EINVAL, "mint_set_radii: radii must be finite and >= 0");
  hipSetDevice(c->cfg.device);
  int rc = c->hip(hipStreamSynchronize(c->stream), "sync");
  if (rc) return rc;
  c->rad.assign(radius, radius + n);
  if (c->many) return mb_set_radii(c);
  const int PS = prm_size(c->NM);
  for (int s = 0; s < c->S; ++s)
    if ((rc = c->hip(hipMemcpy(c->d_prm + (size_t)s * PS + P_RAD(c->NM), radius + (size_t)s * c->NM,
                               (size_t)c->NM * sizeof(double), hipMemcpyHostToDevice),
                     "H2D radii")))
      return rc;
  return MINT_OK;
}

// E = sum_j m_j |v_j|^2/2 + |sum_j m_j v_j|^2/(2M) - sum_j M m_j/|x_j|
//     + sum_{j<k} m_j m_k U(|x_j - x_k|)   (democratic-heliocentric Hamiltonian),
// U(r) = -atan(eps/r)/eps, the potential of the softened pair force.
int mint_energy(void *ctx, double *E) {
  Ctx *c = (Ctx *)ctx;
  if (!c) return MINT_EINVAL;
  if (!c->uploaded) return c->fail(MINT_ESTATE, "mint_energy: nothing uploaded");
  if (!E) return c->fail(MINT_EINVAL, "mint_energy: null output");
  hipSetDevice(c->cfg.device);
  int rc = c->hip(hipStreamSynchronize(c->stream), "sync");
  if (rc) return rc;
  const int S = c->S, NM = c->NM;
  std::vector<double> mx((size_t)S * NM * 3), mv((size_t)S * NM * 3);
  const size_t mb = mx.size() * sizeof(double);
  if ((rc = c->hip(hipMemcpy(mx.data(), c->d_mx, mb, hipMemcpyDeviceToHost), "D2H massive x"))) return rc;
  if ((rc = c->hip(hipMemcpy(mv.data(), c->d_mv, mb, hipMemcpyDeviceToHost), "D2H massive v"))) return rc;
  const double M = c->cfg.central_mass;
  const double eps = std::sqrt(c->cfg.softening2);
  for (int s = 0; s < S; ++s) {
    const double *x = mx.data() + (size_t)s * NM * 3;
    const double *v = mv.data() + (size_t)s * NM * 3;
    const double *m = c->m.data() + (size_t)s * NM;
    double kin = 0.0, pot = 0.0, px = 0.0, py = 0.0, pz = 0.0;
    for (int j = 0; j < NM; ++j) {
      kin += 0.5 * m[j] * dot3h(v + 3 * j, v + 3 * j);
      px = std::fma(m[j], v[3 * j + 0], px);
      py = std::fma(m[j], v[3 * j + 1], py);
      pz = std::fma(m[j], v[3 * j + 2], pz);
      pot -= M * m[j] / std::sqrt(dot3h(x + 3 * j, x + 3 * j));
      for (int k = j + 1; k < NM; ++k) {
        const double d[3] = {x[3 * k] - x[3 * j], x[3 * k + 1] - x[3 * j + 1], x[3 * k + 2] - x[3 * j + 2]};
        // potential of the softened pair force m_j m_k d/((r^2+eps^2) r)
        const double r = std::sqrt(dot3h(d, d));
        pot += m[j] * m[k] * (eps > 0.0 ? -std::atan(eps / r) / eps : -1.0 / r);
      }
    }
    const double p2 = std::fma(px, px, std::fma(py, py, pz * pz));
    E[s] = kin + p2 / (2.0 * M) + pot;
  }
  return MINT_OK;
}

int mint_status(void *ctx, uint32_t *status) {
  Ctx *c = (Ctx *)ctx;
  if (!c) return MINT_EINVAL;
  if (!c->uploaded) return c->fail(MINT_ESTATE, "mint_status: nothing uploaded");
  if (!status) return c->fail(MINT_EINVAL, "mint_status: null output");
  hipSetDevice(c->cfg.device);
  int rc = c->hip(hipStreamSynchronize(c->stream), "sync");
  if (rc) return rc;
  return c->hip(hipMemcpy(status, c->d_status, (size_t)c->S * sizeof(unsigned int), hipMemcpyDeviceToHost), "D2H status");
}

int mint_get_stats(void *ctx, mint_stats *out) {
  Ctx *c = (Ctx *)ctx;
  if (!c) return MINT_EINVAL;
  if (!c->uploaded || !out) return c->fail(MINT_ESTATE, "mint_get_stats: nothing uploaded");
  hipSetDevice(c->cfg.device);
  unsigned long long v[S_COUNT];
  int rc = c->hip(hipStreamSynchronize(c->stream), "sync");
  if (rc) return rc;
  if ((rc = c->hip(hipMemcpy(v, c->d_stats, sizeof(v), hipMemcpyDeviceToHost), "D2H stats"))) return rc;
  out->tp_steps = (int64_t)v[S_TPSTEPS];
  out->kepler_solves = (int64_t)v[S_KEPLER];
  out->newton_iters = (int64_t)v[S_NEWTON];
  out->bs_steps = (int64_t)v[S_BSSTEPS];
  out->bs_rhs_evals = (int64_t)v[S_BSRHS];
  out->deferred = (int64_t)v[S_DEFER];
  out->prebucketed = (int64_t)v[S_PREBUCKET];
  out->idle_lane_steps = (int64_t)v[S_IDLE];
  out->enc_tp_steps = (int64_t)v[S_ENCSTEPS];
  out->massive_groups = (int64_t)v[S_MGROUPS];
  out->massive_group_bodies = (int64_t)v[S_MGBODIES];
  out->overlaps = (int64_t)v[S_OVERLAP];
  return MINT_OK;
}

int mint_get_timing(void *ctx, mint_timing *out) {
  Ctx *c = (Ctx *)ctx;
  if (!c || !out) return MINT_EINVAL;
  *out = c->timing;
  return MINT_OK;
}

int mint_reset_counters(void *ctx) {
  Ctx *c = (Ctx *)ctx;
  if (!c) return MINT_EINVAL;
  c->timing = mint_timing{};
  if (!c->uploaded) return MINT_OK;
  hipSetDevice(c->cfg.device);
  return c->hip(hipMemset(c->d_stats, 0, S_COUNT * sizeof(unsigned long long)), "stats clear");
}

int mint_set_mode(void *ctx, int32_t form, int32_t guess, int32_t use_scorer) {
  Ctx *c = (Ctx *)ctx;
  if (!c) return MINT_EINVAL;
  if (form < 0 || form > 2 || guess < 0 || guess > 2) return c->fail(MINT_EINVAL, "mint_set_mode: bad form or guess");
  if (c->cfg.form != form || c->cfg.kepler_guess != guess || c->cfg.use_scorer != use_scorer) {
    c->drop_graphs();
  }
  c->cfg.form = form;
  c->cfg.kepler_guess = guess;
  c->cfg.use_scorer = use_scorer;
  return MINT_OK;
}

// ---------------------------------------------------------------------------
// Element conversions at the configuration edge (host).

// Kepler's equation E - e sin E = M (elliptic) or e sinh H - H = M
// (hyperbolic) by safeguarded Newton to machine precision.
static double solve_kepler_eq(double Mn, double e) {
  if (e < 1.0) {
    double Mr = std::fmod(Mn, 2.0 * M_PI);
    if (Mr < 0) Mr += 2.0 * M_PI;
    double E = e < 0.8 ? Mr : M_PI;
    for (int i = 0; i < 100; ++i) {
      const double f = E - e * std::sin(E) - Mr;
      const double d = f / (1.0 - e * std::cos(E));
      E -= d;
      if (std::fabs(d) <= 1e-15 * std::max(1.0, std::fabs(E))) break;
    }
    return E;
  }
  double H = std::asinh(Mn / e);
  for (int i = 0; i < 200; ++i) {
    const double f = e * std::sinh(H) - H - Mn;
    const double d = f / (e * std::cosh(H) - 1.0);
    H -= d;
    if (std::fabs(d) <= 1e-15 * std::max(1.0, std::fabs(H))) break;
  }
  return H;
}

int mint_elements_to_cartesian(int64_t n, double mu, const double *inc, const double *node,
                               const double *argp, const double *manom, const double *ecc,
                               const double *sma, double *x, double *v) {
  if (n < 0 || !(mu > 0.0) || !inc || !node || !argp || !manom || !ecc || !sma || !x || !v) return MINT_EINVAL;
  for (int64_t k = 0; k < n; ++k) {
    const double e = ecc[k], a = sma[k];
    const double ci = std::cos(inc[k]), si = std::sin(inc[k]);
    const double co = std::cos(node[k]), so = std::sin(node[k]);
    const double cw = std::cos(argp[k]), sw = std::sin(argp[k]);
    // perifocal unit vectors P, Q
    const double P[3] = {cw * co - sw * so * ci, cw * so + sw * co * ci, sw * si};
    const double Q[3] = {-sw * co - cw * so * ci, -sw * so + cw * co * ci, cw * si};
    double xp, yp, vxp, vyp;
    if (e < 1.0) {
      const double E = solve_kepler_eq(manom[k], e);
      const double cE = std::cos(E), sE = std::sin(E);
      const double b = a * std::sqrt(1.0 - e * e);
      xp = a * (cE - e);
      yp = b * sE;
      const double r = a * (1.0 - e * cE);
      const double nfac = std::sqrt(mu * a) / r;
      vxp = -nfac * sE;
      vyp = nfac * std::sqrt(1.0 - e * e) * cE;
    } else {
      const double aa = std::fabs(a);
      const double H = solve_kepler_eq(manom[k], e);
      const double cH = std::cosh(H), sH = std::sinh(H);
      xp = aa * (e - cH);
      yp = aa * std::sqrt(e * e - 1.0) * sH;
      const double r = aa * (e * cH - 1.0);
      const double nfac = std::sqrt(mu * aa) / r;
      vxp = -nfac * sH;
      vyp = nfac * std::sqrt(e * e - 1.0) * cH;
    }
    for (int d = 0; d < 3; ++d) {
      x[3 * k + d] = xp * P[d] + yp * Q[d];
      v[3 * k + d] = vxp * P[d] + vyp * Q[d];
    }
  }
  return MINT_OK;
}

int mint_cartesian_to_elements(int64_t n, double mu, const double *x, const double *v,
                               double *inc, double *node, double *argp, double *manom,
                               double *ecc, double *sma) {
  if (n < 0 || !(mu > 0.0) || !x || !v || !inc || !node || !argp || !manom || !ecc || !sma) return MINT_EINVAL;
  for (int64_t k = 0; k < n; ++k) {
    const double *r = x + 3 * k, *u = v + 3 * k;
    const double h[3] = {r[1] * u[2] - r[2] * u[1], r[2] * u[0] - r[0] * u[2], r[0] * u[1] - r[1] * u[0]};
    const double hn = std::sqrt(dot3h(h, h));
    const double rn = std::sqrt(dot3h(r, r));
    const double v2 = dot3h(u, u);
    const double rv = dot3h(r, u);
    const double ev[3] = {((v2 - mu / rn) * r[0] - rv * u[0]) / mu, ((v2 - mu / rn) * r[1] - rv * u[1]) / mu,
                          ((v2 - mu / rn) * r[2] - rv * u[2]) / mu};
    const double e = std::sqrt(dot3h(ev, ev));
    const double a = 1.0 / (2.0 / rn - v2 / mu);
    const double i = std::acos(std::max(-1.0, std::min(1.0, h[2] / hn)));
    double Om = std::atan2(h[0], -h[1]);
    if (hn * std::sin(i) == 0.0) Om = 0.0;
    // argument of latitude and true anomaly
    const double nvec[3] = {std::cos(Om), std::sin(Om), 0.0};
    const double ci = std::cos(i), si = std::sin(i);
    const double pvec[3] = {-std::sin(Om) * ci, std::cos(Om) * ci, si};
    const double ul = std::atan2(dot3h(r, pvec), dot3h(r, nvec));
    double f;
    if (e > 1e-14) {
      // e cos f = h^2/(mu r) - 1,  e sin f = h (r.v)/(mu r)
      f = std::atan2((hn * rv) / (mu * rn), dot3h(h, h) / (mu * rn) - 1.0);
    } else {
      f = ul;
    }
    double w = ul - f;
    w = std::fmod(w + 4.0 * M_PI, 2.0 * M_PI);
    double Mn;
    if (e < 1.0) {
      const double E = std::atan2(std::sqrt(1.0 - e * e) * std::sin(f), e + std::cos(f));
      Mn = E - e * std::sin(E);
      Mn = std::fmod(Mn + 4.0 * M_PI, 2.0 * M_PI);
    } else {
      const double Hh = 2.0 * std::atanh(std::sqrt((e - 1.0) / (e + 1.0)) * std::tan(0.5 * f));
      Mn = e * std::sinh(Hh) - Hh;
    }
    inc[k] = i;
    node[k] = Om;
    argp[k] = w;
    manom[k] = Mn;
    ecc[k] = e;
    sma[k] = a;
  }
  return MINT_OK;
}

double mint_flops_per_tp_step(int32_t nm, double nbar) {
  return 2.0 * nm * 28.0 + 12.0 + 12.0 + 9.0 * nm + 60.0 + 40.0 * nbar;
}

double mint_flops_per_kick_pass(int64_t n) { return 28.0 * (double)n * (double)(n - 1); }

}  // extern "C"
