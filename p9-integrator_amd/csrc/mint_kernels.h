// mint_kernels.h — gfx950 kernels of the Mercurius step.
//
//  ephemeris_kernel   massive subsystem of each system for C steps; one lane
//                     per massive body; writes the per-step ephemeris records that
//                     every TP phase reads (replaces the cluster-wide reads of
//                     block 0's shared memory, src/sim.cuh:15-26,682-741)
//  tp_step_kernel     one step for one TP per lane (plain and flat forms)
//  tp_main_kernel     C steps per launch with TP state in registers; a TP that
//                     fails the encounter check is appended to the BS queue and
//                     its lane goes idle for the rest of the launch
//  tp_encounter_kernel  consumes the BS queue: each lane takes one deferred TP
//                     and carries it to the end of the launch window
//  tp_prebucket_kernel  builds a window's queue head: TPs deferred in the
//                     previous window (carried) and, with the scorer, TPs
//                     likely to meet a massive body within the window
//
// A TP deferred mid-window is finished by the encounter kernel after the main
// kernel; with a next window in the same call it also starts that window in
// the queue ("carried"), so its tail can run beside the next main kernel.
#pragma once
#include "mint_device.h"
#include "mint_nn.h"

namespace mint {

struct Tile {
  int sys;
  int count;
  long long start;
};

// One deferred TP (64 B).  phase 0: the TP starts the window in the queue; its
// state is read from the TP arrays when the entry is served (s unused);
// phase 1: state after jump 1 of step k0, encounter flagged.
struct QEntry {
  int tp;
  int sys;
  int k0;
  int phase;
  double s[6];
};

struct TPArrays {
  double *x, *y, *z, *vx, *vy, *vz;
};

// stats slots (unsigned long long)
enum {
  S_TPSTEPS = 0,
  S_KEPLER = 1,
  S_NEWTON = 2,
  S_BSSTEPS = 3,
  S_BSRHS = 4,
  S_DEFER = 5,
  S_PREBUCKET = 6,
  S_IDLE = 7,
  S_ENCSTEPS = 8,
  S_MGROUPS = 9,
  S_MGBODIES = 10,
  S_OVERLAP = 11,
  S_COUNT = 12
};

__device__ __forceinline__ void q_put(QEntry *e, int tp, int sys, int k0, int phase, double x,
                                      double y, double z, double vx, double vy, double vz) {
  e->tp = tp;
  e->sys = sys;
  e->k0 = k0;
  e->phase = phase;
  e->s[0] = x;
  e->s[1] = y;
  e->s[2] = z;
  e->s[3] = vx;
  e->s[4] = vy;
  e->s[5] = vz;
}

__device__ __forceinline__ void wave_add(unsigned long long *dst, unsigned int v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  if ((threadIdx.x & 63) == 0 && v != 0u) atomicAdd(dst, (unsigned long long)v);
}

// Wavefront-wide helpers for converged code (every lane of the wavefront
// active).  wave_count: lanes where pred holds.  wave_sum: sum of v over the
// wavefront.  wave_sum_rare: the same, read lane by lane from those with
// v != 0 (cheap when they are few).  wave_add_uniform: add a wavefront-uniform
// v to a global counter once.
__device__ __forceinline__ unsigned int wave_count(bool pred) {
  return (unsigned int)__builtin_popcountll(__builtin_amdgcn_ballot_w64(pred));
}

__device__ __forceinline__ unsigned int wave_sum(unsigned int v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  return __builtin_amdgcn_readfirstlane(v);
}

__device__ __forceinline__ unsigned int wave_sum_rare(unsigned int v) {
  unsigned long long m = __builtin_amdgcn_ballot_w64(v != 0u);
  unsigned int t = 0;
  while (m != 0ull) {
    const int l = __builtin_ctzll(m);
    t += (unsigned int)__builtin_amdgcn_readlane((int)v, l);
    m &= m - 1ull;
  }
  return t;
}

__device__ __forceinline__ void wave_add_uniform(unsigned long long *dst, unsigned int v) {
  if ((threadIdx.x & 63) == 0 && v != 0u) atomicAdd(dst, (unsigned long long)v);
}

__device__ __forceinline__ bool finite6(double a, double b, double c, double d, double e,
                                        double f) {
  return isfinite(a) && isfinite(b) && isfinite(c) && isfinite(d) && isfinite(e) && isfinite(f);
}

// Kepler-MLP input features of one drift (FP32).  tau = dt sqrt(mu/r0^3).
__device__ __forceinline__ void kepler_features(const KepStart &k, double mu, double dt,
                                                float *f, double &tau) {
  tau = dt * sqrt(mu / (k.r0 * (k.r0 * k.r0)));
  f[0] = (log2f((float)tau) + 6.0f) * 0.25f;
  f[1] = (float)(k.zeta0 / mu);
  f[2] = (float)(k.eta0 / sqrt(mu * k.r0));
  float ra = (float)(k.r0 * k.beta / mu);
  f[3] = fminf(fmaxf(ra, -2.0f), 2.0f);
}

// Learned correction factor 1 + c tau^3 of the series guess.  Out of line:
// the network's weights and activations stay out of the register budget of
// the step loops that call it (the call is taken only above the gate).
static __device__ __noinline__ double learned_guess_factor(double r0, double eta0, double beta, double zeta0,
                                                    double mu, double dt, const float *__restrict__ nnw) {
  KepStart k;
  k.r0 = r0;
  k.ir0 = 0.0;
  k.eta0 = eta0;
  k.beta = beta;
  k.zeta0 = zeta0;
  float f[4];
  double tau;
  kepler_features(k, mu, dt, f, tau);
  const double c = (double)mlp_kepler_fwd(nnw, f);
  return fma(c, tau * (tau * tau), 1.0);
}

// Starting guess of the TP Kepler solve.  GUESS: 0 crude, 1 series, 2 learned
// (series guess times (1 + c tau^3), c from the MLP, where tau > KM_TAU_GATE).
// gate_ir0 (StepConst) bounds 1/r0 from below where the gate can open, with a
// margin of 1e-6 over the rounding of the gate's product, so the one
// comparison in front of it decides nothing and spares the product for every
// distant body.
template <int GUESS>
__device__ __forceinline__ double tp_guess(const KepStart &k, double mu, double dt,
                                           double gate_ir0, const float *__restrict__ nnw) {
  if (GUESS == 0) return guess_crude(k, dt);
  double s = guess_series(k, dt);
  if (GUESS == 2 && k.ir0 > gate_ir0) {
    // gate tau = dt sqrt(mu / r0^3) > KM_TAU_GATE, tested without sqrt or division
    const double ir3 = k.ir0 * (k.ir0 * k.ir0);
    if ((dt * dt) * (mu * ir3) > KM_TAU_GATE * KM_TAU_GATE)
      s = s * learned_guess_factor(k.r0, k.eta0, k.beta, k.zeta0, mu, dt, nnw);
  }
  return s;
}

// TP drift over the step: Kepler unless flagged.  Returns Newton iterations.
template <int GUESS>
__device__ __forceinline__ int tp_kepler(double &x, double &y, double &z, double &vx, double &vy,
                                         double &vz, const StepConst &sc,
                                         const float *__restrict__ nnw, uint32_t &st) {
  const KepStart k = kep_start(x, y, z, vx, vy, vz, sc.mu);
  const double s0 = tp_guess<GUESS>(k, sc.mu, sc.h, sc.gate_ir0, nnw);
  bool ok;
  const int it = kepler_drift_from(k, s0, x, y, z, vx, vy, vz, sc.mu, sc.h, ok);
  if (!ok) st |= ST_NEWTON;
  return it;
}

// ---------------------------------------------------------------------------
// Massive subsystem, lane per body: a group of EPH_L(NM) lanes per system
// (the body count rounded up to a power of two), lane j holding body j, so a
// step's five Kepler drifts, kicks and pair checks run side by side instead of
// in one lane.  Each lane forms exactly what iteration j of the body loops of
// accel_massive, momentum_q and the pair check forms, in the same order, with
// the other bodies' positions and velocities taken from their lanes by
// shuffles; the records are therefore the same bits.  Lanes beyond NM shadow
// the last body (nothing of theirs is written).
template <int NM>
constexpr int EPH_L = NM <= 1 ? 1 : (NM <= 2 ? 2 : (NM <= 4 ? 4 : 8));

// Per-lane copies of body j's parameters, read once per launch (held in
// registers across the step loop): partner masses and the pair changeover
// fields max/min over (j, k), and the squared sums of physical radii.
template <int NM>
struct LaneParams {
  double m[NM], prclo[NM], prcwinv[NM], prc2[NM], sr2[NM];
};

template <int NM>
__device__ __forceinline__ void lane_params(const double *__restrict__ prm, int j, LaneParams<NM> &p) {
#pragma unroll
  for (int k = 0; k < NM; ++k) {
    p.m[k] = prm[P_M(NM) + k];
    p.prclo[k] = prm[P_PRCLO(NM) + j * NM + k];
    p.prcwinv[k] = prm[P_PRCWINV(NM) + j * NM + k];
    p.prc2[k] = prm[P_PRC2(NM) + j * NM + k];
    const double sr = prm[P_RAD(NM) + j] + prm[P_RAD(NM) + k];
    p.sr2[k] = sr * sr;
  }
}

// Kick acceleration of body j (this lane) from the bodies of its lane group
// (lanes base .. base+NM-1): iteration j of accel_massive.
template <int NM>
__device__ __forceinline__ void accel_lane(const double *x, int base, int j, const LaneParams<NM> &p,
                                           double eps2, double *a) {
  double dx[NM], dy[NM], dz[NM], num[NM], den[NM], w[NM];
#pragma unroll
  for (int k = 0; k < NM; ++k) {
    const double xk = __shfl(x[0], base + k, 64);
    const double yk = __shfl(x[1], base + k, 64);
    const double zk = __shfl(x[2], base + k, 64);
    dx[k] = xk - x[0];
    dy[k] = yk - x[1];
    dz[k] = zk - x[2];
    const double r2 = dot3(dx[k], dy[k], dz[k], dx[k], dy[k], dz[k]);
    kick_nd(r2, eps2, p.m[k], p.prclo[k], p.prcwinv[k], num[k], den[k]);
    if (k == j) {  // the body's own slot is inert
      dx[k] = 0.0;
      dy[k] = 0.0;
      dz[k] = 0.0;
      num[k] = 0.0;
      den[k] = 1.0;
    }
  }
  batched_div<NM>(num, den, w);
  double ax = 0.0, ay = 0.0, az = 0.0;
#pragma unroll
  for (int k = 0; k < NM; ++k) {
    ax = fma(w[k], dx[k], ax);
    ay = fma(w[k], dy[k], ay);
    az = fma(w[k], dz[k], az);
  }
  a[0] = ax;
  a[1] = ay;
  a[2] = az;
}

// Q = (sum_k m_k v_k) / M over the lane group, the order of momentum_q.
template <int NM>
__device__ __forceinline__ void momentum_lane(const double *v, int base, const LaneParams<NM> &p, double M,
                                              double *Q) {
  double px = 0.0, py = 0.0, pz = 0.0;
#pragma unroll
  for (int k = 0; k < NM; ++k) {
    const double m = p.m[k];
    px = fma(m, __shfl(v[0], base + k, 64), px);
    py = fma(m, __shfl(v[1], base + k, 64), py);
    pz = fma(m, __shfl(v[2], base + k, 64), pz);
  }
  Q[0] = px / M;
  Q[1] = py / M;
  Q[2] = pz / M;
}

// Encounter groups of one system, from its bodies' states S6[NM][6] (x, v)
// and adjacency rows adj[NM] (bit k of adj[j]: pair (j, k) inside its
// changeover radius): each connected component, members in index order, is
// drifted by one BS as one system, in place; *ingrp receives the bodies in a
// group.  The component walk and the BS are those of the one-lane form.  Out
// of line: taken only when some pair of the system is that close.  Its
// arrays (a group's states and members, and the BS's) live in the caller's
// workspace ws of massive_groups_ws<NM>() doubles (LDS in the ephemeris
// kernel), so that no private memory is needed.
template <int NM>
constexpr int massive_groups_ws() {
  return 7 * NM + bs_group_ws<NM>();  // sv (6 NM), the member indices (NM), the BS's
}

template <int NM>
__device__ __noinline__ uint32_t massive_groups(double *S6, const uint32_t *adj, const double *__restrict__ prm,
                                                StepConst sc, unsigned long long *stats, uint32_t *ingrp,
                                                double *ws) {
  uint32_t st = ST_MASSIVE_ENC, in = 0u;
  unsigned int ng = 0, nb = 0, nr = 0;
  for (int r = 0; r < NM; ++r) {
    if (adj[r] == 0u || ((in >> r) & 1u)) continue;
    uint32_t comp = 1u << r, front = comp;
    while (front) {
      uint32_t nxt = 0u;
      for (int j = 0; j < NM; ++j)
        if ((front >> j) & 1u) nxt |= adj[j];
      front = nxt & ~comp;
      comp |= nxt;
    }
    in |= comp;
    double *sv = ws;
    int *idx = (int *)(ws + 6 * NM);
    int g = 0;
    for (int j = 0; j < NM; ++j) {
      if (!((comp >> j) & 1u)) continue;
      for (int c = 0; c < 6; ++c) sv[6 * g + c] = S6[6 * j + c];
      idx[g++] = j;
    }
    st |= bs_drift_group<NM>(sv, g, idx, sc.h, prm, sc.mu, sc.eps2, sc.tol, nr, ws + 7 * NM);
    for (int q = 0; q < g; ++q)
      for (int c = 0; c < 6; ++c) S6[6 * idx[q] + c] = sv[6 * q + c];
    ++ng;
    nb += (unsigned int)g;
  }
  if (stats) {
    atomicAdd(stats + S_MGROUPS, (unsigned long long)ng);
    atomicAdd(stats + S_MGBODIES, (unsigned long long)nb);
  }
  *ingrp = in;
  return st;
}

// The massive subsystem of systems [s_begin, s_end) for C steps: the
// per-step ephemeris records every TP phase reads (replaces the cluster-wide
// reads of block 0's shared memory, src/sim.cuh:15-26,682-741), from the
// start state (ix, iv); the state at the end of the C steps goes into (mx, mv)
// when mx is not null (ix = mx: in place) and into (nx, nv) when nx is not
// null (the next launch window's start, which the queue head of that window
// reads).  One-wave blocks of 64 / EPH_L systems.
template <int NM>
__global__ void __launch_bounds__(64) ephemeris_kernel(int s_begin, int s_end, int C,
                                                       const double *__restrict__ prm_all,
                                                       const double *ix, const double *iv,
                                                       double *mx, double *mv,
                                                       double *__restrict__ xs0,
                                                       double *__restrict__ rec,
                                                       double *__restrict__ nx,
                                                       double *__restrict__ nv,
                                                       unsigned int *__restrict__ status,
                                                       unsigned long long *__restrict__ stats,
                                                       StepConst sc, double M) {
  constexpr int L = EPH_L<NM>;
  constexpr int SPW = 64 / L;  // systems per wave
  constexpr int REC = rec_size(NM);
  __shared__ double sh_state[SPW][NM][6];
  __shared__ uint32_t sh_adj[SPW][NM];
  __shared__ uint32_t sh_grp[SPW];
  __shared__ double sh_ws[SPW][massive_groups_ws<NM>()];  // massive_groups' arrays, per system
  const int lane = (int)threadIdx.x;
  const int gs = lane / L, j = lane % L, base = gs * L;
  const int s_raw = s_begin + (int)blockIdx.x * SPW + gs;
  const bool sys_ok = s_raw < s_end;
  const int s = sys_ok ? s_raw : s_end - 1;  // a group past the end shadows the last system
  const bool own = sys_ok && j < NM;
  const int jb = j < NM ? j : NM - 1;
  const double *prm = prm_all + (size_t)s * prm_size(NM);
  LaneParams<NM> lp;
  lane_params<NM>(prm, jb, lp);
  double x[3], v[3], a[3], Q[3];
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    x[c] = ix[((size_t)s * NM + jb) * 3 + c];
    v[c] = iv[((size_t)s * NM + jb) * 3 + c];
    if (own) xs0[((size_t)s * NM + jb) * 3 + c] = x[c];
  }
  accel_lane<NM>(x, base, jb, lp, sc.eps2, a);
  uint32_t st = 0;
  for (int k = 0; k < C; ++k) {
    double *r = rec + ((size_t)s * C + k) * REC;
    // pair checks of encounter_groups on the step-start positions (the
    // separations jump 1 leaves, but for the rounding of its common
    // translation): the separation of a pair is formed as (higher index) -
    // (lower index) in both of its lanes; overlaps are counted by the lower
    // index's lane
    uint32_t adj = 0u;
    unsigned int ov = 0;
#pragma unroll
    for (int q = 0; q < NM; ++q) {
      const double xq = __shfl(x[0], base + q, 64);
      const double yq = __shfl(x[1], base + q, 64);
      const double zq = __shfl(x[2], base + q, 64);
      const bool hi = q > jb;
      const double dx = hi ? xq - x[0] : x[0] - xq;
      const double dy = hi ? yq - x[1] : x[1] - yq;
      const double dz = hi ? zq - x[2] : x[2] - zq;
      const double r2 = dot3(dx, dy, dz, dx, dy, dz);
      if (hi && r2 < lp.sr2[q]) ++ov;
      if (q != jb && r2 < lp.prc2[q]) adj |= 1u << q;
    }
    if (!own) {
      adj = 0u;
      ov = 0;
    }
    if (ov) {
      st |= ST_OVERLAP;
      if (stats) atomicAdd(stats + S_OVERLAP, (unsigned long long)ov);
    }
#pragma unroll
    for (int c = 0; c < 3; ++c) v[c] = fma(sc.hh, a[c], v[c]);
    momentum_lane<NM>(v, base, lp, M, Q);
#pragma unroll
    for (int c = 0; c < 3; ++c) x[c] = fma(sc.hh, Q[c], x[c]);
    if (own) {
#pragma unroll
      for (int c = 0; c < 3; ++c) {
        r[R_Y(NM) + 3 * jb + c] = x[c];
        r[R_W(NM) + 3 * jb + c] = v[c];
      }
      if (jb == 0) {
        r[R_Q1(NM) + 0] = Q[0];
        r[R_Q1(NM) + 1] = Q[1];
        r[R_Q1(NM) + 2] = Q[2];
      }
    }
    bool ingrp = false;
    if (__builtin_amdgcn_ballot_w64(adj != 0u) != 0ull) {  // some system of this wave has a close pair
      const unsigned long long gb = __builtin_amdgcn_ballot_w64(adj != 0u);
      const bool sys_enc = ((gb >> base) & ((L == 64 ? ~0ull : (1ull << L) - 1ull))) != 0ull;
      if (own) {
#pragma unroll
        for (int c = 0; c < 3; ++c) {
          sh_state[gs][jb][c] = x[c];
          sh_state[gs][jb][3 + c] = v[c];
        }
        sh_adj[gs][jb] = adj;
      }
      __syncthreads();
      if (own && jb == 0) {
        uint32_t in = 0u;
        if (sys_enc)
          st |= massive_groups<NM>(&sh_state[gs][0][0], &sh_adj[gs][0], prm, sc, stats, &in, &sh_ws[gs][0]);
        sh_grp[gs] = in;
      }
      __syncthreads();
      if (own) {
#pragma unroll
        for (int c = 0; c < 3; ++c) {
          x[c] = sh_state[gs][jb][c];
          v[c] = sh_state[gs][jb][3 + c];
        }
        ingrp = ((sh_grp[gs] >> jb) & 1u) != 0u;
      }
      __syncthreads();
    }
    if (!ingrp) {
      bool ok;
      kepler_drift_series(x[0], x[1], x[2], v[0], v[1], v[2], sc.mu, sc.h, ok);
      if (!ok) st |= ST_NEWTON;
    }
    momentum_lane<NM>(v, base, lp, M, Q);
#pragma unroll
    for (int c = 0; c < 3; ++c) x[c] = fma(sc.hh, Q[c], x[c]);
    if (own && jb == 0) {
      r[R_Q2(NM) + 0] = Q[0];
      r[R_Q2(NM) + 1] = Q[1];
      r[R_Q2(NM) + 2] = Q[2];
    }
    accel_lane<NM>(x, base, jb, lp, sc.eps2, a);
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      v[c] = fma(sc.hh, a[c], v[c]);
      if (own) r[R_XE(NM) + 3 * jb + c] = x[c];
    }
  }
  if (!own) return;
  bool fin = true;
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    if (mx) {
      mx[((size_t)s * NM + jb) * 3 + c] = x[c];
      mv[((size_t)s * NM + jb) * 3 + c] = v[c];
    }
    if (nx) {
      nx[((size_t)s * NM + jb) * 3 + c] = x[c];
      nv[((size_t)s * NM + jb) * 3 + c] = v[c];
    }
    fin = fin && isfinite(x[c]) && isfinite(v[c]);
  }
  if (!fin) st |= ST_NAN;
  if (st) atomicOr(&status[s], st);
}

// ---------------------------------------------------------------------------
// Perturber cache: for each system of a list, each step of the window, each
// substep point and each massive body, a Kepler drift of the body's
// drift-start state (Y, W) from the ephemeris.  The lists hold the systems
// whose TPs are in the encounter queue (built with the queue, each system
// once), so the grid is fixed and small and its work scales with the flagged
// systems only: beside the main kernel a sparse grid over every system would
// wait for dispatch slots block by block.
template <int NM>
__global__ void __launch_bounds__(256) pcache_kernel(const int *__restrict__ list,
                                                     const unsigned int *__restrict__ count, int C,
                                                     const double *__restrict__ rec,
                                                     double *__restrict__ pc, StepConst sc) {
  constexpr int REC = rec_size(NM);
  const long long n = (long long)(*count) * C * PC_PTS * NM;
  const long long stride = (long long)gridDim.x * blockDim.x;
  for (long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x; t < n; t += stride) {
    const int j = (int)(t % NM);
    long long u = t / NM;
    const int m = (int)(u % PC_PTS) + 1;
    u /= PC_PTS;
    const int k = (int)(u % C);
    const int s = list[u / C];
    const double *r = rec + ((size_t)s * C + k) * REC;
    double px = r[R_Y(NM) + 3 * j], py = r[R_Y(NM) + 3 * j + 1], pz = r[R_Y(NM) + 3 * j + 2];
    double qx = r[R_W(NM) + 3 * j], qy = r[R_W(NM) + 3 * j + 1], qz = r[R_W(NM) + 3 * j + 2];
    const double tau = (double)m * (sc.h / (double)PC_PTS);
    bool ok;
    kepler_drift_series(px, py, pz, qx, qy, qz, sc.mu, tau, ok);
    double *o = pc + (((size_t)s * C + k) * PC_PTS + (m - 1)) * 3 * NM + 3 * j;
    o[0] = px;
    o[1] = py;
    o[2] = pz;
  }
}

// Flag system s for the perturber cache (need[s]: 0 none, 1 queue head,
// 2 mid-window deferral) and list it the first time it is flagged.
__device__ __forceinline__ void flag_system(unsigned int *need, unsigned int v, int s, int *list,
                                           unsigned int *count) {
  if (need[s] == 0u && atomicCAS(&need[s], 0u, v) == 0u) list[atomicAdd(count, 1u)] = s;
}

// ---------------------------------------------------------------------------
// One full step of one TP (plain / flat forms).  The ephemeris buffer holds
// one record per system (C = 1).
template <int NM, int GUESS>
__global__ void __launch_bounds__(256) tp_step_kernel(long long i0, long long i1,
                                                      const int *__restrict__ tp_sys,
                                                      TPArrays tp,
                                                      const double *__restrict__ prm_all,
                                                      const double *__restrict__ xs0,
                                                      const double *__restrict__ rec,
                                                      StepConst sc,
                                                      const float *__restrict__ nnw,
                                                      unsigned int *__restrict__ status,
                                                      unsigned long long *__restrict__ stats) {
  constexpr int REC = rec_size(NM);
  const long long i = i0 + (long long)blockIdx.x * blockDim.x + threadIdx.x;
  unsigned int n_kep = 0, n_newton = 0, n_bs = 0, n_rhs = 0, n_step = 0;
  if (i < i1) {
    const int s = tp_sys[i];
    const double *prm = prm_all + (size_t)s * prm_size(NM);
    const double *r = rec + (size_t)s * REC;
    double x = tp.x[i], y = tp.y[i], z = tp.z[i];
    double vx = tp.vx[i], vy = tp.vy[i], vz = tp.vz[i];
    uint32_t st = 0;
    double ax, ay, az;
    accel_tp<NM>(x, y, z, xs0 + (size_t)s * 3 * NM, prm, sc.eps2, ax, ay, az);
    vx = fma(sc.hh, ax, vx);
    vy = fma(sc.hh, ay, vy);
    vz = fma(sc.hh, az, vz);
    x = fma(sc.hh, r[R_Q1(NM) + 0], x);
    y = fma(sc.hh, r[R_Q1(NM) + 1], y);
    z = fma(sc.hh, r[R_Q1(NM) + 2], z);
    if (encounter_tp<NM>(x, y, z, r + R_Y(NM), prm)) {
      double yv[6] = {x, y, z, vx, vy, vz};
      double Rt[KMAX * 6];
      int nr = 0;
      st |= bs_drift_tp<NM>(yv, sc.h, r + R_Y(NM), r + R_W(NM), nullptr, prm, sc.mu, sc.eps2, sc.tol, nr,
                            Rt, 1);
      x = yv[0]; y = yv[1]; z = yv[2]; vx = yv[3]; vy = yv[4]; vz = yv[5];
      n_bs = 1;
      n_rhs = (unsigned)nr;
    } else {
      n_newton = (unsigned)tp_kepler<GUESS>(x, y, z, vx, vy, vz, sc, nnw, st);
      n_kep = 1;
    }
    x = fma(sc.hh, r[R_Q2(NM) + 0], x);
    y = fma(sc.hh, r[R_Q2(NM) + 1], y);
    z = fma(sc.hh, r[R_Q2(NM) + 2], z);
    accel_tp<NM>(x, y, z, r + R_XE(NM), prm, sc.eps2, ax, ay, az);
    vx = fma(sc.hh, ax, vx);
    vy = fma(sc.hh, ay, vy);
    vz = fma(sc.hh, az, vz);
    tp.x[i] = x; tp.y[i] = y; tp.z[i] = z;
    tp.vx[i] = vx; tp.vy[i] = vy; tp.vz[i] = vz;
    if (!finite6(x, y, z, vx, vy, vz)) st |= ST_NAN;
    if (st) atomicOr(&status[s], st);
    n_step = 1;
  }
  if (stats) {
    wave_add(stats + S_TPSTEPS, n_step);
    wave_add(stats + S_KEPLER, n_kep);
    wave_add(stats + S_NEWTON, n_newton);
    wave_add(stats + S_BSSTEPS, n_bs);
    wave_add(stats + S_BSRHS, n_rhs);
  }
}

// ---------------------------------------------------------------------------
// Scorer features of one (TP, massive body) pair at the start of a launch
// window of duration T (the body's position and velocity at the window
// start, before its first half kick): distance, approach rate and closest approach in units
// of r_c, and Hill-zone size ratios.  Every feature is clamped to a fixed
// range (f0, f3 in [0, 8], f1 in [-4, 4], f2, f4 in [0, 4], f5 in [0, 2]),
// which is what lets the host bound the network's output over a feature box.  Differences are taken in FP64
// (positions are hundreds of AU); the rest is FP32 with the hardware square
// root, reciprocal square root and reciprocal (about one ulp each): the scorer
// only ranks pairs, and the exact distance check decides every step.
// scorer_tp_r: the TP's heliocentric distance in FP32, the same for all its
// pairs; scorer_features_d takes the pair's FP64 differences (body - TP) as
// formed by the caller.
__device__ __forceinline__ float scorer_tp_r(double x, double y, double z) {
  const float xf = (float)x, yf = (float)y, zf = (float)z;
  return __builtin_amdgcn_sqrtf(fmaf(xf, xf, fmaf(yf, yf, zf * zf)));
}

__device__ __forceinline__ void scorer_features_d(double ddx, double ddy, double ddz, double dux, double duy,
                                                  double duz, double X, double Y, double Z, float rt, float rcf,
                                                  double T, float *f) {
  const float dx = (float)ddx, dy = (float)ddy, dz = (float)ddz;
  const float ux = (float)dux, uy = (float)duy, uz = (float)duz;
  const float Tf = (float)T, irc = __builtin_amdgcn_rcpf(rcf);
  const float r2 = fmaf(dx, dx, fmaf(dy, dy, dz * dz));
  const float u2 = fmaf(ux, ux, fmaf(uy, uy, uz * uz));
  const float du = fmaf(dx, ux, fmaf(dy, uy, dz * uz));
  const float r = __builtin_amdgcn_sqrtf(r2);
  float tmin = u2 > 0.0f ? -du * __builtin_amdgcn_rcpf(u2) : 0.0f;
  tmin = fminf(fmaxf(tmin, 0.0f), Tf);
  const float mx = fmaf(ux, tmin, dx), my = fmaf(uy, tmin, dy), mz = fmaf(uz, tmin, dz);
  const float rmin = __builtin_amdgcn_sqrtf(fmaf(mx, mx, fmaf(my, my, mz * mz)));
  const float Xf = (float)X, Yf = (float)Y, Zf = (float)Z;
  const float iR = __builtin_amdgcn_rsqf(fmaf(Xf, Xf, fmaf(Yf, Yf, Zf * Zf)));
  f[0] = fminf(r * irc, 64.0f) * 0.125f;
  f[1] = fmaxf(fminf(((du * __builtin_amdgcn_rcpf(r)) * irc) * Tf, 16.0f), -16.0f) * 0.25f;
  f[2] = fminf((__builtin_amdgcn_sqrtf(u2) * Tf) * irc, 16.0f) * 0.25f;
  f[3] = fminf(rmin * irc, 64.0f) * 0.125f;
  f[4] = fminf(rcf * iR, 1.0f) * 4.0f;
  f[5] = fminf(rt * iR, 8.0f) * 0.25f;
}

__device__ __forceinline__ void scorer_features(double x, double y, double z, double vx, double vy,
                                                double vz, double X, double Y, double Z, double VX,
                                                double VY, double VZ, float rcf, double T,
                                                float *f) {
  scorer_features_d(X - x, Y - y, Z - z, VX - vx, VY - vy, VZ - vz, X, Y, Z, scorer_tp_r(x, y, z), rcf, T, f);
}

// r_c of pair body j in FP32, from r_c^2 (the scorer's unit of length).
__device__ __forceinline__ float scorer_rc(double rc2) { return __builtin_amdgcn_sqrtf((float)rc2); }

// Whether the scorer features of a pair are certain to have both distance
// features above g (f0, f3 > g: distance and straight-line closest approach
// over the window more than 8 g r_c), from the pair's differences (body - TP)
// in position d and velocity u.  Sufficient, without a square root:
// r^2 >= 2 (A^2 + (|u| T)^2) implies r >= A + |u| T, and |d + u t| >= |d| -
// |u| T over the window; A = 8 g r_c and every term carry a margin of 1e-3,
// far more than the FP32 roundings of the features (about 1e-6 of the
// distances involved) can take away.  Formed in FP64 from the body's values
// as they are (uniform operands where a wavefront's TPs share a system), so
// no conversion is spent on a pair that turns out far.  A pair it does not
// rule out is checked on its formed features; the scorer only decides which
// TPs start the window in the queue, so the test changes no result either
// way.  A2 = (8.008 g)^2 r_c^2, uT2 = (1.001 T)^2.
__device__ __forceinline__ bool scorer_far(double dx, double dy, double dz, double ux, double uy, double uz,
                                           double A2, double uT2) {
  const double r2 = dot3(dx, dy, dz, dx, dy, dz);
  const double u2 = dot3(ux, uy, uz, ux, uy, uz);
  return r2 * 0.999 >= 2.0 * fma(u2, uT2, A2);
}


// One TP per lane: registers are capped for four waves per SIMD (128 VGPRs;
// the few that spill are off the step loop's common path), which measured
// faster than three waves with no spills (five waves spill ~50-65 and ran
// 40 % slower).  Two or four TPs per lane keep the compiler's allocation.
template <int NM, int P, int GUESS>
__global__ void __launch_bounds__(512) __attribute__((amdgpu_waves_per_eu(P == 1 ? 4 : 1))) tp_main_kernel(
    const Tile *__restrict__ tiles, int ntiles, int tiles_per_xcd, TPArrays tp,
    const double *__restrict__ prm_all, const double *__restrict__ xs0,
    const double *__restrict__ rec, int C, StepConst sc, const float *__restrict__ nnw,
    const unsigned char *__restrict__ pre, QEntry *__restrict__ q,
    unsigned int *__restrict__ qcount, unsigned int *__restrict__ need, int *__restrict__ slist,
    unsigned char *__restrict__ carry, unsigned int *__restrict__ status,
    unsigned long long *__restrict__ stats, long long *__restrict__ kstamp) {
  constexpr int REC = rec_size(NM);
  // kstamp (launch timing): each wavefront's first lane stamps the
  // constant-rate wall clock at entry and exit into its own pair of slots
  const long long wslot = ((long long)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
  if (kstamp && (threadIdx.x & 63) == 0) kstamp[2 * wslot] = wall_clock64();
  // Blocks are dispatched round-robin over the 8 XCDs; give each XCD a
  // contiguous run of tiles so the tiles of one system share an L2.
  const int b = blockIdx.x;
  const int t_idx = (b & 7) * tiles_per_xcd + (b >> 3);
  if (t_idx >= ntiles) {
    if (kstamp && (threadIdx.x & 63) == 0) kstamp[2 * wslot + 1] = wall_clock64();
    return;
  }
  const Tile tile = tiles[t_idx];
  const int s = tile.sys;
  const double *prm = prm_all + (size_t)s * prm_size(NM);
  const double *xsys = xs0 + (size_t)s * 3 * NM;
  const double *rsys = rec + (size_t)s * C * REC;

  double X[P][3], V[P][3], A[P][3];
  bool act[P];
  int idx[P];  // TPs per context < 2^31
  // Work counters and status bits are kept per wavefront in scalar registers
  // and updated where the lanes have reconverged after each step, from
  // ballots: nothing per lane is carried through the step loop, so nothing
  // has to be saved around the learned guess's call.  Every step the kernel
  // completes is one Kepler solve; the lane-steps it leaves idle are C per
  // TP slot minus those steps; a TP leaves the kernel early only by deferral.
  unsigned int n_step = 0, n_newton_extra = 0, st_wave = 0;
#pragma unroll
  for (int p = 0; p < P; ++p) {
    const int loc = p * (int)blockDim.x + (int)threadIdx.x;
    act[p] = loc < tile.count;
    idx[p] = (int)(tile.start + loc);
    if (act[p]) {
      X[p][0] = tp.x[idx[p]]; X[p][1] = tp.y[idx[p]]; X[p][2] = tp.z[idx[p]];
      V[p][0] = tp.vx[idx[p]]; V[p][1] = tp.vy[idx[p]]; V[p][2] = tp.vz[idx[p]];
    }
  }
  if (pre != nullptr) {
    // TPs at the head of this window's queue (carried or picked by the scorer)
    // are the encounter kernel's for the whole window
#pragma unroll
    for (int p = 0; p < P; ++p)
      if (act[p] && pre[idx[p]]) act[p] = false;
  }
  unsigned int n_start = 0;
  // far[p]: the last kick found every pair beyond its far radius, so the next
  // encounter check cannot fire (P_FAR2) and is skipped
  bool far[P];
#pragma unroll
  for (int p = 0; p < P; ++p) {
    n_start += wave_count(act[p]);
    far[p] = false;
    if (act[p])
      accel_tp<NM>(X[p][0], X[p][1], X[p][2], xsys, prm, sc.eps2, A[p][0], A[p][1], A[p][2], far[p]);
  }

  // The half step in a vector register: beside a jump momentum (a scalar
  // register) it would take a second scalar operand, which one VOP3
  // instruction cannot encode, so each jump would first copy one of them.
  double hh = sc.hh;
  asm volatile("" : "+v"(hh));
  for (int k = 0; k < C; ++k) {
    // The step's jump momenta are loaded up front (uniform, so one round of
    // scalar loads whose latency overlaps the kick; loaded where each value
    // is used, the loads would sit behind the divergent branches and each
    // wait would be exposed).  The check positions are loaded where the
    // check runs, which most steps skip (far): held from the top, they take
    // scalar registers across the step for nothing (1 % slower at 128-step
    // windows).  The second jump momentum and the kick positions are loaded
    // after the Kepler drift: held across it, the momentum costs 1 % and the
    // positions overflow the scalar registers into vector-lane spills (1.3 %
    // slower).
    const double *rg = rsys + (size_t)__builtin_amdgcn_readfirstlane(k) * REC;
    double r[REC];
#pragma unroll
    for (int i = 0; i < REC; ++i) r[i] = rg[i];
#pragma unroll
    for (int p = 0; p < P; ++p) {
      unsigned int extra = 0;  // Newton iterations beyond the first
      uint32_t stb = 0;
      if (act[p]) {
        double x = X[p][0], y = X[p][1], z = X[p][2];
        double vx = fma(sc.hh, A[p][0], V[p][0]);
        double vy = fma(sc.hh, A[p][1], V[p][1]);
        double vz = fma(sc.hh, A[p][2], V[p][2]);
        x = fma(hh, r[R_Q1(NM) + 0], x);
        y = fma(hh, r[R_Q1(NM) + 1], y);
        z = fma(hh, r[R_Q1(NM) + 2], z);
        bool enc = false;
        if (!far[p]) enc = encounter_tp<NM>(x, y, z, rg + R_Y(NM), prm);
        if (enc) {
          // after the queue head (qcount[0], final before this kernel), by
          // the count of mid-window deferrals (qcount[4])
          const unsigned int slot = qcount[0] + atomicAdd(qcount + 4, 1u);
          q_put(q + slot, (int)idx[p], s, k, 1, x, y, z, vx, vy, vz);
          flag_system(need, 2u, s, slist, qcount + 3);
          if (carry != nullptr) carry[idx[p]] = 1;  // queued for the next window too
          act[p] = false;
        } else {
          extra = (unsigned)tp_kepler<GUESS>(x, y, z, vx, vy, vz, sc, nnw, stb) - 1u;
          double xe[3 * NM];
#pragma unroll
          for (int i = 0; i < 3 * NM; ++i) xe[i] = rg[R_XE(NM) + i];
          x = fma(hh, rg[R_Q2(NM) + 0], x);
          y = fma(hh, rg[R_Q2(NM) + 1], y);
          z = fma(hh, rg[R_Q2(NM) + 2], z);
          double ax, ay, az;
          accel_tp<NM>(x, y, z, xe, prm, sc.eps2, ax, ay, az, far[p]);
          A[p][0] = ax; A[p][1] = ay; A[p][2] = az;
          X[p][0] = x; X[p][1] = y; X[p][2] = z;
          V[p][0] = fma(sc.hh, ax, vx);
          V[p][1] = fma(sc.hh, ay, vy);
          V[p][2] = fma(sc.hh, az, vz);
        }
      }
      n_step += wave_count(act[p]);
      n_newton_extra += wave_sum_rare(extra);
      if (wave_count(stb != 0) != 0) st_wave |= ST_NEWTON;
    }
  }
  uint32_t st = 0;
  unsigned int n_end = 0;
#pragma unroll
  for (int p = 0; p < P; ++p) {
    n_end += wave_count(act[p]);
    if (!act[p]) continue;
    tp.x[idx[p]] = X[p][0]; tp.y[idx[p]] = X[p][1]; tp.z[idx[p]] = X[p][2];
    tp.vx[idx[p]] = V[p][0]; tp.vy[idx[p]] = V[p][1]; tp.vz[idx[p]] = V[p][2];
    if (!finite6(X[p][0], X[p][1], X[p][2], V[p][0], V[p][1], V[p][2])) st |= ST_NAN;
  }
  if ((threadIdx.x & 63) == 0) st |= st_wave;
  if (st) atomicOr(&status[s], st);
  if (stats) {
    unsigned int slots = 0;
#pragma unroll
    for (int p = 0; p < P; ++p) slots += (p * (int)blockDim.x + (int)threadIdx.x < tile.count) ? 1u : 0u;
    wave_add_uniform(stats + S_TPSTEPS, n_step);
    wave_add_uniform(stats + S_KEPLER, n_step);
    wave_add_uniform(stats + S_NEWTON, n_step + n_newton_extra);
    wave_add_uniform(stats + S_DEFER, n_start - n_end);
    wave_add_uniform(stats + S_IDLE, wave_sum(slots) * (unsigned)C - n_step);
  }
  if (kstamp && (threadIdx.x & 63) == 0) kstamp[2 * wslot + 1] = wall_clock64();
}

// ---------------------------------------------------------------------------
// Queue head of a launch window: one lane per TP.  A TP carried from the
// previous window (deferred there mid-window) starts this window in the
// encounter queue; so does, with scorer weights, a TP whose largest pair
// logit exceeds the threshold (learned pre-bucketing).  Queued TPs are
// skipped by the main kernel.  The exact distance check still decides every
// step inside the queue, so results do not change.  Carry flags are cleared
// here; a carried TP's state is not read (its previous window may still be
// finishing), the queue entry reads it when served.
//
// The network runs over a packed list: each lane forms the features of its
// NM pairs, the wavefront gathers the pairs to be scored into LDS, and its 64
// lanes score the list, one pair each per pass.  A pair enters the list only
// if its straight-line closest approach over the window comes within 4 r_c
// (SC_REACH), and not if both distance features are at or above `cut`
// (distance and closest approach at least 8 cut r_c): the host sets cut
// (+inf: none left out) only where an interval bound of the network's logit
// over every such pair lies below the threshold, so that exclusion changes no
// pick.  A sufficient FP64 test at the smaller of the two radii rules most
// pairs out before their features are formed; on the 4096-system sweep about
// one pair in a thousand TPs is then scored.
constexpr int PB_BLOCK = 256;

// Orders a wavefront's LDS writes before its later LDS reads (the lanes of
// one wavefront only: no workgroup barrier).
__device__ __forceinline__ void wave_lds_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

template <int NM>
__global__ void __launch_bounds__(PB_BLOCK) tp_prebucket_kernel(
    long long n, const int *__restrict__ tp_sys, TPArrays tp, const double *__restrict__ prm_all,
    const double *__restrict__ wx, const double *__restrict__ wv, int C, StepConst sc,
    const float *__restrict__ scw, float thresh, float cut, unsigned char *__restrict__ carry,
    unsigned char *__restrict__ pre, QEntry *__restrict__ q, unsigned int *__restrict__ qcount,
    unsigned int *__restrict__ need, int *__restrict__ slist, unsigned long long *__restrict__ stats) {
  // Each wavefront packs and scores its own pairs in its own 64 rows: 6.6 KB
  // of LDS per block (a pass of 64 entries per wavefront, each entry's owner
  // lane, a pick flag per lane) and no workgroup barrier
  __shared__ float sh_f[PB_BLOCK][SC_IN];
  __shared__ unsigned char sh_item[PB_BLOCK];
  __shared__ unsigned char sh_go[PB_BLOCK];
  const int t = (int)threadIdx.x;
  const long long i = (long long)blockIdx.x * PB_BLOCK + t;
  unsigned int n_pre = 0;
  // Every per-lane load is issued here, at once (one round trip instead of a
  // chain of carry flag, system index and state).  The state of a carried TP
  // is loaded too but not used: its previous window may still be writing it.
  const long long ic = i < n ? i : n - 1;
  const bool carried = carry[ic] != 0;
  int s = tp_sys[ic];
  double x = 0.0, y = 0.0, z = 0.0, vx = 0.0, vy = 0.0, vz = 0.0;
  if (scw != nullptr) {
    x = tp.x[ic]; y = tp.y[ic]; z = tp.z[ic];
    vx = tp.vx[ic]; vy = tp.vy[ic]; vz = tp.vz[ic];
  }
  bool go = i < n && carried;
  if (go) carry[i] = 0;
  if (scw != nullptr) {  // uniform: every lane takes part in its wavefront's phases
    sh_go[t] = 0;
    float f[NM][SC_IN];
    unsigned int want = 0u;
    const bool mine = i < n && !go;
    // the pairs of this lane's TP against system sys's bodies
    auto pairs = [&](int sys) {
      const double *prm = prm_all + (size_t)sys * prm_size(NM);
      const double *bx = wx + (size_t)sys * 3 * NM, *bv = wv + (size_t)sys * 3 * NM;
      const double T = sc.h * (double)C;
      const double uT2 = (1.001 * 1.001) * (T * T);
      const double g = (double)fminf(cut, SC_REACH);
      const double a2 = (8.008 * 8.008) * (g * g);
      const float rt = scorer_tp_r(x, y, z);
#pragma unroll
      for (int j = 0; j < NM; ++j) {
        const double rc2 = prm[P_RC2(NM) + j];
        const double X = bx[3 * j], Y = bx[3 * j + 1], Z = bx[3 * j + 2];
        const double dx = X - x, dy = Y - y, dz = Z - z;
        const double ux = bv[3 * j] - vx, uy = bv[3 * j + 1] - vy, uz = bv[3 * j + 2] - vz;
        if (scorer_far(dx, dy, dz, ux, uy, uz, a2 * rc2, uT2)) continue;
        scorer_features_d(dx, dy, dz, ux, uy, uz, X, Y, Z, rt, scorer_rc(rc2), T, f[j]);
        if (f[j][3] <= SC_REACH && !(f[j][0] >= cut && f[j][3] >= cut)) want |= 1u << j;
      }
    };
    // TPs are stored by system, so a wavefront's lanes nearly always share
    // one: then the system's bodies come through scalar loads, once
    const unsigned long long mm = __ballot(mine);
    const int lead = mm ? __ffsll((long long)mm) - 1 : 0;
    const int s0 = __builtin_amdgcn_readlane(s, lead);
    const bool one = mm != 0ull && __ballot(mine && s != s0) == 0ull;
    if (one) {
      if (mine) pairs(s0);
    } else if (mine) {
      pairs(s);
    }
    // slots in the wavefront's list, body by body: pair (lane, j) goes after
    // the pairs of bodies < j and of lower lanes with body j (ballots)
    const unsigned int lane = (unsigned int)t & 63u, wb = (unsigned int)t & ~63u;
    unsigned int slot[NM], wtot = 0u;
#pragma unroll
    for (int j = 0; j < NM; ++j) {
      const unsigned long long bj = __ballot((want >> j) & 1u);
      slot[j] = wtot + __builtin_amdgcn_mbcnt_hi((unsigned int)(bj >> 32),
                                                 __builtin_amdgcn_mbcnt_lo((unsigned int)bj, 0u));
      wtot += (unsigned int)__popcll(bj);
    }
    // passes over the list, 64 entries each (nearly always none or one): a
    // lane writes the entries of its pairs that fall in the pass, lane e
    // scores entry e, and a pair above the threshold flags its owner (a pair
    // above it is what the largest logit above it means; a NaN logit flags
    // nothing either way)
    for (unsigned int p0 = 0; p0 < wtot; p0 += 64u) {
#pragma unroll
      for (int j = 0; j < NM; ++j) {
        if (((want >> j) & 1u) && slot[j] >= p0 && slot[j] < p0 + 64u) {
#pragma unroll
          for (int k = 0; k < SC_IN; ++k) sh_f[wb + slot[j] - p0][k] = f[j][k];
          sh_item[wb + slot[j] - p0] = (unsigned char)lane;
        }
      }
      wave_lds_sync();
      if (lane < wtot - p0) {
        float g[SC_IN];
#pragma unroll
        for (int k = 0; k < SC_IN; ++k) g[k] = sh_f[t][k];
        // scorer_fwd's sums in its order, one hidden unit's 8 weights loaded
        // at a time (all 129 held at once would spill from scalar registers)
        float out = scw[SC_B2];
#pragma unroll 2
        for (int u = 0; u < SC_H; ++u) {
          float a = scw[SC_B1 + u];
#pragma unroll
          for (int k = 0; k < SC_IN; ++k) a = fmaf(scw[SC_W1 + u * SC_IN + k], g[k], a);
          out = fmaf(scw[SC_W2 + u], nn_tanh(a), out);
        }
        if (out > thresh) sh_go[wb + sh_item[t]] = 1;
      }
      wave_lds_sync();
    }
    if (mine) {
      go = sh_go[t] != 0;
      n_pre = go ? 1u : 0u;
    }
  }
  if (i < n) {
    pre[i] = go ? 1 : 0;
    if (go) {
      const unsigned int qs = atomicAdd(qcount, 1u);
      QEntry *e = q + qs;
      e->tp = (int)i;
      e->sys = s;
      e->k0 = 0;
      e->phase = 0;
      flag_system(need, 1u, s, slist, qcount + 1);
    }
  }
  if (stats) wave_add(stats + S_PREBUCKET, n_pre);
}

// ---------------------------------------------------------------------------
// BS queue consumer.  The queue is short (tens to thousands of entries) and
// each entry is a sequential chain of up to C steps.  G lanes serve one entry
// (G = 1, or, for the shortest chain, G = 4 GP: each BS right-hand side split
// over GP perturber lanes and four extrapolation rows run side by side); wave
// w takes entries w*L .. w*L+L-1, one per group of G lanes.  Spread, the entries go over every wave of the grid
// (L = ceil(count / waves)); dense, beside a long main kernel, they are
// packed 64 / G to a wave, so that few SIMDs give up register space to this
// kernel (322 registers a lane) and the main kernel keeps its occupancy.
// Each entry is carried from its step k0 to the end of the window, with the
// exact check deciding Kepler or BS at every step.  The lanes of a group
// compute the same values; the group's first lane writes them.
template <int NM, int GUESS, int G>
__global__ void __launch_bounds__(64) tp_encounter_kernel(
    const QEntry *__restrict__ q, const unsigned int *__restrict__ qcount,
    const unsigned int *__restrict__ qbegin, int dense, TPArrays tp,
    const double *__restrict__ prm_all, const double *__restrict__ xs0,
    const double *__restrict__ rec, const double *__restrict__ pc, int C, StepConst sc,
    const float *__restrict__ nnw, unsigned int *__restrict__ status,
    unsigned long long *__restrict__ stats) {
  constexpr int REC = rec_size(NM);
  // BS extrapolation tableau of each lane, lane-interleaved in LDS (one-wave blocks)
  __shared__ double Rsh[KMAX * 6 * 64];
  double *Rt = Rsh + (threadIdx.x & 63u);
  // G > 1: each group stages a step's ephemeris record and perturber cache
  // in LDS, two steps deep (the next step's is loaded while this one runs)
  constexpr int SST = G > 1 ? REC + pc_size(NM) : 1;
  constexpr int NPF = G > 1 ? (SST + G - 1) / G : 1;
  __shared__ double Sst[G > 1 ? 2 * (64 / G) * SST : 1];
  // G > 1: and the entry's system parameter block, once per entry
  constexpr int SPR = G > 1 ? prm_size(NM) : 1;
  __shared__ double Spr[G > 1 ? (64 / G) * SPR : 1];
  // entries [0, *qcount) (qbegin null), or *qcount entries from *qbegin
  const unsigned int e0 = qbegin ? *qbegin : 0u;
  const unsigned int e1 = qbegin ? e0 + *qcount : *qcount;
  const unsigned int n = e1 > e0 ? e1 - e0 : 0u;
  unsigned int n_kep = 0, n_newton = 0, n_bs = 0, n_rhs = 0, n_step = 0;
  const unsigned int nw = (gridDim.x * blockDim.x) >> 6;
  const unsigned int w = (blockIdx.x * blockDim.x + threadIdx.x) >> 6;
  const unsigned int lane = threadIdx.x & 63u;
  constexpr unsigned int GPW = 64u / (unsigned int)G;  // groups per wave
  // G > 1: GP lanes per BS row (one per massive body, rounded up to a power
  // of two) times R row groups that run the extrapolation rows side by side
  constexpr int GP = G == 1 ? 1 : (NM <= 1 ? 1 : (NM <= 2 ? 2 : (NM <= 4 ? 4 : 8)));
  constexpr int R = G / GP;
  static_assert(G == 1 || G == GP * R, "G is GP lanes times R row groups");
  const unsigned int g = lane / (unsigned int)G, el = lane % (unsigned int)G, ebase = lane - el;
  const unsigned int jl = el % (unsigned int)GP, row = el / (unsigned int)GP;
  const unsigned int lead = el == 0u ? 1u : 0u;
  unsigned int L = dense ? GPW : (n + nw - 1u) / nw;
  L = L < 1u ? 1u : (L > GPW ? GPW : L);
  for (unsigned int e = w * L + g; g < L && e < n; e += nw * L) {
    const QEntry en = q[e0 + e];
    const int s = en.sys;
    const double *prm = prm_all + (size_t)s * prm_size(NM);
    if (G > 1) {
      // the parameter block read from LDS by every right-hand side, kick and
      // check of the entry (the group's lanes write and read it in order)
      double *d = Spr + (size_t)g * SPR;
      for (int i = (int)el; i < SPR; i += G) d[i] = prm[i];
      __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
      __builtin_amdgcn_wave_barrier();
      __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
      prm = d;
    }
    const double *rsys = rec + (size_t)s * C * REC;
    double x, y, z, vx, vy, vz;
    if (en.phase == 0) {  // queued from the window start: the state is in the TP arrays
      x = tp.x[en.tp]; y = tp.y[en.tp]; z = tp.z[en.tp];
      vx = tp.vx[en.tp]; vy = tp.vy[en.tp]; vz = tp.vz[en.tp];
    } else {
      x = en.s[0]; y = en.s[1]; z = en.s[2]; vx = en.s[3]; vy = en.s[4]; vz = en.s[5];
    }
    double ax, ay, az;
    uint32_t st = 0;
    int k = en.k0;
    bool flag = true;
    if (en.phase == 0) {
      const double *xs = (k == 0) ? xs0 + (size_t)s * 3 * NM : rsys + (size_t)(k - 1) * REC + R_XE(NM);
      accel_tp<NM>(x, y, z, xs, prm, sc.eps2, ax, ay, az);
      const double *r = rsys + (size_t)k * REC;
      vx = fma(sc.hh, ax, vx);
      vy = fma(sc.hh, ay, vy);
      vz = fma(sc.hh, az, vz);
      x = fma(sc.hh, r[R_Q1(NM) + 0], x);
      y = fma(sc.hh, r[R_Q1(NM) + 1], y);
      z = fma(sc.hh, r[R_Q1(NM) + 2], z);
      flag = encounter_tp<NM>(x, y, z, r + R_Y(NM), prm);
    }
    // step data of step kk into registers: this lane's share (elements el,
    // el + G, ...) of the record then the perturber cache
    double pf[NPF];
    double *Sg = Sst + (size_t)g * 2 * SST;
    auto fetch = [&](int kk) {
      const double *ra = rsys + (size_t)kk * REC;
      const double *pa = pc + ((size_t)s * C + kk) * pc_size(NM);
#pragma unroll
      for (int u = 0; u < NPF; ++u) {
        const int i = (int)el + u * G;
        pf[u] = i < REC ? ra[i] : (i < SST ? pa[i - REC] : 0.0);
      }
    };
    auto stage = [&](int kk) {
      double *d = Sg + (kk & 1) * SST;
#pragma unroll
      for (int u = 0; u < NPF; ++u) {
        const int i = (int)el + u * G;
        if (i < SST) d[i] = pf[u];
      }
      // the group's lanes read what the others wrote: LDS operations of one
      // wavefront complete in order, and nothing is moved across this point
      __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
      __builtin_amdgcn_wave_barrier();
      __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    };
    if (G > 1) {
      fetch(k);
      stage(k);
      if (k + 1 < C) fetch(k + 1);
    }
    for (;;) {
      const double *r = G > 1 ? Sg + (k & 1) * SST : rsys + (size_t)k * REC;
      const double *pk = G > 1 ? r + REC : pc + ((size_t)s * C + k) * pc_size(NM);
      if (flag) {
        double yv[6] = {x, y, z, vx, vy, vz};
        int nr = 0;
        if (G == 1)
          st |= bs_drift_tp<NM, 1>(yv, sc.h, r + R_Y(NM), r + R_W(NM), pk, prm, sc.mu, sc.eps2, sc.tol,
                                   nr, Rt, 64);
        else
          st |= bs_drift_tp_rows<NM, GP, R>(yv, sc.h, r + R_Y(NM), r + R_W(NM), pk, prm, sc.mu, sc.eps2,
                                            sc.tol, nr, Rt, 64, (int)jl, (int)row, (int)ebase);
        x = yv[0]; y = yv[1]; z = yv[2]; vx = yv[3]; vy = yv[4]; vz = yv[5];
        n_bs += lead;
        n_rhs += lead * (unsigned)nr;
      } else {
        n_newton += lead * (unsigned)tp_kepler<GUESS>(x, y, z, vx, vy, vz, sc, nnw, st);
        n_kep += lead;
      }
      x = fma(sc.hh, r[R_Q2(NM) + 0], x);
      y = fma(sc.hh, r[R_Q2(NM) + 1], y);
      z = fma(sc.hh, r[R_Q2(NM) + 2], z);
      accel_tp<NM>(x, y, z, r + R_XE(NM), prm, sc.eps2, ax, ay, az);
      vx = fma(sc.hh, ax, vx);
      vy = fma(sc.hh, ay, vy);
      vz = fma(sc.hh, az, vz);
      n_step += lead;
      if (++k >= C) break;
      if (G > 1) {
        stage(k);
        if (k + 1 < C) fetch(k + 1);
      }
      const double *rn = G > 1 ? Sg + (k & 1) * SST : rsys + (size_t)k * REC;
      vx = fma(sc.hh, ax, vx);
      vy = fma(sc.hh, ay, vy);
      vz = fma(sc.hh, az, vz);
      x = fma(sc.hh, rn[R_Q1(NM) + 0], x);
      y = fma(sc.hh, rn[R_Q1(NM) + 1], y);
      z = fma(sc.hh, rn[R_Q1(NM) + 2], z);
      flag = encounter_tp<NM>(x, y, z, rn + R_Y(NM), prm);
    }
    if (lead) {
      const long long i = en.tp;
      tp.x[i] = x; tp.y[i] = y; tp.z[i] = z;
      tp.vx[i] = vx; tp.vy[i] = vy; tp.vz[i] = vz;
      if (!finite6(x, y, z, vx, vy, vz)) st |= ST_NAN;
      if (st) atomicOr(&status[s], st);
    }
  }
  if (stats) {
    wave_add(stats + S_TPSTEPS, n_step);
    wave_add(stats + S_KEPLER, n_kep);
    wave_add(stats + S_NEWTON, n_newton);
    wave_add(stats + S_BSSTEPS, n_bs);
    wave_add(stats + S_BSRHS, n_rhs);
    wave_add(stats + S_ENCSTEPS, n_step);
  }
}


// Clears a queue's counters ctl[8] and system flags need[S] (at the start of
// every launch window but a call's first) and, with n > 0, copies n doubles
// of massive state from (sx, sv) to (dx, dv) (at a call's end: the last
// window's end state into the resident buffers, and parity 0's queue cleared
// for the next call's first window).
static __global__ void window_start_kernel(unsigned int *__restrict__ ctl, unsigned int *__restrict__ need, int S,
                                           double *__restrict__ dx, double *__restrict__ dv,
                                           const double *__restrict__ sx, const double *__restrict__ sv,
                                           long long n) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < 8) ctl[i] = 0u;
  if (i < S) need[i] = 0u;
  if (i < n) {
    dx[i] = sx[i];
    dv[i] = sv[i];
  }
}

}  // namespace mint
