// mint_nn_api.hip — learned parts: trace logging, FP32 forward / backward /
// Adam kernels, and the C ABI around them.  Nothing here touches the FP64
// step arithmetic: the Kepler MLP only moves the Newton starting point and the
// scorer only decides which TPs start a launch window in the encounter queue.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#include "mint_api_internal.h"

namespace mint {

int step_impl(Ctx *c, int64_t nsteps);

namespace {

// ---------------------------------------------------------------------------
// Trace kernels

// Kepler trace: for every `stride`-th body (TPs, then massive bodies), the
// features of a drift by dt from its current state and the target
// c = (s* / s_series - 1) / tau^3, s* the converged universal variable.
__global__ void kepler_trace_kernel(const double *__restrict__ x, const double *__restrict__ y,
                                    const double *__restrict__ z, const double *__restrict__ vx,
                                    const double *__restrict__ vy, const double *__restrict__ vz,
                                    long long n, long long stride, double mu, double dt,
                                    float *__restrict__ feat, float *__restrict__ target,
                                    unsigned int *__restrict__ counter, unsigned int cap) {
  const long long j = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  const long long i = j * stride;
  if (i >= n) return;
  const KepStart k = kep_start(x[i], y[i], z[i], vx[i], vy[i], vz[i], mu);
  if (!(k.beta > 0.0)) return;  // bound orbits only
  const double s0 = guess_series(k, dt);
  const KepSol o = kepler_solve(k, s0, dt);
  if (!o.ok) return;
  const double sstar = o.s - o.ds;
  float f[4];
  double tau;
  kepler_features(k, mu, dt, f, tau);
  const double c = (sstar / s0 - 1.0) / (tau * (tau * tau));
  if (!isfinite(c)) return;
  const unsigned int slot = atomicAdd(counter, 1u);
  if (slot >= cap) return;
  for (int q = 0; q < 4; ++q) feat[(size_t)slot * 4 + q] = f[q];
  target[slot] = (float)c;
}

// Encounter trace: for every TP, scorer features of each (TP, body) pair at
// the window start and label = 1 if the exact check fires for that pair at any
// step of the window along the TP's Kepler-drift path; all positive pairs and a
// hashed one-in-64 of the negatives are written.
template <int NM>
__global__ void encounter_trace_kernel(TPArrays tp, const int *__restrict__ tp_sys, long long n,
                                       long long stride, const double *__restrict__ prm_all,
                                       const double *__restrict__ xs0, const double *__restrict__ vs0,
                                       const double *__restrict__ rec,
                                       int C, StepConst sc, float *__restrict__ feat,
                                       float *__restrict__ label, unsigned int *__restrict__ counter,
                                       unsigned int cap, unsigned int win) {
  constexpr int REC = rec_size(NM);
  const long long j = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  const long long i = j * stride;
  if (i >= n) return;
  const int s = tp_sys[i];
  const double *prm = prm_all + (size_t)s * prm_size(NM);
  const double *xsys = xs0 + (size_t)s * 3 * NM;
  const double *rsys = rec + (size_t)s * C * REC;
  double x = tp.x[i], y = tp.y[i], z = tp.z[i], vx = tp.vx[i], vy = tp.vy[i], vz = tp.vz[i];
  float f[NM][SC_IN];
  const double T = sc.h * (double)C;
  const double *vsys = vs0 + (size_t)s * 3 * NM;
  for (int q = 0; q < NM; ++q)
    scorer_features(x, y, z, vx, vy, vz, xsys[3 * q], xsys[3 * q + 1], xsys[3 * q + 2], vsys[3 * q],
                    vsys[3 * q + 1], vsys[3 * q + 2], scorer_rc(prm[P_RC2(NM) + q]), T, f[q]);
  unsigned int hit = 0;
  double ax, ay, az;
  accel_tp<NM>(x, y, z, xsys, prm, sc.eps2, ax, ay, az);
  for (int k = 0; k < C; ++k) {
    const double *r = rsys + (size_t)k * REC;
    vx = fma(sc.hh, ax, vx);
    vy = fma(sc.hh, ay, vy);
    vz = fma(sc.hh, az, vz);
    x = fma(sc.hh, r[R_Q1(NM) + 0], x);
    y = fma(sc.hh, r[R_Q1(NM) + 1], y);
    z = fma(sc.hh, r[R_Q1(NM) + 2], z);
    for (int q = 0; q < NM; ++q) {
      const double dx = r[R_Y(NM) + 3 * q] - x, dy = r[R_Y(NM) + 3 * q + 1] - y,
                   dz = r[R_Y(NM) + 3 * q + 2] - z;
      if (dot3(dx, dy, dz, dx, dy, dz) < prm[P_RC2(NM) + q]) hit |= 1u << q;
    }
    bool ok;
    kepler_drift_series(x, y, z, vx, vy, vz, sc.mu, sc.h, ok);
    x = fma(sc.hh, r[R_Q2(NM) + 0], x);
    y = fma(sc.hh, r[R_Q2(NM) + 1], y);
    z = fma(sc.hh, r[R_Q2(NM) + 2], z);
    accel_tp<NM>(x, y, z, r + R_XE(NM), prm, sc.eps2, ax, ay, az);
    vx = fma(sc.hh, ax, vx);
    vy = fma(sc.hh, ay, vy);
    vz = fma(sc.hh, az, vz);
  }
  for (int q = 0; q < NM; ++q) {
    const bool pos = (hit >> q) & 1u;
    // keep every positive pair and one negative in 64 (hash of TP, body, window)
    unsigned int hsh = (unsigned int)i * 2654435761u ^ (unsigned int)(q + 1) * 40503u ^ win * 2246822519u;
    hsh ^= hsh >> 15;
    hsh *= 2246822519u;
    hsh ^= hsh >> 13;
    if (!pos && (hsh & 63u) != 0u) continue;
    const unsigned int slot = atomicAdd(counter, 1u);
    if (slot >= cap) return;
    for (int u = 0; u < SC_IN; ++u) feat[(size_t)slot * SC_IN + u] = f[q][u];
    label[slot] = pos ? 1.0f : 0.0f;
  }
}

// ---------------------------------------------------------------------------
// Training kernels.  One thread per sample: forward pass with the hidden
// activations kept in registers, then the gradient parameter by parameter:
// each component is summed over the wave with shuffles and added once per
// wave to an LDS accumulator, which is added to the global gradient once per
// block.  Lanes past the end of the batch contribute zero.

__device__ __forceinline__ float wave_sum(float v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  return v;
}

template <int WHICH>
struct Net;

template <>
struct Net<0> {
  static constexpr int NP = KM_NP, IN = KM_IN;
  // out = W3 . tanh(W2 tanh(W1 in + b1) + b2) + b3 ; loss 0.5 (out - t)^2
  __device__ static float grad(const float *__restrict__ w, const float *in, float t, bool live, float *acc) {
    float h1[KM_H], h2[KM_H], d2[KM_H], d1[KM_H];
    for (int i = 0; i < KM_H; ++i) {
      float a = w[KM_B1 + i];
      for (int k = 0; k < KM_IN; ++k) a = fmaf(w[KM_W1 + i * KM_IN + k], in[k], a);
      h1[i] = nn_tanh(a);
    }
    float out = w[KM_B3];
    for (int i = 0; i < KM_H; ++i) {
      float a = w[KM_B2 + i];
      for (int k = 0; k < KM_H; ++k) a = fmaf(w[KM_W2 + i * KM_H + k], h1[k], a);
      h2[i] = nn_tanh(a);
      out = fmaf(w[KM_W3 + i], h2[i], out);
    }
    const float e = live ? out - t : 0.f;
    for (int i = 0; i < KM_H; ++i) d2[i] = e * w[KM_W3 + i] * (1.f - h2[i] * h2[i]);
    for (int k = 0; k < KM_H; ++k) {
      float a = 0.f;
      for (int i = 0; i < KM_H; ++i) a = fmaf(d2[i], w[KM_W2 + i * KM_H + k], a);
      d1[k] = a * (1.f - h1[k] * h1[k]);
    }
    const bool l0 = (threadIdx.x & 63) == 0;
    auto put = [&](int p, float g) {
      g = wave_sum(g);
      if (l0) atomicAdd(&acc[p], g);
    };
    put(KM_B3, e);
    for (int i = 0; i < KM_H; ++i) put(KM_W3 + i, e * h2[i]);
    for (int i = 0; i < KM_H; ++i) {
      put(KM_B2 + i, d2[i]);
      for (int k = 0; k < KM_H; ++k) put(KM_W2 + i * KM_H + k, d2[i] * h1[k]);
    }
    for (int k = 0; k < KM_H; ++k) {
      put(KM_B1 + k, d1[k]);
      for (int q = 0; q < KM_IN; ++q) put(KM_W1 + k * KM_IN + q, d1[k] * in[q]);
    }
    return 0.5f * e * e;
  }
  __device__ static float fwd(const float *__restrict__ w, const float *in) { return mlp_kepler_fwd(w, in); }
};

template <>
struct Net<1> {
  static constexpr int NP = SC_NP, IN = SC_IN;
  // logit = W2 . tanh(W1 in + b1) + b2 ; BCE with logits
  __device__ static float grad(const float *__restrict__ w, const float *in, float t, bool live, float *acc) {
    float h[SC_H];
    float out = w[SC_B2];
    for (int i = 0; i < SC_H; ++i) {
      float a = w[SC_B1 + i];
      for (int k = 0; k < SC_IN; ++k) a = fmaf(w[SC_W1 + i * SC_IN + k], in[k], a);
      h[i] = nn_tanh(a);
      out = fmaf(w[SC_W2 + i], h[i], out);
    }
    const float p = 1.f / (1.f + expf(-out));
    const float loss = live ? fmaxf(out, 0.f) - out * t + log1pf(expf(-fabsf(out))) : 0.f;
    const float e = live ? p - t : 0.f;
    const bool l0 = (threadIdx.x & 63) == 0;
    auto put = [&](int q, float g) {
      g = wave_sum(g);
      if (l0) atomicAdd(&acc[q], g);
    };
    put(SC_B2, e);
    for (int i = 0; i < SC_H; ++i) {
      put(SC_W2 + i, e * h[i]);
      const float d = e * w[SC_W2 + i] * (1.f - h[i] * h[i]);
      put(SC_B1 + i, d);
      for (int k = 0; k < SC_IN; ++k) put(SC_W1 + i * SC_IN + k, d * in[k]);
    }
    return loss;
  }
  __device__ static float fwd(const float *__restrict__ w, const float *in) { return scorer_fwd(w, in); }
};

template <int WHICH>
__global__ void __launch_bounds__(256) nn_grad_kernel(const float *__restrict__ w,
                                                      const float *__restrict__ feat,
                                                      const float *__restrict__ target,
                                                      const int *__restrict__ perm, long long n,
                                                      float *__restrict__ grad,
                                                      float *__restrict__ loss_sum) {
  using N = Net<WHICH>;
  __shared__ float g[N::NP];
  __shared__ float lsum;
  for (int i = threadIdx.x; i < N::NP; i += blockDim.x) g[i] = 0.f;
  if (threadIdx.x == 0) lsum = 0.f;
  __syncthreads();
  const long long j = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  const bool live = j < n;
  float in[N::IN];
  float t = 0.f;
  const long long s = live ? (perm ? perm[j] : j) : 0;
  for (int q = 0; q < N::IN; ++q) in[q] = live ? feat[(size_t)s * N::IN + q] : 0.f;
  if (live) t = target[s];
  const float l = wave_sum(N::grad(w, in, t, live, g));
  if ((threadIdx.x & 63) == 0) atomicAdd(&lsum, l);
  __syncthreads();
  for (int i = threadIdx.x; i < N::NP; i += blockDim.x) atomicAdd(&grad[i], g[i]);
  if (threadIdx.x == 0) atomicAdd(loss_sum, lsum);
}

template <int WHICH>
__global__ void nn_fwd_kernel(const float *__restrict__ w, const float *__restrict__ feat, long long n,
                              float *__restrict__ out) {
  using N = Net<WHICH>;
  const long long j = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= n) return;
  float in[N::IN];
  for (int q = 0; q < N::IN; ++q) in[q] = feat[(size_t)j * N::IN + q];
  out[j] = N::fwd(w, in);
}

// Adam (Kingma & Ba 2015) on the mean gradient grad/nb.
__global__ void adam_kernel(float *__restrict__ w, float *__restrict__ m, float *__restrict__ v,
                            const float *__restrict__ grad, int np, float inv_nb, float lr,
                            float b1, float b2, float eps, float c1, float c2) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= np) return;
  const float g = grad[i] * inv_nb;
  const float mi = b1 * m[i] + (1.f - b1) * g;
  const float vi = b2 * v[i] + (1.f - b2) * g * g;
  m[i] = mi;
  v[i] = vi;
  w[i] -= lr * (mi / c1) / (sqrtf(vi / c2) + eps);
}

__global__ void zero_f32_kernel(float *p, int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = 0.f;
}

// splitmix64 for weight initialisation and shuffles
struct SplitMix {
  unsigned long long s;
  unsigned long long next() {
    unsigned long long z = (s += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
  }
  double uniform() { return (double)(next() >> 11) * (1.0 / 9007199254740992.0); }
};

void init_weights(int which, unsigned long long seed, std::vector<float> &w) {
  SplitMix r{seed};
  auto fill = [&](int off, int fan_out, int fan_in) {
    const double lim = std::sqrt(6.0 / (fan_in + fan_out));
    for (int i = 0; i < fan_out * fan_in; ++i) w[off + i] = (float)((2.0 * r.uniform() - 1.0) * lim);
  };
  if (which == 0) {
    w.assign(KM_NP, 0.f);
    fill(KM_W1, KM_H, KM_IN);
    fill(KM_W2, KM_H, KM_H);
    fill(KM_W3, 1, KM_H);
  } else {
    w.assign(SC_NP, 0.f);
    fill(SC_W1, SC_H, SC_IN);
    fill(SC_W2, 1, SC_H);
  }
}

int nparams(int which) { return which == 0 ? KM_NP : (which == 1 ? SC_NP : -1); }
int nin(int which) { return which == 0 ? KM_IN : SC_IN; }

struct DevBuf {
  void *p = nullptr;
  ~DevBuf() {
    if (p) hipFree(p);
  }
  bool alloc(size_t b) { return hipMalloc(&p, std::max<size_t>(b, 8)) == hipSuccess; }
};

hipError_t launch_grad(int which, const float *w, const float *feat, const float *tgt, const int *perm,
                       long long n, float *grad, float *lsum, hipStream_t st) {
  const unsigned int grid = (unsigned int)((n + 255) / 256);
  if (grid == 0) return hipSuccess;
  if (which == 0)
    nn_grad_kernel<0><<<grid, 256, 0, st>>>(w, feat, tgt, perm, n, grad, lsum);
  else
    nn_grad_kernel<1><<<grid, 256, 0, st>>>(w, feat, tgt, perm, n, grad, lsum);
  return hipGetLastError();
}

std::vector<float> &host_weights(Ctx *c, int which) { return which == 0 ? c->w_kep : c->w_sc; }
float *dev_weights(Ctx *c, int which) { return which == 0 ? c->d_nnw_kep : c->d_nnw_sc; }

// Upper bound of the scorer's logit over a box of its features: interval
// arithmetic through the hidden layer (tanh is increasing), in FP64.
double scorer_box_bound(const std::vector<float> &w, const double *lo, const double *hi) {
  double out = w[SC_B2];
  for (int u = 0; u < SC_H; ++u) {
    double a = w[SC_B1 + u], b = w[SC_B1 + u];
    for (int k = 0; k < SC_IN; ++k) {
      const double wk = w[SC_W1 + u * SC_IN + k];
      a += std::min(wk * lo[k], wk * hi[k]);
      b += std::max(wk * lo[k], wk * hi[k]);
    }
    const double w2 = w[SC_W2 + u];
    out += std::max(w2 * std::tanh(a), w2 * std::tanh(b));
  }
  return out;
}

// The same, tightened by splitting the box along its widest side (at most
// `depth` times on any path) until a part's bound is `target` or below.
double scorer_bound_split(const std::vector<float> &w, double *lo, double *hi, double target, int depth) {
  const double u = scorer_box_bound(w, lo, hi);
  if (u <= target || depth == 0) return u;
  int k = 0;
  for (int q = 1; q < SC_IN; ++q)
    if (hi[q] - lo[q] > hi[k] - lo[k]) k = q;
  const double l = lo[k], h = hi[k], m = 0.5 * (l + h);
  hi[k] = m;
  const double u1 = scorer_bound_split(w, lo, hi, target, depth - 1);
  hi[k] = h;
  lo[k] = m;
  const double u2 = u1 > target ? u1 : scorer_bound_split(w, lo, hi, target, depth - 1);
  lo[k] = l;
  return std::max(u1, u2);
}

}  // namespace

// The scorer's far-pair cutoff for weights w and threshold t: the smallest c
// of a ladder such that every pair with both distance features at or above c
// (distance and straight-line closest approach at least 8 c r_c; the other
// features anywhere in their clamped ranges) has a logit bound below t, with
// a margin far above what the device's FP32 sums and its tanh (about 1e-7
// absolute each) can add; such pairs are left out of the scoring.  *bound
// receives the bound at that c.  No safe c: returns +inf (nothing is left out).
float scorer_cutoff(const std::vector<float> &w, float t, float *bound) {
  static const double ladder[] = {1.5, 2.0, 3.0, 4.0, 6.0, 8.0};
  double wsum = 1.0;
  for (int u = 0; u < SC_H; ++u) wsum += std::fabs(w[SC_W2 + u]);
  const double margin = 1e-3 * wsum;
  *bound = INFINITY;
  for (double c : ladder) {
    double lo[SC_IN] = {c, -4.0, 0.0, c, 0.0, 0.0};
    double hi[SC_IN] = {8.0, 4.0, 4.0, 8.0, 4.0, 2.0};
    const double b = scorer_bound_split(w, lo, hi, (double)t - margin, 12) + margin;
    if (b < (double)t) {
      *bound = (float)b;
      return (float)c;
    }
  }
  return INFINITY;
}

}  // namespace mint

using namespace mint;

extern "C" {

int mint_nn_nparams(int32_t which) { return nparams(which); }

int mint_nn_get_weights(void *ctx, int32_t which, float *w) {
  Ctx *c = (Ctx *)ctx;
  if (!c || !w || nparams(which) < 0) return MINT_EINVAL;
  auto &hw = host_weights(c, which);
  if (hw.empty()) return c->fail(MINT_ESTATE, "mint_nn_get_weights: network has no weights");
  std::memcpy(w, hw.data(), hw.size() * sizeof(float));
  return MINT_OK;
}

int mint_nn_set_weights(void *ctx, int32_t which, const float *w) {
  Ctx *c = (Ctx *)ctx;
  if (!c || !w || nparams(which) < 0) return MINT_EINVAL;
  auto &hw = host_weights(c, which);
  hw.assign(w, w + nparams(which));
  if (which == 1) c->sc_cut = scorer_cutoff(hw, c->score_thresh, &c->sc_sat_bound);
  c->drop_graphs();
  if (c->uploaded && dev_weights(c, which)) {  // (the many-body engine keeps no network buffers)
    hipSetDevice(c->cfg.device);
    return c->hip(hipMemcpy(dev_weights(c, which), hw.data(), hw.size() * sizeof(float), hipMemcpyHostToDevice),
                  "H2D weights");
  }
  return MINT_OK;
}

int mint_set_score_threshold(void *ctx, float t) {
  Ctx *c = (Ctx *)ctx;
  if (!c) return MINT_EINVAL;
  c->score_thresh = t;
  if (!c->w_sc.empty()) c->sc_cut = scorer_cutoff(c->w_sc, t, &c->sc_sat_bound);
  c->drop_graphs();
  return MINT_OK;
}

int mint_scorer_skip(void *ctx, int32_t mode, float *bound, float *cutoff, int32_t *active) {
  Ctx *c = (Ctx *)ctx;
  if (!c) return MINT_EINVAL;
  if (mode != -1 && mode != 0 && mode != 1) return c->fail(MINT_EINVAL, "mint_scorer_skip: mode must be -1, 0 or 1");
  if (mode >= 0 && mode != c->scorer_skip_mode) {
    c->scorer_skip_mode = mode;
    c->drop_graphs();
  }
  if (bound) *bound = c->sc_sat_bound;
  if (cutoff) *cutoff = c->sc_cut;
  if (active) *active = c->scorer_skip_on() ? 1 : 0;
  return MINT_OK;
}

int mint_trace_kepler(void *ctx, int64_t nsteps, int64_t cap, float *feat, float *target, int64_t *n_out) {
  Ctx *c = (Ctx *)ctx;
  if (!c) return MINT_EINVAL;
  if (!c->uploaded) return c->fail(MINT_ESTATE, "mint_trace_kepler: nothing uploaded");
  if (c->many) return c->fail(MINT_EINVAL, "mint_trace_kepler: needs test particles (at most 8 massive bodies)");
  if (cap <= 0 || cap > 0x7fffffff || !feat || !target || !n_out || nsteps < 1)
    return c->fail(MINT_EINVAL, "mint_trace_kepler: bad arguments");
  hipSetDevice(c->cfg.device);
  DevBuf df, dt, dc;
  if (!df.alloc((size_t)cap * 4 * sizeof(float)) || !dt.alloc((size_t)cap * sizeof(float)) || !dc.alloc(4))
    return c->fail(MINT_ENOMEM, "mint_trace_kepler: allocation");
  hipMemset(dc.p, 0, 4);
  // sample points: up to 16 windows spread over nsteps
  const int nwin = (int)std::min<int64_t>(16, nsteps);
  const long long per_win = std::max<long long>(1, cap / nwin);
  const long long nb = c->Ntp + (long long)c->S * c->NM;
  const long long stride = std::max<long long>(1, nb / per_win);
  int64_t done = 0;
  for (int w = 0; w < nwin; ++w) {
    const int64_t upto = nsteps * (w + 1) / nwin;
    if (upto > done) {
      const int rc = step_impl(c, upto - done);
      if (rc) return rc;
      done = upto;
    }
    // TPs
    if (c->Ntp > 0) {
      const long long cnt = (c->Ntp + stride - 1) / stride;
      kepler_trace_kernel<<<(unsigned)((cnt + 255) / 256), 256, 0, c->stream>>>(
          c->d_tp[0], c->d_tp[1], c->d_tp[2], c->d_tp[3], c->d_tp[4], c->d_tp[5], c->Ntp, stride,
          c->cfg.central_mass, c->cfg.dt, (float *)df.p, (float *)dt.p, (unsigned int *)dc.p, (unsigned)cap);
    }
    // massive bodies (SoA views of the xyz-interleaved arrays)
    const long long nm = (long long)c->S * c->NM;
    kepler_trace_kernel<<<(unsigned)((nm + 255) / 256), 256, 0, c->stream>>>(
        c->d_mx, c->d_mx + 1, c->d_mx + 2, c->d_mv, c->d_mv + 1, c->d_mv + 2, 3 * nm, 3,
        c->cfg.central_mass, c->cfg.dt, (float *)df.p, (float *)dt.p, (unsigned int *)dc.p, (unsigned)cap);
    int rc = c->hip(hipGetLastError(), "kepler trace launch");
    if (rc) return rc;
  }
  hipStreamSynchronize(c->stream);
  unsigned int cnt = 0;
  hipMemcpy(&cnt, dc.p, 4, hipMemcpyDeviceToHost);
  const int64_t n = std::min<int64_t>(cnt, cap);
  hipMemcpy(feat, df.p, (size_t)n * 4 * sizeof(float), hipMemcpyDeviceToHost);
  hipMemcpy(target, dt.p, (size_t)n * sizeof(float), hipMemcpyDeviceToHost);
  *n_out = n;
  return c->hip(hipGetLastError(), "kepler trace");
}

int mint_trace_encounter(void *ctx, int64_t nsteps, int64_t cap, float *feat, float *label, int64_t *n_out) {
  Ctx *c = (Ctx *)ctx;
  if (!c) return MINT_EINVAL;
  if (!c->uploaded) return c->fail(MINT_ESTATE, "mint_trace_encounter: nothing uploaded");
  if (c->many) return c->fail(MINT_EINVAL, "mint_trace_encounter: needs test particles (at most 8 massive bodies)");
  if (cap <= 0 || cap > 0x7fffffff || !feat || !label || !n_out || nsteps < 1)
    return c->fail(MINT_EINVAL, "mint_trace_encounter: bad arguments");
  if (c->NM != 5 && c->NM != 2)
    return c->fail(MINT_EINVAL, "mint_trace_encounter: built for 2 or 5 massive bodies");
  hipSetDevice(c->cfg.device);
  DevBuf df, dl, dc;
  if (!df.alloc((size_t)cap * SC_IN * sizeof(float)) || !dl.alloc((size_t)cap * sizeof(float)) || !dc.alloc(4))
    return c->fail(MINT_ENOMEM, "mint_trace_encounter: allocation");
  hipMemset(dc.p, 0, 4);
  int C = c->cfg.chunk_steps > 0 ? c->cfg.chunk_steps : (c->tuned_C > 0 ? c->tuned_C : 64);
  int rc = c->ensure_rec(C);
  if (rc) return rc;
  const int nwin = (int)std::max<int64_t>(1, nsteps / C);
  const long long stride = 1;
  const StepConst sc = c->step_const();
  for (int w = 0; w < nwin; ++w) {
    // fill the ephemeris for the coming window without advancing: run the
    // ephemeris on a copy of the massive state
    DevBuf cmx, cmv;
    const size_t mb = (size_t)c->S * c->NM * 3 * sizeof(double);
    if (!cmx.alloc(mb) || !cmv.alloc(mb)) return c->fail(MINT_ENOMEM, "trace copy");
    hipMemcpyAsync(cmx.p, c->d_mx, mb, hipMemcpyDeviceToDevice, c->stream);
    hipMemcpyAsync(cmv.p, c->d_mv, mb, hipMemcpyDeviceToDevice, c->stream);
    rc = c->hip(c->L->ephemeris(c->stream, 0, c->S, C, c->d_prm, (const double *)cmx.p, (const double *)cmv.p,
                                (double *)cmx.p, (double *)cmv.p, c->d_xs0,
                                c->d_rec, nullptr, nullptr, c->d_status, nullptr, sc, c->cfg.central_mass),
                "trace ephemeris");
    if (rc) return rc;
    const long long cnt = (c->Ntp + stride - 1) / stride;
    if (cnt > 0) {
      if (c->NM == 5)
        encounter_trace_kernel<5><<<(unsigned)((cnt + 255) / 256), 256, 0, c->stream>>>(
            c->tp_arrays(), c->d_tp_sys, c->Ntp, stride, c->d_prm, c->d_xs0, c->d_mv, c->d_rec, C, sc,
            (float *)df.p, (float *)dl.p, (unsigned int *)dc.p, (unsigned)cap, (unsigned)w);
      else
        encounter_trace_kernel<2><<<(unsigned)((cnt + 255) / 256), 256, 0, c->stream>>>(
            c->tp_arrays(), c->d_tp_sys, c->Ntp, stride, c->d_prm, c->d_xs0, c->d_mv, c->d_rec, C, sc,
            (float *)df.p, (float *)dl.p, (unsigned int *)dc.p, (unsigned)cap, (unsigned)w);
      if ((rc = c->hip(hipGetLastError(), "encounter trace launch"))) return rc;
    }
    hipStreamSynchronize(c->stream);
    // now advance the real state by the window
    if ((rc = step_impl(c, C))) return rc;
  }
  hipStreamSynchronize(c->stream);
  unsigned int cnt = 0;
  hipMemcpy(&cnt, dc.p, 4, hipMemcpyDeviceToHost);
  const int64_t n = std::min<int64_t>(cnt, cap);
  hipMemcpy(feat, df.p, (size_t)n * SC_IN * sizeof(float), hipMemcpyDeviceToHost);
  hipMemcpy(label, dl.p, (size_t)n * sizeof(float), hipMemcpyDeviceToHost);
  *n_out = n;
  return c->hip(hipGetLastError(), "encounter trace");
}

int mint_nn_forward(void *ctx, int32_t which, int64_t n, const float *feat, float *out) {
  Ctx *c = (Ctx *)ctx;
  if (!c || nparams(which) < 0 || n < 0 || (n > 0 && (!feat || !out))) return MINT_EINVAL;
  auto &hw = host_weights(c, which);
  if (hw.empty()) return c->fail(MINT_ESTATE, "mint_nn_forward: network has no weights");
  hipSetDevice(c->cfg.device);
  DevBuf dw, df, dout;
  const int ni = nin(which);
  if (!dw.alloc(hw.size() * 4) || !df.alloc((size_t)n * ni * 4) || !dout.alloc((size_t)n * 4))
    return c->fail(MINT_ENOMEM, "mint_nn_forward: allocation");
  hipMemcpy(dw.p, hw.data(), hw.size() * 4, hipMemcpyHostToDevice);
  hipMemcpy(df.p, feat, (size_t)n * ni * 4, hipMemcpyHostToDevice);
  const unsigned grid = (unsigned)((n + 255) / 256);
  if (grid) {
    if (which == 0)
      nn_fwd_kernel<0><<<grid, 256, 0, c->stream>>>((float *)dw.p, (float *)df.p, n, (float *)dout.p);
    else
      nn_fwd_kernel<1><<<grid, 256, 0, c->stream>>>((float *)dw.p, (float *)df.p, n, (float *)dout.p);
  }
  int rc = c->hip(hipGetLastError(), "nn forward");
  if (rc) return rc;
  hipStreamSynchronize(c->stream);
  return c->hip(hipMemcpy(out, dout.p, (size_t)n * 4, hipMemcpyDeviceToHost), "D2H nn out");
}

int mint_nn_loss_grad(void *ctx, int32_t which, int64_t n, const float *feat, const float *target,
                      float *loss, float *grad) {
  Ctx *c = (Ctx *)ctx;
  if (!c || nparams(which) < 0 || n <= 0 || !feat || !target || !loss || !grad) return MINT_EINVAL;
  auto &hw = host_weights(c, which);
  if (hw.empty()) return c->fail(MINT_ESTATE, "mint_nn_loss_grad: network has no weights");
  hipSetDevice(c->cfg.device);
  const int np = nparams(which), ni = nin(which);
  DevBuf dw, df, dt, dg, dl;
  if (!dw.alloc(np * 4) || !df.alloc((size_t)n * ni * 4) || !dt.alloc((size_t)n * 4) || !dg.alloc(np * 4) ||
      !dl.alloc(4))
    return c->fail(MINT_ENOMEM, "mint_nn_loss_grad: allocation");
  hipMemcpy(dw.p, hw.data(), np * 4, hipMemcpyHostToDevice);
  hipMemcpy(df.p, feat, (size_t)n * ni * 4, hipMemcpyHostToDevice);
  hipMemcpy(dt.p, target, (size_t)n * 4, hipMemcpyHostToDevice);
  hipMemset(dg.p, 0, np * 4);
  hipMemset(dl.p, 0, 4);
  int rc = c->hip(launch_grad(which, (float *)dw.p, (float *)df.p, (float *)dt.p, nullptr, n, (float *)dg.p,
                              (float *)dl.p, c->stream),
                  "nn grad");
  if (rc) return rc;
  hipStreamSynchronize(c->stream);
  std::vector<float> g(np);
  float l = 0.f;
  hipMemcpy(g.data(), dg.p, np * 4, hipMemcpyDeviceToHost);
  hipMemcpy(&l, dl.p, 4, hipMemcpyDeviceToHost);
  for (int i = 0; i < np; ++i) grad[i] = g[i] / (float)n;
  *loss = l / (float)n;
  return c->hip(hipGetLastError(), "nn grad");
}

int mint_train(void *ctx, int32_t which, int64_t n, const float *feat, const float *target, int32_t epochs,
               int32_t batch, float lr, uint64_t seed, float *loss_out) {
  Ctx *c = (Ctx *)ctx;
  if (!c || nparams(which) < 0 || n <= 0 || !feat || !target || epochs <= 0 || batch <= 0 || !(lr > 0.f))
    return MINT_EINVAL;
  if (n > 0x7fffffff) return c->fail(MINT_EINVAL, "mint_train: too many samples");
  hipSetDevice(c->cfg.device);
  const int np = nparams(which), ni = nin(which);
  std::vector<float> &hw = host_weights(c, which);
  if (hw.empty()) init_weights(which, seed, hw);
  DevBuf dw, dm, dv, df, dt, dg, dl, dp;
  if (!dw.alloc(np * 4) || !dm.alloc(np * 4) || !dv.alloc(np * 4) || !df.alloc((size_t)n * ni * 4) ||
      !dt.alloc((size_t)n * 4) || !dg.alloc(np * 4) || !dl.alloc(4) || !dp.alloc((size_t)n * 4))
    return c->fail(MINT_ENOMEM, "mint_train: allocation");
  hipMemcpy(dw.p, hw.data(), np * 4, hipMemcpyHostToDevice);
  hipMemset(dm.p, 0, np * 4);
  hipMemset(dv.p, 0, np * 4);
  hipMemcpy(df.p, feat, (size_t)n * ni * 4, hipMemcpyHostToDevice);
  hipMemcpy(dt.p, target, (size_t)n * 4, hipMemcpyHostToDevice);
  SplitMix r{seed ^ 0x5eedULL};
  std::vector<int> perm((size_t)n);
  for (int64_t i = 0; i < n; ++i) perm[i] = (int)i;
  const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
  long long t = 0;
  for (int ep = 0; ep < epochs; ++ep) {
    for (int64_t i = n - 1; i > 0; --i) std::swap(perm[i], perm[r.next() % (unsigned long long)(i + 1)]);
    hipMemcpy(dp.p, perm.data(), (size_t)n * 4, hipMemcpyHostToDevice);
    hipMemset(dl.p, 0, 4);
    for (int64_t b0 = 0; b0 < n; b0 += batch) {
      const int64_t nb = std::min<int64_t>(batch, n - b0);
      zero_f32_kernel<<<(np + 255) / 256, 256, 0, c->stream>>>((float *)dg.p, np);
      launch_grad(which, (float *)dw.p, (float *)df.p, (float *)dt.p, (int *)dp.p + b0, nb, (float *)dg.p,
                  (float *)dl.p, c->stream);
      ++t;
      const float c1 = 1.f - std::pow(b1, (float)t), c2 = 1.f - std::pow(b2, (float)t);
      adam_kernel<<<(np + 255) / 256, 256, 0, c->stream>>>((float *)dw.p, (float *)dm.p, (float *)dv.p,
                                                           (float *)dg.p, np, 1.f / (float)nb, lr, b1, b2,
                                                           eps, c1, c2);
    }
    int rc = c->hip(hipGetLastError(), "train launch");
    if (rc) return rc;
    float l = 0.f;
    hipMemcpy(&l, dl.p, 4, hipMemcpyDeviceToHost);
    if (loss_out) loss_out[ep] = l / (float)n;
  }
  hipStreamSynchronize(c->stream);
  hipMemcpy(hw.data(), dw.p, np * 4, hipMemcpyDeviceToHost);
  c->drop_graphs();
  if (c->uploaded && dev_weights(c, which))
    hipMemcpy(dev_weights(c, which), hw.data(), np * 4, hipMemcpyHostToDevice);
  return c->hip(hipGetLastError(), "train");
}

}  // extern "C"
