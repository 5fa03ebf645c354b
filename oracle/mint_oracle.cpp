// mint_oracle.cpp — CPU FP64 oracle of the Mercurius step.
//
// TEST INFRASTRUCTURE ONLY.  This file is the checker for the GPU library: it
// is loaded by tests/, by __graft_entry__.smoke() and by bench.py's CPU
// baseline leg, and by nothing in the product.  It is a plain, scalar,
// array-of-structures restatement of the reference step (src/sim.cuh:679-742
// of at2005/p9-integrator) with the documented semantic decisions applied
// (universal-variable Kepler drift, r_c = n_hill Hill radii fixed at set-up,
// tolerance Bulirsch-Stoer on flagged bodies only, moving perturbers, one
// changeover per pair, massive encounter groups integrated together, sums
// over massive bodies in fixed 16-body chunks).  It is written independently
// of the GPU code but
// follows the same stated operation order (every fma is explicit, the file is
// built with -ffp-contract=off), so the two agree to the last bit.
//
// Pinning (see tests/test_oracle.py): Stumpff series vs closed forms, Kepler
// drift vs a 40-digit universal-variable solve (mpmath root finder, closed-form
// Stumpff functions, Lagrange identity f*gdot - fdot*g = 1 checked there),
// energy and angular momentum of pure Kepler drifts, changeover identities K(0)=0, K(1)=1, K(y)+K(1-y)=1, BS on pure
// Kepler vs the analytic drift, second-order energy convergence of the whole
// step, and the hand-built encounter threshold cases.
#include <cmath>
#include <cstdint>
#include <cstring>
#include <thread>
#include <vector>

namespace {

struct V3 {
  double x, y, z;
};

inline V3 sub(const V3 &a, const V3 &b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
// dot(a,b) = fma(ax,bx, fma(ay,by, az*bz))
inline double dot(const V3 &a, const V3 &b) { return std::fma(a.x, b.x, std::fma(a.y, b.y, a.z * b.z)); }

struct Body {
  V3 x, v;
};

struct Params {
  double h, mu, eps2, tol;
  int guess;  // 0 crude, 1 series
};

constexpr int KMAX = 8;
constexpr int BS_MAXDEPTH = 24;
constexpr int NEWTON_MAXIT = 40;

struct Stats {
  int64_t tp_steps = 0, kepler = 0, newton = 0, bs_steps = 0, bs_rhs = 0;
  int64_t mgroups = 0, mgroup_bodies = 0, overlaps = 0;
  int64_t mkepler = 0, mnewton = 0;  // massive bodies' Kepler drifts and their Newton iterations
};

// --- Stumpff functions (Danby 1988 §6.9), replacing the Kepler-equation
// solve of src/sim.cuh:90-126 -------------------------------------------------
void stumpff(double z, double c[4]) {
  static const double C2[7] = {0.5, -0.041666666666666664, 0.001388888888888889, -2.48015873015873e-05,
                               2.755731922398589e-07, -2.08767569878681e-09, 1.1470745597729725e-11};
  static const double C3[7] = {0.16666666666666666, -0.008333333333333333, 0.0001984126984126984,
                               -2.7557319223985893e-06, 2.505210838544172e-08, -1.6059043836821613e-10,
                               7.647163731819816e-13};
  int n = 0;
  while (std::fabs(z) > 0.1 && n < 64) {
    z *= 0.25;
    n++;
  }
  // truncated by |z| band: degree 2 (|z| <= 1e-5), 3 (<= 1e-3), else 6
  const double az = std::fabs(z);
  const int top = az <= 1e-5 ? 2 : (az <= 1e-3 ? 3 : 6);
  double s2 = C2[top], s3 = C3[top];
  for (int k = top - 1; k >= 0; --k) {
    s2 = std::fma(s2, z, C2[k]);
    s3 = std::fma(s3, z, C3[k]);
  }
  double c0 = std::fma(-z, s2, 1.0), c1 = std::fma(-z, s3, 1.0), c2 = s2, c3 = s3;
  while (n-- > 0) {
    const double a3 = std::fma(c0, c3, c2) * 0.25;
    const double a2 = (c1 * c1) * 0.5;
    const double a1 = c0 * c1;
    const double a0 = std::fma(2.0 * c0, c0, -1.0);
    c0 = a0;
    c1 = a1;
    c2 = a2;
    c3 = a3;
  }
  c[0] = c0;
  c[1] = c1;
  c[2] = c2;
  c[3] = c3;
}

// --- universal-variable Kepler drift (replaces the element round trip of
// src/sim.cuh:704-713) ---------------------------------------------------------
int kepler_drift(Body &b, double mu, double dt, int guess, bool &ok, double *s_root = nullptr) {
  const double r0 = std::sqrt(dot(b.x, b.x));
  const double inv_r0 = 1.0 / r0;
  const double v2 = dot(b.v, b.v);
  const double eta0 = dot(b.x, b.v);
  const double beta = std::fma(2.0 * mu, inv_r0, -v2);
  const double zeta0 = std::fma(-beta, r0, mu);
  double s;
  if (guess == 0) {
    s = dt * inv_r0;
  } else {
    // s = u (1 + a u + b u^2): third-order inversion of the Kepler equation
    const double u = dt * inv_r0;
    const double e = eta0 * inv_r0;
    const double a = -0.5 * e;
    const double bb = std::fma(0.5 * e, e, -(zeta0 * inv_r0) * (1.0 / 6.0));
    s = u * std::fma(u, std::fma(u, bb, a), 1.0);
  }
  double G0 = 0, G1 = 0, G2 = 0, G3 = 0, ds = 0;
  int it = 0;
  ok = false;
  // bracket from the monotonicity of F (F' = r > 0, F(0) = -dt)
  double lo = dt >= 0.0 ? 0.0 : -INFINITY, hi = dt >= 0.0 ? INFINITY : 0.0;
  for (;;) {
    ++it;
    double c[4];
    const double ss = s * s;
    stumpff(beta * ss, c);
    G0 = c[0];
    G1 = s * c[1];
    G2 = ss * c[2];
    G3 = ss * (s * c[3]);
    const double F = std::fma(r0, s, std::fma(eta0, G2, std::fma(zeta0, G3, -dt)));
    const double rp = std::fma(eta0, G1, std::fma(zeta0, G2, r0));
    const double rpp = std::fma(eta0, G0, zeta0 * G1);
    ds = F / rp;
    const double as = std::fabs(s);
    if (std::fabs(ds) <= 2e-6 * as && std::fabs(rpp) * (ds * ds) <= 4.4e-16 * (std::fabs(rp) * as)) {
      ok = true;
      break;
    }
    if (it >= NEWTON_MAXIT) break;
    if (F < 0.0) lo = s; else hi = s;
    double next = s - ds;
    if (!(next > lo && next < hi)) {
      // Laguerre-Conway step with n = 5, then bisection
      const double disc = std::fabs(std::fma(16.0 * rp, rp, -(20.0 * F * rpp)));
      next = s - (5.0 * F) / (rp + std::copysign(std::sqrt(disc), rp));
      if (!(next > lo && next < hi)) next = 0.5 * (lo + hi);
    }
    s = next;
  }
  if (s_root) *s_root = s - ds;
  // carry G1..G3 from s to s - ds (G_k' = G_{k-1}, G_0' = -beta G_1)
  const double hs = -ds;
  const double q = 0.5 * (hs * hs);
  const double t0 = G0 * q, t1 = G1 * q;
  const double g1 = G1 + std::fma(G0, hs, -(beta * t1));
  const double g2 = G2 + std::fma(G1, hs, t0);
  const double g3 = G3 + std::fma(G2, hs, t1);
  const double r = std::fma(eta0, g1, std::fma(zeta0, g2, r0));
  const double inv_r = 1.0 / r;
  const double fm1 = -(mu * g2) * inv_r0;
  const double g = std::fma(-mu, g3, dt);
  const double fdot = -(mu * g1) * (inv_r0 * inv_r);
  const double gm1 = -(mu * g2) * inv_r;
  const V3 x = b.x, v = b.v;
  b.x.x = x.x + std::fma(fm1, x.x, g * v.x);
  b.x.y = x.y + std::fma(fm1, x.y, g * v.y);
  b.x.z = x.z + std::fma(fm1, x.z, g * v.z);
  b.v.x = v.x + std::fma(fdot, x.x, gm1 * v.x);
  b.v.y = v.y + std::fma(fdot, x.y, gm1 * v.y);
  b.v.z = v.z + std::fma(fdot, x.z, gm1 * v.z);
  return it;
}

// --- changeover (src/sim.cuh:166-186): y = (r - 0.1 rc)/(0.9 rc) ---------------
struct Rc {
  double lo, winv, r2;  // 0.1 rc, 1/(0.9 rc), rc^2
};
Rc make_rc(double rc) { return {0.1 * rc, 1.0 / (0.9 * rc), rc * rc}; }

// Changeover of a massive pair: r_c = max(r_c,j, r_c,k).  Each field of
// make_rc is a monotone, correctly rounded function of r_c, so the pair values
// are the max/min of the two bodies' fields exactly (0.1 r_c and r_c^2
// increase with r_c, 1/(0.9 r_c) decreases).
Rc pair_rc(const Rc &a, const Rc &b) {
  return {a.lo > b.lo ? a.lo : b.lo, a.winv < b.winv ? a.winv : b.winv, a.r2 > b.r2 ? a.r2 : b.r2};
}

// Sums over massive bodies run in chunks of CH bodies (index order): each
// chunk's sum starts from zero, and the chunk sums are folded by halving
// strides (fold_chunks).  Pair weights are divided in groups of DG bodies with
// one division each.
constexpr int CH = 16;
constexpr int DG = 8;

// K m / ((r^2+eps^2) r) as a numerator / denominator pair
void kick_nd(double r2, double eps2, double m, const Rc &c, double &num, double &den) {
  const double r = std::sqrt(r2);
  const double D = (r2 + eps2) * r;
  const double y = (r - c.lo) * c.winv;
  num = m;
  den = D;
  if (y <= 0.0) {
    num = 0.0;
  } else if (y < 1.0) {
    const double y2 = y * y;
    num = m * y2;
    den = D * std::fma(2.0, y2, std::fma(-2.0, y, 1.0));
  }
}

double kick_w(double r2, double eps2, double m, const Rc &c) {
  double num, den;
  kick_nd(r2, eps2, m, c, num, den);
  return num / den;
}

// num[i] / den[i] for all i with one division: prefix products P, inv = 1/P_last,
// walk back w_i = num_i (inv P_{i-1}), inv *= den_i, w_0 = num_0 inv
constexpr int MAXB = 8;  // massive bodies per system with test particles; division group size

void batched_div(int n, const double *num, const double *den, double *w) {
  double P[MAXB];
  P[0] = den[0];
  for (int i = 1; i < n; ++i) P[i] = P[i - 1] * den[i];
  double inv = 1.0 / P[n - 1];
  for (int i = n - 1; i >= 1; --i) {
    w[i] = num[i] * (inv * P[i - 1]);
    inv = inv * den[i];
  }
  w[0] = num[0] * inv;
}

// (1-K) m / ((r^2+eps^2) r)
double close_w(double r2, double eps2, double m, const Rc &c) {
  const double r = std::sqrt(r2);
  const double y = (r - c.lo) * c.winv;
  if (y >= 1.0) return 0.0;
  const double D = (r2 + eps2) * r;
  if (y <= 0.0) return m / D;
  const double om = 1.0 - y;
  return (m * (om * om)) / (D * std::fma(2.0, y * y, std::fma(-2.0, y, 1.0)));
}

struct Sys {
  int nm;
  std::vector<double> m;
  std::vector<Rc> rc_tp;   // [j]; massive pairs use pair_rc(rc_tp[j], rc_tp[k])
  std::vector<double> rad; // physical radii (empty: not recorded)
  std::vector<Body> mb;    // massive, heliocentric x, DH v
  std::vector<Body> tp;
};

// kick acceleration of a TP (body_interaction_kick, src/sim.cuh:327-382)
V3 accel_tp(const V3 &x, const std::vector<V3> &X, const Sys &s, double eps2) {
  V3 d[MAXB];
  double num[MAXB] = {}, den[MAXB] = {}, w[MAXB] = {};
  for (int j = 0; j < s.nm; ++j) {
    d[j] = sub(X[j], x);
    kick_nd(dot(d[j], d[j]), eps2, s.m[j], s.rc_tp[j], num[j], den[j]);
  }
  batched_div(s.nm, num, den, w);
  V3 a{0, 0, 0};
  for (int j = 0; j < s.nm; ++j) {
    a.x = std::fma(w[j], d[j].x, a.x);
    a.y = std::fma(w[j], d[j].y, a.y);
    a.z = std::fma(w[j], d[j].z, a.z);
  }
  return a;
}

// Kick acceleration of massive body i from the others (body_interaction_kick,
// src/sim.cuh:327-382, over massive partners, self excluded by index).  In a
// division group the body itself is an inert slot (numerator 0, denominator 1,
// separation 0), which leaves every other weight and the sum unchanged.
// The fold of chunk sums a[0..n): with P = n rounded up to a power of two and
// a[l] = +0 for l >= n, slot l < h takes a[l] + a[l + h] for h = P/2, ..., 1;
// the result is a[0] (one chunk: its own sum).  This is the order of a
// wavefront's shuffle reduction over one chunk per lane.
V3 fold_chunks(std::vector<V3> a) {
  size_t P = 1;
  while (P < a.size()) P <<= 1;
  a.resize(P, V3{0.0, 0.0, 0.0});
  for (size_t h = P / 2; h >= 1; h /= 2)
    for (size_t l = 0; l < h; ++l) a[l] = {a[l].x + a[l + h].x, a[l].y + a[l + h].y, a[l].z + a[l + h].z};
  return a[0];
}

V3 accel_one(int i, const std::vector<V3> &X, const Sys &s, double eps2) {
  const int nm = s.nm;
  std::vector<V3> part;
  for (int c0 = 0; c0 < nm; c0 += CH) {
    const int c1 = std::min(c0 + CH, nm);
    V3 p{0, 0, 0};
    for (int g0 = c0; g0 < c1; g0 += DG) {
      const int n = std::min(g0 + DG, c1) - g0;
      V3 d[DG];
      double num[DG] = {}, den[DG] = {}, w[DG] = {};
      for (int q = 0; q < n; ++q) {
        const int k = g0 + q;
        if (k == i) {
          d[q] = {0, 0, 0};
          num[q] = 0.0;
          den[q] = 1.0;
        } else {
          d[q] = sub(X[k], X[i]);
          kick_nd(dot(d[q], d[q]), eps2, s.m[k], pair_rc(s.rc_tp[i], s.rc_tp[k]), num[q], den[q]);
        }
      }
      batched_div(n, num, den, w);
      for (int q = 0; q < n; ++q) {
        p.x = std::fma(w[q], d[q].x, p.x);
        p.y = std::fma(w[q], d[q].y, p.y);
        p.z = std::fma(w[q], d[q].z, p.z);
      }
    }
    part.push_back(p);
  }
  return fold_chunks(part);
}

std::vector<V3> accel_massive(const std::vector<V3> &X, const Sys &s, double eps2) {
  std::vector<V3> A(s.nm);
  for (int j = 0; j < s.nm; ++j) A[j] = accel_one(j, X, s, eps2);
  return A;
}

// Q = sum m_j v_j / M (main_body_kinetic, src/sim.cuh:384-406), chunked sum
V3 momentum_q(const Sys &s, double M) {
  std::vector<V3> part;
  for (int c0 = 0; c0 < s.nm; c0 += CH) {
    V3 p{0, 0, 0};
    for (int j = c0; j < std::min(c0 + CH, s.nm); ++j) {
      p.x = std::fma(s.m[j], s.mb[j].v.x, p.x);
      p.y = std::fma(s.m[j], s.mb[j].v.y, p.y);
      p.z = std::fma(s.m[j], s.mb[j].v.z, p.z);
    }
    part.push_back(p);
  }
  const V3 q = fold_chunks(part);
  return {q.x / M, q.y / M, q.z / M};
}

// --- close-encounter RHS and Bulirsch-Stoer (src/sim.cuh:410-551) -------------
typedef double State6[6];

void rhs_tp(double tau, const double *y, const std::vector<Body> &pert0, const Sys &s, const Params &p,
            double *f) {
  f[0] = y[3];
  f[1] = y[4];
  f[2] = y[5];
  const V3 x{y[0], y[1], y[2]};
  const double r2 = dot(x, x);
  const double r = std::sqrt(r2);
  const double w0 = p.mu / (r2 * r);
  V3 a{-(w0 * x.x), -(w0 * x.y), -(w0 * x.z)};
  for (int j = 0; j < s.nm; ++j) {
    Body pb = pert0[j];
    if (tau != 0.0) {
      bool ok;
      kepler_drift(pb, p.mu, tau, 1, ok);
    }
    const V3 d = sub(pb.x, x);
    const double w = close_w(dot(d, d), p.eps2, s.m[j], s.rc_tp[j]);
    a.x = std::fma(w, d.x, a.x);
    a.y = std::fma(w, d.y, a.y);
    a.z = std::fma(w, d.z, a.z);
  }
  f[3] = a.x;
  f[4] = a.y;
  f[5] = a.z;
}

// Generic Gragg + Aitken-Neville BS over [0,h] on an n-vector; rhs(t, y, f).
template <class RHS>
uint32_t bs_drift(std::vector<double> &y, double h, const std::vector<double> &scale_idx_unused, RHS rhs,
                  const std::vector<int> &scale_group, int ngroups, double tol, int64_t &nrhs) {
  (void)scale_idx_unused;
  const size_t n = y.size();
  uint32_t st = 0;
  double t = 0.0, H = h;
  int depth = 0;
  std::vector<double> f0(n), cur(n), z0(n), z1(n), f(n), sc(ngroups);
  std::vector<std::vector<double>> R(KMAX, std::vector<double>(n));
  while (t < h) {
    if (H > h - t) H = h - t;
    rhs(t, y.data(), f0.data());
    nrhs += 1;
    // scale per group: |x| of positions, |v| of velocities
    for (int g = 0; g < ngroups; ++g) {
      const int b = g / 2, vel = g % 2;
      const double *q = y.data() + 6 * b + 3 * vel;
      sc[g] = std::sqrt(std::fma(q[0], q[0], std::fma(q[1], q[1], q[2] * q[2])));
    }
    bool conv = false;
    for (int k = 0; k < KMAX; ++k) {
      const int N = 2 << k;
      const double hs = H / (double)N, h2 = 2.0 * hs;
      for (size_t i = 0; i < n; ++i) {
        z0[i] = y[i];
        z1[i] = std::fma(hs, f0[i], y[i]);
      }
      for (int mm = 1; mm < N; ++mm) {
        rhs(t + (double)mm * hs, z1.data(), f.data());
        for (size_t i = 0; i < n; ++i) {
          const double z2 = std::fma(h2, f[i], z0[i]);
          z0[i] = z1[i];
          z1[i] = z2;
        }
      }
      rhs(t + H, z1.data(), f.data());
      nrhs += N;
      for (size_t i = 0; i < n; ++i) cur[i] = 0.5 * (z1[i] + std::fma(hs, f[i], z0[i]));
      for (int j = 1; j <= k; ++j) {
        const double fac = (double)((1ll << (2 * j)) - 1);
        for (size_t i = 0; i < n; ++i) {
          const double prev = R[j - 1][i];
          R[j - 1][i] = cur[i];
          cur[i] = cur[i] + (cur[i] - prev) / fac;
        }
      }
      R[k] = cur;
      if (k >= 1) {
        double err = 0.0;
        for (size_t i = 0; i < n; ++i) {
          const double e = std::fabs(cur[i] - R[k - 1][i]) / sc[scale_group[i]];
          err = e > err ? e : err;
        }
        if (err <= tol) {
          conv = true;
          break;
        }
      }
    }
    if (conv || depth >= BS_MAXDEPTH) {
      y = cur;
      t += H;
      if (!conv) {
        // halving budget spent: flag, and the rest of the interval in one step
        st |= 4u;
        H = h - t;
      }
    } else {
      H *= 0.5;
      ++depth;
    }
  }
  return st;
}

std::vector<int> groups_for(int nbodies) {
  std::vector<int> g(6 * nbodies);
  for (int i = 0; i < 6 * nbodies; ++i) g[i] = 2 * (i / 6) + ((i % 6) < 3 ? 0 : 1);
  return g;
}

uint32_t bs_tp(Body &b, const std::vector<Body> &pert0, const Sys &s, const Params &p, int64_t &nrhs) {
  std::vector<double> y{b.x.x, b.x.y, b.x.z, b.v.x, b.v.y, b.v.z};
  static const std::vector<int> grp = groups_for(1);
  auto rhs = [&](double tau, const double *yy, double *f) { rhs_tp(tau, yy, pert0, s, p, f); };
  const uint32_t st = bs_drift(y, p.h, {}, rhs, grp, 2, p.tol, nrhs);
  b.x = {y[0], y[1], y[2]};
  b.v = {y[3], y[4], y[5]};
  return st;
}

// BS drift of one massive encounter group (members in index order): each
// member feels the Sun and the (1-K)-weighted pulls of the other members.
uint32_t bs_group(Sys &s, const std::vector<int> &mem, const Params &p, int64_t &nrhs) {
  const int g = (int)mem.size();
  std::vector<double> y(6 * g);
  for (int a = 0; a < g; ++a) {
    const Body &b = s.mb[mem[a]];
    double *q = y.data() + 6 * a;
    q[0] = b.x.x; q[1] = b.x.y; q[2] = b.x.z; q[3] = b.v.x; q[4] = b.v.y; q[5] = b.v.z;
  }
  const std::vector<int> grp = groups_for(g);
  auto rhs = [&](double, const double *yy, double *f) {
    for (int a = 0; a < g; ++a) {
      const int j = mem[a];
      const double *q = yy + 6 * a;
      f[6 * a + 0] = q[3];
      f[6 * a + 1] = q[4];
      f[6 * a + 2] = q[5];
      const V3 x{q[0], q[1], q[2]};
      const double r2 = dot(x, x);
      const double r = std::sqrt(r2);
      const double w0 = p.mu / (r2 * r);
      V3 acc{-(w0 * x.x), -(w0 * x.y), -(w0 * x.z)};
      for (int b = 0; b < g; ++b) {
        if (b == a) continue;
        const int k = mem[b];
        const V3 d{yy[6 * b] - x.x, yy[6 * b + 1] - x.y, yy[6 * b + 2] - x.z};
        const double w = close_w(dot(d, d), p.eps2, s.m[k], pair_rc(s.rc_tp[j], s.rc_tp[k]));
        acc.x = std::fma(w, d.x, acc.x);
        acc.y = std::fma(w, d.y, acc.y);
        acc.z = std::fma(w, d.z, acc.z);
      }
      f[6 * a + 3] = acc.x;
      f[6 * a + 4] = acc.y;
      f[6 * a + 5] = acc.z;
    }
  };
  const uint32_t st = bs_drift(y, p.h, {}, rhs, grp, 2 * g, p.tol, nrhs);
  for (int a = 0; a < g; ++a) {
    const double *q = y.data() + 6 * a;
    s.mb[mem[a]].x = {q[0], q[1], q[2]};
    s.mb[mem[a]].v = {q[3], q[4], q[5]};
  }
  return st;
}

// Massive encounter groups: the connected components of the graph whose edges
// are the pairs inside their changeover radius (|x_k - x_j|^2 < r_c^2 of the
// pair).  Groups are listed by their smallest member, members in index order;
// bodies outside every edge are absent.  Also counts the pairs closer than the
// sum of their physical radii.
std::vector<std::vector<int>> encounter_groups(const Sys &s, int64_t &overlaps) {
  const int nm = s.nm;
  std::vector<int> parent(nm);
  std::vector<char> inedge(nm, 0);
  for (int j = 0; j < nm; ++j) parent[j] = j;
  auto find = [&](int a) {
    while (parent[a] != a) a = parent[a];
    return a;
  };
  const bool radii = !s.rad.empty();
  for (int j = 0; j < nm; ++j)
    for (int k = j + 1; k < nm; ++k) {
      const V3 d = sub(s.mb[k].x, s.mb[j].x);
      const double r2 = dot(d, d);
      if (radii) {
        const double sr = s.rad[j] + s.rad[k];
        if (r2 < sr * sr) ++overlaps;
      }
      if (r2 < pair_rc(s.rc_tp[j], s.rc_tp[k]).r2) {
        inedge[j] = inedge[k] = 1;
        const int a = find(j), b = find(k);
        if (a < b) parent[b] = a;
        if (b < a) parent[a] = b;
      }
    }
  std::vector<std::vector<int>> groups;
  std::vector<int> gid(nm, -1);
  for (int j = 0; j < nm; ++j) {
    if (!inedge[j]) continue;
    const int r = find(j);
    if (gid[r] < 0) {
      gid[r] = (int)groups.size();
      groups.emplace_back();
    }
    groups[gid[r]].push_back(j);
  }
  return groups;
}

// --- one Mercurius step of one system (src/sim.cuh:679-742) -------------------
// hits (trace mode): per TP of the system, OR of the bodies whose exact check
// fired; the TP then takes its Kepler drift whatever the check says (the
// library's encounter-trace definition).
uint32_t step(Sys &s, const Params &p, double M, Stats &stt, uint32_t *hits = nullptr) {
  const int nm = s.nm;
  const double hh = 0.5 * p.h;
  uint32_t st = 0;
  std::vector<V3> X(nm);
  for (int j = 0; j < nm; ++j) X[j] = s.mb[j].x;
  // massive encounter groups from the separations at step start: jump 1
  // translates every massive body by the same (h/2) Q1, so the pair
  // separations it leaves are these up to the rounding of the translation
  // (close_encounter_p, src/sim.cuh:589-613, tests them after the jump)
  int64_t ov = 0;
  const std::vector<std::vector<int>> groups = encounter_groups(s, ov);
  // 1. kick 1/2 (massive and TPs, positions at step start)
  const std::vector<V3> Am = accel_massive(X, s, p.eps2);
  for (Body &t : s.tp) {
    const V3 a = accel_tp(t.x, X, s, p.eps2);
    t.v.x = std::fma(hh, a.x, t.v.x);
    t.v.y = std::fma(hh, a.y, t.v.y);
    t.v.z = std::fma(hh, a.z, t.v.z);
  }
  for (int j = 0; j < nm; ++j) {
    s.mb[j].v.x = std::fma(hh, Am[j].x, s.mb[j].v.x);
    s.mb[j].v.y = std::fma(hh, Am[j].y, s.mb[j].v.y);
    s.mb[j].v.z = std::fma(hh, Am[j].z, s.mb[j].v.z);
  }
  // 2. jump 1/2 with the momentum after the kick
  const V3 Q1 = momentum_q(s, M);
  for (Body &b : s.mb) {
    b.x.x = std::fma(hh, Q1.x, b.x.x);
    b.x.y = std::fma(hh, Q1.y, b.x.y);
    b.x.z = std::fma(hh, Q1.z, b.x.z);
  }
  for (Body &t : s.tp) {
    t.x.x = std::fma(hh, Q1.x, t.x.x);
    t.x.y = std::fma(hh, Q1.y, t.x.y);
    t.x.z = std::fma(hh, Q1.z, t.x.z);
  }
  // 3. TP encounter flags on post-jump positions; 4. drift
  const std::vector<Body> pert0 = s.mb;
  size_t ti = 0;
  for (Body &t : s.tp) {
    bool flag = false;
    uint32_t hb = 0;
    for (int j = 0; j < nm; ++j) {
      const V3 d = sub(pert0[j].x, t.x);
      const bool in = dot(d, d) < s.rc_tp[j].r2;
      flag = flag || in;
      if (in) hb |= 1u << j;
    }
    if (hits) {
      hits[ti] |= hb;
      flag = false;
    }
    ++ti;
    if (flag) {
      st |= bs_tp(t, pert0, s, p, stt.bs_rhs);
      stt.bs_steps++;
    } else {
      bool ok;
      stt.newton += kepler_drift(t, p.mu, p.h, p.guess, ok);
      stt.kepler++;
      if (!ok) st |= 2u;
    }
    stt.tp_steps++;
  }
  stt.overlaps += ov;
  std::vector<char> ingroup(nm, 0);
  int64_t nr = 0;
  for (const std::vector<int> &g : groups) {
    st |= 8u;
    st |= bs_group(s, g, p, nr);
    for (int j : g) ingroup[j] = 1;
    stt.mgroups++;
    stt.mgroup_bodies += (int64_t)g.size();
  }
  if (ov) st |= 16u;
  for (int j = 0; j < nm; ++j) {
    if (ingroup[j]) continue;
    bool ok;
    stt.mnewton += kepler_drift(s.mb[j], p.mu, p.h, 1, ok);
    stt.mkepler++;
    if (!ok) st |= 2u;
  }
  // 5. jump 1/2 with the momentum after the drift
  const V3 Q2 = momentum_q(s, M);
  for (Body &b : s.mb) {
    b.x.x = std::fma(hh, Q2.x, b.x.x);
    b.x.y = std::fma(hh, Q2.y, b.x.y);
    b.x.z = std::fma(hh, Q2.z, b.x.z);
  }
  for (Body &t : s.tp) {
    t.x.x = std::fma(hh, Q2.x, t.x.x);
    t.x.y = std::fma(hh, Q2.y, t.x.y);
    t.x.z = std::fma(hh, Q2.z, t.x.z);
  }
  // 6. kick 1/2 with the new positions
  for (int j = 0; j < nm; ++j) X[j] = s.mb[j].x;
  const std::vector<V3> Am2 = accel_massive(X, s, p.eps2);
  for (Body &t : s.tp) {
    const V3 a = accel_tp(t.x, X, s, p.eps2);
    t.v.x = std::fma(hh, a.x, t.v.x);
    t.v.y = std::fma(hh, a.y, t.v.y);
    t.v.z = std::fma(hh, a.z, t.v.z);
  }
  for (int j = 0; j < nm; ++j) {
    s.mb[j].v.x = std::fma(hh, Am2[j].x, s.mb[j].v.x);
    s.mb[j].v.y = std::fma(hh, Am2[j].y, s.mb[j].v.y);
    s.mb[j].v.z = std::fma(hh, Am2[j].z, s.mb[j].v.z);
  }
  return st;
}

}  // namespace

extern "C" {

struct OrcParams {
  double dt, central_mass, softening2, bs_tol;
  int32_t guess;
  int32_t threads;
};

struct OrcStats {
  int64_t tp_steps, kepler_solves, newton_iters, bs_steps, bs_rhs_evals;
  int64_t massive_groups, massive_group_bodies, overlaps;
  int64_t massive_kepler_solves, massive_newton_iters;
};

int orc_stumpff(double z, double *c) {
  stumpff(z, c);
  return 0;
}

// drift one body; returns Newton iterations, writes the converged root
int orc_kepler_drift(double *x, double *v, double mu, double dt, int guess, double *s_root) {
  Body b{{x[0], x[1], x[2]}, {v[0], v[1], v[2]}};
  bool ok;
  const int it = kepler_drift(b, mu, dt, guess, ok, s_root);
  x[0] = b.x.x; x[1] = b.x.y; x[2] = b.x.z;
  v[0] = b.v.x; v[1] = b.v.y; v[2] = b.v.z;
  return ok ? it : -it;
}

double orc_kick_weight(double r2, double eps2, double m, double rc) { return kick_w(r2, eps2, m, make_rc(rc)); }
double orc_close_weight(double r2, double eps2, double m, double rc) { return close_w(r2, eps2, m, make_rc(rc)); }

// Hill-radius changeover radii, same definition as the library's set-up.
void orc_hill_radii(int64_t n, double M, double n_hill, const double *m, const double *x, const double *v,
                    double *rc) {
  for (int64_t i = 0; i < n; ++i) {
    const V3 xx{x[3 * i], x[3 * i + 1], x[3 * i + 2]}, vv{v[3 * i], v[3 * i + 1], v[3 * i + 2]};
    const double r = std::sqrt(dot(xx, xx));
    const double inva = 2.0 / r - dot(vv, vv) / M;
    const double a = inva > 0.0 ? 1.0 / inva : r;
    rc[i] = n_hill * (a * std::cbrt(m[i] / (3.0 * M)));
  }
}

// BS drift of one TP against NM perturbers (state at drift start), for tests.
int orc_bs_tp(double *y6, int nm, const double *m, const double *rc, const double *px, const double *pv,
              double h, double mu, double eps2, double tol, int64_t *nrhs) {
  Sys s;
  s.nm = nm;
  s.m.assign(m, m + nm);
  for (int j = 0; j < nm; ++j) s.rc_tp.push_back(make_rc(rc[j]));
  std::vector<Body> pert(nm);
  for (int j = 0; j < nm; ++j) pert[j] = {{px[3 * j], px[3 * j + 1], px[3 * j + 2]}, {pv[3 * j], pv[3 * j + 1], pv[3 * j + 2]}};
  Params p{h, mu, eps2, tol, 1};
  Body b{{y6[0], y6[1], y6[2]}, {y6[3], y6[4], y6[5]}};
  int64_t nr = 0;
  const uint32_t st = bs_tp(b, pert, s, p, nr);
  y6[0] = b.x.x; y6[1] = b.x.y; y6[2] = b.x.z; y6[3] = b.v.x; y6[4] = b.v.y; y6[5] = b.v.z;
  if (nrhs) *nrhs = nr;
  return (int)st;
}

// Integrate S systems for nsteps.  In/out state is heliocentric (x, v);
// velocities are moved to democratic-heliocentric form before the first step
// and back after the last (src/sim.cuh:555-587,677,748).  rc[S*NM] are the
// changeover radii.  threads > 1 spreads systems over std::thread.
static int run_impl(const OrcParams *op, int32_t S, int32_t NM, const int64_t *tp_off, const double *m,
                    const double *rc, const double *rad, double *mx, double *mv, double *tx, double *ty, double *tz,
                    double *tvx, double *tvy, double *tvz, int64_t nsteps, uint32_t *status, OrcStats *stats,
                    uint32_t *hits);

int orc_run(const OrcParams *op, int32_t S, int32_t NM, const int64_t *tp_off, const double *m, const double *rc,
            const double *rad, double *mx, double *mv, double *tx, double *ty, double *tz, double *tvx,
            double *tvy, double *tvz, int64_t nsteps, uint32_t *status, OrcStats *stats) {
  return run_impl(op, S, NM, tp_off, m, rc, rad, mx, mv, tx, ty, tz, tvx, tvy, tvz, nsteps, status, stats, nullptr);
}

// Encounter-trace labels: nsteps steps in which every TP follows its Kepler
// drift; hits[i] receives the OR over the steps of the bodies whose exact
// check fired for TP i (zeroed here).  The state is advanced like orc_run's.
int orc_trace_hits(const OrcParams *op, int32_t S, int32_t NM, const int64_t *tp_off, const double *m,
                   const double *rc, double *mx, double *mv, double *tx, double *ty, double *tz, double *tvx,
                   double *tvy, double *tvz, int64_t nsteps, uint32_t *hits) {
  std::memset(hits, 0, sizeof(uint32_t) * (size_t)tp_off[S]);
  return run_impl(op, S, NM, tp_off, m, rc, nullptr, mx, mv, tx, ty, tz, tvx, tvy, tvz, nsteps, nullptr, nullptr,
                  hits);
}

// One Kepler-trace sample (the library's trace definition) of a drift by dt
// from the DH state (x, v): features f[4] (log2 of the dimensionless step
// tau = dt sqrt(mu / r0^3), zeta0 / mu, eta0 / sqrt(mu r0), r0 beta / mu
// clamped to +-2) and target c = (s* / s_series - 1) / tau^3, s* the root of
// this oracle's own converged solve from the series guess.  Returns 0 where
// the trace takes no sample (unbound orbit, no convergence, non-finite c).
int orc_kepler_sample(const double *x, const double *v, double mu, double dt, float *f, float *c) {
  Body b{{x[0], x[1], x[2]}, {v[0], v[1], v[2]}};
  const double r0 = std::sqrt(dot(b.x, b.x));
  const double inv_r0 = 1.0 / r0;
  const double v2 = dot(b.v, b.v);
  const double eta0 = dot(b.x, b.v);
  const double beta = std::fma(2.0 * mu, inv_r0, -v2);
  const double zeta0 = std::fma(-beta, r0, mu);
  if (!(beta > 0.0)) return 0;
  const double u = dt * inv_r0;
  const double e = eta0 * inv_r0;
  const double s0 = u * std::fma(u, std::fma(u, std::fma(0.5 * e, e, -(zeta0 * inv_r0) * (1.0 / 6.0)), -0.5 * e), 1.0);
  bool ok = false;
  double sroot = 0.0;
  kepler_drift(b, mu, dt, 1, ok, &sroot);
  if (!ok) return 0;
  const double tau = dt * std::sqrt(mu / (r0 * (r0 * r0)));
  const double cc = (sroot / s0 - 1.0) / (tau * (tau * tau));
  if (!std::isfinite(cc)) return 0;
  f[0] = (std::log2((float)tau) + 6.0f) * 0.25f;
  f[1] = (float)(zeta0 / mu);
  f[2] = (float)(eta0 / std::sqrt(mu * r0));
  const float ra = (float)(r0 * beta / mu);
  f[3] = std::fmin(std::fmax(ra, -2.0f), 2.0f);
  *c = (float)cc;
  return 1;
}

static int run_impl(const OrcParams *op, int32_t S, int32_t NM, const int64_t *tp_off, const double *m,
                    const double *rc, const double *rad, double *mx, double *mv, double *tx, double *ty, double *tz,
                    double *tvx, double *tvy, double *tvz, int64_t nsteps, uint32_t *status, OrcStats *stats,
                    uint32_t *hits) {
  if (NM > MAXB && tp_off[S] > 0) return -1;  // test particles need at most MAXB massive bodies
  const double M = op->central_mass;
  const Params p{op->dt, M, op->softening2, op->bs_tol, op->guess};
  const int nthr = op->threads > 1 ? op->threads : 1;
  std::vector<Stats> tstats(nthr);
  auto work = [&](int tid) {
    for (int si = tid; si < S; si += nthr) {
      Sys s;
      s.nm = NM;
      s.m.assign(m + (size_t)si * NM, m + (size_t)(si + 1) * NM);
      const double *rcs = rc + (size_t)si * NM;
      for (int j = 0; j < NM; ++j) s.rc_tp.push_back(make_rc(rcs[j]));
      if (rad) s.rad.assign(rad + (size_t)si * NM, rad + (size_t)(si + 1) * NM);
      // DH velocities: v - (sum m v)/(M + sum m)
      double px = 0, py = 0, pz = 0, msum = 0;
      for (int j = 0; j < NM; ++j) {
        const double *vv = mv + ((size_t)si * NM + j) * 3;
        px = std::fma(s.m[j], vv[0], px);
        py = std::fma(s.m[j], vv[1], py);
        pz = std::fma(s.m[j], vv[2], pz);
        msum = msum + s.m[j];
      }
      const double den = M + msum;
      const V3 sh{px / den, py / den, pz / den};
      for (int j = 0; j < NM; ++j) {
        const double *xx = mx + ((size_t)si * NM + j) * 3, *vv = mv + ((size_t)si * NM + j) * 3;
        s.mb.push_back({{xx[0], xx[1], xx[2]}, {vv[0] - sh.x, vv[1] - sh.y, vv[2] - sh.z}});
      }
      for (int64_t i = tp_off[si]; i < tp_off[si + 1]; ++i)
        s.tp.push_back({{tx[i], ty[i], tz[i]}, {tvx[i] - sh.x, tvy[i] - sh.y, tvz[i] - sh.z}});
      uint32_t st = 0;
      for (int64_t n = 0; n < nsteps; ++n) st |= step(s, p, M, tstats[tid], hits ? hits + tp_off[si] : nullptr);
      // back to heliocentric: v + (sum m v_dh)/M
      double qx = 0, qy = 0, qz = 0;
      for (int j = 0; j < NM; ++j) {
        qx = std::fma(s.m[j], s.mb[j].v.x, qx);
        qy = std::fma(s.m[j], s.mb[j].v.y, qy);
        qz = std::fma(s.m[j], s.mb[j].v.z, qz);
      }
      const V3 rs{qx / M, qy / M, qz / M};
      bool fin = true;
      for (int j = 0; j < NM; ++j) {
        double *xx = mx + ((size_t)si * NM + j) * 3, *vv = mv + ((size_t)si * NM + j) * 3;
        xx[0] = s.mb[j].x.x; xx[1] = s.mb[j].x.y; xx[2] = s.mb[j].x.z;
        vv[0] = s.mb[j].v.x + rs.x; vv[1] = s.mb[j].v.y + rs.y; vv[2] = s.mb[j].v.z + rs.z;
      }
      int64_t i = tp_off[si];
      for (const Body &t : s.tp) {
        tx[i] = t.x.x; ty[i] = t.x.y; tz[i] = t.x.z;
        tvx[i] = t.v.x + rs.x; tvy[i] = t.v.y + rs.y; tvz[i] = t.v.z + rs.z;
        fin = fin && std::isfinite(t.x.x) && std::isfinite(t.x.y) && std::isfinite(t.x.z) && std::isfinite(t.v.x) &&
              std::isfinite(t.v.y) && std::isfinite(t.v.z);
        ++i;
      }
      for (const Body &b : s.mb)
        fin = fin && std::isfinite(b.x.x) && std::isfinite(b.x.y) && std::isfinite(b.x.z) && std::isfinite(b.v.x) &&
              std::isfinite(b.v.y) && std::isfinite(b.v.z);
      if (!fin) st |= 1u;
      if (status) status[si] = st;
    }
  };
  if (nthr == 1) {
    work(0);
  } else {
    std::vector<std::thread> th;
    for (int t = 0; t < nthr; ++t) th.emplace_back(work, t);
    for (auto &t : th) t.join();
  }
  if (stats) {
    std::memset(stats, 0, sizeof(*stats));
    for (const Stats &s : tstats) {
      stats->tp_steps += s.tp_steps;
      stats->kepler_solves += s.kepler;
      stats->newton_iters += s.newton;
      stats->bs_steps += s.bs_steps;
      stats->bs_rhs_evals += s.bs_rhs;
      stats->massive_groups += s.mgroups;
      stats->massive_group_bodies += s.mgroup_bodies;
      stats->overlaps += s.overlaps;
      stats->massive_kepler_solves += s.mkepler;
      stats->massive_newton_iters += s.mnewton;
    }
  }
  return 0;
}

// Energy of the massive bodies from heliocentric (x, v):
// E = sum m v_dh^2/2 + |sum m v_dh|^2/(2M) - sum M m/r + sum_{j<k} m_j m_k U(r_jk),
// with v_dh = v - (sum m v)/(M + sum m) and U the pair potential whose force is
// the kick's m_j m_k d/((r^2 + eps^2) r): U(r) = -atan(eps/r)/eps (-1/r when
// eps = 0).
double pair_potential(double r, double eps) { return eps > 0.0 ? -std::atan(eps / r) / eps : -1.0 / r; }

void orc_energy(int32_t S, int32_t NM, const double *m, const double *mx, const double *mv, double M, double eps2,
                double *E) {
  const double eps = std::sqrt(eps2);
  for (int s = 0; s < S; ++s) {
    const double *mm = m + (size_t)s * NM;
    double px = 0, py = 0, pz = 0, msum = 0;
    for (int j = 0; j < NM; ++j) {
      const double *vv = mv + ((size_t)s * NM + j) * 3;
      px = std::fma(mm[j], vv[0], px);
      py = std::fma(mm[j], vv[1], py);
      pz = std::fma(mm[j], vv[2], pz);
      msum = msum + mm[j];
    }
    const double den = M + msum;
    const V3 sh{px / den, py / den, pz / den};
    double kin = 0, pot = 0, qx = 0, qy = 0, qz = 0;
    std::vector<V3> x(NM), v(NM);
    for (int j = 0; j < NM; ++j) {
      const double *xx = mx + ((size_t)s * NM + j) * 3, *vv = mv + ((size_t)s * NM + j) * 3;
      x[j] = {xx[0], xx[1], xx[2]};
      v[j] = {vv[0] - sh.x, vv[1] - sh.y, vv[2] - sh.z};
    }
    for (int j = 0; j < NM; ++j) {
      kin += 0.5 * mm[j] * dot(v[j], v[j]);
      qx = std::fma(mm[j], v[j].x, qx);
      qy = std::fma(mm[j], v[j].y, qy);
      qz = std::fma(mm[j], v[j].z, qz);
      pot -= M * mm[j] / std::sqrt(dot(x[j], x[j]));
      for (int k = j + 1; k < NM; ++k) {
        const V3 d = sub(x[k], x[j]);
        pot += mm[j] * mm[k] * pair_potential(std::sqrt(dot(d, d)), eps);
      }
    }
    const double p2 = std::fma(qx, qx, std::fma(qy, qy, qz * qz));
    E[s] = kin + p2 / (2.0 * M) + pot;
  }
}

// Kick accelerations of NM massive bodies (heliocentric x, body-major xyz).
void orc_accel_massive(int32_t NM, const double *m, const double *rc, const double *x, double eps2, double *a) {
  Sys s;
  s.nm = NM;
  s.m.assign(m, m + NM);
  for (int j = 0; j < NM; ++j) s.rc_tp.push_back(make_rc(rc[j]));
  std::vector<V3> X(NM);
  for (int j = 0; j < NM; ++j) X[j] = {x[3 * j], x[3 * j + 1], x[3 * j + 2]};
  for (int j = 0; j < NM; ++j) {
    const V3 q = accel_one(j, X, s, eps2);
    a[3 * j] = q.x;
    a[3 * j + 1] = q.y;
    a[3 * j + 2] = q.z;
  }
}

// Kick acceleration of a test particle at x from NM massive bodies at X
// (body-major xyz), changeover radii rc[NM] (body_interaction_kick for a
// massless body, src/sim.cuh:327-382).
int orc_accel_tp(int32_t NM, const double *m, const double *rc, const double *X, const double *x, double eps2,
                 double *a) {
  if (NM < 1 || NM > MAXB) return -1;
  Sys s;
  s.nm = NM;
  s.m.assign(m, m + NM);
  for (int j = 0; j < NM; ++j) s.rc_tp.push_back(make_rc(rc[j]));
  std::vector<V3> P(NM);
  for (int j = 0; j < NM; ++j) P[j] = {X[3 * j], X[3 * j + 1], X[3 * j + 2]};
  const V3 q = accel_tp({x[0], x[1], x[2]}, P, s, eps2);
  a[0] = q.x;
  a[1] = q.y;
  a[2] = q.z;
  return 0;
}

// Encounter groups of NM massive bodies at positions x: gid[j] = group number
// (groups numbered by smallest member) or -1.  Returns the number of groups;
// *overlaps counts pairs closer than the sum of their radii (rad may be null).
int orc_encounter_groups(int32_t NM, const double *rc, const double *rad, const double *x, int32_t *gid,
                         int64_t *overlaps) {
  Sys s;
  s.nm = NM;
  s.m.assign(NM, 1.0);
  for (int j = 0; j < NM; ++j) s.rc_tp.push_back(make_rc(rc[j]));
  if (rad) s.rad.assign(rad, rad + NM);
  s.mb.resize(NM);
  for (int j = 0; j < NM; ++j) s.mb[j].x = {x[3 * j], x[3 * j + 1], x[3 * j + 2]};
  int64_t ov = 0;
  const std::vector<std::vector<int>> g = encounter_groups(s, ov);
  for (int j = 0; j < NM; ++j) gid[j] = -1;
  for (size_t q = 0; q < g.size(); ++q)
    for (int j : g[q]) gid[j] = (int32_t)q;
  if (overlaps) *overlaps = ov;
  return (int)g.size();
}

}  // extern "C"
