// mint_device.h — FP64 device numerics of the Mercurius step for gfx950.
//
// Every function here is plain IEEE arithmetic: +, -, *, /, sqrt and explicit
// fma, compiled with -ffp-contract=off, so the same operation sequence gives
// the same bits on the host.  No transcendental function is called on the
// step path (the Stumpff functions are a fixed-length series with argument
// quartering), which is what lets a host FP64 restatement match bit for bit.
//
// Reference functions each piece replaces are cited as file:line of
// at2005/p9-integrator.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace mint {

constexpr int KMAX = 8;            // BS rows, substeps N_k = 2^(k+1) = 2..256
constexpr int BS_MAXDEPTH = 24;    // interval halvings before giving up (then the rest in one step)
constexpr int NEWTON_MAXIT = 40;   // Kepler Newton iteration cap

constexpr uint32_t ST_NAN = 1u, ST_NEWTON = 2u, ST_BS_DEPTH = 4u, ST_MASSIVE_ENC = 8u;
constexpr uint32_t ST_OVERLAP = 16u, ST_LIMIT = 32u;

// Sums over massive bodies run in chunks of MB_CHUNK bodies in index order:
// each chunk's sum starts from zero and the chunk sums are folded by halving
// strides (slot l takes slot l + h for h = P/2, ..., 1 over the chunk count
// rounded up to a power of two P, missing chunks +0; one chunk: its own sum).  Kick weights are divided in groups of MB_DIVG bodies with one
// division per group (batched_div), the body itself an inert slot.  Systems of
// up to 8 massive bodies are one chunk and one division group.
constexpr int MB_CHUNK = 16;
constexpr int MB_DIVG = 8;

// Per-system parameter block (doubles), NM massive bodies:
//   [0,NM)           m_j
//   [NM,2NM)         0.1 r_c,j            TP-body changeover, r_c = r_c,j
//   [2NM,3NM)        1/(0.9 r_c,j)
//   [3NM,4NM)        r_c,j^2
//   [4NM, 4NM+NM^2)          pair 0.1 r_c   massive-massive, r_c = max(r_c,j, r_c,k)
//   [4NM+NM^2, 4NM+2NM^2)    pair 1/(0.9 r_c)
//   [4NM+2NM^2, 4NM+3NM^2)   pair r_c^2
//   [4NM+3NM^2, 5NM+3NM^2)   physical radius R_j (0 = none)
__host__ __device__ constexpr int prm_size(int nm) { return ((6 * nm + 3 * nm * nm) + 1) & ~1; }
__host__ __device__ constexpr int P_M(int) { return 0; }
__host__ __device__ constexpr int P_RCLO(int nm) { return nm; }
__host__ __device__ constexpr int P_RCWINV(int nm) { return 2 * nm; }
__host__ __device__ constexpr int P_RC2(int nm) { return 3 * nm; }
__host__ __device__ constexpr int P_PRCLO(int nm) { return 4 * nm; }
__host__ __device__ constexpr int P_PRCWINV(int nm) { return 4 * nm + nm * nm; }
__host__ __device__ constexpr int P_PRC2(int nm) { return 4 * nm + 2 * nm * nm; }
__host__ __device__ constexpr int P_RAD(int nm) { return 4 * nm + 3 * nm * nm; }
// Far radius r_c^2 (1 + 1e-3) of body j.  A TP whose kick finds it farther
// than this (i) has K = 1 whatever the rounding of y = (r - 0.1 r_c) /
// (0.9 r_c) (the computed y exceeds 1 by ~5e-4), so the kick skips forming
// y; and (ii) is outside r_c at the next step's encounter check: that check
// measures the same separation after both ends moved by the same jump, which
// changes it by a few ulps of the positions (far below 5e-4 r_c for bodies
// within 1e8 AU and r_c above 2e-4 AU).
__host__ __device__ constexpr int P_FAR2(int nm) { return 5 * nm + 3 * nm * nm; }

// Perturber cache: for every system and step, the massive positions
// Kepler-drifted from (Y, W) by tau = m (h/16), m = 1..16, the substep times of
// BS rows N = 2, 4, 8, 16 over the undivided step.  A BS right-hand side at one
// of these times reads them instead of solving Kepler's equation (the same
// function of the same inputs, so the same bits).  [s][k][m-1][3 NM] doubles.
constexpr int PC_PTS = 16;
__host__ __device__ constexpr int pc_size(int nm) { return PC_PTS * 3 * nm; }

// Ephemeris record of one step (doubles): what every TP phase of that step
// reads from the massive subsystem (the per-phase reads of block 0's shared
// memory in src/sim.cuh:682-741).
//   [0,3)  Q1 = (sum m_j v_j)/M after the first kick      (jump 1)
//   [3,6)  Q2 = (sum m_j v_j)/M after the drift            (jump 2)
//   [6, 6+3NM)       Y: massive positions after jump 1     (encounter test, BS)
//   [6+3NM, 6+6NM)   W: massive velocities at drift start  (BS perturber drift)
//   [6+6NM, 6+9NM)   Xe: massive positions after jump 2    (closing kick)
__host__ __device__ constexpr int rec_size(int nm) { return ((6 + 9 * nm) + 1) & ~1; }
__host__ __device__ constexpr int R_Q1(int) { return 0; }
__host__ __device__ constexpr int R_Q2(int) { return 3; }
__host__ __device__ constexpr int R_Y(int) { return 6; }
__host__ __device__ constexpr int R_W(int nm) { return 6 + 3 * nm; }
__host__ __device__ constexpr int R_XE(int nm) { return 6 + 6 * nm; }

struct StepConst {
  double h;     // step
  double hh;    // h/2
  double mu;    // G*M
  double gate_ir0;  // 1/r0 at or below which the learned Kepler guess's gate stays shut
  double eps2;  // pair softening
  double tol;   // BS tolerance
};

// Square root of a positive finite x well above 2^-767 (a squared distance in
// AU always is): the correctly rounded result by the same reciprocal-square-
// root refinement the compiler uses for sqrt, without its rescaling of inputs
// below 2^-767.  Zero and infinity give NaN.
__device__ __forceinline__ double sqrt_dist2_pos(double x) {
  const double y = __builtin_amdgcn_rsq(x);
  double g = x * y, h = y * 0.5;
  const double r = fma(-h, g, 0.5);
  g = fma(g, r, g);
  h = fma(h, r, h);
  double d = fma(-g, g, x);
  g = fma(d, h, g);
  d = fma(-g, g, x);
  return fma(d, h, g);
}

// Square root of a squared distance: sqrt_dist2_pos, with zero and infinity
// returned as they are.
__device__ __forceinline__ double sqrt_dist2(double x) {
  const double g = sqrt_dist2_pos(x);
  return (x == 0.0 || x == INFINITY) ? x : g;
}

// Quotient a / b of operands well inside the normal range (|a|, |b| and the
// quotient between about 2^-900 and 2^900, as every quotient on the step path
// is): the correctly rounded result by the same Newton-Raphson refinement of
// the hardware reciprocal the compiler uses for a / b, without its operand
// rescaling and special-value fix-up, which change nothing there.
__device__ __forceinline__ double div_normal(double a, double b) {
  double y = __builtin_amdgcn_rcp(b);
  double e = fma(-b, y, 1.0);
  y = fma(y, e, y);
  e = fma(-b, y, 1.0);
  y = fma(y, e, y);
  const double q = a * y;
  const double r = fma(-b, q, a);
  return fma(r, y, q);
}

__device__ __forceinline__ double dot3(double ax, double ay, double az, double bx, double by,
                                       double bz) {
  return fma(ax, bx, fma(ay, by, az * bz));
}

// Coefficients of the series bands above |z| = 1e-5, highest degree first:
// c2's (1/12!, ..., 1/6! with alternating signs), c3's (1/13!, ..., 1/7!),
// then c2's and c3's degree-1 terms -1/4! and -1/5!.  Read from constant
// memory where a band needs them, so that kernels looping over drifts do not
// hold them in vector registers.
static __constant__ double k_stumpff_hi[12] = {
    1.1470745597729725e-11, -2.08767569878681e-09, 2.755731922398589e-07, -2.48015873015873e-05,
    0.001388888888888889,   7.647163731819816e-13, -1.6059043836821613e-10, 2.505210838544172e-08,
    -2.7557319223985893e-06, 0.0001984126984126984, -0.041666666666666664, -0.008333333333333333};

// c0..c3 by the series alone, for |z| <= 0.1.
__device__ __forceinline__ void stumpff_series(double z, double &c0, double &c1, double &c2,
                                               double &c3) {
  // Series length by |z| band: the first omitted term is below 3e-19 of the
  // sum in every band (|z| <= 1e-5: degree 2, <= 1e-3: degree 3, else 6).
  // The degree-2 band (every drift of a distant body) is formed first and
  // without a branch; the others replace it.
  const double az = fabs(z);
  double t2 = fma(0.001388888888888889, z, -0.041666666666666664);
  double t3 = fma(0.0001984126984126984, z, -0.008333333333333333);
  if (!(az <= 1e-5)) {
    // the table's address laundered here, so its loads stay in this branch
    const double *tb = k_stumpff_hi;
    asm volatile("" : "+s"(tb));
    if (az <= 1e-3) {
      t2 = fma(tb[3], z, tb[4]);
      t3 = fma(tb[8], z, tb[9]);
    } else {
      t2 = fma(tb[0], z, tb[1]);
      t2 = fma(t2, z, tb[2]);
      t2 = fma(t2, z, tb[3]);
      t2 = fma(t2, z, tb[4]);
      t3 = fma(tb[5], z, tb[6]);
      t3 = fma(t3, z, tb[7]);
      t3 = fma(t3, z, tb[8]);
      t3 = fma(t3, z, tb[9]);
    }
    t2 = fma(t2, z, tb[10]);
    t3 = fma(t3, z, tb[11]);
  }
  t2 = fma(t2, z, 0.5);
  t3 = fma(t3, z, 0.16666666666666666);
  c2 = t2;
  c3 = t3;
  c1 = fma(-z, c3, 1.0);
  c0 = fma(-z, c2, 1.0);
}

// Stumpff functions c0..c3(z) (Danby 1988 §6.9; Stiefel & Scheifele 1971):
// quarter z until |z| <= 0.1, sum c2 and c3 as 7-term series, then undo the
// quartering with c0(4z)=2c0^2-1, c1(4z)=c0c1, c2(4z)=c1^2/2,
// c3(4z)=(c2+c0c3)/4.  Replaces the element round trip and the fixed
// 8-iteration Kepler-equation solve of src/sim.cuh:90-126,188-324.
// |z| <= 0.1 (and NaN) goes to the series directly.
__device__ __forceinline__ void stumpff(double z, double &c0, double &c1, double &c2, double &c3) {
  if (!(fabs(z) > 0.1)) {
    stumpff_series(z, c0, c1, c2, c3);
    return;
  }
  int n = 0;
  while (fabs(z) > 0.1 && n < 64) {
    z *= 0.25;
    ++n;
  }
  stumpff_series(z, c0, c1, c2, c3);
  for (; n > 0; --n) {
    const double n3 = fma(c0, c3, c2) * 0.25;
    const double n2 = (c1 * c1) * 0.5;
    const double n1 = c0 * c1;
    const double n0 = fma(2.0 * c0, c0, -1.0);
    c0 = n0;
    c1 = n1;
    c2 = n2;
    c3 = n3;
  }
}

// Quantities of the universal-variable Kepler problem at the drift start.
struct KepStart {
  double r0, ir0, eta0, beta, zeta0;  // ir0 = 1/r0
};

__device__ __forceinline__ KepStart kep_start(double x, double y, double z, double vx, double vy,
                                              double vz, double mu) {
  KepStart k;
  // r0 = 0 or infinity gives a NaN 1/r0 (so a NaN drift) with either square
  // root, so the bare refinement is enough
  k.r0 = sqrt_dist2_pos(dot3(x, y, z, x, y, z));
  k.ir0 = div_normal(1.0, k.r0);
  const double v2 = dot3(vx, vy, vz, vx, vy, vz);
  k.eta0 = dot3(x, y, z, vx, vy, vz);
  k.beta = fma(2.0 * mu, k.ir0, -v2);
  k.zeta0 = fma(-k.beta, k.r0, mu);
  return k;
}

// Starting guesses for the universal variable s.
__device__ __forceinline__ double guess_crude(const KepStart &k, double dt) { return dt * k.ir0; }
// Third-order series inversion of dt = r0 s + eta0 s^2/2 + zeta0 s^3/6 + O(s^4):
// s = u (1 + a u + b u^2), u = dt/r0, a = -eta0/(2 r0), b = eta0^2/(2 r0^2) - zeta0/(6 r0).
__device__ __forceinline__ double guess_series(const KepStart &k, double dt) {
  const double u = dt * k.ir0;
  const double e = k.eta0 * k.ir0;
  const double a = -0.5 * e;
  const double b = fma(0.5 * e, e, -(k.zeta0 * k.ir0) * (1.0 / 6.0));
  return u * fma(u, fma(u, b, a), 1.0);
}

// Universal-variable Kepler drift by dt about mu (Danby 1988 §6.9; the
// G-function form of Mikkola & Innanen 1999).  Newton on
//   F(s) = r0 s + eta0 G2 + zeta0 G3 - dt,  F' = r(s) = r0 + eta0 G1 + zeta0 G2,
// stopping when the predicted post-update error |F''| ds^2 / (2 F') is below
// 2^-52 |s| and |ds| <= 2e-6 |s|; the G-functions are then carried to s - ds by
// a second-order Taylor step (G_k' = G_{k-1}) instead of a fresh evaluation.
// A step that leaves the bracket kept from the signs of F is replaced by a
// Laguerre-Conway step, or by bisection.
// Returns the Newton iteration count; ok=false if the cap was hit.
struct KepSol {
  double G0, G1, G2, G3;  // G-functions at s
  double ds;              // last Newton correction; the root is s - ds
  double s;
  int it;
  bool ok;
};

// One Newton evaluation at s: G-functions, F, F', F'' and the correction ds;
// true when the stopping rule holds.
__device__ __forceinline__ bool kepler_eval(const KepStart &k, double s, double dt, KepSol &o,
                                            double &F, double &rp) {
  double c0, c1, c2, c3;
  const double ss = s * s;
  stumpff(k.beta * ss, c0, c1, c2, c3);
  o.G0 = c0;
  o.G1 = s * c1;
  o.G2 = ss * c2;
  o.G3 = ss * (s * c3);
  F = fma(k.r0, s, fma(k.eta0, o.G2, fma(k.zeta0, o.G3, -dt)));
  rp = fma(k.eta0, o.G1, fma(k.zeta0, o.G2, k.r0));
  const double rpp = fma(k.eta0, o.G0, k.zeta0 * o.G1);
  o.ds = div_normal(F, rp);
  const double as = fabs(s);
  return fabs(o.ds) <= 2e-6 * as && fabs(rpp) * (o.ds * o.ds) <= 4.4e-16 * (fabs(rp) * as);
}

// The first iteration is written out (it is the last one for almost every
// drift), the rest loop.
__device__ __forceinline__ KepSol kepler_solve(const KepStart &k, double s, double dt) {
  KepSol o;
  double F, rp;
  o.it = 1;
  o.ok = kepler_eval(k, s, dt, o, F, rp);
  if (!o.ok) {
    // F is increasing (F' = r > 0) and F(0) = -dt, so the root lies in (lo, hi).
    double lo = dt >= 0.0 ? 0.0 : -INFINITY;
    double hi = dt >= 0.0 ? INFINITY : 0.0;
    while (o.it < NEWTON_MAXIT) {
      if (F < 0.0)
        lo = s;
      else
        hi = s;
      double sn = s - o.ds;
      if (!(sn > lo && sn < hi)) {
        // Newton left the bracket: Laguerre-Conway step (n = 5, Conway 1986),
        // then bisection if that also leaves it.
        const double rpp = fma(k.eta0, o.G0, k.zeta0 * o.G1);
        const double d2 = fabs(fma(16.0 * rp, rp, -(20.0 * F * rpp)));
        const double den = rp + copysign(sqrt(d2), rp);
        sn = s - (5.0 * F) / den;
        if (!(sn > lo && sn < hi)) sn = 0.5 * (lo + hi);
      }
      s = sn;
      ++o.it;
      if (kepler_eval(k, s, dt, o, F, rp)) {
        o.ok = true;
        break;
      }
    }
  }
  o.s = s;
  return o;
}

__device__ __forceinline__ int kepler_drift_from(const KepStart &k, double s, double &x, double &y,
                                                 double &z, double &vx, double &vy, double &vz,
                                                 double mu, double dt, bool &ok) {
  const KepSol o = kepler_solve(k, s, dt);
  ok = o.ok;
  const double hs = -o.ds;
  const double hh2 = 0.5 * (hs * hs);
  const double t0 = o.G0 * hh2;
  const double t1 = o.G1 * hh2;
  const double G1n = o.G1 + fma(o.G0, hs, -(k.beta * t1));
  const double G2n = o.G2 + fma(o.G1, hs, t0);
  const double G3n = o.G3 + fma(o.G2, hs, t1);
  const double r = fma(k.eta0, G1n, fma(k.zeta0, G2n, k.r0));
  const double ir = div_normal(1.0, r);
  const double fm1 = -(mu * G2n) * k.ir0;
  const double g = fma(-mu, G3n, dt);
  const double fdot = -(mu * G1n) * (k.ir0 * ir);
  const double gm1 = -(mu * G2n) * ir;
  const double nx = x + fma(fm1, x, g * vx);
  const double ny = y + fma(fm1, y, g * vy);
  const double nz = z + fma(fm1, z, g * vz);
  const double nvx = vx + fma(fdot, x, gm1 * vx);
  const double nvy = vy + fma(fdot, y, gm1 * vy);
  const double nvz = vz + fma(fdot, z, gm1 * vz);
  x = nx;
  y = ny;
  z = nz;
  vx = nvx;
  vy = nvy;
  vz = nvz;
  return o.it;
}

// Kepler drift with the series guess (massive bodies, BS perturbers).
__device__ __forceinline__ int kepler_drift_series(double &x, double &y, double &z, double &vx,
                                                   double &vy, double &vz, double mu, double dt,
                                                   bool &ok) {
  const KepStart k = kep_start(x, y, z, vx, vy, vz, mu);
  return kepler_drift_from(k, guess_series(k, dt), x, y, z, vx, vy, vz, mu, dt, ok);
}

// Changeover-weighted pair acceleration factor (src/sim.cuh:166-186,327-382):
// y = (r - 0.1 r_c)/(0.9 r_c); K = 0 (y<=0), y^2/(2y^2-2y+1) (0<y<1), 1 (y>=1).
// The kick weight is K * m / ((r^2 + eps^2) r) = num / den.
// r = |r2|^(1/2) is passed in.
__device__ __forceinline__ void kick_nd_r(double r2, double r, double eps2, double m, double rclo,
                                          double rcwinv, double &num, double &den) {
  const double D = (r2 + eps2) * r;
  const double yv = (r - rclo) * rcwinv;
  num = m;
  den = D;
  if (yv <= 0.0) {
    num = 0.0;
  } else if (yv < 1.0) {
    const double y2 = yv * yv;
    num = m * y2;
    den = D * fma(2.0, y2, fma(-2.0, yv, 1.0));
  }
}

__device__ __forceinline__ void kick_nd(double r2, double eps2, double m, double rclo,
                                        double rcwinv, double &num, double &den) {
  kick_nd_r(r2, sqrt_dist2(r2), eps2, m, rclo, rcwinv, num, den);
}

// kick_nd with the square root of a positive finite r2.  Where the weights go
// on to one batched division (accel_tp), r2 = 0 or infinity gives NaN weights
// with either root (the reciprocal of a zero or infinite product is NaN), so
// the zero and infinity cases of sqrt_dist2 change nothing there.
__device__ __forceinline__ void kick_nd_pos(double r2, double eps2, double m, double rclo,
                                            double rcwinv, double &num, double &den) {
  kick_nd_r(r2, sqrt_dist2_pos(r2), eps2, m, rclo, rcwinv, num, den);
}

// w_i = num_i / den_i for n terms with one division (batched reciprocal):
// P_i = den_0 ... den_i; inv = 1/P_{n-1}; for i = n-1..1: w_i = num_i (inv
// P_{i-1}), inv *= den_i; w_0 = num_0 inv.
template <int N>
__device__ __forceinline__ void batched_div(const double *num, const double *den, double *w) {
  double P[N];
  P[0] = den[0];
#pragma unroll
  for (int i = 1; i < N; ++i) P[i] = P[i - 1] * den[i];
  double inv = div_normal(1.0, P[N - 1]);
#pragma unroll
  for (int i = N - 1; i >= 1; --i) {
    w[i] = num[i] * (inv * P[i - 1]);
    inv = inv * den[i];
  }
  w[0] = num[0] * inv;
}

// Close-encounter weight (1 - K) m / ((r^2 + eps^2) r) = (1-y)^2/(2y^2-2y+1) ...
// zero for y >= 1 (replaces the double-counting weight of src/sim.cuh:357-361).
__device__ __forceinline__ double close_weight(double r2, double eps2, double m, double rclo,
                                               double rcwinv) {
  const double r = sqrt_dist2(r2);
  const double yv = (r - rclo) * rcwinv;
  if (yv >= 1.0) return 0.0;
  const double D = (r2 + eps2) * r;
  double num = m, den = D;
  if (yv > 0.0) {
    const double om = 1.0 - yv;
    num = m * (om * om);
    den = D * fma(2.0, yv * yv, fma(-2.0, yv, 1.0));
  }
  return div_normal(num, den);
}

// close_weight formed without branches: both forms are evaluated and
// selected (the same values, NaN inputs included), so that it interleaves
// with independent work instead of splitting the wavefront's instruction
// stream at divergent branches.
__device__ __forceinline__ double close_weight_sel(double r2, double eps2, double m, double rclo,
                                                   double rcwinv) {
  const double r = sqrt_dist2(r2);
  const double yv = (r - rclo) * rcwinv;
  const double D = (r2 + eps2) * r;
  const double om = 1.0 - yv;
  const double nb = m * (om * om), db = D * fma(2.0, yv * yv, fma(-2.0, yv, 1.0));
  const bool band = yv > 0.0;
  const double w = div_normal(band ? nb : m, band ? db : D);
  return yv >= 1.0 ? 0.0 : w;
}

// TP kick acceleration from the NM massive bodies at positions X (body-major
// xyz).  Replaces body_interaction_kick for a massless body (src/sim.cuh:327-382).
// far: every pair is beyond its far radius (see P_FAR2).
template <int NM>
__device__ __forceinline__ void accel_tp(double x, double y, double z,
                                         const double *__restrict__ X,
                                         const double *__restrict__ prm, double eps2,
                                         double &ax, double &ay, double &az, bool &far) {
  double dx[NM], dy[NM], dz[NM], num[NM], den[NM], w[NM];
  far = true;
#pragma unroll
  for (int j = 0; j < NM; ++j) {
    dx[j] = X[3 * j + 0] - x;
    dy[j] = X[3 * j + 1] - y;
    dz[j] = X[3 * j + 2] - z;
    const double r2 = dot3(dx[j], dy[j], dz[j], dx[j], dy[j], dz[j]);
    if (r2 > prm[P_FAR2(NM) + j]) {  // K = 1: the weight kick_nd forms there
      num[j] = prm[P_M(NM) + j];
      den[j] = (r2 + eps2) * sqrt_dist2_pos(r2);
    } else {
      far = false;
      kick_nd_pos(r2, eps2, prm[P_M(NM) + j], prm[P_RCLO(NM) + j], prm[P_RCWINV(NM) + j], num[j],
                  den[j]);
    }
  }
  batched_div<NM>(num, den, w);
  ax = 0.0;
  ay = 0.0;
  az = 0.0;
#pragma unroll
  for (int j = 0; j < NM; ++j) {
    ax = fma(w[j], dx[j], ax);
    ay = fma(w[j], dy[j], ay);
    az = fma(w[j], dz[j], az);
  }
}

template <int NM>
__device__ __forceinline__ void accel_tp(double x, double y, double z,
                                         const double *__restrict__ X,
                                         const double *__restrict__ prm, double eps2,
                                         double &ax, double &ay, double &az) {
  bool far;
  accel_tp<NM>(x, y, z, X, prm, eps2, ax, ay, az, far);
}

// The exact distance check: is the TP within r_c,j of any massive body?
// (close_encounter_p, src/sim.cuh:589-613, on squared distances.)
template <int NM>
__device__ __forceinline__ bool encounter_tp(double x, double y, double z,
                                             const double *__restrict__ Y,
                                             const double *__restrict__ prm) {
  // every pair, no short-circuit branches: r2 < rc2 exactly when r2 - rc2 < 0
  // (the difference of two unequal finite doubles is never rounded to zero),
  // so the smallest difference decides; a NaN pair never counts, as in the
  // comparison (fmin drops it)
  double m = INFINITY;
#pragma unroll
  for (int j = 0; j < NM; ++j) {
    const double dx = Y[3 * j + 0] - x;
    const double dy = Y[3 * j + 1] - y;
    const double dz = Y[3 * j + 2] - z;
    const double r2 = dot3(dx, dy, dz, dx, dy, dz);
    m = fmin(m, r2 - prm[P_RC2(NM) + j]);
  }
  return m < 0.0;
}

// Right-hand side of the close-encounter problem of one TP at time tau into
// the drift: Kepler force of the (unsoftened) Sun plus (1-K)-weighted pulls of
// the massive bodies, each on its own Kepler orbit from (Y_j, W_j).
// (update_my_velocity_total, src/sim.cuh:410-443, with moving perturbers.)
//
// G > 1: G lanes evaluate the right-hand side of one TP together (lanes
// base .. base+G-1 of the wave, this lane is base+jl, all holding the same
// TP state): lane jl places perturber jl and forms its pull, then every lane
// adds the NM pulls in body order from shuffles, so each lane gets the value
// the serial loop (G = 1) computes, bit for bit.
// Index 1..PC_PTS of the perturber-cache point at time tau into the step, or
// 0 when tau is not exactly one of them.
__device__ __forceinline__ int cache_point(double tau, double h) {
  const double hs16 = h / (double)PC_PTS;
  int ci = (int)(tau * ((double)PC_PTS / h) + 0.5);
  if (!(ci >= 1 && ci <= PC_PTS && (double)ci * hs16 == tau)) ci = 0;
  return ci;
}

// Position of perturber j at time tau into the drift: the cache point when
// there is one, else its Kepler drift from (Y_j, W_j).
template <int NM>
__device__ __forceinline__ void perturber_pos(int j, double tau, int ci, const double *__restrict__ Y,
                                              const double *__restrict__ W,
                                              const double *__restrict__ P4, double mu, double &px,
                                              double &py, double &pz) {
  px = Y[3 * j + 0];
  py = Y[3 * j + 1];
  pz = Y[3 * j + 2];
  if (ci > 0 && P4 != nullptr) {
    px = P4[(ci - 1) * 3 * NM + 3 * j + 0];
    py = P4[(ci - 1) * 3 * NM + 3 * j + 1];
    pz = P4[(ci - 1) * 3 * NM + 3 * j + 2];
  } else if (tau != 0.0) {
    double qx = W[3 * j + 0], qy = W[3 * j + 1], qz = W[3 * j + 2];
    bool ok;
    kepler_drift_series(px, py, pz, qx, qy, qz, mu, tau, ok);
  }
}

// (1-K)-weighted pull of perturber j at (px, py, pz) on the TP at (x, y, z)
// (SEL: the weight formed without branches, same values).
template <int NM, bool SEL = false>
__device__ __forceinline__ void pull_at(int j, double px, double py, double pz, double x, double y,
                                        double z, const double *__restrict__ prm, double eps2,
                                        double &w, double &dx, double &dy, double &dz) {
  dx = px - x;
  dy = py - y;
  dz = pz - z;
  const double d2 = dot3(dx, dy, dz, dx, dy, dz);
  if (SEL)
    w = close_weight_sel(d2, eps2, prm[P_M(NM) + j], prm[P_RCLO(NM) + j], prm[P_RCWINV(NM) + j]);
  else
    w = close_weight(d2, eps2, prm[P_M(NM) + j], prm[P_RCLO(NM) + j], prm[P_RCWINV(NM) + j]);
}

template <int NM, bool SEL = false>
__device__ __forceinline__ void perturber_pull(int j, double tau, double x, double y, double z,
                                               const double *__restrict__ Y,
                                               const double *__restrict__ W,
                                               const double *__restrict__ P4, int ci,
                                               const double *__restrict__ prm, double mu,
                                               double eps2, double &w, double &dx, double &dy,
                                               double &dz) {
  double px, py, pz;
  perturber_pos<NM>(j, tau, ci, Y, W, P4, mu, px, py, pz);
  pull_at<NM, SEL>(j, px, py, pz, x, y, z, prm, eps2, w, dx, dy, dz);
}

// Grouped right-hand side (G lanes per TP, see rhs_tp) with this lane's
// perturber already placed at (px, py, pz).
template <int NM>
__device__ __forceinline__ void rhs_tp_at(const double *yv, double px, double py, double pz,
                                          const double *__restrict__ prm, double mu, double eps2,
                                          double *f, int jl, int base) {
  const double x = yv[0], y = yv[1], z = yv[2];
  f[0] = yv[3];
  f[1] = yv[4];
  f[2] = yv[5];
  const double r2 = dot3(x, y, z, x, y, z);
  const double r = sqrt_dist2(r2);
  const double w0 = div_normal(mu, r2 * r);
  double ax = -(w0 * x), ay = -(w0 * y), az = -(w0 * z);
  double w, dx, dy, dz;
  pull_at<NM, true>(jl < NM ? jl : NM - 1, px, py, pz, x, y, z, prm, eps2, w, dx, dy, dz);
#pragma unroll
  for (int j = 0; j < NM; ++j) {
    const double wj = __shfl(w, base + j, 64);
    ax = fma(wj, __shfl(dx, base + j, 64), ax);
    ay = fma(wj, __shfl(dy, base + j, 64), ay);
    az = fma(wj, __shfl(dz, base + j, 64), az);
  }
  f[3] = ax;
  f[4] = ay;
  f[5] = az;
}

template <int NM, int G = 1>
__device__ __forceinline__ void rhs_tp(double tau, const double *yv, const double *__restrict__ Y,
                                       const double *__restrict__ W,
                                       const double *__restrict__ P4, double h,
                                       const double *__restrict__ prm, double mu, double eps2,
                                       double *f, int jl = 0, int base = 0) {
  const double x = yv[0], y = yv[1], z = yv[2];
  f[0] = yv[3];
  f[1] = yv[4];
  f[2] = yv[5];
  const double r2 = dot3(x, y, z, x, y, z);
  const double r = sqrt_dist2(r2);
  const double w0 = div_normal(mu, r2 * r);
  double ax = -(w0 * x), ay = -(w0 * y), az = -(w0 * z);
  const int ci = cache_point(tau, h);
  if (G == 1) {
#pragma unroll
    for (int j = 0; j < NM; ++j) {
      double w, dx, dy, dz;
      perturber_pull<NM>(j, tau, x, y, z, Y, W, P4, ci, prm, mu, eps2, w, dx, dy, dz);
      ax = fma(w, dx, ax);
      ay = fma(w, dy, ay);
      az = fma(w, dz, az);
    }
  } else {
    double w, dx, dy, dz;
    perturber_pull<NM, true>(jl < NM ? jl : NM - 1, tau, x, y, z, Y, W, P4, ci, prm, mu, eps2, w, dx, dy, dz);
#pragma unroll
    for (int j = 0; j < NM; ++j) {
      const double wj = __shfl(w, base + j, 64);
      ax = fma(wj, __shfl(dx, base + j, 64), ax);
      ay = fma(wj, __shfl(dy, base + j, 64), ay);
      az = fma(wj, __shfl(dz, base + j, 64), az);
    }
  }
  f[3] = ax;
  f[4] = ay;
  f[5] = az;
}

// Gragg's modified midpoint rule with N (even) substeps over [t0, t0+H]
// (Hairer, Norsett & Wanner, Solving ODEs I, §II.9), on y = (x, v).
// Replaces modified_midpoint (src/sim.cuh:445-500), whose half-updated
// velocity does not have an even error expansion.
template <int NM, int G = 1>
__device__ __forceinline__ void gragg_tp(const double *y0, const double *f0, double t0, double H,
                                      int N, const double *__restrict__ Y,
                                      const double *__restrict__ W,
                                      const double *__restrict__ P4, double h,
                                      const double *__restrict__ prm, double mu, double eps2,
                                      double *out, int &nrhs, int jl = 0, int base = 0) {
  const double hs = H / (double)N;
  const double h2 = 2.0 * hs;
  double z0[6], z1[6], f[6];
#pragma unroll
  for (int i = 0; i < 6; ++i) {
    z0[i] = y0[i];
    z1[i] = fma(hs, f0[i], y0[i]);
  }
  if (G == 1) {
    for (int m = 1; m < N; ++m) {
      rhs_tp<NM, G>(t0 + (double)m * hs, z1, Y, W, P4, h, prm, mu, eps2, f, jl, base);
#pragma unroll
      for (int i = 0; i < 6; ++i) {
        const double z2 = fma(h2, f[i], z0[i]);
        z0[i] = z1[i];
        z1[i] = z2;
      }
    }
    rhs_tp<NM, G>(t0 + H, z1, Y, W, P4, h, prm, mu, eps2, f, jl, base);
  } else {
    // the lane's perturber is placed one substep ahead (a cache load or a
    // drift, independent of the TP state), so its latency overlaps the
    // current right-hand side
    const int j = jl < NM ? jl : NM - 1;
    double px, py, pz;
    const double t1 = N > 1 ? t0 + hs : t0 + H;
    perturber_pos<NM>(j, t1, cache_point(t1, h), Y, W, P4, mu, px, py, pz);
    for (int m = 1; m < N; ++m) {
      const double tn = m + 1 < N ? t0 + (double)(m + 1) * hs : t0 + H;
      double nx, ny, nz;
      perturber_pos<NM>(j, tn, cache_point(tn, h), Y, W, P4, mu, nx, ny, nz);
      rhs_tp_at<NM>(z1, px, py, pz, prm, mu, eps2, f, jl, base);
#pragma unroll
      for (int i = 0; i < 6; ++i) {
        const double z2 = fma(h2, f[i], z0[i]);
        z0[i] = z1[i];
        z1[i] = z2;
      }
      px = nx;
      py = ny;
      pz = nz;
    }
    rhs_tp_at<NM>(z1, px, py, pz, prm, mu, eps2, f, jl, base);
  }
  nrhs += N;
#pragma unroll
  for (int i = 0; i < 6; ++i) out[i] = 0.5 * (z1[i] + fma(hs, f[i], z0[i]));
}

__device__ __forceinline__ double bs_fac(int j) {
  // 4^j - 1 for the doubling sequence (Richardson factor of src/sim.cuh:522-544)
  return (double)((1ll << (2 * j)) - 1);
}

// Bulirsch-Stoer drift of one TP over the step h: Gragg rows N = 2,4,...,256,
// Aitken-Neville extrapolation in H^2, error |T_kk - T_k,k-1| scaled by |x0|,
// |v0| against tol; on failure the interval is halved.  Replaces the fixed
// 3-row richardson_extrapolation of src/sim.cuh:502-551.  Returns status bits.
// The extrapolation tableau element (row, i) lives at Rt[(row * 6 + i) * rs]:
// a private array (rs = 1) or a lane's column of an LDS block (rs = lanes).
// G > 1: the G lanes of a group run the same step together (see rhs_tp).
//
// bs_extrapolate: row k of the extrapolation.  cur holds row k's Gragg value
// on entry and T_kk on return; the tableau is updated in place.  Returns the
// scaled error |T_kk - T_k,k-1| (0 for k = 0).
__device__ __forceinline__ double bs_extrapolate(int k, double *cur, double *Rt, int rs, double xs,
                                                 double vs) {
  for (int j = 1; j <= k; ++j) {
    const double fac = bs_fac(j);
#pragma unroll
    for (int i = 0; i < 6; ++i) {
      double *slot = Rt + ((j - 1) * 6 + i) * rs;
      const double prev = *slot;
      *slot = cur[i];
      cur[i] = cur[i] + div_normal(cur[i] - prev, fac);
    }
  }
#pragma unroll
  for (int i = 0; i < 6; ++i) Rt[(k * 6 + i) * rs] = cur[i];
  double err = 0.0;
  if (k >= 1) {
#pragma unroll
    for (int i = 0; i < 6; ++i) {
      const double e = div_normal(fabs(cur[i] - Rt[((k - 1) * 6 + i) * rs]), i < 3 ? xs : vs);
      err = e > err ? e : err;
    }
  }
  return err;
}

template <int NM, int G = 1>
__device__ __forceinline__ uint32_t bs_drift_tp(double *yv, double h, const double *__restrict__ Y,
                                                const double *__restrict__ W,
                                                const double *__restrict__ P4,
                                                const double *__restrict__ prm, double mu,
                                                double eps2, double tol, int &nrhs, double *Rt,
                                                int rs, int jl = 0, int base = 0) {
  uint32_t st = 0;
  double t = 0.0, H = h;
  int depth = 0;
  while (t < h) {
    if (H > h - t) H = h - t;
    double f0[6];
    rhs_tp<NM, G>(t, yv, Y, W, P4, h, prm, mu, eps2, f0, jl, base);
    nrhs += 1;
    const double xs = sqrt_dist2(dot3(yv[0], yv[1], yv[2], yv[0], yv[1], yv[2]));
    const double vs = sqrt_dist2(dot3(yv[3], yv[4], yv[5], yv[3], yv[4], yv[5]));
    double cur[6];
    bool conv = false;
    for (int k = 0; k < KMAX; ++k) {
      gragg_tp<NM, G>(yv, f0, t, H, 2 << k, Y, W, P4, h, prm, mu, eps2, cur, nrhs, jl, base);
      const double err = bs_extrapolate(k, cur, Rt, rs, xs, vs);
      if (k >= 1 && err <= tol) {
        conv = true;
        break;
      }
    }
    if (conv || depth >= BS_MAXDEPTH) {
      for (int i = 0; i < 6; ++i) yv[i] = cur[i];
      t += H;
      if (!conv) {
        // the halving budget is spent: flag it and take the rest of the
        // interval in one step, so the work stays bounded
        st |= ST_BS_DEPTH;
        H = h - t;
      }
    } else {
      H *= 0.5;
      ++depth;
    }
  }
  return st;
}

// bs_drift_tp for an entry served by R row groups of GP lanes each (lanes
// ebase .. ebase + R GP - 1; this lane is perturber lane jl of row group
// row): the Gragg rows kb .. kb+R-1 of one attempt run side by side, one per
// group, in lockstep by substep, and as each row completes every lane of the
// entry takes it from its group (a shuffle), extends the extrapolation and
// tests convergence, leaving at the first converged row.  Every row and every
// extrapolation step is formed as in bs_drift_tp and in the same order, so
// the result and the right-hand-side count (rows up to the converged one) are
// the same; the latency of an attempt drops from 2 + 4 + ... + N_k substeps
// to N_k.
template <int NM, int GP, int R>
__device__ __forceinline__ uint32_t bs_drift_tp_rows(double *yv, double h,
                                                     const double *__restrict__ Y,
                                                     const double *__restrict__ W,
                                                     const double *__restrict__ P4,
                                                     const double *__restrict__ prm, double mu,
                                                     double eps2, double tol, int &nrhs,
                                                     double *Rt, int rs, int jl, int row, int ebase) {
  const int base = ebase + row * GP;  // this row group's first lane
  const int j = jl < NM ? jl : NM - 1;
  uint32_t st = 0;
  double t = 0.0, H = h;
  int depth = 0;
  while (t < h) {
    if (H > h - t) H = h - t;
    double f0[6];
    rhs_tp<NM, GP>(t, yv, Y, W, P4, h, prm, mu, eps2, f0, jl, base);
    nrhs += 1;
    const double xs = sqrt_dist2(dot3(yv[0], yv[1], yv[2], yv[0], yv[1], yv[2]));
    const double vs = sqrt_dist2(dot3(yv[3], yv[4], yv[5], yv[3], yv[4], yv[5]));
    double cur[6];
    bool conv = false;
    for (int kb = 0; kb < KMAX && !conv; kb += R) {
      const int klast = kb + R - 1 < KMAX - 1 ? kb + R - 1 : KMAX - 1;
      const int km = kb + row;                        // this group's row
      const int Nm = km <= klast ? (2 << km) : 0;     // 0: no row for this group
      const int Nmax = 2 << klast;
      const double hs = Nm > 0 ? H / (double)Nm : 0.0;
      const double h2 = 2.0 * hs;
      double z0[6], z1[6], f[6], mine[6];
#pragma unroll
      for (int i = 0; i < 6; ++i) {
        z0[i] = yv[i];
        z1[i] = fma(hs, f0[i], yv[i]);
        mine[i] = 0.0;
      }
      // the lane's perturber one substep ahead, as in gragg_tp
      double px = 0.0, py = 0.0, pz = 0.0;
      if (Nm > 0) {
        const double t1 = t + hs;
        perturber_pos<NM>(j, t1, cache_point(t1, h), Y, W, P4, mu, px, py, pz);
      }
      int k = kb;  // next row to extrapolate
      for (int m = 1; m <= Nmax; ++m) {
        // one right-hand side for every row still running (the rows that
        // finish at this substep included), then each row's own update
        if (m <= Nm) {
          double nx = 0.0, ny = 0.0, nz = 0.0;
          if (m < Nm) {
            const double tn = m + 1 < Nm ? t + (double)(m + 1) * hs : t + H;
            perturber_pos<NM>(j, tn, cache_point(tn, h), Y, W, P4, mu, nx, ny, nz);
          }
          rhs_tp_at<NM>(z1, px, py, pz, prm, mu, eps2, f, jl, base);
          if (m < Nm) {
#pragma unroll
            for (int i = 0; i < 6; ++i) {
              const double z2 = fma(h2, f[i], z0[i]);
              z0[i] = z1[i];
              z1[i] = z2;
            }
            px = nx;
            py = ny;
            pz = nz;
          } else {
#pragma unroll
            for (int i = 0; i < 6; ++i) mine[i] = 0.5 * (z1[i] + fma(hs, f[i], z0[i]));
          }
        }
        if (m == (2 << k)) {  // row k has just been completed by its group
          const int src = ebase + (k - kb) * GP;
#pragma unroll
          for (int i = 0; i < 6; ++i) cur[i] = __shfl(mine[i], src, 64);
          nrhs += 2 << k;
          const double err = bs_extrapolate(k, cur, Rt, rs, xs, vs);
          if (k >= 1 && err <= tol) conv = true;
          ++k;
          if (conv) break;
        }
      }
    }
    if (conv || depth >= BS_MAXDEPTH) {
      for (int i = 0; i < 6; ++i) yv[i] = cur[i];
      t += H;
      if (!conv) {
        st |= ST_BS_DEPTH;
        H = h - t;
      }
    } else {
      H *= 0.5;
      ++depth;
    }
  }
  return st;
}

// ---------------------------------------------------------------------------
// Massive subsystem (NM bodies of one system, state body-major xyz)
// ---------------------------------------------------------------------------

// Kick acceleration of every massive body from the others, pair changeover
// radius max(r_c,j, r_c,k) (body_interaction_kick over massive partners,
// src/sim.cuh:327-382, with self-exclusion by body index: the body's own slot
// has numerator 0, denominator 1 and separation 0, which changes no bit of the
// other weights or of the sum).
template <int NM>
__device__ __forceinline__ void accel_massive(const double *X, const double *__restrict__ prm,
                                              double eps2, double *A) {
  static_assert(NM <= MB_DIVG, "one division group");
#pragma unroll
  for (int j = 0; j < NM; ++j) {
    double dx[NM], dy[NM], dz[NM], num[NM], den[NM], w[NM];
#pragma unroll
    for (int k = 0; k < NM; ++k) {
      if (k == j) {
        dx[k] = 0.0;
        dy[k] = 0.0;
        dz[k] = 0.0;
        num[k] = 0.0;
        den[k] = 1.0;
        continue;
      }
      dx[k] = X[3 * k + 0] - X[3 * j + 0];
      dy[k] = X[3 * k + 1] - X[3 * j + 1];
      dz[k] = X[3 * k + 2] - X[3 * j + 2];
      const double r2 = dot3(dx[k], dy[k], dz[k], dx[k], dy[k], dz[k]);
      kick_nd(r2, eps2, prm[P_M(NM) + k], prm[P_PRCLO(NM) + j * NM + k],
              prm[P_PRCWINV(NM) + j * NM + k], num[k], den[k]);
    }
    batched_div<NM>(num, den, w);
    double ax = 0.0, ay = 0.0, az = 0.0;
#pragma unroll
    for (int k = 0; k < NM; ++k) {
      ax = fma(w[k], dx[k], ax);
      ay = fma(w[k], dy[k], ay);
      az = fma(w[k], dz[k], az);
    }
    A[3 * j + 0] = ax;
    A[3 * j + 1] = ay;
    A[3 * j + 2] = az;
  }
}

// Q = (sum_j m_j v_j) / M  (main_body_kinetic's momentum sum, src/sim.cuh:384-406)
template <int NM>
__device__ __forceinline__ void momentum_q(const double *V, const double *__restrict__ prm,
                                           double M, double *Q) {
  double px = 0.0, py = 0.0, pz = 0.0;
#pragma unroll
  for (int j = 0; j < NM; ++j) {
    const double m = prm[P_M(NM) + j];
    px = fma(m, V[3 * j + 0], px);
    py = fma(m, V[3 * j + 1], py);
    pz = fma(m, V[3 * j + 2], pz);
  }
  Q[0] = px / M;
  Q[1] = py / M;
  Q[2] = pz / M;
}

// Close-encounter RHS of one encounter group (g members, body indices idx[],
// state [g][6]): the Sun plus the (1-K)-weighted pulls of the other members.
template <int NM>
__device__ __forceinline__ void rhs_group(const double *s, int g, const int *idx,
                                          const double *__restrict__ prm, double mu, double eps2,
                                          double *f) {
  for (int a = 0; a < g; ++a) {
    const int j = idx[a];
    const double x = s[6 * a + 0], y = s[6 * a + 1], z = s[6 * a + 2];
    f[6 * a + 0] = s[6 * a + 3];
    f[6 * a + 1] = s[6 * a + 4];
    f[6 * a + 2] = s[6 * a + 5];
    const double r2 = dot3(x, y, z, x, y, z);
    const double r = sqrt_dist2(r2);
    const double w0 = div_normal(mu, r2 * r);
    double ax = -(w0 * x), ay = -(w0 * y), az = -(w0 * z);
    for (int b = 0; b < g; ++b) {
      if (b == a) continue;
      const int k = idx[b];
      const double dx = s[6 * b + 0] - x, dy = s[6 * b + 1] - y, dz = s[6 * b + 2] - z;
      const double d2 = dot3(dx, dy, dz, dx, dy, dz);
      const double w = close_weight(d2, eps2, prm[P_M(NM) + k], prm[P_PRCLO(NM) + j * NM + k],
                                    prm[P_PRCWINV(NM) + j * NM + k]);
      ax = fma(w, dx, ax);
      ay = fma(w, dy, ay);
      az = fma(w, dz, az);
    }
    f[6 * a + 3] = ax;
    f[6 * a + 4] = ay;
    f[6 * a + 5] = az;
  }
}

template <int NM>
__device__ __noinline__ void gragg_group(const double *y0, const double *f0, double H, int N, int g,
                                         const int *idx, const double *__restrict__ prm,
                                         double mu, double eps2, double *out, double *ws) {
  constexpr int D = 6 * NM;
  const int n = 6 * g;
  const double hs = H / (double)N;
  const double h2 = 2.0 * hs;
  double *z0 = ws, *z1 = ws + D, *f = ws + 2 * D;  // the caller's workspace (3 D doubles)
  for (int i = 0; i < n; ++i) {
    z0[i] = y0[i];
    z1[i] = fma(hs, f0[i], y0[i]);
  }
  for (int m = 1; m < N; ++m) {
    rhs_group<NM>(z1, g, idx, prm, mu, eps2, f);
    for (int i = 0; i < n; ++i) {
      const double z2 = fma(h2, f[i], z0[i]);
      z0[i] = z1[i];
      z1[i] = z2;
    }
  }
  rhs_group<NM>(z1, g, idx, prm, mu, eps2, f);
  for (int i = 0; i < n; ++i) out[i] = 0.5 * (z1[i] + fma(hs, f[i], z0[i]));
}

// Workspace of bs_drift_group, in doubles: f0, cur (D each), the error
// scales (2 NM), the extrapolation tableau (KMAX D) and gragg_group's (3 D).
template <int NM>
constexpr int bs_group_ws() {
  return (5 + KMAX) * 6 * NM + 2 * NM;
}

// BS drift of one massive encounter group over h.  Error scale per member:
// |x_j|, |v_j| at the start of each interval.  Its arrays live in the
// caller's workspace ws (bs_group_ws<NM>() doubles, LDS in the ephemeris
// kernel), so that no private memory is needed.
template <int NM>
__device__ __noinline__ uint32_t bs_drift_group(double *s, int g, const int *idx, double h,
                                                const double *__restrict__ prm, double mu,
                                                double eps2, double tol, unsigned int &nrhs, double *ws) {
  constexpr int D = 6 * NM;
  const int n = 6 * g;
  double *f0 = ws, *cur = ws + D, *sc = ws + 2 * D, *R = ws + 2 * D + 2 * NM;
  double *gws = R + KMAX * D;
  uint32_t st = 0;
  double t = 0.0, H = h;
  int depth = 0;
  while (t < h) {
    if (H > h - t) H = h - t;
    rhs_group<NM>(s, g, idx, prm, mu, eps2, f0);
    ++nrhs;
    for (int a = 0; a < g; ++a) {
      sc[2 * a] = sqrt_dist2(dot3(s[6 * a], s[6 * a + 1], s[6 * a + 2], s[6 * a], s[6 * a + 1], s[6 * a + 2]));
      sc[2 * a + 1] =
          sqrt_dist2(dot3(s[6 * a + 3], s[6 * a + 4], s[6 * a + 5], s[6 * a + 3], s[6 * a + 4], s[6 * a + 5]));
    }
    bool conv = false;
    for (int k = 0; k < KMAX; ++k) {
      gragg_group<NM>(s, f0, H, 2 << k, g, idx, prm, mu, eps2, cur, gws);
      nrhs += 2u << k;
      for (int j = 1; j <= k; ++j) {
        const double fac = bs_fac(j);
        for (int i = 0; i < n; ++i) {
          const double prev = R[(j - 1) * D + i];
          R[(j - 1) * D + i] = cur[i];
          cur[i] = cur[i] + div_normal(cur[i] - prev, fac);
        }
      }
      for (int i = 0; i < n; ++i) R[k * D + i] = cur[i];
      if (k >= 1) {
        double err = 0.0;
        for (int i = 0; i < n; ++i) {
          const int b = i / 6, c = i % 6;
          const double e = div_normal(fabs(cur[i] - R[(k - 1) * D + i]), sc[2 * b + (c < 3 ? 0 : 1)]);
          err = e > err ? e : err;
        }
        if (err <= tol) {
          conv = true;
          break;
        }
      }
    }
    if (conv || depth >= BS_MAXDEPTH) {
      for (int i = 0; i < n; ++i) s[i] = cur[i];
      t += H;
      if (!conv) {
        // the halving budget is spent: flag it and take the rest of the
        // interval in one step, so the work stays bounded
        st |= ST_BS_DEPTH;
        H = h - t;
      }
    } else {
      H *= 0.5;
      ++depth;
    }
  }
  return st;
}

}  // namespace mint
