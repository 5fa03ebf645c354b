// mint_mb.hip — the many-body engine: systems of more than 8 massive bodies
// and no test particles (the all-massive planetary-embryo configuration).
//
// It integrates the same massive-subsystem step as the lane-per-system
// ephemeris kernel, operation for operation, so the two agree to the last bit
// where both apply; what changes is the parallel shape.  One step is two
// launches per system set (blockIdx.y = system):
//
//   mb_drift_kernel  momentum after the opening kick and jump 1 (each
//                    workgroup forms them), then the Kepler drift of every
//                    body in no pair (one lane each); meanwhile each group
//                    workgroup turns the step's pair list into encounter
//                    groups (connected components, listed by smallest member)
//                    in its LDS and drifts its share of them by Bulirsch-
//                    Stoer (one wavefront per group: row groups of lanes, one
//                    member per lane, run the Gragg rows side by side;
//                    positions exchanged by wavefront shuffles).
//   mb_close_kernel  momentum after the drift, jump 2, O(N^2) kick
//                    accelerations and the closing kick (fused with the next
//                    step's opening kick, which sees the same positions), and
//                    the next step's exact pair check of every pair (j < k)
//                    against its changeover radius on the same separations:
//                    pairs inside are appended to the next step's pair list.
//
// The pair check is taken on a step's start positions, before jump 1: the
// jump translates every body by the same (h/2) Q1, so it leaves the pair
// separations as they are up to the rounding of the translation (the oracle
// and the ephemeris kernel check the same separations).
//
// Reference: body_interaction_kick with every body massive (src/sim.cuh:327-382),
// main_body_kinetic (src/sim.cuh:384-406), close_encounter_p
// (src/sim.cuh:589-613) and the BS drift of flagged bodies
// (src/sim.cuh:410-551,699-730).  The reference integrates each body against
// frozen partners; here a group's members are integrated together.
//
// Kick sums: body i's sum over the 16 partners of one chunk is formed from
// zero in partner order, two division groups of 8 with one division each
// (batched_div), the body itself and bodies past N as inert slots; two
// threads share a (body, chunk) slot, one per division group, the second
// taking the first's partial sum on; the chunk sums are then folded by
// halving strides (fold_lanes).  The momentum sum uses the same chunks and
// the same fold.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <vector>

#include "mint_api_internal.h"

namespace mint {
namespace {

struct MBShape {
  int N;    // bodies per system
  int nch;  // 16-body chunks
  int TI;   // bodies per workgroup (open/close kernels)
  int CS;   // chunk slots per workgroup; TI * CS = MB_BLOCK; a window of CS
            // chunks of partner data is staged in LDS at a time
  int P;    // nch rounded up to a power of two (the chunk fold's width)
};

constexpr int MB_BLOCK = 256;
constexpr int MB_MAXCS = 64;   // chunk slots: a staged window holds up to 64 x 16 bodies
constexpr int MB_DRIFT_GROUP_BLOCKS = 16;  // workgroups that walk the group list (two wavefronts each)
constexpr int MB_GRAPH_STEPS = 64;

MBShape mb_shape(int N) {
  MBShape s;
  s.N = N;
  s.nch = (N + MB_CHUNK - 1) / MB_CHUNK;
  int cs = 1;
  while (cs < s.nch && cs < MB_MAXCS) cs <<= 1;
  s.CS = cs;
  s.TI = MB_BLOCK / cs;
  int p = 1;
  while (p < s.nch) p <<= 1;
  s.P = p;  // = CS up to 64 chunks (1024 bodies), 128 beyond
  return s;
}

// body parameters of one system, [5][N]
struct BP {
  const double *m, *lo, *winv, *rc2, *rad;
};

__device__ __forceinline__ BP bp_of(const double *bp_all, int s, int N) {
  const double *b = bp_all + (size_t)s * 5 * N;
  return BP{b, b + N, b + 2 * N, b + 3 * N, b + 4 * N};
}

// Chunk sums a[0..nch) are folded by halving strides: with P = nch rounded
// up to a power of two and a[l] = +0 for l >= nch, slot l < h takes
// a[l] + a[l + h] for h = P/2, ..., 1, and the fold is a[0] (one chunk: its
// own sum).  Lanes [0, G) of a lane group hold a[l] (and, for P = 128, have
// already added a[l + 64]); G = min(P, 64) divides 64.  The fold ends in the
// group's first lane.
//
// DPP moves (gfx9 data-parallel primitives; all lanes of the wavefront
// active): dpp64<C> returns lane l's view of x through the DPP control C,
// 0 where the source lane is outside the row.  Controls used: 0x101..0x108
// row_shl:n (lane l reads l + n within its 16-lane row), 0x111..0x118
// row_shr:n, 0x142 row_bcast:15, 0x143 row_bcast:31, 0xB1 / 0x4E quad_perm
// [1,0,3,2] / [2,3,0,1], 0x141 row_half_mirror, 0x140 row_mirror.
template <int C>
__device__ __forceinline__ double dpp64(double x) {
  const int lo = __builtin_amdgcn_update_dpp(0, __double2loint(x), C, 0xf, 0xf, true);
  const int hi = __builtin_amdgcn_update_dpp(0, __double2hiint(x), C, 0xf, 0xf, true);
  return __hiloint2double(hi, lo);
}

__device__ __forceinline__ void fold_lanes(double &x, double &y, double &z, int G) {
  if (G > 32) {
    x = x + __shfl_down(x, 32, 64);
    y = y + __shfl_down(y, 32, 64);
    z = z + __shfl_down(z, 32, 64);
  }
  if (G > 16) {
    x = x + __shfl_down(x, 16, 64);
    y = y + __shfl_down(y, 16, 64);
    z = z + __shfl_down(z, 16, 64);
  }
  if (G > 8) {
    x = x + dpp64<0x108>(x);
    y = y + dpp64<0x108>(y);
    z = z + dpp64<0x108>(z);
  }
  if (G > 4) {
    x = x + dpp64<0x104>(x);
    y = y + dpp64<0x104>(y);
    z = z + dpp64<0x104>(z);
  }
  if (G > 2) {
    x = x + dpp64<0x102>(x);
    y = y + dpp64<0x102>(y);
    z = z + dpp64<0x102>(z);
  }
  if (G > 1) {
    x = x + dpp64<0x101>(x);
    y = y + dpp64<0x101>(y);
    z = z + dpp64<0x101>(z);
  }
}

// kick_nd formed without branches: both changeover forms are evaluated and
// selected (the same values kick_nd gives, NaN inputs included).
__device__ __forceinline__ void kick_nd_sel(double r2, double eps2, double m, double rclo, double rcwinv,
                                            double &num, double &den) {
  const double r = sqrt_dist2(r2);
  const double D = (r2 + eps2) * r;
  const double yv = (r - rclo) * rcwinv;
  const double y2 = yv * yv;
  const double nb = m * y2, db = D * fma(2.0, y2, fma(-2.0, yv, 1.0));
  const bool band = yv > 0.0 && yv < 1.0;
  num = yv <= 0.0 ? 0.0 : (band ? nb : m);
  den = band ? db : D;
}

// Staged partner windows in LDS carry one double of padding after each
// 16-body chunk.  The 16 chunk slots of a wavefront read the same offset in 16
// consecutive chunks; unpadded, those chunks start 48 (positions) or 16 (per-
// body scalars) doubles apart, on only two LDS bank positions.  With chunk
// strides of 49 and 17 doubles the 16 chunks start on 16 distinct even banks.
__device__ __forceinline__ int lds3(int k) { return 3 * k + (k >> 4); }  // position x of window body k
__device__ __forceinline__ int lds1(int k) { return k + (k >> 4); }      // scalar of window body k

// Bodies a thread stages per window (windows of at most MB_MAXCS chunks).
// The first window's data is loaded into registers before the momentum sum,
// so that both sets of loads are in flight together; the jump is applied
// when the values go to LDS, once the momentum is known.
// Threads of a kick workgroup: two per (body, chunk) slot, one per division
// group of the chunk's 16 partners.
constexpr int MB_NT = 2 * MB_BLOCK;
constexpr int MB_PF = MB_MAXCS * MB_CHUNK / MB_NT;

// Chunk t's momentum sum over a staged window of nw bodies (sv in the lds3
// layout, sm in the lds1 layout): from zero, in index order.  Slots past the
// window (not staged) are read and their terms selected away, without
// branches; eight bodies' loads at a time.
__device__ __forceinline__ void chunk_momentum(const double *sv, const double *sm, int t, int nw, double &px,
                                               double &py, double &pz) {
  px = 0.0;
  py = 0.0;
  pz = 0.0;
#pragma unroll
  for (int h = 0; h < MB_CHUNK; h += MB_DIVG) {
    double mm[MB_DIVG], vx[MB_DIVG], vy[MB_DIVG], vz[MB_DIVG];
#pragma unroll
    for (int q = 0; q < MB_DIVG; ++q) {
      const int k = t * MB_CHUNK + h + q;
      mm[q] = sm[lds1(k)];
      vx[q] = sv[lds3(k) + 0];
      vy[q] = sv[lds3(k) + 1];
      vz[q] = sv[lds3(k) + 2];
    }
#pragma unroll
    for (int q = 0; q < MB_DIVG; ++q) {
      const bool in = t * MB_CHUNK + h + q < nw;
      px = in ? fma(mm[q], vx[q], px) : px;
      py = in ? fma(mm[q], vy[q], py) : py;
      pz = in ? fma(mm[q], vz[q], pz) : pz;
    }
  }
}

// Q = (sum_j m_j v_j)/M by the whole workgroup.  Masses and velocities pass
// through LDS a window of WB bodies at a time (sv in the lds3 layout, sm in
// the lds1 layout; the first window comes from the caller's registers pm, pv,
// loaded by body index k = threadIdx.x + u MB_NT, so every load is
// coalesced); lane c of wavefront 0 forms the window's chunk c sum (from zero,
// index order), and the chunk sums are folded (fold_lanes).  At most two
// windows: N <= 2048 = 2 WB.  Qs[3] receives Q; the caller's next write to sv
// or sm follows the barrier this ends with.
__device__ void wg_momentum(const double *__restrict__ V, const double *__restrict__ m, const MBShape &sh,
                            double M, const double (&pm)[MB_PF], const double (&pv)[MB_PF][3], double *sv,
                            double *sm, double *Qs) {
  const int t = (int)threadIdx.x, N = sh.N, WB = sh.CS * MB_CHUNK;
  double c0x = 0.0, c0y = 0.0, c0z = 0.0, c1x = 0.0, c1y = 0.0, c1z = 0.0;
  for (int b0 = 0, pass = 0; b0 < N; b0 += WB, ++pass) {
    const int nw = min(WB, N - b0);
    if (pass > 0) __syncthreads();  // wavefront 0 has read the previous window
    if (pass == 0) {
#pragma unroll
      for (int u = 0; u < MB_PF; ++u) {
        const int k = t + u * MB_NT;
        if (k < nw) {
          sv[lds3(k) + 0] = pv[u][0];
          sv[lds3(k) + 1] = pv[u][1];
          sv[lds3(k) + 2] = pv[u][2];
          sm[lds1(k)] = pm[u];
        }
      }
    } else {
      for (int k = t; k < nw; k += MB_NT) {
        const int kk = b0 + k;
        sv[lds3(k) + 0] = V[3 * kk + 0];
        sv[lds3(k) + 1] = V[3 * kk + 1];
        sv[lds3(k) + 2] = V[3 * kk + 2];
        sm[lds1(k)] = m[kk];
      }
    }
    __syncthreads();
    if (t * MB_CHUNK < nw) {  // t < CS <= 64: wavefront 0
      double px, py, pz;
      chunk_momentum(sv, sm, t, nw, px, py, pz);
      if (pass == 0) {
        c0x = px;
        c0y = py;
        c0z = pz;
      } else {
        c1x = px;
        c1y = py;
        c1z = pz;
      }
    }
  }
  if (t < 64) {
    if (sh.P > 64) {  // first level of the fold: chunk l + 64 (or +0)
      c0x = c0x + c1x;
      c0y = c0y + c1y;
      c0z = c0z + c1z;
    }
    fold_lanes(c0x, c0y, c0z, sh.P < 64 ? sh.P : 64);
    if (t == 0) {
      Qs[0] = c0x / M;
      Qs[1] = c0y / M;
      Qs[2] = c0z / M;
    }
  }
  __syncthreads();
}

// wg_momentum by one wavefront (a one-wave workgroup): the same windows, chunk
// sums and fold, so the same Q, which every lane receives.  Each window's
// velocities and masses are loaded coalesced (lane l takes bodies l, l + 64,
// ...) into sv and sm (3 WB + CS and WB + CS doubles of LDS).  Called by all
// 64 lanes; ends with a barrier, after which the caller may reuse the LDS.
__device__ void wave_momentum(const double *__restrict__ V, const double *__restrict__ m, const MBShape &sh,
                              double M, double *sv, double *sm, double &q0, double &q1, double &q2) {
  const int t = (int)threadIdx.x & 63, N = sh.N, WB = sh.CS * MB_CHUNK;
  double c0x = 0.0, c0y = 0.0, c0z = 0.0, c1x = 0.0, c1y = 0.0, c1z = 0.0;
  for (int b0 = 0, pass = 0; b0 < N; b0 += WB, ++pass) {
    const int nw = min(WB, N - b0);
    if (pass > 0) wave_lds_sync();  // the previous window has been read
    // every load of the window in flight at once (a window holds at most
    // MB_MAXCS x 16 = 1024 bodies, 16 a lane)
    constexpr int U = MB_MAXCS * MB_CHUNK / 64;
    double lv[U][3], lm[U];
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const int k = t + 64 * u, kk = b0 + (k < nw ? k : 0);
      lv[u][0] = V[3 * kk + 0];
      lv[u][1] = V[3 * kk + 1];
      lv[u][2] = V[3 * kk + 2];
      lm[u] = m[kk];
    }
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const int k = t + 64 * u;
      if (k < nw) {
        sv[lds3(k) + 0] = lv[u][0];
        sv[lds3(k) + 1] = lv[u][1];
        sv[lds3(k) + 2] = lv[u][2];
        sm[lds1(k)] = lm[u];
      }
    }
    wave_lds_sync();
    if (t * MB_CHUNK < nw) {
      double px, py, pz;
      chunk_momentum(sv, sm, t, nw, px, py, pz);
      if (pass == 0) {
        c0x = px;
        c0y = py;
        c0z = pz;
      } else {
        c1x = px;
        c1y = py;
        c1z = pz;
      }
    }
  }
  if (sh.P > 64) {  // first level of the fold: chunk l + 64 (or +0)
    c0x = c0x + c1x;
    c0y = c0y + c1y;
    c0z = c0z + c1z;
  }
  fold_lanes(c0x, c0y, c0z, sh.P < 64 ? sh.P : 64);
  q0 = __shfl(c0x, 0, 64) / M;
  q1 = __shfl(c0y, 0, 64) / M;
  q2 = __shfl(c0z, 0, 64) / M;
  wave_lds_sync();
}

// ---------------------------------------------------------------------------
// Kicks.  INIT: positions XA as they are, velocities VA, one kick (the opening
// kick of a call's first step).  Otherwise positions XB + (h/2) Q2 (jump 2)
// are stored to XA, velocities VB, and kicks = 1 (closing kick of a call's
// last step) or 2 (closing kick, then the next step's opening kick).
//
// check: the next step's pair check (close_encounter_p, src/sim.cuh:589-613)
// rides on the kick's separations, which are those of the next step's start
// positions XA: every pair j < k inside its changeover radius (r^2 <
// max(r_c,j, r_c,k)^2) is appended to the pair list (count ctl[par]) and both
// bodies flagged; pairs closer than the sum of their physical radii are
// counted.  The kick loop only marks the slots with r^2 below the system's
// bound lim (its largest r_c^2, and the overlap bound with its largest
// radius; mbmax = [S][2]: largest r_c^2, largest radius), which every pair
// that passes either exact test is below (rounding is monotonic); the marked
// slots are then tested exactly.
template <bool INIT>
__global__ void __launch_bounds__(MB_NT)
    mb_close_kernel(MBShape sh, const double *__restrict__ bp_all, const double *__restrict__ mbmax_all,
                    double *__restrict__ xa_all, double *__restrict__ va_all, const double *__restrict__ xb_all,
                    const double *__restrict__ vb_all, int kicks, int last, int check, int par,
                    int *__restrict__ flag_all, int *__restrict__ edge_all, unsigned int *__restrict__ ctl_all,
                    unsigned int *__restrict__ status, unsigned long long *__restrict__ stats, StepConst sc,
                    double M) {
  extern __shared__ double lds[];
  const int N = sh.N, s = blockIdx.y;
  const BP bp = bp_of(bp_all, s, N);
  int *flag = flag_all + (size_t)s * N;
  int *edge = edge_all + (size_t)s * MB_EMAX * 2;
  unsigned int *npairs = ctl_all + 4 * s + par;
  double *XA = xa_all + (size_t)s * N * 3;
  double *VA = va_all + (size_t)s * N * 3;
  const double *XB = xb_all + (size_t)s * N * 3;
  const double *VB = vb_all + (size_t)s * N * 3;
  const int WB = sh.CS * MB_CHUNK;                   // bodies per staged window
  double *part = lds;                                // [nch][TI][3]
  double *Qs = part + (size_t)sh.nch * sh.TI * 3;    // [3]
  double *sx = Qs + 3;                               // partner positions, lds3 layout
  double *sm = sx + 3 * WB + sh.CS;                  // per-body scalars, lds1 layout
  double *slo = sm + WB + sh.CS, *swv = slo + WB + sh.CS;
  const double *X = INIT ? XA : XB;
  const double *V = INIT ? VA : VB;
  // thread = (chunk slot cs, half hf, body ti): the two halves of a slot
  // take the chunk's two division groups of eight partners
  const int ti = (int)threadIdx.x % sh.TI, rs = (int)threadIdx.x / sh.TI;
  const int cs = rs >> 1, hf = rs & 1;
  const int i = (int)blockIdx.x * sh.TI + ti;
  const bool act = i < N;
  unsigned int ov = 0;  // overlapping pairs found by this lane
  unsigned int mark0 = 0u, mark1 = 0u;  // marked slots of this lane's chunk in windows 0 and 1
  // this lane's body and the first window, loaded ahead of the momentum sum
  double xi = 0.0, yi = 0.0, zi = 0.0, loi = 0.0, wvi = 0.0, vx = 0.0, vy = 0.0, vz = 0.0;
  double rc2i = 0.0, radi = 0.0, lim = -1.0;  // lim < 0: no slot is marked
  if (act) {
    xi = X[3 * i + 0];
    yi = X[3 * i + 1];
    zi = X[3 * i + 2];
    loi = bp.lo[i];
    wvi = bp.winv[i];
    if (check) {
      rc2i = bp.rc2[i];
      radi = bp.rad[i];
      const double so = radi + mbmax_all[2 * s + 1];
      const double ol = so * so, rm = mbmax_all[2 * s];
      lim = rm > ol ? rm : ol;
    }
    if (rs == 0) {
      vx = V[3 * i + 0];
      vy = V[3 * i + 1];
      vz = V[3 * i + 2];
    }
  }
  double pf[MB_PF][6], pm[MB_PF], pv[MB_PF][3];
#pragma unroll
  for (int u = 0; u < MB_PF; ++u) {
    const int k = (int)threadIdx.x + u * MB_NT;
    if (k < WB && k < N) {
      pf[u][0] = X[3 * k + 0];
      pf[u][1] = X[3 * k + 1];
      pf[u][2] = X[3 * k + 2];
      pf[u][3] = bp.m[k];
      pf[u][4] = bp.lo[k];
      pf[u][5] = bp.winv[k];
      pm[u] = pf[u][3];
      if (!INIT) {
        pv[u][0] = V[3 * k + 0];
        pv[u][1] = V[3 * k + 1];
        pv[u][2] = V[3 * k + 2];
      }
    }
  }
  double q0 = 0.0, q1 = 0.0, q2 = 0.0;
  if (!INIT) {
    // the velocities pass through the partner window's LDS (sx, sm), which
    // the window loop below fills after the barrier wg_momentum ends with
    wg_momentum(VB, bp.m, sh, M, pm, pv, sx, sm, Qs);
    q0 = Qs[0];
    q1 = Qs[1];
    q2 = Qs[2];
    xi = fma(sc.hh, q0, xi);
    yi = fma(sc.hh, q1, yi);
    zi = fma(sc.hh, q2, zi);
  }
  auto stage = [&](int k, double xk, double yk, double zk, double mk, double lok, double wvk) {
    if (!INIT) {
      xk = fma(sc.hh, q0, xk);
      yk = fma(sc.hh, q1, yk);
      zk = fma(sc.hh, q2, zk);
    }
    sx[lds3(k) + 0] = xk;
    sx[lds3(k) + 1] = yk;
    sx[lds3(k) + 2] = zk;
    sm[lds1(k)] = mk;
    slo[lds1(k)] = lok;
    swv[lds1(k)] = wvk;
  };
  for (int w0 = 0; w0 < N; w0 += WB) {
    __syncthreads();
    const int nw = min(WB, N - w0);
    if (w0 == 0) {
#pragma unroll
      for (int u = 0; u < MB_PF; ++u) {
        const int k = (int)threadIdx.x + u * MB_NT;
        if (k < nw) stage(k, pf[u][0], pf[u][1], pf[u][2], pf[u][3], pf[u][4], pf[u][5]);
      }
    } else {
      for (int k = threadIdx.x; k < nw; k += blockDim.x) {
        const int kk = w0 + k;
        stage(k, X[3 * kk + 0], X[3 * kk + 1], X[3 * kk + 2], bp.m[kk], bp.lo[kk], bp.winv[kk]);
      }
    }
    __syncthreads();
    // this thread's chunk and division group: partners c 16 + hf 8 + q; every
    // thread reaches the barrier below, whether or not its slot is in use
    const int c = w0 / MB_CHUNK + cs;
    const bool slot = c < sh.nch;
    double dx[MB_DIVG], dy[MB_DIVG], dz[MB_DIVG], w[MB_DIVG];
    if (slot && act) {
      double num[MB_DIVG], den[MB_DIVG];
      unsigned int mark = 0u;  // slots of this group below lim
#pragma unroll
      for (int q = 0; q < MB_DIVG; ++q) {
        // every slot is formed and the inert ones (the body itself, slots
        // past N, whose window entries are not staged) are selected away,
        // so that the eight pairs' chains interleave without branches
        const int k = c * MB_CHUNK + hf * MB_DIVG + q;
        const int kl = k - w0, k3 = lds3(kl), k1 = lds1(kl);
        const bool live = k < N && k != i;
        const double ex = sx[k3 + 0] - xi, ey = sx[k3 + 1] - yi, ez = sx[k3 + 2] - zi;
        const double lok = slo[k1], wvk = swv[k1];
        const double r2 = dot3(ex, ey, ez, ex, ey, ez);
        mark |= (r2 < lim ? 1u : 0u) << (hf * MB_DIVG + q);
        double nq, dq;
        kick_nd_sel(r2, sc.eps2, sm[k1], loi > lok ? loi : lok, wvi < wvk ? wvi : wvk, nq, dq);
        dx[q] = live ? ex : 0.0;
        dy[q] = live ? ey : 0.0;
        dz[q] = live ? ez : 0.0;
        num[q] = live ? nq : 0.0;
        den[q] = live ? dq : 1.0;
      }
      batched_div<MB_DIVG>(num, den, w);
      if (w0 == 0)
        mark0 = mark;
      else
        mark1 = mark;
    } else {
#pragma unroll
      for (int q = 0; q < MB_DIVG; ++q) {
        dx[q] = 0.0;
        dy[q] = 0.0;
        dz[q] = 0.0;
        w[q] = 0.0;
      }
    }
    // the chunk sum from zero in partner order: the first group's terms (half
    // 0), then, on from its partial through the chunk's slot in part, the
    // second's (half 1)
    double *pp = part + ((size_t)(slot ? c : 0) * sh.TI + ti) * 3;
    double px = 0.0, py = 0.0, pz = 0.0;
    if (slot && hf == 0) {
#pragma unroll
      for (int q = 0; q < MB_DIVG; ++q) {
        px = fma(w[q], dx[q], px);
        py = fma(w[q], dy[q], py);
        pz = fma(w[q], dz[q], pz);
      }
      pp[0] = px;
      pp[1] = py;
      pp[2] = pz;
    }
    __syncthreads();
    if (slot && hf == 1) {
      px = pp[0];
      py = pp[1];
      pz = pp[2];
#pragma unroll
      for (int q = 0; q < MB_DIVG; ++q) {
        px = fma(w[q], dx[q], px);
        py = fma(w[q], dy[q], py);
        pz = fma(w[q], dz[q], pz);
      }
      pp[0] = px;
      pp[1] = py;
      pp[2] = pz;
    }
  }
  __syncthreads();
  if ((int)threadIdx.x < MB_BLOCK) {  // the first MB_BLOCK threads: whole wavefronts
    // fold of body b's chunk sums by lanes [b G, (b + 1) G); TI G = MB_BLOCK
    // (G = min(P, 64) = CS); the folds go to acc[TI][3] over the spent window
    const int G = sh.P < 64 ? sh.P : 64;
    const int b = (int)threadIdx.x / G, l = (int)threadIdx.x % G;
    double ax = 0.0, ay = 0.0, az = 0.0;
    if (l < sh.nch) {
      const double *pp = part + ((size_t)l * sh.TI + b) * 3;
      ax = pp[0];
      ay = pp[1];
      az = pp[2];
    }
    if (sh.P > 64) {  // first level of the fold: chunk l + 64 (or +0)
      const bool in = l + 64 < sh.nch;
      const double *pp = part + ((size_t)(in ? l + 64 : 0) * sh.TI + b) * 3;
      ax = ax + (in ? pp[0] : 0.0);
      ay = ay + (in ? pp[1] : 0.0);
      az = az + (in ? pp[2] : 0.0);
    }
    fold_lanes(ax, ay, az, G);
    if (l == 0) {
      sx[3 * b + 0] = ax;
      sx[3 * b + 1] = ay;
      sx[3 * b + 2] = az;
    }
  }
  __syncthreads();
  if (rs == 0 && act) {
    const double ax = sx[3 * ti + 0], ay = sx[3 * ti + 1], az = sx[3 * ti + 2];
    vx = fma(sc.hh, ax, vx);
    vy = fma(sc.hh, ay, vy);
    vz = fma(sc.hh, az, vz);
    if (kicks == 2) {
      vx = fma(sc.hh, ax, vx);
      vy = fma(sc.hh, ay, vy);
      vz = fma(sc.hh, az, vz);
    }
    VA[3 * i + 0] = vx;
    VA[3 * i + 1] = vy;
    VA[3 * i + 2] = vz;
    if (!INIT) {
      XA[3 * i + 0] = xi;
      XA[3 * i + 1] = yi;
      XA[3 * i + 2] = zi;
    }
    if (last && !(isfinite(xi) && isfinite(yi) && isfinite(zi) && isfinite(vx) && isfinite(vy) &&
                  isfinite(vz)))
      atomicOr(status + s, ST_NAN);
  }
  // the marked slots' exact tests, pairs i < k by the lower index's lane,
  // after the kicks (so that nothing of the kick loop is held for them): the
  // partner's position formed as it was staged, the separation as the kick
  // formed it (partner minus own body)
  for (int wi = 0; wi < 2; ++wi) {
    unsigned int mark = wi == 0 ? mark0 : mark1;
    const int c = wi * sh.CS + cs;
    while (mark != 0u) {
      const int k = c * MB_CHUNK + __ffs(mark) - 1;
      mark &= mark - 1u;
      if (k <= i || k >= N) continue;
      double xk = X[3 * k + 0], yk = X[3 * k + 1], zk = X[3 * k + 2];
      if (!INIT) {
        xk = fma(sc.hh, q0, xk);
        yk = fma(sc.hh, q1, yk);
        zk = fma(sc.hh, q2, zk);
      }
      const double ex = xk - xi, ey = yk - yi, ez = zk - zi;
      const double r2 = dot3(ex, ey, ez, ex, ey, ez);
      const double sr = radi + bp.rad[k];
      if (r2 < sr * sr) ++ov;
      const double rc2k = bp.rc2[k];
      if (r2 < (rc2i > rc2k ? rc2i : rc2k)) {
        const unsigned int slot = atomicAdd(npairs, 1u);
        if (slot < (unsigned int)MB_EMAX) {
          edge[2 * slot] = i;
          edge[2 * slot + 1] = k;
        }
        flag[i] = 1;
        flag[k] = 1;
      }
    }
  }
  if (ov) atomicOr(status + s, ST_OVERLAP);
  if (stats) wave_add(stats + S_OVERLAP, ov);
}

// ---------------------------------------------------------------------------
// Encounter groups from the pair list, built in LDS by one wavefront (the
// workgroup of a group worker).  The bodies in pairs are marked in a bitmap of
// the system's N <= 2048 bodies (64 words, one per lane); a scan of the word
// popcounts ranks them, which gives the sorted compact list fl[F] and each
// pair's ends as compact indices.  Min-label propagation over the pairs then
// runs until no label changes, so every body carries the compact index of its
// group's smallest member (the root).  Groups are numbered by root; a group
// starts after the members of the groups with smaller roots, and a member's
// place in its group is the number of members before it (all of which lie
// between the root and itself in the compact order).  Outputs gst[G+1] and
// gm[F] (body indices); returns G, and sets big when a group has more than
// MB_GMAX members.  LDS (ints): bm[64], pre[64], and fl, L, ce, cnt, gst, gm
// of 2 MB_EMAX each (gst one more).
__device__ __forceinline__ int wave_incl_scan(int v, int) {
  // inclusive scans within 16-lane rows (row_shr 1, 2, 4, 8, zero fill),
  // then each row's total added to the rows after it (row_bcast 15, 31)
  v += __builtin_amdgcn_update_dpp(0, v, 0x111, 0xf, 0xf, true);
  v += __builtin_amdgcn_update_dpp(0, v, 0x112, 0xf, 0xf, true);
  v += __builtin_amdgcn_update_dpp(0, v, 0x114, 0xf, 0xf, true);
  v += __builtin_amdgcn_update_dpp(0, v, 0x118, 0xf, 0xf, true);
  v += __builtin_amdgcn_update_dpp(0, v, 0x142, 0xa, 0xf, false);
  v += __builtin_amdgcn_update_dpp(0, v, 0x143, 0xc, 0xf, false);
  return v;
}

static_assert(2 * MB_EMAX >= 128, "the first 64 pairs are read ahead of the pair count");

// e0, e1: the lane's entries lane and lane + 64 of the pair list, loaded by
// the caller alongside the pair count (the first 64 pairs).
__device__ int build_groups(int N, int ne, const int *__restrict__ edge, int e0, int e1,
                            unsigned int *bm, int *pre, int *fl, int *L, int *ce, int *cnt, int *gst,
                            int *gm, bool &big) {
  const int lane = (int)threadIdx.x & 63;  // one wavefront
  const int n2 = 2 * ne, nwd = (N + 31) >> 5;
  bm[lane] = 0u;
  wave_lds_sync();
  for (int p = lane; p < n2; p += 64) {
    const int b = p < 64 ? e0 : (p < 128 ? e1 : edge[p]);
    ce[p] = b;
    atomicOr(&bm[b >> 5], 1u << (b & 31));
  }
  wave_lds_sync();
  const unsigned int word = lane < nwd ? bm[lane] : 0u;
  const int pc = __popc(word);
  const int inc = wave_incl_scan(pc, lane);
  pre[lane] = inc - pc;
  const int F = __shfl(inc, 63, 64);
  {
    unsigned int w = word;
    int r = inc - pc;
    while (w) {
      const int b = __ffs(w) - 1;
      fl[r++] = lane * 32 + b;
      w &= w - 1u;
    }
  }
  wave_lds_sync();
  for (int p = lane; p < n2; p += 64) {
    const int b = ce[p];
    ce[p] = pre[b >> 5] + __popc(bm[b >> 5] & ((1u << (b & 31)) - 1u));
  }
  for (int c = lane; c < F; c += 64) {
    L[c] = c;
    cnt[c] = 0;
  }
  wave_lds_sync();
  for (;;) {
    int changed = 0;
    for (int e = lane; e < ne; e += 64) {
      const int a = ce[2 * e], b = ce[2 * e + 1];
      const int la = L[a], lb = L[b];
      if (la < lb) {
        atomicMin(&L[b], la);
        changed = 1;
      } else if (lb < la) {
        atomicMin(&L[a], lb);
        changed = 1;
      }
    }
    wave_lds_sync();
    if (!__any(changed)) break;
  }
  if (F <= 64) {
    // one compact body per lane: the groups in root order by ballots (a
    // group of at most 64 members is never too big)
    const int Lc = lane < F ? L[lane] : -1;
    unsigned long long roots = __ballot(lane < F && Lc == lane);
    int start = 0, g = 0;
    while (roots) {
      const int r = __ffsll((long long)roots) - 1;
      roots &= roots - 1ull;
      const unsigned long long mem = __ballot(Lc == r);
      if (Lc == r) gm[start + __popcll(mem & ((1ull << lane) - 1ull))] = fl[lane];
      if (lane == 0) gst[g] = start;
      start += __popcll(mem);
      ++g;
    }
    if (lane == 0) gst[g] = F;
    big = false;
    wave_lds_sync();
    return g;
  }
  for (int c = lane; c < F; c += 64) atomicAdd(&cnt[L[c]], 1);
  wave_lds_sync();
  // scan over the compact bodies of (root, size at root), PER a lane
  const int PER = (F + 63) >> 6;
  const int c0 = lane * PER;
  int sr = 0, sz = 0;
  bool mybig = false;
  for (int u = 0; u < PER; ++u) {
    const int c = c0 + u;
    if (c < F && L[c] == c) {
      sr += 1;
      sz += cnt[c];
      mybig = mybig || cnt[c] > MB_GMAX;
    }
  }
  const int ir = wave_incl_scan(sr, lane), iz = wave_incl_scan(sz, lane);
  const int G = __shfl(ir, 63, 64);
  int xr = ir - sr, xz = iz - sz;
  wave_lds_sync();
  for (int u = 0; u < PER; ++u) {
    const int c = c0 + u;
    if (c < F && L[c] == c) {
      const int size = cnt[c];
      gst[xr++] = xz;
      cnt[c] = xz;  // group start, read by the members below
      xz += size;
    }
  }
  wave_lds_sync();
  for (int c = lane; c < F; c += 64) {
    const int r = L[c];
    int rank = 0;
    for (int q = r; q < c; ++q) rank += (L[q] == r);
    gm[cnt[r] + rank] = fl[c];
  }
  if (lane == 0) gst[G] = F;
  big = __any(mybig);
  wave_lds_sync();
  return G;
}

// ---------------------------------------------------------------------------
// Bulirsch-Stoer drift of one encounter group by one wavefront.  The wave is
// split into row groups of GP lanes (GP: the group size rounded up to a power
// of two); lane a of a row group holds member a's state (6 doubles), its mass
// and changeover fields, and each right-hand side reads the other members of
// its row group by wavefront shuffles.  The Gragg rows of an attempt run side
// by side, rows kb .. kb+R-1 one per row group (R = min(64 / GP, KMAX)), in
// lockstep by substep; as row k completes, every lane takes member a's row-k
// value from that row group and extends its own copy of the extrapolation
// tableau, so the tableau, the error (the max over members, which does not
// depend on the order it is taken in) and the convergence test are the same
// in every row group.  Every row and extrapolation step is formed as in the
// single-lane group BS and in the same order (same bits and right-hand-side
// counts); the latency of an attempt drops from 2 + 4 + ... + N_k substeps to
// N_k while k < R.  Once the halving budget is spent the rest of the interval
// is taken in one step and flagged, as in the single-lane group BS.
//
// GM > 0 (GM = GP - 1, row groups of 2 or 4 lanes): the partners of member a
// are taken in GM slots, slot s holding member s (s < a) or s + 1, so in
// member order without the member itself; their masses and pair changeover
// fields are read once per drift, every slot is formed without branches and
// slots past the group are selected away, so that the partners' chains and
// the Sun term interleave.  GM = 0: any group size, one member after the
// other.  Same values either way.
template <int GM>
__device__ uint32_t bs_group_wave(double *y, int a, int GP, int gs, double h, double mu, double eps2,
                                  double tol, double mj, double loj, double wvj, unsigned int &nrhs) {
  const int lane = (int)(threadIdx.x & 63u);
  const int row = lane / GP, base = row * GP;
  const int R = 64 / GP < KMAX ? 64 / GP : KMAX;
  const bool act = a < gs;
  constexpr int NS = GM > 0 ? GM : 1;
  int psrc[NS];
  bool plive[NS];
  double pmb[NS], plo[NS], pwv[NS];
  if constexpr (GM > 0) {
#pragma unroll
    for (int q = 0; q < GM; ++q) {
      const int b = q < a ? q : q + 1;
      psrc[q] = base + b;
      plive[q] = b < gs;
      const double lob = __shfl(loj, psrc[q], 64), wvb = __shfl(wvj, psrc[q], 64);
      pmb[q] = __shfl(mj, psrc[q], 64);
      plo[q] = loj > lob ? loj : lob;
      pwv[q] = wvj < wvb ? wvj : wvb;
    }
  }
  auto rhs = [&](const double *z, double *f) {
    const double x = z[0], yy = z[1], zz = z[2];
    f[0] = z[3];
    f[1] = z[4];
    f[2] = z[5];
    const double r2 = dot3(x, yy, zz, x, yy, zz);
    const double r = sqrt_dist2(r2);
    const double w0 = div_normal(mu, r2 * r);
    double ax = -(w0 * x), ay = -(w0 * yy), az = -(w0 * zz);
    if constexpr (GM > 0) {
#pragma unroll
      for (int q = 0; q < GM; ++q) {
        const double xb = __shfl(x, psrc[q], 64), yb = __shfl(yy, psrc[q], 64), zb = __shfl(zz, psrc[q], 64);
        const double dx = xb - x, dy = yb - yy, dz = zb - zz;
        const double d2 = dot3(dx, dy, dz, dx, dy, dz);
        const double w = close_weight_sel(d2, eps2, pmb[q], plo[q], pwv[q]);
        ax = plive[q] ? fma(w, dx, ax) : ax;
        ay = plive[q] ? fma(w, dy, ay) : ay;
        az = plive[q] ? fma(w, dz, az) : az;
      }
    } else {
      for (int b = 0; b < gs; ++b) {
        const int src = base + b;
        const double xb = __shfl(x, src, 64), yb = __shfl(yy, src, 64), zb = __shfl(zz, src, 64);
        const double mb = __shfl(mj, src, 64), lob = __shfl(loj, src, 64), wvb = __shfl(wvj, src, 64);
        if (b == a) continue;
        const double dx = xb - x, dy = yb - yy, dz = zb - zz;
        const double d2 = dot3(dx, dy, dz, dx, dy, dz);
        const double w = close_weight(d2, eps2, mb, loj > lob ? loj : lob, wvj < wvb ? wvj : wvb);
        ax = fma(w, dx, ax);
        ay = fma(w, dy, ay);
        az = fma(w, dz, az);
      }
    }
    f[3] = act ? ax : 0.0;
    f[4] = act ? ay : 0.0;
    f[5] = act ? az : 0.0;
  };
  uint32_t st = 0;
  double t = 0.0, H = h;
  int depth = 0;
  while (t < h) {
    if (H > h - t) H = h - t;
    double f0[6];
    rhs(y, f0);
    ++nrhs;
    const double scx = sqrt_dist2(dot3(y[0], y[1], y[2], y[0], y[1], y[2]));
    const double scv = sqrt_dist2(dot3(y[3], y[4], y[5], y[3], y[4], y[5]));
    double Rt[KMAX][6], cur[6], z0[6], z1[6], mine[6];
    double hs = 0.0, h2 = 0.0;
    int Nm = 0, m = 0, kb = 0;
    bool conv = false;
#pragma unroll
    for (int k = 0; k < KMAX; ++k) {
      if (k == kb) {  // rows kb .. kb+R-1 start, one per row group
        const int km = kb + row;
        Nm = (row < R && km < KMAX) ? (2 << km) : 0;  // 0: no row for this group
        hs = Nm > 0 ? H / (double)Nm : 0.0;
        h2 = 2.0 * hs;
#pragma unroll
        for (int c = 0; c < 6; ++c) {
          z0[c] = y[c];
          z1[c] = fma(hs, f0[c], y[c]);
          mine[c] = 0.0;
        }
        m = 0;
      }
      // substeps up to row k's last: one right-hand side for every row still
      // running, then each row's own update (a row group's lanes share a row)
      const int Ns = 2 << k;
      while (m < Ns) {
        ++m;
        if (m <= Nm) {
          double f[6];
          rhs(z1, f);
          if (m < Nm) {
#pragma unroll
            for (int c = 0; c < 6; ++c) {
              const double z2 = fma(h2, f[c], z0[c]);
              z0[c] = z1[c];
              z1[c] = z2;
            }
          } else {
#pragma unroll
            for (int c = 0; c < 6; ++c) mine[c] = 0.5 * (z1[c] + fma(hs, f[c], z0[c]));
          }
        }
      }
      const int src = (k - kb) * GP + a;  // member a in the row group that ran row k
#pragma unroll
      for (int c = 0; c < 6; ++c) cur[c] = __shfl(mine[c], src, 64);
      nrhs += (unsigned int)Ns;
#pragma unroll
      for (int j = 1; j <= k; ++j) {
        const double fac = bs_fac(j);
#pragma unroll
        for (int c = 0; c < 6; ++c) {
          const double prev = Rt[j - 1][c];
          Rt[j - 1][c] = cur[c];
          cur[c] = cur[c] + div_normal(cur[c] - prev, fac);
        }
      }
#pragma unroll
      for (int c = 0; c < 6; ++c) Rt[k][c] = cur[c];
      if (k >= 1) {
        double err = 0.0;
        if (act) {
#pragma unroll
          for (int c = 0; c < 6; ++c) {
            const double e = div_normal(fabs(cur[c] - Rt[k - 1][c]), c < 3 ? scx : scv);
            err = e > err ? e : err;
          }
        }
        // the largest error over the wavefront (exact in any order): within
        // rows by DPP (pairs, quads, half rows), then across rows
        double o = dpp64<0xB1>(err);
        err = o > err ? o : err;
        o = dpp64<0x4E>(err);
        err = o > err ? o : err;
        o = dpp64<0x141>(err);
        err = o > err ? o : err;
        o = dpp64<0x140>(err);
        err = o > err ? o : err;
        o = __shfl_xor(err, 16, 64);
        err = o > err ? o : err;
        o = __shfl_xor(err, 32, 64);
        err = o > err ? o : err;
        if (err <= tol) {
          conv = true;
          break;
        }
      }
      if (k == kb + R - 1) kb += R;
    }
    if (conv || depth >= BS_MAXDEPTH) {
#pragma unroll
      for (int c = 0; c < 6; ++c) y[c] = cur[c];
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

// Drift, in workgroups of two wavefronts.  Q1, the momentum after the opening
// kick, gives jump 1: a body's drift starts from XA + (h/2) Q1.  Workgroups
// [0, nK): each wavefront forms Q1 (wave_momentum, in its own half of the
// LDS) and drifts its 64 bodies that are in no pair on their Kepler orbits
// (clearing the flags for the next step's check).  The others: wavefront 0
// forms Q1 while wavefront 1 builds the encounter groups from the step's
// pair list (count ctl[par], written by the previous kick), each in its own
// LDS; then every wavefront takes every 2 nG-th group.  The other pair count,
// ctl[par ^ 1], is emptied here for the closing kick, which appends the next
// step's pairs there.  LDS (drift_lds): Q1 (4 doubles), region A (the
// momentum staging), region B (build_groups' arrays, or the second
// wavefront's momentum staging in a Kepler workgroup).
constexpr int MB_DRIFT_BLOCK = 128;

__device__ __forceinline__ size_t mom_doubles(const MBShape &sh) {
  return 4 * (size_t)sh.CS * MB_CHUNK + 2 * (size_t)sh.CS;
}

__global__ void __launch_bounds__(MB_DRIFT_BLOCK)
    mb_drift_kernel(MBShape sh, int nK, int nG, int par, const double *__restrict__ bp_all,
                    const double *__restrict__ xa_all, const double *__restrict__ va_all,
                    double *__restrict__ xb_all, double *__restrict__ vb_all,
                    int *__restrict__ flag_all, const int *__restrict__ edge_all,
                    unsigned int *__restrict__ ctl_all, unsigned int *__restrict__ status,
                    unsigned long long *__restrict__ stats, StepConst sc, double M) {
  extern __shared__ double dlds[];
  const int N = sh.N, s = blockIdx.y;
  const int wv = (int)threadIdx.x >> 6, lane = (int)threadIdx.x & 63;
  const BP bp = bp_of(bp_all, s, N);
  const double *XA = xa_all + (size_t)s * N * 3;
  const double *VA = va_all + (size_t)s * N * 3;
  double *XB = xb_all + (size_t)s * N * 3;
  double *VB = vb_all + (size_t)s * N * 3;
  double *sq = dlds;                           // Q1
  double *regA = dlds + 4;                     // momentum staging
  double *regB = regA + mom_doubles(sh);       // group arrays
  const size_t WBC = 3 * (size_t)sh.CS * MB_CHUNK + sh.CS;  // staged velocities (lds3 layout)
  if (blockIdx.x == 0 && threadIdx.x == 0) ctl_all[4 * s + (par ^ 1)] = 0u;
  if ((int)blockIdx.x < nK) {
    int *flag = flag_all + (size_t)s * N;
    const int i = (int)blockIdx.x * MB_DRIFT_BLOCK + (int)threadIdx.x;
    const int ic = i < N ? i : N - 1;
    // the body's own loads, in flight with the momentum sum's
    const int fl = flag[ic];
    double x = XA[3 * ic + 0], y = XA[3 * ic + 1], z = XA[3 * ic + 2];
    double vx = VA[3 * ic + 0], vy = VA[3 * ic + 1], vz = VA[3 * ic + 2];
    double *mr = wv == 0 ? regA : regB;
    double q0, q1, q2;
    wave_momentum(VA, bp.m, sh, M, mr, mr + WBC, q0, q1, q2);
    if (i >= N) return;
    if (fl) {
      flag[i] = 0;
      return;
    }
    x = fma(sc.hh, q0, x);
    y = fma(sc.hh, q1, y);
    z = fma(sc.hh, q2, z);
    bool ok;
    kepler_drift_series(x, y, z, vx, vy, vz, sc.mu, sc.h, ok);
    if (!ok) atomicOr(status + s, ST_NEWTON);
    XB[3 * i + 0] = x;
    XB[3 * i + 1] = y;
    XB[3 * i + 2] = z;
    VB[3 * i + 0] = vx;
    VB[3 * i + 1] = vy;
    VB[3 * i + 2] = vz;
    return;
  }
  const int *edge = edge_all + (size_t)s * MB_EMAX * 2;
  // the pair count and the first 64 pairs (2 MB_EMAX entries are there to
  // read; those past the count are not used) in flight together
  const int e0 = edge[lane], e1 = edge[lane + 64];
  const int ne = (int)ctl_all[4 * s + par];
  const bool lead = (int)blockIdx.x == nK && threadIdx.x == 64;  // the first builder's first lane
  if (ne == 0) return;
  if (ne > MB_EMAX) {
    if (lead) atomicOr(status + s, ST_LIMIT);
    return;
  }
  int *ilds = (int *)regB;
  unsigned int *bm = (unsigned int *)ilds;
  int *pre = ilds + 64, *fl = pre + 64, *L = fl + 2 * MB_EMAX, *ce = L + 2 * MB_EMAX;
  int *cnt = ce + 2 * MB_EMAX, *gst = cnt + 2 * MB_EMAX, *gm = gst + 2 * MB_EMAX + 1;
  if (wv == 0) {
    double q0, q1, q2;
    wave_momentum(VA, bp.m, sh, M, regA, regA + WBC, q0, q1, q2);
    if (lane == 0) {
      sq[0] = q0;
      sq[1] = q1;
      sq[2] = q2;
    }
  } else {
    bool big = false;
    const int G = build_groups(N, ne, edge, e0, e1, bm, pre, fl, L, ce, cnt, gst, gm, big);
    if (lane == 0) ((int *)(sq + 3))[0] = G;
    if (lead) {
      atomicOr(status + s, big ? (ST_MASSIVE_ENC | ST_LIMIT) : ST_MASSIVE_ENC);
      if (stats) {
        atomicAdd(stats + S_MGROUPS, (unsigned long long)G);
        atomicAdd(stats + S_MGBODIES, (unsigned long long)gst[G]);
      }
    }
  }
  __syncthreads();
  const double q[3] = {sq[0], sq[1], sq[2]};
  const int G = ((const int *)(sq + 3))[0];
  unsigned int nrhs = 0;
  uint32_t st = 0;
  const int nw = 2 * nG;  // group wavefronts
  for (int g = 2 * ((int)blockIdx.x - nK) + wv; g < G; g += nw) {
    const int g0 = gst[g], gs = gst[g + 1] - g0;
    if (gs > MB_GMAX) continue;
    int GP = 1;  // lanes per row group: the group size rounded up to a power of two
    while (GP < gs) GP <<= 1;
    const int a = lane & (GP - 1);  // member of this lane in its row group
    const int j = a < gs ? gm[g0 + a] : -1;
    // lanes past the group carry a unit position so their own (unused)
    // Sun term stays finite
    double y[6] = {1.0, 0.0, 0.0, 0.0, 0.0, 0.0};
    double mj = 0.0, loj = 0.0, wvj = 0.0;
    if (j >= 0) {
      for (int c = 0; c < 3; ++c) {
        y[c] = fma(sc.hh, q[c], XA[3 * j + c]);
        y[3 + c] = VA[3 * j + c];
      }
      mj = bp.m[j];
      loj = bp.lo[j];
      wvj = bp.winv[j];
    }
    if (GP == 2)
      st |= bs_group_wave<1>(y, a, GP, gs, sc.h, sc.mu, sc.eps2, sc.tol, mj, loj, wvj, nrhs);
    else if (GP == 4)
      st |= bs_group_wave<3>(y, a, GP, gs, sc.h, sc.mu, sc.eps2, sc.tol, mj, loj, wvj, nrhs);
    else
      st |= bs_group_wave<0>(y, a, GP, gs, sc.h, sc.mu, sc.eps2, sc.tol, mj, loj, wvj, nrhs);
    if (j >= 0 && lane < GP)  // every row group ends with the same state; the first writes it
      for (int c = 0; c < 3; ++c) {
        XB[3 * j + c] = y[c];
        VB[3 * j + c] = y[3 + c];
      }
  }
  if (st && lane == 0) atomicOr(status + s, st);
}

// Padded staged window (lds3 / lds1 layouts): 3 WB + CS doubles of positions,
// WB + CS per scalar array.  At the 2048-body limit (nch 128, TI 4, CS 64) the
// kick kernel takes 63,512 bytes, within 64 KiB.
size_t close_lds(const MBShape &sh) {
  const size_t WB = (size_t)sh.CS * MB_CHUNK;
  return ((size_t)sh.nch * sh.TI * 3 + 3 + (3 * WB + sh.CS) + 3 * (WB + sh.CS)) * sizeof(double);
}

// The drift's LDS: Q1 (4 doubles), region A (a momentum staging: 3 WB + CS
// and WB + CS doubles), region B (the group building's arrays, or a second
// momentum staging), up to 84 KB at the 2048-body limit (gfx950 gives a
// workgroup up to 160 KB).
size_t drift_lds(const MBShape &sh) {
  const size_t groups = ((size_t)64 + 64 + 12 * MB_EMAX + 1) * sizeof(int);
  const size_t WB = (size_t)sh.CS * MB_CHUNK;
  const size_t mom = (4 * WB + 2 * (size_t)sh.CS) * sizeof(double);
  return 4 * sizeof(double) + mom + std::max(groups, mom);
}

// The kicks of one launch; check: with the pair check of the next step, whose
// pairs are counted in parity par's counter.
int launch_close(Ctx *c, bool init, int kicks, int last, int check, int par) {
  const MBShape sh = mb_shape(c->NM);
  const dim3 grid((unsigned int)((sh.N + sh.TI - 1) / sh.TI), (unsigned int)c->S);
  const StepConst sc = c->step_const();
  unsigned long long *stats = (c->cfg.collect_stats & 1) ? c->d_stats : nullptr;
  if (init)
    mb_close_kernel<true><<<grid, MB_NT, close_lds(sh), c->stream>>>(
        sh, c->d_bp, c->d_mbmax, c->d_mx, c->d_mv, c->d_mxb, c->d_mvb, kicks, last, check, par, c->d_mflag,
        c->d_medge, c->d_mctl, c->d_status, stats, sc, c->cfg.central_mass);
  else
    mb_close_kernel<false><<<grid, MB_NT, close_lds(sh), c->stream>>>(
        sh, c->d_bp, c->d_mbmax, c->d_mx, c->d_mv, c->d_mxb, c->d_mvb, kicks, last, check, par, c->d_mflag,
        c->d_medge, c->d_mctl, c->d_status, stats, sc, c->cfg.central_mass);
  return c->hip(hipGetLastError(), "many-body kick launch");
}

int launch_drift(Ctx *c, int par) {
  const MBShape sh = mb_shape(c->NM);
  const int nK = (sh.N + MB_DRIFT_BLOCK - 1) / MB_DRIFT_BLOCK;
  const int nG = MB_DRIFT_GROUP_BLOCKS;
  const dim3 grid((unsigned int)(nK + nG), (unsigned int)c->S);
  const size_t lds = drift_lds(sh);  // allowed above 64 KB by mb_alloc
  unsigned long long *stats = (c->cfg.collect_stats & 1) ? c->d_stats : nullptr;
  mb_drift_kernel<<<grid, MB_DRIFT_BLOCK, lds, c->stream>>>(sh, nK, nG, par, c->d_bp, c->d_mx,
                                                          c->d_mv, c->d_mxb, c->d_mvb, c->d_mflag,
                                                          c->d_medge, c->d_mctl, c->d_status, stats,
                                                          c->step_const(), c->cfg.central_mass);
  return c->hip(hipGetLastError(), "many-body drift launch");
}

// Step k of a call (par = k mod 2: its pairs were counted in ctl[par] by the
// kick before it): the drift, then the kicks with the next step's pair check
// unless it is the call's last step.
int enqueue_step(Ctx *c, int kicks, int last, int par, KTimer *kt) {
  int rc;
  if (kt) kt->start();
  if ((rc = launch_drift(c, par))) return rc;
  if (kt) kt->stop(c->timing.ms_mb_drift, c->timing.n_mb_drift);
  if (kt) kt->start();
  if ((rc = launch_close(c, false, kicks, last, last ? 0 : 1, par ^ 1))) return rc;
  if (kt) kt->stop(c->timing.ms_mb_close, c->timing.n_mb_close);
  return MINT_OK;
}

int mb_graph(Ctx *c, hipGraphExec_t *out) {
  const GraphSig sig = c->graph_sig();
  if (c->mb_graph && c->mb_graph_steps == MB_GRAPH_STEPS && c->mb_sig.same(sig)) {
    *out = c->mb_graph;
    return MINT_OK;
  }
  if (c->mb_graph) {  // captured with other constants (a checkpoint of another dt, say)
    hipGraphExecDestroy(c->mb_graph);
    c->mb_graph = nullptr;
    c->mb_graph_steps = 0;
  }
  int rc = c->hip(hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal), "begin capture");
  if (rc) return rc;
  // replays start at even steps of a call (MB_GRAPH_STEPS is even)
  for (int k = 0; k < MB_GRAPH_STEPS && !rc; ++k) rc = enqueue_step(c, 2, 0, k & 1, nullptr);
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
  if ((rc = c->hip(hipGraphUpload(ex, c->stream), "graph upload")) ||
      (rc = c->hip(hipStreamSynchronize(c->stream), "graph upload"))) {
    hipGraphExecDestroy(ex);
    return rc;
  }
  c->mb_graph = ex;
  c->mb_graph_steps = MB_GRAPH_STEPS;
  c->mb_sig = sig;
  *out = ex;
  return MINT_OK;
}

int upload_params(Ctx *c) {
  const int S = c->S, N = c->NM;
  std::vector<double> bp((size_t)S * 5 * N), mx((size_t)S * 2, 0.0);
  for (int s = 0; s < S; ++s) {
    double *b = bp.data() + (size_t)s * 5 * N;
    for (int j = 0; j < N; ++j) {
      const double rc = c->rc[(size_t)s * N + j];
      b[j] = c->m[(size_t)s * N + j];
      b[N + j] = 0.1 * rc;
      b[2 * N + j] = 1.0 / (0.9 * rc);
      b[3 * N + j] = rc * rc;
      b[4 * N + j] = c->rad.empty() ? 0.0 : c->rad[(size_t)s * N + j];
      // the pair check's bounds: the largest r_c^2 and radius of the system
      mx[2 * s] = std::max(mx[2 * s], b[3 * N + j]);
      mx[2 * s + 1] = std::max(mx[2 * s + 1], b[4 * N + j]);
    }
  }
  int rc = c->hip(hipMemcpy(c->d_bp, bp.data(), bp.size() * sizeof(double), hipMemcpyHostToDevice),
                  "H2D body parameters");
  if (rc) return rc;
  return c->hip(hipMemcpy(c->d_mbmax, mx.data(), mx.size() * sizeof(double), hipMemcpyHostToDevice),
                "H2D pair-check bounds");
}

}  // namespace

int mb_alloc(Ctx *c) {
  const int S = c->S, N = c->NM;
  const size_t b3 = (size_t)S * N * 3 * sizeof(double);
  int rc;
  if ((rc = c->dalloc((void **)&c->d_bp, (size_t)S * 5 * N * sizeof(double), "body parameters"))) return rc;
  if ((rc = c->dalloc((void **)&c->d_mxb, b3, "post-drift positions"))) return rc;
  if ((rc = c->dalloc((void **)&c->d_mvb, b3, "post-drift velocities"))) return rc;
  if ((rc = c->dalloc((void **)&c->d_mbmax, (size_t)S * 2 * sizeof(double), "pair-check bounds"))) return rc;
  if ((rc = c->dalloc((void **)&c->d_mflag, (size_t)S * N * sizeof(int), "encounter flags"))) return rc;
  if ((rc = c->dalloc((void **)&c->d_medge, (size_t)S * MB_EMAX * 2 * sizeof(int), "encounter pairs"))) return rc;
  if ((rc = c->dalloc((void **)&c->d_mctl, (size_t)S * 4 * sizeof(unsigned int), "group counters"))) return rc;
  // the drift's workgroups take more than 64 KB of LDS (allowed here for the
  // largest system any context may hold, so contexts of every size launch)
  return c->hip(hipFuncSetAttribute((const void *)mb_drift_kernel, hipFuncAttributeMaxDynamicSharedMemorySize,
                                    (int)drift_lds(mb_shape(MB_MAXN))),
                "many-body drift LDS size");
}

int mb_reset(Ctx *c) {
  const int S = c->S, N = c->NM;
  int rc;
  if ((rc = c->hip(hipMemset(c->d_mflag, 0, (size_t)S * N * sizeof(int)), "flags clear"))) return rc;
  if ((rc = c->hip(hipMemset(c->d_mctl, 0, (size_t)S * 4 * sizeof(unsigned int)), "counters clear"))) return rc;
  return upload_params(c);
}

int mb_set_radii(Ctx *c) { return upload_params(c); }

int run_many(Ctx *c, int64_t nsteps) {
  int rc;
  KTimer kt(c);
  KTimer *ktp = timing_on(c) ? &kt : nullptr;
  // both pair counts empty (the first kick counts step 0's pairs in ctl[0])
  if ((rc = c->hip(hipMemsetAsync(c->d_mctl, 0, (size_t)c->S * 4 * sizeof(unsigned int), c->stream),
                   "pair counts clear")))
    return rc;
  if (ktp) ktp->start();
  if ((rc = launch_close(c, true, 1, 0, 1, 0))) return rc;
  if (ktp) ktp->stop(c->timing.ms_mb_close, c->timing.n_mb_close);
  int64_t rest = nsteps - 1, k = 0;
  if (c->cfg.use_graph && !ktp && rest >= MB_GRAPH_STEPS) {
    hipGraphExec_t ex = nullptr;
    if ((rc = mb_graph(c, &ex))) return rc;
    for (; rest >= MB_GRAPH_STEPS; rest -= MB_GRAPH_STEPS, k += MB_GRAPH_STEPS)
      if ((rc = c->hip(hipGraphLaunch(ex, c->stream), "graph launch"))) return rc;
  }
  for (; rest > 0; --rest, ++k)
    if ((rc = enqueue_step(c, 2, 0, (int)(k & 1), ktp))) return rc;
  if ((rc = enqueue_step(c, 1, 1, (int)(k & 1), ktp))) return rc;
  if ((rc = c->hip(hipStreamSynchronize(c->stream), "many-body step"))) return rc;
  std::vector<unsigned int> st((size_t)c->S);
  if ((rc = c->hip(hipMemcpy(st.data(), c->d_status, st.size() * sizeof(unsigned int), hipMemcpyDeviceToHost),
                   "D2H status")))
    return rc;
  for (int s = 0; s < c->S; ++s)
    if (st[s] & ST_LIMIT)
      return c->fail(MINT_ELIMIT, "many-body engine: system " + std::to_string(s) + " had more than " +
                                      std::to_string(MB_EMAX) + " encounter pairs or an encounter group of more "
                                      "than " + std::to_string(MB_GMAX) + " bodies in one step; its state is not valid");
  return MINT_OK;
}

}  // namespace mint
