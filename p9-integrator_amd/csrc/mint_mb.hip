// mint_mb.hip — the many-body engine: systems of more than 8 massive bodies
// and no test particles (the all-massive planetary-embryo configuration).
//
// It integrates the same massive-subsystem step as the lane-per-system
// ephemeris kernel, operation for operation, so the two agree to the last bit
// where both apply; what changes is the parallel shape.  One step is three
// launches per system set (blockIdx.y = system):
//
//   mb_open_kernel   momentum after the opening kick, jump 1, and the exact
//                    pair check of every pair (j < k) against its changeover
//                    radius; pairs inside are appended to an edge list.  The
//                    last workgroup to finish turns the edges into encounter
//                    groups (connected components, listed by smallest member).
//   mb_drift_kernel  Kepler drift of every body in no group (one lane each);
//                    Bulirsch-Stoer drift of each group (one wavefront each,
//                    one member per lane, positions exchanged through LDS).
//   mb_close_kernel  momentum after the drift, jump 2, O(N^2) kick
//                    accelerations and the closing kick (fused with the next
//                    step's opening kick, which sees the same positions).
//
// Reference: body_interaction_kick with every body massive (src/sim.cuh:327-382),
// main_body_kinetic (src/sim.cuh:384-406), close_encounter_p
// (src/sim.cuh:589-613) and the BS drift of flagged bodies
// (src/sim.cuh:410-551,699-730).  The reference integrates each body against
// frozen partners; here a group's members are integrated together.
//
// Kick sums: lane (i, chunk) accumulates the 16 partners of one chunk from
// zero, two division groups of 8 with one division each (batched_div), the
// body itself and bodies past N as inert slots; the chunk sums are then added
// in chunk order.  The momentum sum uses the same chunks.
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
  int CS;   // chunk slots per workgroup; TI * CS = MB_BLOCK
};

constexpr int MB_BLOCK = 256;
constexpr int MB_DRIFT_GROUP_BLOCKS = 32;  // workgroups that walk the group list
constexpr int MB_GRAPH_STEPS = 64;

MBShape mb_shape(int N) {
  MBShape s;
  s.N = N;
  s.nch = (N + MB_CHUNK - 1) / MB_CHUNK;
  int cs = 1;
  while (cs < s.nch && cs < MB_BLOCK) cs <<= 1;
  s.CS = cs;
  s.TI = MB_BLOCK / cs;
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

// Q = (sum_j m_j v_j)/M by the whole workgroup: chunk sums (from zero, index
// order) in part[], then added in chunk order by one thread.
__device__ void wg_momentum(const double *__restrict__ V, const double *__restrict__ m, int N,
                            int nch, double M, double *part, double *Q) {
  for (int c = threadIdx.x; c < nch; c += blockDim.x) {
    double px = 0.0, py = 0.0, pz = 0.0;
    const int k1 = min(N, (c + 1) * MB_CHUNK);
    for (int k = c * MB_CHUNK; k < k1; ++k) {
      const double mk = m[k];
      px = fma(mk, V[3 * k + 0], px);
      py = fma(mk, V[3 * k + 1], py);
      pz = fma(mk, V[3 * k + 2], pz);
    }
    part[3 * c + 0] = px;
    part[3 * c + 1] = py;
    part[3 * c + 2] = pz;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    double qx = part[0], qy = part[1], qz = part[2];
    for (int c = 1; c < nch; ++c) {
      qx = qx + part[3 * c + 0];
      qy = qy + part[3 * c + 1];
      qz = qz + part[3 * c + 2];
    }
    Q[0] = qx / M;
    Q[1] = qy / M;
    Q[2] = qz / M;
  }
  __syncthreads();
}

// ---------------------------------------------------------------------------
// Kicks.  INIT: positions XA as they are, velocities VA, one kick (the opening
// kick of a call's first step).  Otherwise positions XB + (h/2) Q2 (jump 2)
// are stored to XA, velocities VB, and kicks = 1 (closing kick of a call's
// last step) or 2 (closing kick, then the next step's opening kick).
template <bool INIT>
__global__ void __launch_bounds__(MB_BLOCK)
    mb_close_kernel(MBShape sh, const double *__restrict__ bp_all, double *__restrict__ xa_all,
                    double *__restrict__ va_all, const double *__restrict__ xb_all,
                    const double *__restrict__ vb_all, int kicks, int last,
                    unsigned int *__restrict__ status, StepConst sc, double M) {
  extern __shared__ double lds[];
  const int N = sh.N, s = blockIdx.y;
  const BP bp = bp_of(bp_all, s, N);
  double *XA = xa_all + (size_t)s * N * 3;
  double *VA = va_all + (size_t)s * N * 3;
  const double *XB = xb_all + (size_t)s * N * 3;
  const double *VB = vb_all + (size_t)s * N * 3;
  double *part = lds;                            // [nch][TI][3]
  double *qpart = lds + (size_t)sh.nch * sh.TI * 3;  // [nch][3]
  double *Qs = qpart + 3 * sh.nch;               // [3]
  const double *X = INIT ? XA : XB;
  double q0 = 0.0, q1 = 0.0, q2 = 0.0;
  if (!INIT) {
    wg_momentum(VB, bp.m, N, sh.nch, M, qpart, Qs);
    q0 = Qs[0];
    q1 = Qs[1];
    q2 = Qs[2];
  }
  const int ti = (int)threadIdx.x % sh.TI, cs = (int)threadIdx.x / sh.TI;
  const int i = (int)blockIdx.x * sh.TI + ti;
  const bool act = i < N;
  double xi = 0.0, yi = 0.0, zi = 0.0, loi = 0.0, wvi = 0.0;
  if (act) {
    xi = X[3 * i + 0];
    yi = X[3 * i + 1];
    zi = X[3 * i + 2];
    if (!INIT) {
      xi = fma(sc.hh, q0, xi);
      yi = fma(sc.hh, q1, yi);
      zi = fma(sc.hh, q2, zi);
    }
    loi = bp.lo[i];
    wvi = bp.winv[i];
  }
  for (int c = cs; c < sh.nch; c += sh.CS) {
    double px = 0.0, py = 0.0, pz = 0.0;
    if (act) {
#pragma unroll
      for (int hg = 0; hg < MB_CHUNK / MB_DIVG; ++hg) {
        double dx[MB_DIVG], dy[MB_DIVG], dz[MB_DIVG], num[MB_DIVG], den[MB_DIVG], w[MB_DIVG];
#pragma unroll
        for (int q = 0; q < MB_DIVG; ++q) {
          const int k = c * MB_CHUNK + hg * MB_DIVG + q;
          dx[q] = 0.0;
          dy[q] = 0.0;
          dz[q] = 0.0;
          num[q] = 0.0;
          den[q] = 1.0;
          if (k < N && k != i) {
            double xk = X[3 * k + 0], yk = X[3 * k + 1], zk = X[3 * k + 2];
            if (!INIT) {
              xk = fma(sc.hh, q0, xk);
              yk = fma(sc.hh, q1, yk);
              zk = fma(sc.hh, q2, zk);
            }
            dx[q] = xk - xi;
            dy[q] = yk - yi;
            dz[q] = zk - zi;
            const double r2 = dot3(dx[q], dy[q], dz[q], dx[q], dy[q], dz[q]);
            const double lok = bp.lo[k], wvk = bp.winv[k];
            kick_nd(r2, sc.eps2, bp.m[k], loi > lok ? loi : lok, wvi < wvk ? wvi : wvk, num[q],
                    den[q]);
          }
        }
        batched_div<MB_DIVG>(num, den, w);
#pragma unroll
        for (int q = 0; q < MB_DIVG; ++q) {
          px = fma(w[q], dx[q], px);
          py = fma(w[q], dy[q], py);
          pz = fma(w[q], dz[q], pz);
        }
      }
    }
    double *pp = part + ((size_t)c * sh.TI + ti) * 3;
    pp[0] = px;
    pp[1] = py;
    pp[2] = pz;
  }
  __syncthreads();
  if (cs == 0 && act) {
    double ax = part[3 * ti + 0], ay = part[3 * ti + 1], az = part[3 * ti + 2];
    for (int c = 1; c < sh.nch; ++c) {
      const double *pp = part + ((size_t)c * sh.TI + ti) * 3;
      ax = ax + pp[0];
      ay = ay + pp[1];
      az = az + pp[2];
    }
    const double *V = INIT ? VA : VB;
    double vx = V[3 * i + 0], vy = V[3 * i + 1], vz = V[3 * i + 2];
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
}

// ---------------------------------------------------------------------------
// Encounter groups from the edge list, by the last workgroup of the open
// kernel.  Labels: min-label propagation over the edges until no label
// changes, so every member ends with its group's smallest index.  Then one
// block-wide exclusive scan over the bodies of (flagged, root, size at root)
// gives each flagged body its slot in the compact list, each root its group
// number and the group's start in the member list; members keep index order.
// LDS: Lb[N], Fb[N], Cb[N], E[2 MB_EMAX], scr[64] (ints).
__device__ void build_groups(int N, unsigned int ne, const int *__restrict__ edge,
                             unsigned int *ctl, int *__restrict__ gstart, int *__restrict__ gmem,
                             int *Lb, int *Fb, int *Cb, int *E, int *scr,
                             unsigned int *status_s, unsigned long long *stats) {
  const int tid = threadIdx.x, bd = blockDim.x;
  for (int i = tid; i < N; i += bd) {
    Lb[i] = i;
    Fb[i] = 0;
    Cb[i] = 0;
  }
  __syncthreads();
  for (int e = tid; e < (int)ne; e += bd) {
    const int a = edge[2 * e], b = edge[2 * e + 1];
    E[2 * e] = a;
    E[2 * e + 1] = b;
    Fb[a] = 1;
    Fb[b] = 1;
  }
  __syncthreads();
  for (;;) {
    if (tid == 0) scr[0] = 0;
    __syncthreads();
    for (int e = tid; e < (int)ne; e += bd) {
      const int a = E[2 * e], b = E[2 * e + 1];
      const int la = Lb[a], lb = Lb[b];
      if (la < lb) {
        atomicMin(&Lb[b], la);
        scr[0] = 1;
      } else if (lb < la) {
        atomicMin(&Lb[a], lb);
        scr[0] = 1;
      }
    }
    __syncthreads();
    const int changed = scr[0];
    __syncthreads();
    if (!changed) break;
  }
  for (int i = tid; i < N; i += bd)
    if (Fb[i]) atomicAdd(&Cb[Lb[i]], 1);
  __syncthreads();
  // block-wide exclusive scan of (flagged, root, size at root), PER bodies a thread
  const int PER = (N + bd - 1) / bd;
  const int e0 = tid * PER;
  int sf = 0, sr = 0, sz = 0;
  for (int u = 0; u < PER; ++u) {
    const int e = e0 + u;
    if (e < N && Fb[e]) {
      sf += 1;
      if (Lb[e] == e) {
        sr += 1;
        sz += Cb[e];
      }
    }
  }
  const int lane = tid & 63, wv = tid >> 6, nw = bd >> 6;
  int inf = sf, inr = sr, inz = sz;
  for (int off = 1; off < 64; off <<= 1) {
    const int tf = __shfl_up(inf, off, 64), tr = __shfl_up(inr, off, 64), tz = __shfl_up(inz, off, 64);
    if (lane >= off) {
      inf += tf;
      inr += tr;
      inz += tz;
    }
  }
  if (lane == 63) {
    scr[1 + 3 * wv] = inf;
    scr[2 + 3 * wv] = inr;
    scr[3 + 3 * wv] = inz;
  }
  __syncthreads();
  int bf = 0, br = 0, bz = 0, tf = 0, tr = 0;
  for (int w = 0; w < nw; ++w) {
    if (w < wv) {
      bf += scr[1 + 3 * w];
      br += scr[2 + 3 * w];
      bz += scr[3 + 3 * w];
    }
    tf += scr[1 + 3 * w];
    tr += scr[2 + 3 * w];
  }
  int xf = bf + inf - sf, xr = br + inr - sr, xz = bz + inz - sz;
  int *fl = E;  // the edges are no longer needed: compact flagged list
  bool big = false;
  __syncthreads();
  for (int u = 0; u < PER; ++u) {
    const int e = e0 + u;
    if (e < N && Fb[e]) {
      fl[xf++] = e;
      if (Lb[e] == e) {
        const int size = Cb[e];
        big = big || size > MB_GMAX;
        gstart[xr++] = xz;
        Cb[e] = xz;  // group start, read by the members below
        xz += size;
      }
    }
  }
  __syncthreads();
  for (int e = tid; e < N; e += bd) {
    if (!Fb[e]) continue;
    const int r = Lb[e];
    int rank = 0;
    for (int p = 0; fl[p] != e; ++p) rank += (Lb[fl[p]] == r);
    gmem[Cb[r] + rank] = e;
  }
  if (big) atomicOr(status_s, ST_LIMIT);
  if (tid == 0) {
    gstart[tr] = tf;
    ctl[2] = (unsigned int)tr;
    ctl[3] = (unsigned int)tf;
    if (tr > 0) atomicOr(status_s, ST_MASSIVE_ENC);
    if (stats) {
      atomicAdd(stats + S_MGROUPS, (unsigned long long)tr);
      atomicAdd(stats + S_MGBODIES, (unsigned long long)tf);
    }
  }
}

// Momentum after the opening kick, jump 1 (stored to Y by the body's owner)
// and the exact pair check (close_encounter_p, src/sim.cuh:589-613) of every
// pair j < k against r_c^2 of the pair.  Also counts pairs closer than the
// sum of their physical radii.
__global__ void __launch_bounds__(MB_BLOCK)
    mb_open_kernel(MBShape sh, const double *__restrict__ bp_all, const double *__restrict__ xa_all,
                   const double *__restrict__ va_all, double *__restrict__ y_all,
                   int *__restrict__ flag_all, int *__restrict__ edge_all,
                   unsigned int *__restrict__ ctl_all, int *__restrict__ gstart_all,
                   int *__restrict__ gmem_all, unsigned int *__restrict__ status,
                   unsigned long long *__restrict__ stats, StepConst sc, double M) {
  extern __shared__ double lds[];
  __shared__ int is_last;
  const int N = sh.N, s = blockIdx.y;
  const BP bp = bp_of(bp_all, s, N);
  const double *XA = xa_all + (size_t)s * N * 3;
  const double *VA = va_all + (size_t)s * N * 3;
  double *Y = y_all + (size_t)s * N * 3;
  int *flag = flag_all + (size_t)s * N;
  int *edge = edge_all + (size_t)s * MB_EMAX * 2;
  unsigned int *ctl = ctl_all + 4 * s;
  double *qpart = lds;
  double *Qs = lds + 3 * sh.nch;
  wg_momentum(VA, bp.m, N, sh.nch, M, qpart, Qs);
  const double q0 = Qs[0], q1 = Qs[1], q2 = Qs[2];
  const int ti = (int)threadIdx.x % sh.TI, cs = (int)threadIdx.x / sh.TI;
  const int i = (int)blockIdx.x * sh.TI + ti;
  const bool act = i < N;
  unsigned int ov = 0;
  if (act) {
    const double xi = fma(sc.hh, q0, XA[3 * i + 0]);
    const double yi = fma(sc.hh, q1, XA[3 * i + 1]);
    const double zi = fma(sc.hh, q2, XA[3 * i + 2]);
    if (cs == 0) {
      Y[3 * i + 0] = xi;
      Y[3 * i + 1] = yi;
      Y[3 * i + 2] = zi;
    }
    const double rc2i = bp.rc2[i], radi = bp.rad[i];
    for (int c = cs; c < sh.nch; c += sh.CS) {
      const int k1 = min(N, (c + 1) * MB_CHUNK);
      for (int k = max(c * MB_CHUNK, i + 1); k < k1; ++k) {
        const double dx = fma(sc.hh, q0, XA[3 * k + 0]) - xi;
        const double dy = fma(sc.hh, q1, XA[3 * k + 1]) - yi;
        const double dz = fma(sc.hh, q2, XA[3 * k + 2]) - zi;
        const double r2 = dot3(dx, dy, dz, dx, dy, dz);
        const double sr = radi + bp.rad[k];
        if (r2 < sr * sr) ++ov;
        const double rc2k = bp.rc2[k];
        if (r2 < (rc2i > rc2k ? rc2i : rc2k)) {
          const unsigned int slot = atomicAdd(ctl + 0, 1u);
          if (slot < (unsigned int)MB_EMAX) {
            edge[2 * slot] = i;
            edge[2 * slot + 1] = k;
          }
          flag[i] = 1;
          flag[k] = 1;
        }
      }
    }
  }
  if (ov) atomicOr(status + s, ST_OVERLAP);
  if (stats) wave_add(stats + S_OVERLAP, ov);
  // the last workgroup of this system builds the groups
  __threadfence();
  __syncthreads();
  if (threadIdx.x == 0) is_last = (atomicAdd(ctl + 1, 1u) == gridDim.x - 1);
  __syncthreads();
  if (!is_last) return;
  __threadfence();
  const unsigned int ne = atomicAdd(ctl + 0, 0u);
  if (ne == 0 || ne > (unsigned int)MB_EMAX) {
    if (threadIdx.x == 0) {
      if (ne > (unsigned int)MB_EMAX) atomicOr(status + s, ST_LIMIT);
      ctl[2] = 0u;
      ctl[3] = 0u;
      ctl[0] = 0u;
      ctl[1] = 0u;
    }
    return;
  }
  int *ib = (int *)(lds + 3 * sh.nch + 3);
  build_groups(N, ne, edge, ctl, gstart_all + (size_t)s * (N + 1), gmem_all + (size_t)s * N, ib,
               ib + N, ib + 2 * N, ib + 3 * N, ib + 3 * N + 2 * MB_EMAX, status + s, stats);
  __syncthreads();
  if (threadIdx.x == 0) {
    ctl[0] = 0u;
    ctl[1] = 0u;
  }
}

// ---------------------------------------------------------------------------
// Bulirsch-Stoer drift of one encounter group by one wavefront: lane a holds
// member a's state (6 doubles) and tableau; each right-hand side publishes the
// members' positions in LDS.  Same operations, per member and component, as
// the single-lane group BS; the error is the max over members, which does not
// depend on the order it is taken in.
__device__ uint32_t bs_group_wave(double *y, int a, int gs, double h, double mu, double eps2,
                                  double tol, const double *pm, const double *plo,
                                  const double *pwv, double *pos, unsigned int &nrhs) {
  const bool act = a < gs;
  auto rhs = [&](const double *z, double *f) {
    pos[3 * a + 0] = z[0];
    pos[3 * a + 1] = z[1];
    pos[3 * a + 2] = z[2];
    __syncthreads();
    f[0] = z[3];
    f[1] = z[4];
    f[2] = z[5];
    f[3] = 0.0;
    f[4] = 0.0;
    f[5] = 0.0;
    if (act) {
      const double x = z[0], yy = z[1], zz = z[2];
      const double r2 = dot3(x, yy, zz, x, yy, zz);
      const double r = sqrt(r2);
      const double w0 = mu / (r2 * r);
      double ax = -(w0 * x), ay = -(w0 * yy), az = -(w0 * zz);
      const double loj = plo[a], wvj = pwv[a];
      for (int b = 0; b < gs; ++b) {
        if (b == a) continue;
        const double dx = pos[3 * b + 0] - x, dy = pos[3 * b + 1] - yy, dz = pos[3 * b + 2] - zz;
        const double d2 = dot3(dx, dy, dz, dx, dy, dz);
        const double lok = plo[b], wvk = pwv[b];
        const double w = close_weight(d2, eps2, pm[b], loj > lok ? loj : lok, wvj < wvk ? wvj : wvk);
        ax = fma(w, dx, ax);
        ay = fma(w, dy, ay);
        az = fma(w, dz, az);
      }
      f[3] = ax;
      f[4] = ay;
      f[5] = az;
    }
    __syncthreads();
  };
  uint32_t st = 0;
  double t = 0.0, H = h;
  int depth = 0;
  while (t < h) {
    if (H > h - t) H = h - t;
    double f0[6];
    rhs(y, f0);
    ++nrhs;
    const double scx = sqrt(dot3(y[0], y[1], y[2], y[0], y[1], y[2]));
    const double scv = sqrt(dot3(y[3], y[4], y[5], y[3], y[4], y[5]));
    double R[KMAX][6], cur[6];
    bool conv = false;
#pragma unroll
    for (int k = 0; k < KMAX; ++k) {
      const int Ns = 2 << k;
      const double hs = H / (double)Ns, h2 = 2.0 * hs;
      double z0[6], z1[6], f[6];
#pragma unroll
      for (int c = 0; c < 6; ++c) {
        z0[c] = y[c];
        z1[c] = fma(hs, f0[c], y[c]);
      }
      for (int m = 1; m < Ns; ++m) {
        rhs(z1, f);
#pragma unroll
        for (int c = 0; c < 6; ++c) {
          const double z2 = fma(h2, f[c], z0[c]);
          z0[c] = z1[c];
          z1[c] = z2;
        }
      }
      rhs(z1, f);
      nrhs += (unsigned int)Ns;
#pragma unroll
      for (int c = 0; c < 6; ++c) cur[c] = 0.5 * (z1[c] + fma(hs, f[c], z0[c]));
#pragma unroll
      for (int j = 1; j <= k; ++j) {
        const double fac = bs_fac(j);
#pragma unroll
        for (int c = 0; c < 6; ++c) {
          const double prev = R[j - 1][c];
          R[j - 1][c] = cur[c];
          cur[c] = cur[c] + (cur[c] - prev) / fac;
        }
      }
#pragma unroll
      for (int c = 0; c < 6; ++c) R[k][c] = cur[c];
      if (k >= 1) {
        double err = 0.0;
        if (act) {
#pragma unroll
          for (int c = 0; c < 6; ++c) {
            const double e = fabs(cur[c] - R[k - 1][c]) / (c < 3 ? scx : scv);
            err = e > err ? e : err;
          }
        }
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) {
          const double o = __shfl_xor(err, off, 64);
          err = o > err ? o : err;
        }
        if (err <= tol) {
          conv = true;
          break;
        }
      }
    }
    if (conv || depth >= BS_MAXDEPTH) {
      if (!conv) st |= ST_BS_DEPTH;
#pragma unroll
      for (int c = 0; c < 6; ++c) y[c] = cur[c];
      t += H;
    } else {
      H *= 0.5;
      ++depth;
    }
  }
  return st;
}

// Drift: workgroups [0, nK) drift the bodies in no group on their Kepler
// orbits (and clear the flags for the next step); the rest walk the group list.
__global__ void __launch_bounds__(64)
    mb_drift_kernel(MBShape sh, int nK, int nG, const double *__restrict__ bp_all,
                    const double *__restrict__ y_all, const double *__restrict__ va_all,
                    double *__restrict__ xb_all, double *__restrict__ vb_all,
                    int *__restrict__ flag_all, const unsigned int *__restrict__ ctl_all,
                    const int *__restrict__ gstart_all, const int *__restrict__ gmem_all,
                    unsigned int *__restrict__ status, unsigned long long *__restrict__ stats,
                    StepConst sc) {
  const int N = sh.N, s = blockIdx.y;
  const double *Y = y_all + (size_t)s * N * 3;
  const double *VA = va_all + (size_t)s * N * 3;
  double *XB = xb_all + (size_t)s * N * 3;
  double *VB = vb_all + (size_t)s * N * 3;
  if ((int)blockIdx.x < nK) {
    int *flag = flag_all + (size_t)s * N;
    const int i = (int)blockIdx.x * 64 + (int)threadIdx.x;
    if (i >= N) return;
    if (flag[i]) {
      flag[i] = 0;
      return;
    }
    double x = Y[3 * i + 0], y = Y[3 * i + 1], z = Y[3 * i + 2];
    double vx = VA[3 * i + 0], vy = VA[3 * i + 1], vz = VA[3 * i + 2];
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
  __shared__ double pos[3 * 64], pm[64], plo[64], pwv[64];
  const BP bp = bp_of(bp_all, s, N);
  const unsigned int *ctl = ctl_all + 4 * s;
  const int *gstart = gstart_all + (size_t)s * (N + 1);
  const int *gmem = gmem_all + (size_t)s * N;
  const int G = (int)ctl[2];
  const int a = threadIdx.x;
  unsigned int nrhs = 0;
  uint32_t st = 0;
  for (int g = (int)blockIdx.x - nK; g < G; g += nG) {
    const int g0 = gstart[g], gs = gstart[g + 1] - g0;
    if (gs > MB_GMAX) continue;
    const int j = a < gs ? gmem[g0 + a] : -1;
    double y[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
    pm[a] = 0.0;
    plo[a] = 0.0;
    pwv[a] = 0.0;
    if (j >= 0) {
      for (int c = 0; c < 3; ++c) {
        y[c] = Y[3 * j + c];
        y[3 + c] = VA[3 * j + c];
      }
      pm[a] = bp.m[j];
      plo[a] = bp.lo[j];
      pwv[a] = bp.winv[j];
    }
    __syncthreads();
    st |= bs_group_wave(y, a, gs, sc.h, sc.mu, sc.eps2, sc.tol, pm, plo, pwv, pos, nrhs);
    if (j >= 0)
      for (int c = 0; c < 3; ++c) {
        XB[3 * j + c] = y[c];
        VB[3 * j + c] = y[3 + c];
      }
    __syncthreads();
  }
  if (st && a == 0) atomicOr(status + s, st);
}

size_t close_lds(const MBShape &sh) {
  return ((size_t)sh.nch * sh.TI * 3 + (size_t)sh.nch * 3 + 3) * sizeof(double);
}

size_t open_lds(const MBShape &sh) {
  return ((size_t)sh.nch * 3 + 3) * sizeof(double) +
         ((size_t)3 * sh.N + 2 * MB_EMAX + 64) * sizeof(int);
}

int launch_close(Ctx *c, bool init, int kicks, int last) {
  const MBShape sh = mb_shape(c->NM);
  const dim3 grid((unsigned int)((sh.N + sh.TI - 1) / sh.TI), (unsigned int)c->S);
  const StepConst sc = c->step_const();
  if (init)
    mb_close_kernel<true><<<grid, MB_BLOCK, close_lds(sh), c->stream>>>(
        sh, c->d_bp, c->d_mx, c->d_mv, c->d_mxb, c->d_mvb, kicks, last, c->d_status, sc,
        c->cfg.central_mass);
  else
    mb_close_kernel<false><<<grid, MB_BLOCK, close_lds(sh), c->stream>>>(
        sh, c->d_bp, c->d_mx, c->d_mv, c->d_mxb, c->d_mvb, kicks, last, c->d_status, sc,
        c->cfg.central_mass);
  return c->hip(hipGetLastError(), "many-body kick launch");
}

int launch_open(Ctx *c) {
  const MBShape sh = mb_shape(c->NM);
  const dim3 grid((unsigned int)((sh.N + sh.TI - 1) / sh.TI), (unsigned int)c->S);
  unsigned long long *stats = (c->cfg.collect_stats & 1) ? c->d_stats : nullptr;
  mb_open_kernel<<<grid, MB_BLOCK, open_lds(sh), c->stream>>>(
      sh, c->d_bp, c->d_mx, c->d_mv, c->d_my, c->d_mflag, c->d_medge, c->d_mctl, c->d_gstart,
      c->d_gmem, c->d_status, stats, c->step_const(), c->cfg.central_mass);
  return c->hip(hipGetLastError(), "many-body encounter launch");
}

int launch_drift(Ctx *c) {
  const MBShape sh = mb_shape(c->NM);
  const int nK = (sh.N + 63) / 64;
  const dim3 grid((unsigned int)(nK + MB_DRIFT_GROUP_BLOCKS), (unsigned int)c->S);
  unsigned long long *stats = (c->cfg.collect_stats & 1) ? c->d_stats : nullptr;
  mb_drift_kernel<<<grid, 64, 0, c->stream>>>(sh, nK, MB_DRIFT_GROUP_BLOCKS, c->d_bp, c->d_my, c->d_mv,
                                              c->d_mxb, c->d_mvb, c->d_mflag, c->d_mctl, c->d_gstart,
                                              c->d_gmem, c->d_status, stats, c->step_const());
  return c->hip(hipGetLastError(), "many-body drift launch");
}

int enqueue_step(Ctx *c, int kicks, int last, KTimer *kt) {
  int rc;
  if (kt) kt->start();
  if ((rc = launch_open(c))) return rc;
  if (kt) kt->stop(c->timing.ms_mb_open, c->timing.n_mb_open);
  if (kt) kt->start();
  if ((rc = launch_drift(c))) return rc;
  if (kt) kt->stop(c->timing.ms_mb_drift, c->timing.n_mb_drift);
  if (kt) kt->start();
  if ((rc = launch_close(c, false, kicks, last))) return rc;
  if (kt) kt->stop(c->timing.ms_mb_close, c->timing.n_mb_close);
  return MINT_OK;
}

int mb_graph(Ctx *c, hipGraphExec_t *out) {
  if (c->mb_graph && c->mb_graph_steps == MB_GRAPH_STEPS) {
    *out = c->mb_graph;
    return MINT_OK;
  }
  int rc = c->hip(hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal), "begin capture");
  if (rc) return rc;
  for (int k = 0; k < MB_GRAPH_STEPS && !rc; ++k) rc = enqueue_step(c, 2, 0, nullptr);
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
  c->mb_graph = ex;
  c->mb_graph_steps = MB_GRAPH_STEPS;
  *out = ex;
  return MINT_OK;
}

int upload_params(Ctx *c) {
  const int S = c->S, N = c->NM;
  std::vector<double> bp((size_t)S * 5 * N);
  for (int s = 0; s < S; ++s) {
    double *b = bp.data() + (size_t)s * 5 * N;
    for (int j = 0; j < N; ++j) {
      const double rc = c->rc[(size_t)s * N + j];
      b[j] = c->m[(size_t)s * N + j];
      b[N + j] = 0.1 * rc;
      b[2 * N + j] = 1.0 / (0.9 * rc);
      b[3 * N + j] = rc * rc;
      b[4 * N + j] = c->rad.empty() ? 0.0 : c->rad[(size_t)s * N + j];
    }
  }
  return c->hip(hipMemcpy(c->d_bp, bp.data(), bp.size() * sizeof(double), hipMemcpyHostToDevice),
                "H2D body parameters");
}

}  // namespace

int mb_alloc(Ctx *c) {
  const int S = c->S, N = c->NM;
  const size_t b3 = (size_t)S * N * 3 * sizeof(double);
  int rc;
  if ((rc = c->dalloc((void **)&c->d_bp, (size_t)S * 5 * N * sizeof(double), "body parameters"))) return rc;
  if ((rc = c->dalloc((void **)&c->d_mxb, b3, "post-drift positions"))) return rc;
  if ((rc = c->dalloc((void **)&c->d_mvb, b3, "post-drift velocities"))) return rc;
  if ((rc = c->dalloc((void **)&c->d_my, b3, "drift-start positions"))) return rc;
  if ((rc = c->dalloc((void **)&c->d_mflag, (size_t)S * N * sizeof(int), "encounter flags"))) return rc;
  if ((rc = c->dalloc((void **)&c->d_medge, (size_t)S * MB_EMAX * 2 * sizeof(int), "encounter pairs"))) return rc;
  if ((rc = c->dalloc((void **)&c->d_mctl, (size_t)S * 4 * sizeof(unsigned int), "group counters"))) return rc;
  if ((rc = c->dalloc((void **)&c->d_gstart, (size_t)S * (N + 1) * sizeof(int), "group starts"))) return rc;
  if ((rc = c->dalloc((void **)&c->d_gmem, (size_t)S * N * sizeof(int), "group members"))) return rc;
  return MINT_OK;
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
  if (ktp) ktp->start();
  if ((rc = launch_close(c, true, 1, 0))) return rc;
  if (ktp) ktp->stop(c->timing.ms_mb_close, c->timing.n_mb_close);
  int64_t rest = nsteps - 1;
  if (c->cfg.use_graph && !ktp && rest >= MB_GRAPH_STEPS) {
    hipGraphExec_t ex = nullptr;
    if ((rc = mb_graph(c, &ex))) return rc;
    for (; rest >= MB_GRAPH_STEPS; rest -= MB_GRAPH_STEPS)
      if ((rc = c->hip(hipGraphLaunch(ex, c->stream), "graph launch"))) return rc;
  }
  for (; rest > 0; --rest)
    if ((rc = enqueue_step(c, 2, 0, ktp))) return rc;
  if ((rc = enqueue_step(c, 1, 1, ktp))) return rc;
  if ((rc = c->hip(hipStreamSynchronize(c->stream), "many-body step"))) return rc;
  std::vector<unsigned int> st((size_t)c->S);
  if ((rc = c->hip(hipMemcpy(st.data(), c->d_status, st.size() * sizeof(unsigned int), hipMemcpyDeviceToHost),
                   "D2H status")))
    return rc;
  for (int s = 0; s < c->S; ++s)
    if (st[s] & ST_LIMIT)
      return c->fail(MINT_ELIMIT, "many-body engine: system " + std::to_string(s) +
                                      " had more than 2048 encounter pairs or an encounter group of more "
                                      "than 64 bodies in one step; its state is not valid");
  return MINT_OK;
}

}  // namespace mint
