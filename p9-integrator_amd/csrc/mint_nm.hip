// mint_nm.hip — kernel instantiations for one massive-body count, MINT_NM.
#include "mint_launch.h"

#ifndef MINT_NM
#error "build with -DMINT_NM=<1..8>"
#endif

namespace mint {
namespace {

constexpr int NMV = MINT_NM;

hipError_t ephemeris_l(hipStream_t st, int s_begin, int s_end, int C, const double *prm,
                       const double *ix, const double *iv, double *mx, double *mv, double *xs0,
                       double *rec, double *nx, double *nv,
                       unsigned int *status, unsigned long long *stats, StepConst sc, double M) {
  const int n = s_end - s_begin;
  if (n <= 0) return hipSuccess;
  constexpr int spw = 64 / EPH_L<NMV>;  // systems per one-wave block
  ephemeris_kernel<NMV><<<(n + spw - 1) / spw, 64, 0, st>>>(s_begin, s_end, C, prm, ix, iv, mx, mv, xs0, rec,
                                                            nx, nv, status, stats, sc, M);
  return hipGetLastError();
}

hipError_t tp_step_l(hipStream_t st, int guess, int block, long long i0, long long i1,
                     const int *tp_sys, TPArrays tp, const double *prm, const double *xs0,
                     const double *rec, StepConst sc, const float *nnw, unsigned int *status,
                     unsigned long long *stats) {
  const long long n = i1 - i0;
  if (n <= 0) return hipSuccess;
  if (block <= 0 || block > 256 || (block & 63)) return hipErrorInvalidValue;
  const unsigned int grid = (unsigned int)((n + block - 1) / block);
  switch (guess) {
    case 0:
      tp_step_kernel<NMV, 0><<<grid, block, 0, st>>>(i0, i1, tp_sys, tp, prm, xs0, rec, sc, nnw,
                                                     status, stats);
      break;
    case 1:
      tp_step_kernel<NMV, 1><<<grid, block, 0, st>>>(i0, i1, tp_sys, tp, prm, xs0, rec, sc, nnw,
                                                     status, stats);
      break;
    case 2:
      tp_step_kernel<NMV, 2><<<grid, block, 0, st>>>(i0, i1, tp_sys, tp, prm, xs0, rec, sc, nnw,
                                                     status, stats);
      break;
    default:
      return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

template <int P, int G>
hipError_t main_pg(hipStream_t st, int block, int nblocks, const Tile *tiles, int ntiles,
                   int tiles_per_xcd, TPArrays tp, const double *prm, const double *xs0,
                   const double *rec, int C, StepConst sc, const float *nnw,
                   const unsigned char *pre, QEntry *q, unsigned int *qcount,
                   unsigned int *need, int *slist, unsigned char *carry, unsigned int *status,
                   unsigned long long *stats, long long *kstamp) {
  tp_main_kernel<NMV, P, G><<<nblocks, block, 0, st>>>(tiles, ntiles, tiles_per_xcd, tp, prm, xs0,
                                                       rec, C, sc, nnw, pre, q, qcount, need,
                                                       slist, carry, status, stats, kstamp);
  return hipGetLastError();
}

hipError_t tp_main_l(hipStream_t st, int guess, int P, int block, int nblocks, const Tile *tiles,
                     int ntiles, int tiles_per_xcd, TPArrays tp, const double *prm,
                     const double *xs0, const double *rec, int C, StepConst sc, const float *nnw,
                     const unsigned char *pre, QEntry *q, unsigned int *qcount,
                     unsigned int *need, int *slist, unsigned char *carry, unsigned int *status,
                     unsigned long long *stats, long long *kstamp) {
  if (block <= 0 || block > 512 || (block & 63)) return hipErrorInvalidValue;
#define MINT_MAIN(PP, GG)                                                                       \
  if (P == PP && guess == GG)                                                                   \
    return main_pg<PP, GG>(st, block, nblocks, tiles, ntiles, tiles_per_xcd, tp, prm, xs0, rec, \
                           C, sc, nnw, pre, q, qcount, need, slist, carry, status, stats, kstamp);
  MINT_MAIN(1, 0) MINT_MAIN(1, 1) MINT_MAIN(1, 2)
  MINT_MAIN(2, 0) MINT_MAIN(2, 1) MINT_MAIN(2, 2)
  MINT_MAIN(4, 0) MINT_MAIN(4, 1) MINT_MAIN(4, 2)
#undef MINT_MAIN
  return hipErrorInvalidValue;
}

// lanes per queue entry in the grouped encounter kernel: one per perturber,
// rounded up to a power of two, times four row groups of the BS extrapolation
constexpr int ENC_G = 4 * (NMV <= 1 ? 1 : (NMV <= 2 ? 2 : (NMV <= 4 ? 4 : 8)));

template <int G>
hipError_t encounter_g(hipStream_t st, int guess, int grid, const QEntry *q, const unsigned int *qcount,
                       const unsigned int *qbegin, int dense, TPArrays tp, const double *prm,
                       const double *xs0, const double *rec, const double *pc, int C, StepConst sc,
                       const float *nnw, unsigned int *status, unsigned long long *stats) {
  switch (guess) {
    case 0:
      tp_encounter_kernel<NMV, 0, G><<<grid, 64, 0, st>>>(q, qcount, qbegin, dense, tp, prm, xs0, rec, pc, C,
                                                          sc, nnw, status, stats);
      break;
    case 1:
      tp_encounter_kernel<NMV, 1, G><<<grid, 64, 0, st>>>(q, qcount, qbegin, dense, tp, prm, xs0, rec, pc, C,
                                                          sc, nnw, status, stats);
      break;
    case 2:
      tp_encounter_kernel<NMV, 2, G><<<grid, 64, 0, st>>>(q, qcount, qbegin, dense, tp, prm, xs0, rec, pc, C,
                                                          sc, nnw, status, stats);
      break;
    default:
      return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

hipError_t tp_encounter_l(hipStream_t st, int guess, int grid, int block, const QEntry *q,
                          const unsigned int *qcount, const unsigned int *qbegin, int dense, int grouped,
                          TPArrays tp, const double *prm, const double *xs0, const double *rec,
                          const double *pc, int C, StepConst sc, const float *nnw,
                          unsigned int *status, unsigned long long *stats) {
  if (block != 64) return hipErrorInvalidValue;  // one wave per block: the LDS tableau is per wave
  if (grouped)
    return encounter_g<ENC_G>(st, guess, grid, q, qcount, qbegin, dense, tp, prm, xs0, rec, pc, C, sc, nnw,
                              status, stats);
  return encounter_g<1>(st, guess, grid, q, qcount, qbegin, dense, tp, prm, xs0, rec, pc, C, sc, nnw, status,
                        stats);
}

hipError_t tp_prebucket_l(hipStream_t st, long long n, const int *tp_sys, TPArrays tp,
                          const double *prm, const double *wx, const double *wv, int C,
                          StepConst sc, const float *scw, float thresh, float cut, unsigned char *carry,
                          unsigned char *pre, QEntry *q, unsigned int *qcount, unsigned int *need,
                          int *slist, unsigned long long *stats) {
  if (n <= 0) return hipSuccess;
  tp_prebucket_kernel<NMV><<<(unsigned)((n + PB_BLOCK - 1) / PB_BLOCK), PB_BLOCK, 0, st>>>(
      n, tp_sys, tp, prm, wx, wv, C, sc, scw, thresh, cut, carry, pre, q, qcount, need, slist, stats);
  return hipGetLastError();
}

hipError_t pcache_l(hipStream_t st, const int *list, const unsigned int *count, int C,
                    const double *rec, double *pc, StepConst sc) {
  // fixed grid, grid-stride over the listed systems' work items
  pcache_kernel<NMV><<<1024, 256, 0, st>>>(list, count, C, rec, pc, sc);
  return hipGetLastError();
}

}  // namespace

#define MINT_CAT2(a, b) a##b
#define MINT_CAT(a, b) MINT_CAT2(a, b)
const Launchers *MINT_CAT(launchers_nm, MINT_NM)() {
  static const Launchers L = {ephemeris_l, tp_step_l, tp_main_l, tp_encounter_l, tp_prebucket_l,
                              pcache_l};
  return &L;
}

}  // namespace mint
