// mint_launch.h — host-side launch table.  One translation unit per massive
// count NM (mint_nm.hip built with -DMINT_NM=1..8) instantiates the kernels
// and exposes plain launch functions, so the host code never sees templates.
#pragma once
#include <hip/hip_runtime.h>
#include "mint_kernels.h"

namespace mint {

struct Launchers {
  hipError_t (*ephemeris)(hipStream_t st, int s_begin, int s_end, int C, const double *prm,
                          const double *ix, const double *iv, double *mx, double *mv, double *xs0,
                          double *rec, double *nx, double *nv,
                          unsigned int *status, unsigned long long *stats, StepConst sc,
                          double M);
  hipError_t (*tp_step)(hipStream_t st, int guess, int block, long long i0, long long i1,
                        const int *tp_sys, TPArrays tp, const double *prm, const double *xs0,
                        const double *rec, StepConst sc, const float *nnw,
                        unsigned int *status, unsigned long long *stats);
  hipError_t (*tp_main)(hipStream_t st, int guess, int P, int block, int nblocks,
                        const Tile *tiles, int ntiles, int tiles_per_xcd, TPArrays tp,
                        const double *prm, const double *xs0, const double *rec, int C,
                        StepConst sc, const float *nnw, const unsigned char *pre,
                        QEntry *q, unsigned int *qcount, unsigned int *need, int *slist,
                        unsigned char *carry, unsigned int *status,
                        unsigned long long *stats, long long *kstamp);
  hipError_t (*tp_encounter)(hipStream_t st, int guess, int grid, int block, const QEntry *q,
                             const unsigned int *qcount, const unsigned int *qbegin, int dense,
                             int grouped, TPArrays tp,
                             const double *prm, const double *xs0, const double *rec,
                             const double *pc, int C, StepConst sc, const float *nnw,
                             unsigned int *status, unsigned long long *stats);
  hipError_t (*tp_prebucket)(hipStream_t st, long long n, const int *tp_sys, TPArrays tp,
                             const double *prm, const double *wx, const double *wv, int C,
                             StepConst sc, const float *scw, float thresh, float cut,
                             unsigned char *carry, unsigned char *pre, QEntry *q,
                             unsigned int *qcount, unsigned int *need, int *slist,
                             unsigned long long *stats);
  hipError_t (*pcache)(hipStream_t st, const int *list, const unsigned int *count, int C,
                       const double *rec, double *pc, StepConst sc);
};

const Launchers *launchers_nm1();
const Launchers *launchers_nm2();
const Launchers *launchers_nm3();
const Launchers *launchers_nm4();
const Launchers *launchers_nm5();
const Launchers *launchers_nm6();
const Launchers *launchers_nm7();
const Launchers *launchers_nm8();

inline const Launchers *launchers_for(int nm) {
  switch (nm) {
    case 1: return launchers_nm1();
    case 2: return launchers_nm2();
    case 3: return launchers_nm3();
    case 4: return launchers_nm4();
    case 5: return launchers_nm5();
    case 6: return launchers_nm6();
    case 7: return launchers_nm7();
    case 8: return launchers_nm8();
    default: return nullptr;
  }
}

constexpr int MAX_NM = 8;

}  // namespace mint
