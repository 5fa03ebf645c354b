// mint_noalloc.hip — mint_step_noalloc: the flat per-step schedule with its
// per-step scratch in a fixed device array of the library, so that the call
// allocates nothing (for callers that cannot: code under graph capture,
// serving loops, harnesses that forbid allocation).
//
// Scratch layout for a batch of S systems of NM massive bodies: the
// step-start massive positions [S][NM][3], then one ephemeris record per
// system [S][rec_size(NM)].  Both are written by the ephemeris launch of each
// step before any kernel reads them, so nothing needs clearing between calls.
#include <hip/hip_runtime.h>

#include <mutex>
#include <string>

#include "mint_api_internal.h"

namespace mint {
namespace {
constexpr size_t SCRATCH_DOUBLES = (size_t)1 << 20;  // 8 MiB on every GPU the library runs on
}  // namespace
}  // namespace mint

// The array's name is a host-side stand-in: host code uses only the device
// address hipGetSymbolAddress returns for it on each device.
static __device__ __align__(256) double scratch[mint::SCRATCH_DOUBLES];

namespace mint {
namespace {

constexpr int MAX_DEVICES = 64;
std::mutex g_mu;
double *g_scratch[MAX_DEVICES] = {};
std::string g_scratch_err[MAX_DEVICES];

}  // namespace

// The scratch array's address on `device` (current on the calling thread),
// looked up once per device.  Called by mint_create.
void noalloc_init(int device) {
  if (device < 0 || device >= MAX_DEVICES) return;
  std::lock_guard<std::mutex> lock(g_mu);
  if (g_scratch[device]) return;
  void *p = nullptr;
  const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
  if (e != hipSuccess || !p) {
    g_scratch_err[device] = std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e);
    return;
  }
  g_scratch[device] = (double *)p;
}

}  // namespace mint

using namespace mint;

extern "C" int mint_step_noalloc(void *ctx, int64_t nsteps) {
  Ctx *c = (Ctx *)ctx;
  if (!c) return MINT_EINVAL;
  if (!c->uploaded) return c->fail(MINT_ESTATE, "mint_step_noalloc: nothing uploaded");
  if (c->many) return c->fail(MINT_EINVAL, "mint_step_noalloc: batches of more than 8 massive bodies per system run on "
                                           "the many-body engine, which this entry point does not serve");
  if (nsteps < 0) return c->fail(MINT_EINVAL, "mint_step_noalloc: negative step count");
  if (c->cfg.kepler_guess == MINT_GUESS_LEARNED && c->w_kep.empty())
    return c->fail(MINT_ESTATE, "mint_step_noalloc: learned guess selected but the Kepler MLP has no weights");
  const int dev = c->cfg.device;
  double *base = nullptr;
  std::string err;
  {
    std::lock_guard<std::mutex> lock(g_mu);
    if (dev >= 0 && dev < MAX_DEVICES) {
      base = g_scratch[dev];
      err = g_scratch_err[dev];
    }
  }
  if (!base) return c->fail(MINT_EHIP, "mint_step_noalloc: no scratch on this device (" + err + ")");
  const size_t nx = (size_t)c->S * c->NM * 3;
  const size_t need = nx + (size_t)c->S * rec_size(c->NM);
  if (need > SCRATCH_DOUBLES)
    return c->fail(MINT_ELIMIT, "mint_step_noalloc: the batch needs " + std::to_string(need) +
                                    " doubles of scratch, more than the " + std::to_string(SCRATCH_DOUBLES) +
                                    " (8 MiB) the library holds per GPU");
  if (nsteps == 0) return MINT_OK;
  double *xs0 = base, *rec = base + nx;
  hipSetDevice(dev);
  const StepConst sc = c->step_const();
  const TPArrays tp = c->tp_arrays();
  unsigned long long *stats = (c->cfg.collect_stats & 1) ? c->d_stats : nullptr;
  int rc = c->hip(hipEventRecord(c->ev0, c->stream), "event");
  for (int64_t k = 0; k < nsteps && !rc; ++k) {
    rc = c->hip(c->L->ephemeris(c->stream, 0, c->S, 1, c->d_prm, c->d_mx, c->d_mv, c->d_mx, c->d_mv, xs0, rec, nullptr, nullptr,
                                c->d_status, stats, sc, c->cfg.central_mass),
                "ephemeris launch");
    if (!rc)
      rc = c->hip(c->L->tp_step(c->stream, c->cfg.kepler_guess, 256, 0, c->Ntp, c->d_tp_sys, tp, c->d_prm, xs0, rec,
                                sc, c->d_nnw_kep, c->d_status, stats),
                  "step launch");
  }
  if (rc) {
    hipStreamSynchronize(c->stream);
    return rc;
  }
  if ((rc = c->hip(hipEventRecord(c->ev1, c->stream), "event"))) return rc;
  if ((rc = c->hip(hipEventSynchronize(c->ev1), "step sync"))) return rc;
  float ms = 0.f;
  hipEventElapsedTime(&ms, c->ev0, c->ev1);
  c->timing.ms_total += ms;
  return c->hip(hipGetLastError(), "step");
}
