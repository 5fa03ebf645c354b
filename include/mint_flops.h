/* mint_flops.h — the algorithmic work and bytes the step kernels are priced
 * with (the roofline figures of the benchmark line).  Counting rule: add,
 * sub, mul, div and sqrt 1 FLOP each, fma 2, compare 0.  No counterpart in
 * the reference, which reports no throughput. */
#ifndef MINT_FLOPS_H
#define MINT_FLOPS_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* FLOP per test-particle step with nm massive bodies and nbar Newton
 * iterations per Kepler drift: 2 nm 28 (two changeover half-kicks, 28 per
 * pair) + 12 (two velocity updates) + 12 (two jumps) + 9 nm (exact encounter
 * check) + 60 + 40 nbar (universal-variable drift: set-up and finish, then
 * each Newton iteration).  Bulirsch-Stoer work is counted apart, as
 * right-hand-side evaluations.  nbar is the plain form's mean on the
 * workload measured, the same for every form. */
double mint_flops_per_tp_step(int32_t nm, double nbar);

/* FLOP of one massive kick-acceleration pass over n bodies: 28 per ordered
 * pair, n (n - 1) pairs. */
double mint_flops_per_kick_pass(int64_t n);

/* HBM bytes per test-particle step of the per-step forms: the SoA6 state
 * read and written once. */
#define MINT_BYTES_PER_TP_STEP_PLAIN 96

#ifdef __cplusplus
}
#endif
#endif /* MINT_FLOPS_H */
