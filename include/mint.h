/*
 * mint.h — C ABI of the MI355X (gfx950) Mercurius step library, libmint.so.
 *
 * Units: G = 1, masses in solar masses, lengths in AU, time in years.
 * All state is FP64.  A "system" is one independent sweep member: a central
 * mass M (the Sun) at the heliocentric origin, n_massive massive bodies and a
 * variable number of massless test particles (TPs).
 *
 * The reference has no plugin or FFI layer; its boundary is the kernel
 *   mercurius_solver(ω[], M[], e[], a[], i[], Ω[], m[], pos_out[], N_m, batch, dt)
 *   (src/sim.cuh:615-625), launched by cudaLaunchKernelEx from main
 *   (src/mercurius.cu:45-65,184-198),
 * with host types Body/Sweep/Sim (src/simutils.cuh:19-87).  Each entry point
 * below cites the part of that interface it replaces.
 *
 * Conventions:
 *   - every call returns 0 on success and a negative MINT_E* code on failure;
 *     mint_last_error() returns the message of the last failure on a context;
 *   - the context owns all device memory; the caller owns host buffers, which
 *     are copied in and out and never retained after a call returns;
 *   - a context is used by one host thread at a time and issues all work on
 *     its own HIP stream; calls return after that work has finished.
 */
#ifndef MINT_H
#define MINT_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define MINT_ABI_VERSION 2

/* error codes */
#define MINT_OK 0
#define MINT_EINVAL (-1)   /* bad argument or shape */
#define MINT_EHIP (-2)     /* HIP runtime failure */
#define MINT_ENOMEM (-3)   /* device or host allocation failed */
#define MINT_ESTATE (-4)   /* call out of order (e.g. step before upload) */
#define MINT_EIO (-5)      /* file or parse error (config, weights, cache) */
#define MINT_ELIMIT (-6)   /* a many-body capacity limit was exceeded (see MINT_ST_LIMIT) */

/* Drift forms (how the TP population is scheduled on the GPU). */
#define MINT_FORM_PLAIN 0       /* one launch per step, a few systems per launch, in-place BS */
#define MINT_FORM_FLAT 1        /* one launch per step over all systems, in-place BS */
#define MINT_FORM_BATCHED 2     /* multi-step register-resident kernel + BS encounter queue, hipGraph */

/* Kepler-solver starting guesses for the TP drift. */
#define MINT_GUESS_CRUDE 0      /* s0 = dt / r0 */
#define MINT_GUESS_SERIES 1     /* third-order series in dt */
#define MINT_GUESS_LEARNED 2    /* FP32 MLP correction on top of the series guess */

/* Per-system status bits (mint_status). */
#define MINT_ST_NAN 1u              /* a non-finite value appeared in TP or massive state */
#define MINT_ST_NEWTON 2u           /* a Kepler solve hit the iteration cap */
#define MINT_ST_BS_DEPTH 4u         /* a Bulirsch-Stoer step hit its subdivision cap */
#define MINT_ST_MASSIVE_ENC 8u      /* massive bodies came within their changeover radius */
#define MINT_ST_OVERLAP 16u         /* two massive bodies closer than the sum of their physical
                                       radii (recorded only: nothing merges, as in the reference) */
#define MINT_ST_LIMIT 32u           /* many-body engine: more encounter pairs or a larger
                                       encounter group than it holds; the state is not valid */

/*
 * Integration parameters.  Replaces the compile-time constants of
 * src/constants.cuh:4-13 and the hard-coded dt of src/mercurius.cu:42-43.
 * mint_config_default() fills the defaults listed on each field.
 */
typedef struct mint_config {
  double dt;             /* step length [yr]; default 0.8219 (src/mercurius.cu:43) */
  double central_mass;   /* M; default 1.0 (src/sim.cuh:401) */
  double n_hill;         /* changeover radius r_c = n_hill * Hill radius; default 3 */
  double softening2;     /* eps^2 of the pair force; default 1e-8 (src/constants.cuh:8) */
  double bs_tol;         /* Bulirsch-Stoer tolerance; default 1e-12 (src/constants.cuh:6) */
  int32_t form;          /* MINT_FORM_*; default MINT_FORM_BATCHED */
  int32_t kepler_guess;  /* MINT_GUESS_*; default MINT_GUESS_SERIES */
  int32_t systems_per_launch; /* plain form only; default 4 (src/constants.cuh:12) */
  int32_t block_size;    /* workgroup size; 0 = autotuned (batched) or 256 (per-step forms) */
  int32_t tps_per_lane;  /* batched: TPs per lane (1, 2 or 4); 0 = autotuned */
  int32_t chunk_steps;   /* batched: steps per register-resident launch; 0 = autotuned */
  int32_t use_scorer;    /* batched: pre-bucket likely encounters with the learned scorer */
  int32_t use_graph;     /* batched: replay chunk launches from a captured hipGraph; default 1 */
  int32_t device;        /* HIP device ordinal; default 0 (the -d flag, src/simutils.cuh:170-173) */
  int32_t collect_stats; /* bit 0: count work into mint_stats; bit 1: per-kernel HIP-event
                            timing into mint_timing (launches without the graph, one stream);
                            bit 2: batched form, time each main-kernel launch of the normal
                            schedule (graph and streams unchanged) into ms_tp_main /
                            n_tp_main, from the device wall clock stamped by every
                            wavefront at entry and exit (last exit - first entry);
                            default 0 */
  int32_t many_body;     /* 1: batches without TPs use the many-body engine at any n_massive
                            (same results; default 0 = only above 8 massive bodies) */
} mint_config;

/*
 * Host-side description of a batch of systems (the CSR layout that replaces
 * the reference's [sweep][body] element arrays, src/sim.cuh:652 and
 * src/mercurius.cu:73-164).  All pointers are host pointers.
 *   tp_offset[S+1]   prefix sums of TP counts; TPs of system s are
 *                    tp_offset[s] .. tp_offset[s+1]-1 in the SoA arrays
 *   m[S*Nm]          massive-body masses (> 1e-12, src/mercurius.cu:9)
 *                    Nm = 1..8 with test particles; any Nm >= 1 without them
 *                    (systems of more than 8 massive bodies run on the
 *                    many-body engine: O(Nm^2) kicks, encounter groups)
 *   mx, mv[S*Nm*3]   heliocentric positions / velocities of massive bodies
 *   tx..tvz[Ntp]     heliocentric TP positions / velocities, SoA
 *   rc[S*Nm]         changeover radii; NULL on upload = computed from n_hill
 */
typedef struct mint_state {
  int32_t n_systems;
  int32_t n_massive;
  const int64_t *tp_offset;
  double *m;
  double *mx;
  double *mv;
  double *tx, *ty, *tz, *tvx, *tvy, *tvz;
  double *rc;
} mint_state;

/* Work counters (collect_stats = 1). */
typedef struct mint_stats {
  int64_t tp_steps;        /* TP steps taken */
  int64_t kepler_solves;   /* TP Kepler drifts */
  int64_t newton_iters;    /* Newton iterations over those drifts */
  int64_t bs_steps;        /* TP steps drifted by Bulirsch-Stoer */
  int64_t bs_rhs_evals;    /* right-hand-side evaluations inside BS */
  int64_t deferred;        /* batched: TP-steps handed to the encounter queue */
  int64_t prebucketed;     /* batched: TPs routed to the queue by the scorer */
  int64_t idle_lane_steps; /* batched: lane-steps left idle in the main kernel */
  int64_t enc_tp_steps;    /* batched: TP steps taken by the encounter-queue kernel */
  int64_t massive_groups;  /* massive encounter groups drifted by Bulirsch-Stoer (group-steps) */
  int64_t massive_group_bodies; /* bodies in those groups (body-steps) */
  int64_t overlaps;        /* massive pairs closer than the sum of their radii (pair-steps) */
} mint_stats;

/* Per-kernel timings from HIP events on the context stream (ms, launches). */
typedef struct mint_timing {
  double ms_ephemeris;  int64_t n_ephemeris;
  double ms_tp_main;    int64_t n_tp_main;
  double ms_encounter;  int64_t n_encounter;
  double ms_total;
  /* many-body engine: drift (with the momentum and jump before it), jump +
     kicks (with the next step's pair check); ms_mb_open / n_mb_open stay 0:
     the pair check has no launch of its own */
  double ms_mb_open;    int64_t n_mb_open;
  double ms_mb_drift;   int64_t n_mb_drift;
  double ms_mb_close;   int64_t n_mb_close;
} mint_timing;

int32_t mint_abi_version(void);
void mint_config_default(mint_config *cfg);

/* Create a context on cfg->device (replaces cudaSetDevice + the cluster launch
 * set-up, src/mercurius.cu:36,45-70). */
int mint_create(const mint_config *cfg, void **ctx);
void mint_destroy(void *ctx);
const char *mint_last_error(void *ctx);

/* Copy a batch to the device, converting velocities to democratic-heliocentric
 * form (replaces the H2D copies src/mercurius.cu:73-164 and the setup
 * conversion src/sim.cuh:665-677).  If st->rc is NULL the changeover radii are
 * computed from the uploaded state (mint_hill_radii) and stored. */
int mint_upload(void *ctx, const mint_state *st);

/* Advance every system by nsteps Mercurius steps of cfg->dt (replaces the
 * per-batch launch loop src/mercurius.cu:179-198 and the step loop body
 * src/sim.cuh:679-742).  Returns after the device work has finished. */
int mint_step(void *ctx, int64_t nsteps);

/* Build what a later mint_step(ctx, nsteps) replays (the batched form's
 * captured launch graph for an nsteps call at the current launch shape and
 * constants) without running anything, so that the host-side capture and
 * instantiation are not part of that call.  No-op for the other forms.  Host
 * work only (no counterpart in the reference, which launches its kernel
 * directly, src/mercurius.cu:184-196). */
int mint_prepare(void *ctx, int64_t nsteps);

/* Allocation-free variant of mint_step (no counterpart in the reference): for
 * callers that cannot allocate device memory (code under HIP graph capture,
 * tight serving loops, harnesses that forbid hipMalloc).  Advances the
 * resident batch by nsteps with the flat per-step schedule (per step one
 * ephemeris launch and one TP step launch over all systems: the kernels and
 * results of mint_step with form MINT_FORM_FLAT) and keeps the per-step
 * scratch (step-start massive positions and one ephemeris record per system)
 * in a fixed 8 MiB device array of the library, held on every GPU the
 * library runs on until the process ends, instead of context buffers.  The
 * batch must fit it: S * (3 n_massive + 2 ceil((6 + 9 n_massive) / 2))
 * doubles at most 2^20 (4096 systems of 5 massive bodies need 274,432),
 * else MINT_ELIMIT; systems of more than 8 massive bodies are refused
 * (MINT_EINVAL).  Calls no hipMalloc or hipFree.  Every call on a GPU shares
 * that GPU's array: calls that overlap on one GPU (from several contexts or
 * threads) race. */
int mint_step_noalloc(void *ctx, int64_t nsteps);

/* Copy the state back, converting velocities to heliocentric form (replaces
 * the teardown src/sim.cuh:744-765 and the D2H copies src/mercurius.cu:203-243).
 * Only the non-NULL arrays of st are written; shapes must match the upload. */
int mint_download(void *ctx, mint_state *st);

/* Binary checkpoint of the device state, bit for bit: the internal
 * (democratic-heliocentric) state, masses, changeover radii, the CSR layout,
 * the step parameters and `step_index`.  Resuming from it continues exactly
 * as if the run had not stopped.  The reference persists osculating elements
 * every 20 launches (src/mercurius.cu:201-249) but never reads them back, and
 * its element round trip is lossy (src/sim.cuh:255-256). */
int mint_save_checkpoint(void *ctx, const char *path, int64_t step_index);
/* Replaces the context's state (and dt, central mass, softening, BS
 * tolerance) by the checkpoint's; *step_index receives the saved index. */
int mint_load_checkpoint(void *ctx, const char *path, int64_t *step_index);

/* Physical radii of the massive bodies, radius[S*Nm] (0 = none), after
 * mint_upload.  They are recorded and checked, never used to merge: pairs
 * closer than the sum of their radii at the encounter check are counted
 * (mint_stats.overlaps) and flagged (MINT_ST_OVERLAP).  The reference keeps
 * no physical radii (its Body, src/simutils.cuh:19-29, has none). */
int mint_set_radii(void *ctx, const double *radius);

/* Total energy of each system's massive bodies (democratic-heliocentric
 * Hamiltonian; the reference has no energy diagnostic). E[S].  The pair
 * potential is the one of the softened kick force m_j m_k d/((r^2+eps^2) r):
 * -m_j m_k atan(eps/r)/eps (Newtonian when softening2 = 0). */
int mint_energy(void *ctx, double *E);

/* OR of status bits seen per system since upload. status[S]. */
int mint_status(void *ctx, uint32_t *status);

int mint_get_stats(void *ctx, mint_stats *out);
int mint_get_timing(void *ctx, mint_timing *out);
int mint_reset_counters(void *ctx);

/* Change kepler_guess / use_scorer / form between calls without re-uploading. */
int mint_set_mode(void *ctx, int32_t form, int32_t kepler_guess, int32_t use_scorer);

/* Diagnostic (no counterpart in the reference): evaluate the device's
 * square root of a squared distance (op 0: out = sqrt(a)), its normal-range
 * quotient (op 1: out = a / b) or the square root the test-particle kick uses
 * for positive finite squared distances (op 2: out = sqrt(a), without the
 * zero and infinity cases) on n inputs on device `device`, so callers can
 * check them against the host's IEEE results.  b may be NULL for ops 0, 2. */
int mint_selftest_arith(int32_t device, int32_t op, int64_t n, const double *a, const double *b,
                        double *out);

/* Changeover radii r_c = n_hill * a * cbrt(m / (3M)) from heliocentric states
 * (replaces fetch_r_crit, src/sim.cuh:133-164, evaluated once instead of per
 * pair).  a is the osculating semi-major axis; unbound orbits use r. */
int mint_hill_radii(int64_t n, double central_mass, double n_hill, const double *m,
                    const double *x, const double *v, double *rc);

/* Osculating elements (i, Ω, ω, M, e, a) -> heliocentric Cartesian for
 * n bodies, mu = G M (replaces cartesian_from_elements, src/sim.cuh:188-239,
 * at the configuration edge).  elements are SoA arrays of length n. */
int mint_elements_to_cartesian(int64_t n, double mu, const double *inc, const double *node,
                               const double *argp, const double *mean_anom, const double *ecc,
                               const double *sma, double *x, double *v);
/* Inverse (replaces elements_from_cartesian, src/sim.cuh:241-324, for output). */
int mint_cartesian_to_elements(int64_t n, double mu, const double *x, const double *v,
                               double *inc, double *node, double *argp, double *mean_anom,
                               double *ecc, double *sma);

/* Autotuner (batched form): time the candidate launch shapes on the uploaded
 * batch (each once untimed, then three timed 256-step probes), keep the one
 * with the fastest median, then the same for block and TPs per lane on calls
 * of one launch window (probes of up to 32 steps, ranked by the slowest of
 * three; only where neither is set in the configuration), and store both in
 * cache_path (NULL = no file).  A shape
 * parameter set in the configuration (block_size, tps_per_lane, chunk_steps
 * > 0) is held at that value.  The state on the device is left unchanged.
 * best[3] = {block, tps/lane, chunk} of calls of more than one window. */
int mint_autotune(void *ctx, const char *cache_path, int32_t *best);

/* The launch shape a batched call of nsteps steps uses: shape[3] = {block,
 * tps/lane, chunk}.  No counterpart in the reference (its launch shape is
 * fixed, src/mercurius.cu:45-65); for reporting and benchmarks. */
int mint_launch_shape(void *ctx, int64_t nsteps, int32_t *shape);

/* Learned parts (FP32).  Weights live in the context.
 * Kepler warm-start MLP: 4 -> 32 -> 32 -> 1, tanh.  Scorer: 6 -> 16 -> 1, tanh/logistic.
 * mint_trace_*: run nsteps while logging samples into a device buffer of
 * capacity `cap`, then copy them out (features, then target or label).
 * Kepler: (features, c) of sampled TP and massive-body drifts.  Encounter:
 * per launch window, every (TP, body) pair with label 1 (the exact check fired
 * in the window) and a hashed one-in-64 of the pairs with label 0. */
int mint_trace_kepler(void *ctx, int64_t nsteps, int64_t cap, float *feat /*cap*4*/,
                      float *target /*cap*/, int64_t *n_out);
int mint_trace_encounter(void *ctx, int64_t nsteps, int64_t cap, float *feat /*cap*6*/,
                         float *label /*cap*/, int64_t *n_out);
/* Train on the GPU with the library's FP32 forward/backward/Adam kernels.
 * which = 0 Kepler MLP, 1 scorer.  loss_out[epochs] receives mean losses. */
int mint_train(void *ctx, int32_t which, int64_t n, const float *feat, const float *target,
               int32_t epochs, int32_t batch, float lr, uint64_t seed, float *loss_out);
/* Forward / backward of one network on n samples (device kernels; for tests).
 * grad[nparams] receives the gradient of 0.5*mean((y - target)^2) (which = 0)
 * or of mean BCE-with-logits (which = 1). */
int mint_nn_forward(void *ctx, int32_t which, int64_t n, const float *feat, float *out);
int mint_nn_loss_grad(void *ctx, int32_t which, int64_t n, const float *feat, const float *target,
                      float *loss, float *grad);
int mint_nn_nparams(int32_t which);
/* Pre-bucket a TP when the scorer's largest pair logit exceeds t (default 0). */
int mint_set_score_threshold(void *ctx, float t);
/* Scorer pairs far from a massive body are left out of the network's work
 * when mode = 1 (the default): the library takes the smallest cutoff c of a
 * ladder (1.5 .. 8, in units of 8 r_c) for which an interval bound of the
 * network's logit over every pair with distance and straight-line closest
 * approach over the window both at least 8 c r_c lies below the threshold,
 * so that no pick can change; mode = 0 scores every pair; mode = -1 leaves
 * the mode as it is.  *bound receives that bound, *cutoff that c (+inf: no
 * safe c, nothing is left out), *active whether pairs are left out now (any
 * output may be NULL).  No counterpart in the reference, which has no scorer. */
int mint_scorer_skip(void *ctx, int32_t mode, float *bound, float *cutoff, int32_t *active);
int mint_nn_get_weights(void *ctx, int32_t which, float *w);
int mint_nn_set_weights(void *ctx, int32_t which, const float *w);

#ifdef __cplusplus
}
#endif
#endif /* MINT_H */
