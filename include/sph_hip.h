/* C-ABI of the MI355X two-loop SPH hot path (libsph_hip.so).
 *
 * Drop-in for the reference's step kernel: the Fortran module call
 *   call iterate(n, store, maxconsenrg)          (playground/src/iterator.f90:44)
 * made from main.f90:219,346 and sts_integrator.f90:101,108, plus the
 * reference's module state it reads (statevars, kernel choice).
 *
 * Conventions (same as the reference):
 *   - store: FP64 particle array, particle-major: particle i's 32 fields at
 *     store[i*ld + (es_* - 1)], i.e. the Fortran column store(:, i+1) with
 *     leading dimension ld >= 32 (const.f90:69-103).  Real particles first,
 *     then fixed ones; periodic ghosts are NOT appended (images are virtual),
 *     so the caller never needs spare capacity.
 *   - state: the 49-slot FP64 statevars vector (const.f90:14-66, state.f90:36).
 *   - kernel_id: 0 = M4, 1 = M6, 2 = gauss (the reference's compile-time
 *     `make kernel=`, Makefile:14,84-86; all three are built here).
 *   - All entry points are synchronous with respect to the host and return an
 *     SPH_* code; a context is used by one host thread at a time.
 */
#ifndef SPH_HIP_H
#define SPH_HIP_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Error codes; each mirrors an abort of the reference. */
#define SPH_OK 0
#define SPH_ERR_NEIGH_TRUNCATED 1 /* neighboursearch.f90:127-132 (count incl. self > maxneibnum) */
#define SPH_ERR_NEG_H 2           /* circuit1.f90:202-204 */
#define SPH_ERR_Q_NAN_OR_NEG 3    /* base_m6.f90:48-53 */
#define SPH_ERR_FLD_ANISO 4       /* circuit2.f90:189,367 */
#define SPH_ERR_BAD_DDW 5         /* circuit2.f90:421-424 */
#define SPH_ERR_PRECONDITION 6    /* preruncheck.f90:13-60 */
#define SPH_ERR_BAD_ARG 7         /* invalid argument (state, sizes, process) */
#define SPH_ERR_HIP 8             /* HIP runtime failure */
#define SPH_ERR_NOMEM 9           /* device allocation failed */
#define SPH_ERR_UNSUPPORTED 10    /* outside the implemented scope (e.g. cylindrical, kernel.f90:137) */
#define SPH_ERR_RAD_NEGATIVE 11   /* rad_exchange.f90:60-63 (energy <= 0 after the exchange) */

typedef struct sph_ctx sph_ctx;

typedef struct {
  int precision;      /* 0 = FP64 throughout (the pair loops take 1/|r| from the hardware
                         reciprocal square root refined to a few ulp); 1 = mixed: FP64
                         positions, separations, Newton update and energy sum, FP32 pair math */
  int tensor_mode;    /* 0 = reference diffusion tensor (difcond*I or difcond*diag(b^2),
                         circuit2.f90:194-206); 1 = kappa_par*bb + kappa_perp*(I-bb) */
  double kappa_perp;  /* perpendicular conductivity for tensor_mode 1 */
  int max_sweeps;     /* Newton sweeps cap; 0 -> 100 (circuit1.f90:114) */
  double newton_tol;  /* relative h tolerance; 0 -> 1e-8 (circuit1.f90:104) FP64, 1e-6 mixed */
  int neigh_cap;      /* neighbour cap incl. self, at most 2^20; 0 -> reference maxneibnum(dim) */
  int device;         /* HIP device ordinal */
  int cell_div;       /* search cells are r_max / cell_div wide (1..4); 0 -> 3.  The
                         neighbour sets do not depend on it, only the search's speed */
} sph_options;

typedef struct {
  int64_t n_real, n_fixed;
  int64_t n_pairs;    /* P = sum_i |N_i| of the last step */
  int64_t loop1_pairs;/* pairs evaluated by loop 1 over all Newton sweeps */
  int sweeps;         /* Newton sweeps of the last step (0 for c1simple) */
  int sorted;         /* 1 if the last step had to reorder particles */
  double energy;      /* FP64 energy sum of loop 2 (circuit2.f90:490-492); the
                         reference computes and discards it */
  double ms_neigh, ms_loop1, ms_loop2, ms_total; /* device time of the last step's phases
                                                    (0 unless sph_set_timing turned timing on) */
  double ms_k_neigh, ms_k_loop1, ms_k_loop2;     /* device time of the neighbour kernel, all
                                                    Newton sweep kernels, the loop-2 kernel */
  int64_t ext_records;  /* records of the search's extended grid (real cells plus the periodic
                           images in ghost cells); 0 when the search read its ranges per lane */
} sph_stats;

/* Fill *opt with defaults (FP64, reference tensor, 100 sweeps, 1e-8). */
int sph_default_options(sph_options *opt);

/* Create a context.  Freezes what the reference freezes at first call
 * (iterator.f90:31-42, circuit1.f90:38-51, circuit2.f90:45-56, BC.f90:46-68):
 * dim, equation set, ddw, adden, artts, au, ics, process, box, bordersize.
 * Replaces: initkernel (kernel.f90:91), initBorders (BC.f90:46). */
int sph_ctx_create(const double state[49], int kernel_id, const sph_options *opt, sph_ctx **out);
int sph_ctx_destroy(sph_ctx *ctx);

/* iterate(n, store, maxconsenrg), iterator.f90:44-179, on a HOST store:
 * n = real + fixed particles; store updated in place exactly on the rows the
 * reference writes (wrapped positions, den, h, om, p, c, dtdx..dtdz, ax..az,
 * dh, du, ddt, dbx..dbz, vfld).  *maxconsenrg = 0 as the reference
 * (circuit2.f90:101); the energy sum is in sph_get_stats. */
int sph_iterate(sph_ctx *ctx, int64_t n, double *store, int64_t ld, const double state[49],
                double *maxconsenrg);

/* Same on a DEVICE store (d_store is device memory, same layout). */
int sph_iterate_device(sph_ctx *ctx, int64_t n, double *d_store, int64_t ld, const double state[49],
                       double *maxconsenrg);

/* Resident use: load a store once, step many times, save.  on_device selects
 * whether `store` is host or device memory. */
int sph_load(sph_ctx *ctx, int64_t n, const double *store, int64_t ld, int on_device);
int sph_save(sph_ctx *ctx, double *store, int64_t ld, int on_device);
int sph_step(sph_ctx *ctx, const double state[49], double *maxconsenrg);

/* Neighbour lists of the last step (replaces xgetneighnumber/xgetneighindex,
 * neighboursearch.f90:221-237, plus getCrossRef, BC.f90:36-44): for real
 * particle i (original order) entries off[i]..off[i+1]-1 hold j = real/fixed
 * index of the neighbour and code = (sx+1)+3(sy+1)+9(sz+1), the periodic image
 * shift (13 = none).  off has n_real+1 entries; cap_pairs bounds j/code. */
int sph_neighbours(sph_ctx *ctx, int64_t *off, int32_t *j, int8_t *code, int64_t cap_pairs,
                   int64_t *n_pairs);

/* No counterpart in the reference: sph_neighbours for callers that cannot
 * allocate device memory (serving loops, harnesses that forbid hipMalloc).
 * The same lists, written to DEVICE memory the caller owns: d_off (n_real + 1
 * entries), d_j and d_code (cap_pairs entries; may be NULL with cap_pairs 0
 * for the offsets alone).  The call allocates nothing: its scratch (list
 * lengths and the prefix sum's storage) is a fixed device array of the
 * library, one per GPU, held until the process ends, which bounds n_real to
 * at most 4,194,304 (larger inputs return SPH_ERR_BAD_ARG).  Calls on one GPU
 * share that array: calls that overlap in time on one GPU (other host
 * threads, other contexts) race and must be serialised by the caller.  When
 * the lists hold more than cap_pairs entries, the first cap_pairs are written,
 * *n_pairs gets the total and SPH_ERR_BAD_ARG is returned.  Synchronous. */
int sph_neighbours_device_noalloc(sph_ctx *ctx, int64_t *d_off, int32_t *d_j, int8_t *d_code, int64_t cap_pairs,
                                  int64_t *n_pairs);

/* checkVarsReady (preruncheck.f90:13-60, called by main.f90:217 before the
 * initial iterate) on the resident store: SPH_ERR_PRECONDITION when mu0 <= 0
 * with MHD (*what = 32), the isotropic conductivity is <= 0 (64), or any
 * loaded particle of non-empty type has density (1), sound speed (2), h (4),
 * pressure (8) or internal energy (16) <= 0; the first failing test in the
 * reference's order.  sph_leapfrog runs it before its initial iterate. */
int sph_prerun_check(sph_ctx *ctx, const double state[49], int *what);

/* One pass of the leapfrog loop around iterate (main.f90:216-362, no STS):
 * dt = min(timestep.calc dt, dtcap) (timestep.f90:40-166), predictor,
 * iterate, corrector, all on the resident state.  The pre-run checks and
 * the initial iterate of main.f90:216-219 run first only when the resident store does not hold the
 * derivatives of its current state (after sph_load, or after a failed step);
 * so N calls after a load are the reference's initial iterate plus N passes. */
int sph_leapfrog(sph_ctx *ctx, const double state[49], double dtcap, double *dt_used);

/* One RKL2 super-time-step of T (sts_integrator.f90:87-124) with s stages on
 * the resident state; s iterates. */
int sph_sts(sph_ctx *ctx, const double state[49], int s, double dt);

/* Implicit gas/radiation energy exchange over dt on the resident state, every
 * loaded particle (real and fixed): update_radenergy(n, store, dt),
 * rad_exchange.f90:9-118, as main.f90:367-369 calls it after the corrector
 * (and after sts_integrate when STS is on). */
int sph_rad_exchange(sph_ctx *ctx, const double state[49], double dt);

/* FP64 RKL2 coefficients (sts_integrator.f90:10-51) for stage j of s:
 * out[0..3] = mu(j), nu(j), mu_hat(j,s), gm_hat(j,s); j >= 1. */
int sph_sts_coeffs(int j, int s, double out[4]);

/* ---- Differentiable anisotropic diffusion (no reference counterpart: the
 * reference has no derivatives).  For particles at rest the loop-2 diffusion
 * term with a global tensor K = (xx, yy, zz, xy, xz, yz) is a linear operator
 * dT/dt = A(K) T (circuit2.f90:376-397 with kta = ktb = K); these entry points
 * apply it, its transpose, and differentiate RKL2 super-steps through it.
 * Vectors are device pointers of n_real doubles in original particle order.
 *
 * sph_diff_prepare: one iterate on the resident particles (neighbours, Newton
 * h, EOS), after which h, rho and Omega are frozen; builds reverse lists.
 * Needs FP64 precision, no fixed particles, a diffusion equation set with a
 * Hessian operator (fab, fw, n2w), no FLD. */
int sph_diff_prepare(sph_ctx *ctx, const double state[49]);
/* d_out = A(K) d_T */
int sph_diff_apply(sph_ctx *ctx, const double K[6], const double *d_T, double *d_out);
/* d_out = A(K)^T d_lam; gA[k] = lam^T dA/dK_k Ya, gB[k] = lam^T dA/dK_k Yb */
int sph_diff_apply_adjoint(sph_ctx *ctx, const double K[6], const double *d_lam, const double *d_Ya,
                           const double *d_Yb, double *d_out, double gA[6], double gB[6]);
/* nsteps RKL2 super-steps of s stages (sts_integrator.f90:87-124) from d_T0
 * into d_T, keeping T every `every` steps as checkpoints for the backward. */
int sph_diff_forward(sph_ctx *ctx, const double K[6], int s, double dt, int nsteps, int every,
                     const double *d_T0, double *d_T);
/* Adjoint of the last sph_diff_forward: given dL/dT_final, dL/dT0 and dL/dK
 * (recomputing each checkpoint segment). */
int sph_diff_backward(sph_ctx *ctx, const double *d_gT, double *d_gT0, double gK[6]);

/* ---- Batched problems (no reference counterpart: the reference runs one
 * ./execute per problem, e.g. 000_anisopaper_v1_gls_kxx.py:44-118).  Many
 * problems, each prepared with sph_diff_prepare on its own context, merged
 * into one operator; every RKL2 stage of every problem is one kernel launch,
 * a super-step is one replayed hipGraph.  The batch copies what it needs: the
 * contexts may be destroyed afterwards. */
typedef struct sph_batch sph_batch;
typedef struct {
  int n_problems;
  int64_t n_total;    /* particles over all problems */
  int64_t n_pairs;    /* pair interactions of one operator application */
  int list_cap;       /* merged neighbour-list capacity */
  int block;          /* stage-kernel block size picked by the autotuner */
  int tune_block[4];  /* candidates ... */
  float tune_ms[4];   /* ... and their time for a 4-stage super-step */
} sph_batch_info;
/* Merge ctxs[0..nprob) (same device, sph_diff_prepare done) with tensors
 * K[6 p .. 6 p + 5] (xx, yy, zz, xy, xz, yz); autotunes the stage launch.
 * precision 0: FP64 pair weights (every problem bitwise its sph_diff_forward);
 * 1: pair weights stored in FP32, products, sums and stage states FP64. */
int sph_batch_create(sph_ctx *const *ctxs, int nprob, const double *K, int precision, sph_batch **out);
int sph_batch_destroy(sph_batch *batch);
/* nsteps RKL2 super-steps of s stages (sts_integrator.f90:87-124), problem p
 * with time step dt[p] (host array).  d_T0 / d_T: device vectors of n_total
 * doubles, problem p's particles (original order) at offset off[p]
 * (sph_batch_offsets).  use_graph: replay one captured super-step per step
 * (1) or launch the stages directly (0); both give the same bits (with
 * precision 0 these are each problem's sph_diff_forward bits). */
int sph_batch_sts(sph_batch *batch, int s, const double *dt, int nsteps, const double *d_T0, double *d_T,
                  int use_graph);
int sph_batch_get_info(sph_batch *batch, sph_batch_info *info);
/* off[0..n_problems]: row offsets of the problems in the batch vectors. */
int sph_batch_offsets(sph_batch *batch, int64_t *off);

int sph_get_stats(sph_ctx *ctx, sph_stats *stats);

/* No counterpart in the reference (this library's own profiling): on = 1 makes
 * every later step record HIP events around its phases and kernels, so that
 * sph_get_stats reports their device times (the ms_* fields); on = 0, the
 * default, leaves those fields 0.  Each event leaves the stream idle for a few
 * microseconds between the kernels it separates (about 40 us per 128^3 step).
 * Any other value returns SPH_ERR_BAD_ARG. */
int sph_set_timing(sph_ctx *ctx, int on);
const char *sph_error_string(int code);

#ifdef __cplusplus
}
#endif
#endif
