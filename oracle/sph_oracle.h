/* CPU restatement of the reference two-loop hot path (FP64, OpenMP).
 *
 * TEST INFRASTRUCTURE ONLY.  This is the checker the HIP path is compared
 * against; it is never linked into, loaded by or called from the product
 * library.  Every function cites the reference file:line it restates.
 *
 * Pinning (what makes this restatement trustworthy), per function:
 *   ghosts + neighbour sets  : golden vectors from the reference Fortran build
 *                              (oracle/_ref, tests/golden), exact set equality;
 *                              brute-force O(N^2) on tiny inputs.
 *   kernels W/dW/ddW         : unit normalisation by quadrature, finite
 *                              differences, continuity (tests/test_oracle_math.py)
 *   loop 1 / EOS / loop 2    : golden vectors (1e-12 relative) for hydro+AV,
 *                              MHD, diffusion with every --ddw, FLD; plus
 *                              invariants (momentum and energy conservation,
 *                              uniform-lattice density, Hessian exactness for
 *                              quadratic T).
 *   full-tensor diffusion    : (extension with kappa_perp, no reference form)
 *                              pinned by reducing to the two reference modes
 *                              (isotropic; axis-aligned b with kappa_perp=0)
 *                              and by rotation invariance.
 *   STS coefficients         : known answers sts_integrator.f90:54-68 and the
 *                              reference harness.
 */
#ifndef SPH_ORACLE_H
#define SPH_ORACLE_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct {
  double kappa_perp;   /* full-tensor diffusion: perpendicular conductivity */
  int tensor_mode;     /* 0 = reference diag(b^2) / isotropic, 1 = full tensor */
  int max_sweeps;      /* Newton sweeps cap (reference: 100) */
  double tol;          /* Newton tolerance (reference: 1e-8) */
  int neigh_cap;       /* 0 = reference maxneibnum(dim) */
} orc_ext;

typedef struct {
  int64_t n_real;
  int64_t n_pairs;
  int64_t *nb_off;     /* n_real+1 */
  int32_t *nb_j;       /* real/fixed index (0-based) of each neighbour */
  int8_t *nb_code;     /* image code (sx+1)+3(sy+1)+9(sz+1) */
  int sweeps;          /* Newton sweeps taken (0 for c1simple) */
  double energy;       /* FP64 energy sum of loop 2 (circuit2.f90:490-492) */
  int64_t n_ghost;
} orc_result;

/* Error codes shared with the product C-ABI. */
#define ORC_OK 0
#define ORC_NEIGH_TRUNCATED 1
#define ORC_NEG_H 2
#define ORC_Q_NAN_OR_NEG 3
#define ORC_FLD_ANISO 4
#define ORC_BAD_DDW 5
#define ORC_BAD_ARG 7
#define ORC_RAD_NEGATIVE 11

/* iterate(n, store, dedt): iterator.f90:44-179.  store is (n_total x 32)
 * row-major (row i == Fortran column store(:, i+1)).  kernel 0=M4 1=M6 2=gauss. */
int orc_iterate(const double *state, int kernel, const orc_ext *ext,
                int64_t n_total, double *store, orc_result *res);
void orc_free_result(orc_result *res);

/* Kernel base functions (base_m6.f90:36-97, base_m4.f90, base_gauss.f90). */
double orc_kf(int kernel, double q);
double orc_kdf(int kernel, double q);
double orc_kddf(int kernel, double q);
double orc_knorm(int kernel, int dim);

/* STS coefficients, sts_integrator.f90:10-51. */
double orc_sts_mu(int j);
double orc_sts_nu(int j);
double orc_sts_mu_hat(int j, int s);
double orc_sts_gm_hat(int j, int s);

/* update_radenergy(n, store, dt), rad_exchange.f90:9-118, on all n rows. */
int orc_rad_exchange(const double *state, int64_t n, double *store, double dt);

#ifdef __cplusplus
}
#endif
#endif
