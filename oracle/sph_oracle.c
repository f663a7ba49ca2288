/* CPU restatement of the reference two-loop hot path, FP64 + OpenMP.
 *
 * TEST INFRASTRUCTURE ONLY: the checker for the HIP path, never part of it.
 * Structure deliberately follows the reference (explicit periodic ghost
 * copies, per-particle neighbour lists, i-centric loops) so that it is an
 * independent check of the device path's virtual-image cell list.
 * Pinning of each function is listed in sph_oracle.h.
 *
 * Build: gcc -O2 -fopenmp -ffp-contract=off -fPIC -shared (see oracle/Makefile);
 * contraction must stay off so that the neighbour test rounds as written.
 */
#include "sph_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

/* store rows, const.f90:69-103 (0-based here) */
enum { RX = 0, RY, RZ, VX, VY, VZ, AX, AY, AZ, M, DEN, H, DH, P, U, DU, T, KAPPA,
       DTDX, DTDY, DTDZ, DDT, BX, BY, BZ, DBX, DBY, DBZ, CS, OM, TYPE, VFLD };
/* statevars slots, const.f90:14-66 (0-based here) */
enum { EC_DIM = 0, EC_EQS = 1, EC_DDW = 2, EC_ICS = 3, EC_ADDEN = 4, EC_ARTTS = 5,
       EC_DISO = 7, EC_DCOND = 8, EC_MUZERO = 9, EC_HFAC = 12, EC_XMIN = 16,
       EC_BORDSIZE = 22, EC_PADDING = 23, EC_REALPN = 24, EC_FIXEDPN = 25,
       EC_GAMMA = 26, EC_PROCESS = 30, EC_AU = 31, EC_EQONHYDRO = 33,
       EC_EQONDIFF = 34, EC_EQONFLD = 35, EC_EQONRADEXCH = 36, EC_EQONSTS = 37,
       EC_MOLMASS = 46 };

#define ESD_FAB 400
#define ESD_N2W 401
#define ESD_FW 402
#define ESD_2NW_DS 403
#define ESD_2NW_SD 404

static const double PI_ = 3.141592653589793;     /* 4*atan(1), const.f90:4 */
static const double LIGHTSPEED = 2.997924e10;    /* const.f90:6 */
static const double RG = 8.314e7;                /* const.f90:7 */
static const double SIGMAB = 5.67051e-5;         /* const.f90:8 */

#define ST(i, f) store[(int64_t)(i) * 32 + (f)]

/* ---------------------------------------------------------------- kernels */
/* x**n with integer n as an exponentiation-by-squaring product. */
static inline double ipow(double x, int n) {
  double r = 1.0;
  int first = 1;
  while (n > 0) {
    if (n & 1) { r = first ? x : r * x; first = 0; }
    n >>= 1;
    if (n) x = x * x;
  }
  return r;
}

double orc_knorm(int kernel, int dim) {
  /* base_m4.f90:110, base_m6.f90:13, base_gauss.f90:204 */
  if (kernel == 0) {
    const double k[3] = {2. / 3., 10. / (7. * PI_), 1. / PI_};
    return k[dim - 1];
  } else if (kernel == 1) {
    const double k[3] = {1. / 120., 7. / (478. * PI_), 1. / (120. * PI_)};
    return k[dim - 1];
  }
  const double k[3] = {1. / sqrt(PI_), 1. / PI_, 1. / pow(PI_, 1.5)};
  return k[dim - 1];
}

static double fwc_of(int kernel, int dim) {
  /* base_m4.f90:111, base_m6.f90:15, base_gauss.f90:206 */
  if (kernel == 0) { const double f[3] = {6., 6.3226, 6.66400}; return f[dim - 1]; }
  if (kernel == 1) { const double f[3] = {4., 9.31505 / 2.25, 9.6429 / 2.25}; return f[dim - 1]; }
  return 4.0;
}

static double krad_of(int kernel) { return kernel == 0 ? 2.0 : (kernel == 1 ? 3.0 : 10.0); }

static int maxneib_of(int kernel, int dim) {
  static const int t[3][3] = {{35, 250, 150}, {50, 10000, 150}, {100, 5000, 50000}};
  return t[kernel][dim - 1];
}

/* Shared error flag for the parallel loops (first error wins). */
static int g_err;
static inline void set_err(int e) {
#pragma omp atomic write
  g_err = e;
}

/* base_m6.f90:36-55, base_m4.f90:133-150, base_gauss.f90:228-241 */
double orc_kf(int kernel, double q) {
  if (!(q >= 0.)) { set_err(ORC_Q_NAN_OR_NEG); return 0.; }
  if (kernel == 1) {
    if (q >= 3.) return 0.;
    if (q >= 2.) return ipow(3. - q, 5);
    if (q >= 1.) return ipow(3. - q, 5) - 6. * ipow(2. - q, 5);
    return ipow(3. - q, 5) - 6. * ipow(2. - q, 5) + 15. * ipow(1. - q, 5);
  } else if (kernel == 0) {
    if (q >= 2.) return 0.;
    if (q >= 1.) return 0.25 * ipow(2. - q, 3);
    return 0.25 * ipow(2. - q, 3) - ipow(1. - q, 3);
  }
  return exp(-q * q);
}

/* base_m6.f90:57-76, base_m4.f90:152-169, base_gauss.f90:243-256 */
double orc_kdf(int kernel, double q) {
  if (!(q >= 0.)) { set_err(ORC_Q_NAN_OR_NEG); return 0.; }
  if (kernel == 1) {
    if (q >= 3.) return 0.;
    if (q >= 2.) return -5. * ipow(3. - q, 4);
    if (q >= 1.) return -5. * ipow(3. - q, 4) + 30. * ipow(2. - q, 4);
    return -5. * ipow(3. - q, 4) + 30. * ipow(2. - q, 4) - 75. * ipow(1. - q, 4);
  } else if (kernel == 0) {
    if (q >= 2.) return 0.;
    if (q >= 1.) return -0.75 * ipow(2. - q, 2);
    return -0.75 * ipow(2. - q, 2) + 3 * ipow(1. - q, 2);
  }
  return -2. * exp(-q * q) * q;
}

/* base_m6.f90:78-97, base_m4.f90:171-188, base_gauss.f90:258-271 */
double orc_kddf(int kernel, double q) {
  if (!(q >= 0.)) { set_err(ORC_Q_NAN_OR_NEG); return 0.; }
  if (kernel == 1) {
    if (q >= 3.) return 0.;
    if (q >= 2.) return 20. * ipow(3. - q, 3);
    if (q >= 1.) return 20. * ipow(3. - q, 3) - 120. * ipow(2. - q, 3);
    return 20. * ipow(3. - q, 3) - 120. * ipow(2. - q, 3) + 300. * ipow(1. - q, 3);
  } else if (kernel == 0) {
    if (q >= 2.) return 0.;
    if (q >= 1.) return 1.5 * (2. - q);
    return 1.5 * (2. - q) - 6 * (1. - q);
  }
  return -2 * exp(-q * q) + 4 * exp(-q * q) * q * q;
}

typedef struct {
  int kernel, dim, ddw;
  double wcv, fwc;
} kctx;

/* kernel.f90:178-187 */
static inline double get_w(const kctx *k, double r, double h) {
  double q = r / h;
  return k->wcv * orc_kf(k->kernel, q) / ipow(h, k->dim);
}
/* kernel.f90:239-249 */
static inline double get_dw_dh(const kctx *k, double r, double h) {
  double q = r / h;
  double f = orc_kf(k->kernel, q), df = orc_kdf(k->kernel, q);
  return -k->wcv / ipow(h, k->dim + 1) * (k->dim * f + q * df);
}
/* kernel.f90:189-197 */
static inline void nw_cart(const kctx *k, const double rab[3], double dr, double h, double o[3]) {
  double q = dr / h;
  double df = orc_kdf(k->kernel, q);
  double hp = ipow(h, k->dim + 2);
  for (int l = 0; l < 3; l++) o[l] = k->wcv * df * rab[l] / hp / q;
}
/* kernel.f90:251-259 */
static inline double fab_cart(const kctx *k, double dr, double h) {
  double q = dr / h;
  return k->wcv * orc_kdf(k->kernel, q) * dr / ipow(h, k->dim + 2) / q;
}
/* kernel.f90:271-279 */
static inline double gab_cart(const kctx *k, double dr, double h) {
  double q = dr / h;
  return -2. * k->wcv * orc_kdf(k->kernel, q) / ipow(h, k->dim + 2) / q;
}
/* kernel.f90:261-269 */
static inline double fw_cart(const kctx *k, double dr, double h) {
  double q = dr / h;
  return k->fwc * k->wcv * orc_kf(k->kernel, q) / ipow(h, k->dim + 2);
}

/* hessian_fab_cart (kernel.f90:460-503), hessian_fw_cart (:505-539),
 * hessian_ddw_cart (:377-417), selected as initkernel (:108-134). */
static void hessian(const kctx *k, const double rab[3], double h, double Hs[3][3]) {
  double r2 = rab[0] * rab[0] + rab[1] * rab[1] + rab[2] * rab[2];
  double rr[3][3];
  for (int a = 0; a < 3; a++)
    for (int b = 0; b < 3; b++) rr[a][b] = rab[a] * rab[b] / r2;
  memset(Hs, 0, sizeof(double) * 9);
  double dr = sqrt(r2);
  int dim = k->dim;
  if (k->ddw == ESD_N2W) {
    double cstart = k->wcv / ipow(h, dim + 2);
    double q = dr / h;
    double ddf = orc_kddf(k->kernel, q), df = orc_kdf(k->kernel, q);
    double dfq = df / q;
    for (int a = 0; a < dim; a++)
      for (int b = 0; b < dim; b++)
        Hs[a][b] = (a == b) ? cstart * (ddf * rr[a][a] + dfq * (1 - rr[a][a]))
                            : cstart * (ddf * rr[a][b] - dfq * rr[a][b]);
  } else {
    double fab = (k->ddw == ESD_FW) ? fw_cart(k, dr, h) : gab_cart(k, dr, h);
    fab = 0.5 * fab;
    for (int a = 0; a < dim; a++)
      for (int b = 0; b < dim; b++)
        Hs[a][b] = (a == b) ? ((dim + 2) * rr[a][a] - 1) * fab : (dim + 2) * rr[a][b] * fab;
  }
}

static inline double dot3(const double *a, const double *b) { return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]; }

/* -------------------------------------------------------------- geometry */
typedef struct { int64_t n, cap; int64_t *a; } ilist;
static void ipush(ilist *l, int64_t v) {
  if (l->n == l->cap) { l->cap = l->cap ? 2 * l->cap : 64; l->a = realloc(l->a, sizeof(int64_t) * l->cap); }
  l->a[l->n++] = v;
}

typedef struct {
  int64_t ng;
  double *pos;      /* 3*ng */
  int64_t *cref;    /* ng: real/fixed index */
} ghosts_t;

/* reflectRealParticlesPeriodically, BC.f90:102-194: wrap escaped real
 * particles along the axes in `axes` (bit a set = axis a). */
static void wrap_real(double *store, int64_t nreal, int dim, const double *emin,
                      const double *emax, int axes) {
  for (int64_t i = 0; i < nreal; i++) {
    double dxmin[3], dxmax[3];
    for (int a = 0; a < 3; a++) {
      dxmin[a] = emin[a] - ST(i, RX + a);
      dxmax[a] = ST(i, RX + a) - emax[a];
    }
    for (int a = 0; a < dim; a++) {
      if (!(axes & (1 << a))) continue;
      if (dxmax[a] > 0) ST(i, RX + a) = emin[a] + dxmax[a];
      if (dxmin[a] > 0) ST(i, RX + a) = emax[a] - dxmin[a];
    }
  }
}

/* findInsideBorderParticles (BC.f90:196-309) followed by createPeriodicBorder
 * for each axis set in `axes` in x, y, z order (BC.f90:311-475; hydroshock
 * calls it for y then z, iterator.f90:107-108, which yields the same lists). */
static void make_ghosts(const double *store, int64_t ntot, int dim, const double *emin,
                        const double *emax, double bs, double padding, int axes, ghosts_t *g) {
  double boxmin[3], boxmax[3], pmin[3], pmax[3];
  for (int a = 0; a < 3; a++) {
    boxmax[a] = emax[a] - bs;
    boxmin[a] = emin[a] + bs;
    pmin[a] = emin[a];
    pmax[a] = emax[a];
    if (padding == 0) { pmin[a] = emin[a] + padding / 10; pmax[a] = emax[a] - padding / 10; }
  }
  ilist ib[3][2];
  memset(ib, 0, sizeof(ib));
  for (int64_t i = 0; i < ntot; i++) {
    int ty = (int)ST(i, TYPE);
    if (!(ty == 1 || ty == 2 || ty == 21)) continue;
    for (int a = 0; a < dim; a++) {
      double x = ST(i, RX + a);
      if (x < boxmin[a]) {
        if (x > pmin[a]) ipush(&ib[a][0], i);
      } else if (x > boxmax[a]) {
        if (x < pmax[a]) ipush(&ib[a][1], i);
      }
    }
  }
  int64_t cap = 1024, ng = 0;
  double *gp = malloc(sizeof(double) * 3 * cap);
  int64_t *cr = malloc(sizeof(int64_t) * cap);
  for (int a = 0; a < dim; a++) {
    if (!(axes & (1 << a))) continue;
    for (int side = 0; side < 2; side++) {
      for (int64_t t = 0; t < ib[a][side].n; t++) {
        int64_t pi = ib[a][side].a[t];
        if (ng == cap) {
          cap *= 2;
          gp = realloc(gp, sizeof(double) * 3 * cap);
          cr = realloc(cr, sizeof(int64_t) * cap);
        }
        double p[3];
        int64_t src;
        if (pi < ntot) { for (int c = 0; c < 3; c++) p[c] = ST(pi, RX + c); src = pi; }
        else { for (int c = 0; c < 3; c++) p[c] = gp[3 * (pi - ntot) + c]; src = cr[pi - ntot]; }
        if (side == 0) p[a] = emax[a] + (p[a] - emin[a]);
        else p[a] = emin[a] + (p[a] - emax[a]);
        int64_t k = ntot + ng;
        for (int c = 0; c < 3; c++) gp[3 * ng + c] = p[c];
        cr[ng] = src;
        ng++;
        /* chain onto the later axes' lists without the padding test */
        for (int b = a + 1; b < dim; b++) {
          if (p[b] < boxmin[b]) ipush(&ib[b][0], k);
          else if (p[b] > boxmax[b]) ipush(&ib[b][1], k);
        }
      }
    }
  }
  for (int a = 0; a < 3; a++) { free(ib[a][0].a); free(ib[a][1].a); }
  g->ng = ng;
  g->pos = gp;
  g->cref = cr;
}

static inline const double *ppos(const double *store, const ghosts_t *g, int64_t ntot, int64_t j) {
  return j < ntot ? &store[j * 32 + RX] : &g->pos[3 * (j - ntot)];
}

/* findneighboursKDT + kdtsearchFiller (neighboursearch.f90:74-194) with the
 * kd-tree's inclusion rule (kdtree2.f90:1673-1686): over all real, fixed and
 * ghost points j != i, include iff ((0+dx^2)+dy^2)+dz^2 <= (krad*h_i)^2,
 * dx = x_j - x_i.  A uniform grid replaces the tree (same set); lists are in
 * ascending point index.  Truncation at maxneibnum counts self
 * (neighboursearch.f90:127). */
static int neighbours(const double *store, int64_t nreal, int64_t ntot, const ghosts_t *g,
                      double kr, int cap, int64_t **off_out, int64_t **lst_out) {
  int64_t np = ntot + g->ng;
  double lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
  double rmax = 0;
  for (int64_t j = 0; j < np; j++) {
    const double *p = ppos(store, g, ntot, j);
    for (int a = 0; a < 3; a++) { if (p[a] < lo[a]) lo[a] = p[a]; if (p[a] > hi[a]) hi[a] = p[a]; }
  }
  for (int64_t i = 0; i < nreal; i++) { double r = kr * ST(i, H); if (r > rmax) rmax = r; }
  if (!(rmax > 0)) rmax = 1.0;
  /* every cell edge >= the largest search radius, so the 27 surrounding
   * cells hold every candidate */
  int64_t nc[3];
  double csz[3];
  for (int a = 0; a < 3; a++) {
    double ext = hi[a] - lo[a];
    csz[a] = rmax * 1.0000001;
    if (ext / 512.0 > csz[a]) csz[a] = ext / 512.0 * 1.0000001;
    nc[a] = (int64_t)floor(ext / csz[a]) + 1;
  }
  int64_t ncell = nc[0] * nc[1] * nc[2];
  int64_t *cstart = calloc(ncell + 1, sizeof(int64_t));
  int64_t *cell = malloc(sizeof(int64_t) * np);
  for (int64_t j = 0; j < np; j++) {
    const double *p = ppos(store, g, ntot, j);
    int64_t c[3];
    for (int a = 0; a < 3; a++) {
      c[a] = (int64_t)floor((p[a] - lo[a]) / csz[a]);
      if (c[a] < 0) c[a] = 0;
      if (c[a] >= nc[a]) c[a] = nc[a] - 1;
    }
    cell[j] = (c[2] * nc[1] + c[1]) * nc[0] + c[0];
    cstart[cell[j] + 1]++;
  }
  for (int64_t c = 0; c < ncell; c++) cstart[c + 1] += cstart[c];
  int64_t *fill = malloc(sizeof(int64_t) * ncell);
  memcpy(fill, cstart, sizeof(int64_t) * ncell);
  int64_t *sorted = malloc(sizeof(int64_t) * np);
  for (int64_t j = 0; j < np; j++) sorted[fill[cell[j]]++] = j;  /* ascending j per cell */
  free(fill);

  int64_t *cnt = calloc(nreal + 1, sizeof(int64_t));
  int err = 0;
  /* pass 1 counts, pass 2 fills */
  int64_t *off = NULL, *lst = NULL;
  for (int pass = 0; pass < 2; pass++) {
#pragma omp parallel for schedule(dynamic, 64) reduction(| : err)
    for (int64_t i = 0; i < nreal; i++) {
      const double *pi = &store[i * 32 + RX];
      double rr = kr * store[i * 32 + H];
      double r2 = rr * rr;
      int64_t ci[3];
      for (int a = 0; a < 3; a++) {
        ci[a] = (int64_t)floor((pi[a] - lo[a]) / csz[a]);
        if (ci[a] < 0) ci[a] = 0;
        if (ci[a] >= nc[a]) ci[a] = nc[a] - 1;
      }
      int64_t k = 0;
      int64_t *dst = pass ? &lst[off[i]] : NULL;
      /* candidate cells; within one i the order is irrelevant, the list is
       * sorted afterwards */
      for (int64_t cz = ci[2] - 1; cz <= ci[2] + 1; cz++) {
        if (cz < 0 || cz >= nc[2]) continue;
        for (int64_t cy = ci[1] - 1; cy <= ci[1] + 1; cy++) {
          if (cy < 0 || cy >= nc[1]) continue;
          for (int64_t cx = ci[0] - 1; cx <= ci[0] + 1; cx++) {
            if (cx < 0 || cx >= nc[0]) continue;
            int64_t c = (cz * nc[1] + cy) * nc[0] + cx;
            for (int64_t t = cstart[c]; t < cstart[c + 1]; t++) {
              int64_t j = sorted[t];
              if (j == i) continue;
              const double *pj = ppos(store, g, ntot, j);
              double sd = 0.0, d;
              d = pj[0] - pi[0]; sd = sd + d * d;
              d = pj[1] - pi[1]; sd = sd + d * d;
              d = pj[2] - pi[2]; sd = sd + d * d;
              if (sd <= r2) {
                if (pass) dst[k] = j;
                k++;
              }
            }
          }
        }
      }
      if (!pass) {
        cnt[i] = k;
        if (k + 1 > cap) err |= 1;
      } else {
        /* insertion sort: ascending point index */
        for (int64_t a = 1; a < k; a++) {
          int64_t v = dst[a], b = a - 1;
          while (b >= 0 && dst[b] > v) { dst[b + 1] = dst[b]; b--; }
          dst[b + 1] = v;
        }
      }
    }
    if (!pass) {
      off = malloc(sizeof(int64_t) * (nreal + 1));
      off[0] = 0;
      for (int64_t i = 0; i < nreal; i++) off[i + 1] = off[i] + cnt[i];
      lst = malloc(sizeof(int64_t) * (off[nreal] > 0 ? off[nreal] : 1));
    }
  }
  free(cnt); free(cstart); free(cell); free(sorted);
  *off_out = off;
  *lst_out = lst;
  return err ? ORC_NEIGH_TRUNCATED : ORC_OK;
}

/* ------------------------------------------------------------- the loops */
typedef struct {
  int dim, ddw, adden, artts, diso;
  int hydro, magneto, diff, fld;
  double hfac, dcond, muzero, au, gamma;
  int tensor_mode;
  double kappa_perp;
} flags_t;

/* c1advanced, circuit1.f90:76-224 */
static int loop1_advanced(const kctx *k, const flags_t *F, double *store, int64_t nreal, int64_t ntot,
                          const ghosts_t *g, const int64_t *off, const int64_t *lst,
                          int max_sweeps, double tol, int *sweeps_out) {
  double *hk = malloc(sizeof(double) * ntot), *hkp1 = malloc(sizeof(double) * ntot);
  double *dk = malloc(sizeof(double) * ntot), *dkp1 = malloc(sizeof(double) * ntot);
  double *ok = malloc(sizeof(double) * ntot), *okp1 = malloc(sizeof(double) * ntot);
  double *eps = calloc(ntot, sizeof(double));
  for (int64_t i = 0; i < ntot; i++) { hkp1[i] = ST(i, H); dkp1[i] = ST(i, DEN); okp1[i] = ST(i, OM); }
  for (int64_t i = 0; i < nreal; i++) eps[i] = 1.;
  int iter = 0, dim = F->dim;
  int need_grad = F->diff || F->fld;
  for (;;) {
    double mx = -INFINITY;
    int any = 0;
    for (int64_t i = 0; i < ntot; i++) if (eps[i] > 0) { any = 1; if (eps[i] > mx) mx = eps[i]; }
    if (!(any && mx > tol && iter < max_sweeps)) break;
    iter++;
    memcpy(hk, hkp1, sizeof(double) * ntot);
    memcpy(dk, dkp1, sizeof(double) * ntot);
    memcpy(ok, okp1, sizeof(double) * ntot);
#pragma omp parallel for schedule(dynamic, 64)
    for (int64_t i = 0; i < nreal; i++) {
      if (!(eps[i] > tol)) continue;
      double ma = ST(i, M), ha0 = ST(i, H), ta = ST(i, T);
      const double *ra = &ST(i, RX);
      double ha = hk[i], rhoa = dk[i], oma = ok[i];
      double dtadx[3] = {0, 0, 0};
      double dsum = 0., osum = 0.;
      for (int64_t t = off[i]; t < off[i + 1]; t++) {
        int64_t j = lst[t];
        int64_t rj = j < ntot ? j : g->cref[j - ntot];
        const double *rb = ppos(store, g, ntot, j);
        double r[3] = {ra[0] - rb[0], ra[1] - rb[1], ra[2] - rb[2]};
        double r2 = dot3(r, r);
        double dr = sqrt(r2);
        double mb = ST(rj, M), tb = ST(rj, T);
        double rhob = dk[rj], hb = hk[rj], omb = ok[rj];
        double dwdh = get_dw_dh(k, dr, ha);
        double w = get_w(k, dr, ha);
        dsum = dsum + mb * w;
        osum = osum + mb * dwdh;
        if (need_grad) {
          double nwa[3], nwb[3];
          if (F->ddw == ESD_2NW_SD) {
            nw_cart(k, r, dr, ha, nwa);
            nw_cart(k, r, dr, hb, nwb);
            for (int l = 0; l < 3; l++)
              dtadx[l] = dtadx[l] + mb * (ta * rhoa / oma / rhoa / rhoa * nwa[l] + tb * rhoa / omb / rhob / rhob * nwb[l]);
          } else {
            nw_cart(k, r, dr, ha, nwa);
            for (int l = 0; l < 3; l++) dtadx[l] = dtadx[l] + mb / rhoa / oma * (tb * rhob - ta * rhoa) * nwa[l];
          }
        }
      }
      double w0 = get_w(k, 0., ha), dwdh0 = get_dw_dh(k, 0., ha);
      dsum = dsum + ma * w0;
      osum = osum + ma * dwdh0;
      osum = 1. - osum * (-ha / (dim * dsum));
      double dfdh = -dim * dsum * osum / ha;
      double fh = ma * ipow(F->hfac / ha, dim) - dsum;
      double hn = ha - fh / dfdh;
      if (hn <= 0.) set_err(ORC_NEG_H);
      eps[i] = fabs(hn - ha) / ha0;
      hkp1[i] = hn;
      dkp1[i] = dsum;
      okp1[i] = osum;
      for (int l = 0; l < 3; l++) ST(i, DTDX + l) = dtadx[l];
    }
    if (g_err) break;
  }
  for (int64_t i = 0; i < nreal; i++) { ST(i, H) = hkp1[i]; ST(i, DEN) = dkp1[i]; ST(i, OM) = okp1[i]; }
  *sweeps_out = iter;
  free(hk); free(hkp1); free(dk); free(dkp1); free(ok); free(okp1); free(eps);
  return g_err;
}

/* c1simple, circuit1.f90:227-305, plus the om=1 of c1 (:62) */
static int loop1_simple(const kctx *k, const flags_t *F, double *store, int64_t nreal, int64_t ntot,
                        const ghosts_t *g, const int64_t *off, const int64_t *lst) {
  double *dk = malloc(sizeof(double) * nreal), *hk = malloc(sizeof(double) * nreal);
  double *grad = malloc(sizeof(double) * 3 * nreal);
  int dim = F->dim;
#pragma omp parallel for schedule(dynamic, 64)
  for (int64_t i = 0; i < nreal; i++) {
    const double *ra = &ST(i, RX);
    double ha = ST(i, H), ma = ST(i, M), rhoa = ST(i, DEN), ta = ST(i, T);
    double dsum = 0, dtadx[3] = {0, 0, 0};
    for (int64_t t = off[i]; t < off[i + 1]; t++) {
      int64_t j = lst[t];
      int64_t rj = j < ntot ? j : g->cref[j - ntot];
      const double *rb = ppos(store, g, ntot, j);
      double hb = ST(rj, H), mb = ST(rj, M), rhob = ST(rj, DEN), tb = ST(rj, T);
      double r[3] = {ra[0] - rb[0], ra[1] - rb[1], ra[2] - rb[2]};
      double dr = sqrt(dot3(r, r));
      dsum = dsum + mb * get_w(k, dr, ha);
      if (F->diff) {
        double nwa[3], nwb[3];
        if (F->ddw == ESD_2NW_DS) {
          nw_cart(k, r, dr, ha, nwa);
          for (int l = 0; l < 3; l++) dtadx[l] = dtadx[l] + mb / rhoa * (tb - ta) * nwa[l];
        } else if (F->ddw == ESD_2NW_SD) {
          nw_cart(k, r, dr, ha, nwa);
          nw_cart(k, r, dr, hb, nwb);
          for (int l = 0; l < 3; l++) dtadx[l] = dtadx[l] + mb * (ta / rhoa / rhoa * nwa[l] + tb / rhob / rhob * nwb[l]);
        }
      }
    }
    dsum = dsum + ma * get_w(k, 0., ha);
    dk[i] = dsum;
    hk[i] = F->hfac * pow(ma / dsum, 1. / dim);
    for (int l = 0; l < 3; l++) grad[3 * i + l] = dtadx[l];
  }
  for (int64_t i = 0; i < nreal; i++) {
    ST(i, DEN) = dk[i];
    ST(i, H) = hk[i];
    for (int l = 0; l < 3; l++) ST(i, DTDX + l) = grad[3 * i + l];
  }
  for (int64_t i = 0; i < ntot; i++) ST(i, OM) = 1.;
  free(dk); free(hk); free(grad);
  return g_err;
}

/* eos_adiabatic / eos_isothermal, eos.f90:21-45 */
static void eos(double *store, int64_t nreal, double gamma, int isothermal) {
  double c1 = ST(0, CS);
  for (int64_t i = 0; i < nreal; i++) {
    if (isothermal) {
      ST(i, P) = ST(i, DEN) * c1 * c1;
    } else {
      ST(i, P) = (gamma - 1.) * ST(i, DEN) * ST(i, U);
      ST(i, CS) = sqrt(gamma * ST(i, P) / ST(i, DEN));
    }
  }
}

/* diffusion tensor of one particle: reference modes (circuit2.f90:194-206,
 * 370-374) or the full kappa_par b b + kappa_perp (I - b b) extension */
static void cond_tensor(const flags_t *F, const double ub[3], double kt[3][3]) {
  if (F->diso == 1) {
    memset(kt, 0, sizeof(double) * 9);
    for (int l = 0; l < 3; l++) kt[l][l] = F->dcond;
  } else if (F->tensor_mode == 1) {
    for (int a = 0; a < 3; a++)
      for (int b = 0; b < 3; b++)
        kt[a][b] = F->dcond * ub[a] * ub[b] + F->kappa_perp * ((a == b ? 1.0 : 0.0) - ub[a] * ub[b]);
  } else {
    memset(kt, 0, sizeof(double) * 9);
    for (int l = 0; l < 3; l++) kt[l][l] = F->dcond * ub[l] * ub[l];
  }
}

/* c2, circuit2.f90:58-505 (with art_viscosity :558-575, art_termcond :520-529) */
static int loop2(const kctx *k, const flags_t *F, double *store, int64_t nreal, int64_t ntot,
                 const ghosts_t *g, const int64_t *off, const int64_t *lst, double *energy) {
  double consenrg = 0.;
  int dim = F->dim;
#pragma omp parallel for schedule(dynamic, 64) reduction(+ : consenrg)
  for (int64_t i = 0; i < nreal; i++) {
    double rhoa = ST(i, DEN);
    const double *ra = &ST(i, RX);
    double ba[3] = {ST(i, BX), ST(i, BY), ST(i, BZ)};
    double uba[3] = {0, 0, 0};
    double bb2a = dot3(ba, ba);
    if (bb2a != 0) for (int l = 0; l < 3; l++) uba[l] = ba[l] / sqrt(bb2a);
    double va[3] = {ST(i, VX), ST(i, VY), ST(i, VZ)};
    double ha = ST(i, H), ca = ST(i, CS), pa = ST(i, P), ua = ST(i, U), oma = ST(i, OM);
    double ta = ST(i, T), ma = ST(i, M), kappaa = ST(i, KAPPA);
    double dtadx[3] = {ST(i, DTDX), ST(i, DTDY), ST(i, DTDZ)};
    double dva[3] = {0, 0, 0}, dua = 0, dha = 0, ddta = 0, dba[3] = {0, 0, 0}, drhoadt = 0;
    double kta[3][3] = {{0}}, ktb[3][3] = {{0}}, MPa[3][3] = {{0}};
    if (F->magneto) {
      for (int a = 0; a < 3; a++)
        for (int b = 0; b < 3; b++)
          MPa[a][b] = (a == b) ? (0.5 * bb2a - ba[a] * ba[a]) / F->muzero : -ba[a] * ba[b] / F->muzero;
    }
    if (F->diff) {
      if (F->fld) {
        if (F->diso == 1) {
          double fld_ra = sqrt(dot3(dtadx, dtadx)) / (kappaa * rhoa * ta * rhoa);
          double lambdaa = (2. + fld_ra) / (6. + 3. * fld_ra + fld_ra * fld_ra);
          double fldda = LIGHTSPEED * lambdaa / kappaa / rhoa;
          kta[0][0] = kta[1][1] = kta[2][2] = fldda;
          ST(i, VFLD) = fldda / ((ta * rhoa) / sqrt(dot3(dtadx, dtadx)));
        } else {
          set_err(ORC_FLD_ANISO);
        }
      } else {
        cond_tensor(F, uba, kta);
        if (F->diso == 1) cond_tensor(F, uba, ktb);
      }
    }
    for (int64_t t = off[i]; t < off[i + 1]; t++) {
      int64_t j = lst[t];
      int64_t rj = j < ntot ? j : g->cref[j - ntot];
      const double *rb = ppos(store, g, ntot, j);
      double bb[3] = {ST(rj, BX), ST(rj, BY), ST(rj, BZ)};
      double ubb[3] = {0, 0, 0};
      double bb2 = dot3(bb, bb);
      if (bb2 != 0) for (int l = 0; l < 3; l++) ubb[l] = bb[l] / sqrt(bb2);
      double vb[3] = {ST(rj, VX), ST(rj, VY), ST(rj, VZ)};
      double rhob = ST(rj, DEN), hb = ST(rj, H), cb = ST(rj, CS), pb = ST(rj, P), ub = ST(rj, U);
      double omb = ST(rj, OM), tb = ST(rj, T), mb = ST(rj, M), kappab = ST(rj, KAPPA);
      double dtbdx[3] = {ST(rj, DTDX), ST(rj, DTDY), ST(rj, DTDZ)};
      double qa = 0, qb = 0, qc = 0;
      double rab[3] = {ra[0] - rb[0], ra[1] - rb[1], ra[2] - rb[2]};
      double r2 = dot3(rab, rab);
      double dr = sqrt(r2);
      double vab[3] = {va[0] - vb[0], va[1] - vb[1], va[2] - vb[2]};
      double urab[3] = {rab[0] / dr, rab[1] / dr, rab[2] / dr};
      double nwa[3], nwb[3];
      nw_cart(k, rab, dr, ha, nwa);
      nw_cart(k, rab, dr, hb, nwb);
      if (F->adden) drhoadt = drhoadt + mb * dot3(vab, nwa);
      if (F->artts) {
        double faba = fab_cart(k, dr, ha), fabb = fab_cart(k, dr, hb);
        double dvr = dot3(vab, urab);
        if (dvr < 0) {
          double vsig = 1. * ca + 2. * fabs(dvr);
          qa = -0.5 * rhoa * vsig * dvr;
          vsig = 1. * cb + 2. * fabs(dvr);
          qb = -0.5 * rhob * vsig * dvr;
        }
        for (int l = 0; l < 3; l++)
          dva[l] = dva[l] - mb * (qa * nwa[l] / (rhoa * rhoa) / oma + qb * nwb[l] / (rhob * rhob) / omb);
        if (F->au != -1) {
          double vsigu = sqrt(fabs(pa - pb) / (0.5 * (rhoa + rhob)));
          qc = vsigu * (ua - ub) * 0.5 * (faba / oma / rhoa + fabb / omb / rhob);
          qc = F->au * qc;
        }
        dua = dua + mb * (+dot3(vab, urab) * qa * faba / oma / rhoa + qc + 0.0);
      }
      if (F->hydro) {
        for (int l = 0; l < 3; l++)
          dva[l] = dva[l] - mb * ((pa * nwa[l]) / ((rhoa * rhoa) * oma) + (pb * nwb[l]) / ((rhob * rhob) * omb));
        double pnwa[3] = {pa * nwa[0], pa * nwa[1], pa * nwa[2]};
        dua = dua + mb * dot3(vab, pnwa) / (rhoa * rhoa) / oma;
      }
      if (F->diff) {
        if (F->fld) {
          if (F->diso == 1) {
            double fld_rb = sqrt(dot3(dtbdx, dtbdx)) / (kappab * rhob * rhob * tb);
            double lambdab = (2. + fld_rb) / (6. + 3. * fld_rb + fld_rb * fld_rb);
            double flddb = LIGHTSPEED * lambdab / kappab / rhob;
            memset(ktb, 0, sizeof(ktb));
            ktb[0][0] = ktb[1][1] = ktb[2][2] = flddb;
          } else {
            set_err(ORC_FLD_ANISO);
          }
        } else if (F->diso == 0) {
          cond_tensor(F, ubb, ktb);
        }
        if (F->ddw == ESD_FW || F->ddw == ESD_FAB || F->ddw == ESD_N2W) {
          double Hs[3][3], ktab[3][3];
          for (int a = 0; a < 3; a++)
            for (int b = 0; b < 3; b++) ktab[a][b] = (kta[a][b] + ktb[a][b]) / 2.;
          hessian(k, rab, ha, Hs);
          ddta = ddta + mb / rhob / rhoa * (tb * rhob - ta * rhoa) *
                            (dot3(ktab[0], Hs[0]) + dot3(ktab[1], Hs[1]) + dot3(ktab[2], Hs[2]));
        } else if (F->ddw == ESD_2NW_DS || F->ddw == ESD_2NW_SD) {
          double kda[3], kdb[3];
          for (int l = 0; l < 3; l++) { kda[l] = dot3(kta[l], dtadx); kdb[l] = dot3(ktb[l], dtbdx); }
          if (F->ddw == ESD_2NW_DS) {
            double odda = 1. / oma / rhoa / rhoa, oddb = 1. / omb / rhob / rhob;
            ddta = ddta + mb * (dot3(kda, nwa) * odda + dot3(kdb, nwb) * oddb);
          } else {
            double dd[3] = {kdb[0] - kda[0], kdb[1] - kda[1], kdb[2] - kda[2]};
            ddta = ddta + mb / oma / rhoa / rhoa * (dot3(dd, nwa));
          }
        } else {
          set_err(ORC_BAD_DDW);
        }
      }
      if (F->magneto) {
        double MPb[3][3];
        for (int a = 0; a < 3; a++)
          for (int b = 0; b < 3; b++)
            MPb[a][b] = (a == b) ? (0.5 * bb2 - bb[a] * bb[a]) / F->muzero : -bb[a] * bb[b] / F->muzero;
        double dvat[3], dvbt[3];
        for (int l = 0; l < 3; l++) { dvat[l] = dot3(MPa[l], nwa); dvbt[l] = dot3(MPb[l], nwb); }
        for (int l = 0; l < 3; l++)
          dva[l] = dva[l] - mb * (dvat[l] / ((rhoa * rhoa) * oma) + dvbt[l] / ((rhob * rhob) * omb)) - 0.0;
        double bnwa = dot3(ba, nwa), vnwa = dot3(vab, nwa);
        for (int l = 0; l < 3; l++)
          dba[l] = dba[l] - 1. / (rhoa * oma) * mb * (vab[l] * bnwa - bb[l] * vnwa);
      }
    }
    if (F->adden) dha = -dim * ha / rhoa * drhoadt / oma;
    for (int l = 0; l < 3; l++) ST(i, AX + l) = dva[l];
    ST(i, DH) = dha;
    ST(i, DDT) = ddta;
    for (int l = 0; l < 3; l++) ST(i, DBX + l) = dba[l];
    ST(i, DU) = dua;
    consenrg += ma * (dot3(va, dva) + dua + dot3(ba, dba) / rhoa - 0.5 * dot3(ba, ba) / rhoa / rhoa * drhoadt / oma);
  }
  *energy = consenrg;
  return g_err;
}

/* ---------------------------------------------------------------- iterate */
void orc_free_result(orc_result *res) {
  if (!res) return;
  free(res->nb_off); free(res->nb_j); free(res->nb_code);
  res->nb_off = NULL; res->nb_j = NULL; res->nb_code = NULL;
}

/* iterator.f90:44-179 */
int orc_iterate(const double *state, int kernel, const orc_ext *ext, int64_t ntot,
                double *store, orc_result *res) {
  g_err = 0;
  memset(res, 0, sizeof(*res));
  if (kernel < 0 || kernel > 2) return ORC_BAD_ARG;
  flags_t F;
  memset(&F, 0, sizeof(F));
  F.dim = (int)state[EC_DIM];
  if (F.dim < 1 || F.dim > 3) return ORC_BAD_ARG;
  F.ddw = (int)state[EC_DDW];
  F.adden = (int)state[EC_ADDEN];
  F.artts = (int)state[EC_ARTTS];
  F.diso = (int)state[EC_DISO];
  F.hfac = state[EC_HFAC];
  F.dcond = state[EC_DCOND];
  F.muzero = state[EC_MUZERO];
  F.au = state[EC_AU];
  F.gamma = state[EC_GAMMA];
  F.tensor_mode = ext ? ext->tensor_mode : 0;
  F.kappa_perp = ext ? ext->kappa_perp : 0.0;
  /* getEqComponent, state.f90:117-149 */
  switch ((int)state[EC_EQS]) {
    case 300: F.hydro = 1; break;
    case 302: F.hydro = F.magneto = 1; break;
    case 301: F.diff = 1; break;
    case 306: F.fld = F.hydro = 1; break;
    case 303: F.hydro = F.magneto = F.diff = 1; break;
    case 304: F.hydro = F.diff = 1; break;
    case 307:
      F.diff = (int)state[EC_EQONDIFF] == 1;
      F.hydro = (int)state[EC_EQONHYDRO] == 1;
      F.fld = (int)state[EC_EQONFLD] == 1;
      break;
    default: break;
  }
  int64_t nreal = (int64_t)state[EC_REALPN];
  if (nreal < 0 || nreal > ntot) return ORC_BAD_ARG;
  kctx k = {kernel, F.dim, F.ddw, orc_knorm(kernel, F.dim), fwc_of(kernel, F.dim)};
  double kr = krad_of(kernel);
  int cap = (ext && ext->neigh_cap > 0) ? ext->neigh_cap : maxneib_of(kernel, F.dim);
  int max_sweeps = (ext && ext->max_sweeps > 0) ? ext->max_sweeps : 100;
  double tol = (ext && ext->tol > 0) ? ext->tol : 1e-8;

  double emin[3] = {state[EC_XMIN], state[EC_XMIN + 2], state[EC_XMIN + 4]};
  double emax[3] = {state[EC_XMIN + 1], state[EC_XMIN + 3], state[EC_XMIN + 5]};
  int process = (int)state[EC_PROCESS];
  int ics = (int)state[EC_ICS];
  int wrap_axes = 0, ghost_axes = 0, do_eos = 1, iso = 0, gravity = 0;
  if (process == 601) { wrap_axes = ghost_axes = 7; }
  else if (process == 602) { }
  else if (process == 600) {
    switch (ics) {
      case 102: ghost_axes = 2; break;                           /* shock12 */
      case 103: case 104: case 111: do_eos = 0; break;           /* pulse, ring, fld_gauss */
      case 105: wrap_axes = ghost_axes = 7; iso = 1; break;      /* soundwave */
      case 106: wrap_axes = ghost_axes = 6; break;               /* hydroshock */
      case 101: wrap_axes = ghost_axes = 1; gravity = 1; break;  /* mti */
      case 110: wrap_axes = ghost_axes = 1; gravity = 2; break;  /* mtilowres */
      case 108: wrap_axes = ghost_axes = 7; break;               /* OTvortex */
      default: return ORC_BAD_ARG;
    }
  } else {
    return ORC_BAD_ARG;
  }

  ghosts_t g = {0, NULL, NULL};
  if (wrap_axes) wrap_real(store, nreal, F.dim, emin, emax, wrap_axes);
  if (ghost_axes)
    make_ghosts(store, ntot, F.dim, emin, emax, state[EC_BORDSIZE], state[EC_PADDING], ghost_axes, &g);
  res->n_ghost = g.ng;

  int64_t *off = NULL, *lst = NULL;
  int rc = neighbours(store, nreal, ntot, &g, kr, cap, &off, &lst);
  res->n_real = nreal;
  res->nb_off = off;
  res->n_pairs = off[nreal];
  res->nb_j = malloc(sizeof(int32_t) * (res->n_pairs > 0 ? res->n_pairs : 1));
  res->nb_code = malloc(res->n_pairs > 0 ? res->n_pairs : 1);
  for (int64_t t = 0; t < res->n_pairs; t++) {
    int64_t j = lst[t];
    int64_t rj = j < ntot ? j : g.cref[j - ntot];
    res->nb_j[t] = (int32_t)rj;
    int code = 13;
    if (j >= ntot) {
      const double *pg = &g.pos[3 * (j - ntot)];
      code = 0;
      int mul = 1;
      for (int a = 0; a < 3; a++) {
        int s = pg[a] > ST(rj, RX + a) ? 1 : (pg[a] < ST(rj, RX + a) ? -1 : 0);
        code += (s + 1) * mul;
        mul *= 3;
      }
    }
    res->nb_code[t] = (int8_t)code;
  }
  if (rc == ORC_OK) {
    if (F.adden == 1) rc = loop1_advanced(&k, &F, store, nreal, ntot, &g, off, lst, max_sweeps, tol, &res->sweeps);
    else rc = loop1_simple(&k, &F, store, nreal, ntot, &g, off, lst);
  }
  if (rc == ORC_OK && do_eos) eos(store, nreal, F.gamma, iso);
  if (rc == ORC_OK) rc = loop2(&k, &F, store, nreal, ntot, &g, off, lst, &res->energy);
  /* gravity of the mti set-ups, iterator.f90:143,153-154 */
  if (rc == ORC_OK && gravity == 1)
    for (int64_t i = 0; i < nreal; i++) ST(i, AY) = ST(i, AY) - 1.;
  if (rc == ORC_OK && gravity == 2)
    for (int64_t i = 0; i < nreal; i++) ST(i, AY) = ST(i, AY) - erf(ST(i, RY) * 8.);
  free(lst);
  free(g.pos);
  free(g.cref);
  return rc;
}

/* ------------------------------------------------------------------- STS */
/* sts_integrator.f90:10-51 */
static double sts_b(int j) { return j <= 2 ? 1. / 3. : (j * j + j - 2.) / (2. * j * (j + 1.)); }
static double sts_a(int j) { return 1 - sts_b(j); }
double orc_sts_mu(int j) { return (2. * j - 1.) / j * sts_b(j) / sts_b(j - 1); }
double orc_sts_nu(int j) { return -(j - 1.) / j * sts_b(j) / sts_b(j - 2); }
static double sts_w(int s) { return 4. / (s * s + s - 2.); }
double orc_sts_mu_hat(int j, int s) { return j == 1 ? sts_b(1) * sts_w(s) : orc_sts_mu(j) * sts_w(s); }
double orc_sts_gm_hat(int j, int s) { return -sts_a(j - 1) * orc_sts_mu_hat(j, s); }

/* --------------------------------------------------------- rad exchange */
/* update_radenergy(n, store, dt), rad_exchange.f90:9-66, with
 * solve_internal_energy_implicit (:100-118): per particle (real and fixed,
 * i = 1..n), Newton on
 *   f(u) = u/dt - u0/dt - du - ack (rho (etot - u)/a - (u cv1)^4)
 * from u = u0 while |u - u_old| > 1e-16 and at most 10 iterations; then
 * u <- u_new, t <- etot - u_new.  Returns ORC_RAD_NEGATIVE where the
 * reference stops with "radiation/thermal energy is negative after exchange". */
int orc_rad_exchange(const double *state, int64_t n, double *store, double dt) {
  const double gamma = state[EC_GAMMA], gmw = state[EC_MOLMASS];
  const double a = 4. * SIGMAB / LIGHTSPEED;
  const double cv1 = (gamma - 1.) * gmw / RG;
  int bad = 0;
  for (int64_t i = 0; i < n; i++) {
    const double kappa = ST(i, KAPPA), ack = 4. * SIGMAB * kappa;
    const double rho = ST(i, DEN), u0 = ST(i, U), dudt = ST(i, DU), xi = ST(i, T);
    const double etot = u0 + xi;
    double unew = u0, uold = 2 * u0;
    int iter = 0;
    while (fabs(unew - uold) > 1e-16 && iter < 10) {
      uold = unew;
      iter++;
      double fu = unew / dt - u0 / dt - dudt - ack * (rho * (etot - unew) / a - ipow(unew * cv1, 4));
      double dfu = 1. / dt + ack * (rho / a + 4. * (ipow(unew, 3) * ipow(cv1, 4)));
      unew = unew - fu / dfu;
    }
    ST(i, U) = unew;
    ST(i, T) = etot - unew;
    if (ST(i, T) <= 0 || ST(i, U) <= 0) bad = 1;
  }
  return bad ? ORC_RAD_NEGATIVE : ORC_OK;
}
