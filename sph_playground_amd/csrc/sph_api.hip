// Host side of libsph_hip.so: context, step sequencing and the C-ABI
// declared in include/sph_hip.h.
//
// sph_step follows iterate() (iterator.f90:44-179): for the process/ics chosen
// at context creation it wraps escaped particles, builds the neighbour lists
// (virtual periodic images instead of ghost copies), runs loop 1 (c1advanced
// Newton sweeps or c1simple), the EOS and loop 2.  All particle data stay in
// device memory between calls; the host only sizes the cell grid (one small
// read-back), watches the Newton convergence count and reads error bits.
#include <hip/hip_runtime.h>
#include <hipcub/hipcub.hpp>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <vector>

#include "../../include/sph_hip.h"
#include "sph_kernels.h"

using namespace sph;

namespace sph {
// kernels (sph_kernels.hip)
__global__ void k_wrap_reduce(Fields, const int64_t *, int, int64_t, Box, int, long long *, int *);
__global__ void k_keys(Fields, int, Grid, uint32_t *, uint32_t *, int *);
__global__ void k_check_sorted(const uint32_t *, int, int *);
__global__ void k_reorder(Fields, Fields, const uint32_t *, const int64_t *, int64_t *, int, int);
__global__ void k_cell_start(const uint32_t *, int, uint32_t, int *);
__global__ void k_prep(Fields, const int64_t *, int, int64_t, Box, Grid, int, double4 *, double *, float4 *,
                       uint8_t *, double4 *, double *, uint3 *, FixFrame, float2 *, double *, long long *,
                       long long *, int, long long *, int, int);
template <int DIM, bool EXT>
__global__ void k_neigh(NeighArgs, const double4 *__restrict__, const float4 *__restrict__, uint32_t *__restrict__,
                        int *__restrict__);
template <int DIM, int KT, typename PT, typename V4>
__global__ void k_loop2_fab(Loop2Args<V4>);
template <int DIM, int KT, int GRAD, typename PT>
__global__ void k_loop1(Loop1Args<double4>);
template <int DIM, int KT, int GRAD, typename PT>
__global__ void k_loop1_simple(Loop1Args<double4>, Fields);
template <typename V4>
__global__ void k_post1(Fields, PostSel, const uint8_t *, int, int, int, int, double, double, V4 *, V4 *,
                        V4 *, V4 *, V4 *, double4 *, double *, float2 *, long long *, long long *);
template <int DIM, int KT, typename PT, typename V4, int MODE>
__global__ void k_loop2(Loop2Args<V4>);
__global__ void k_aos_to_soa(const double *, int64_t, int, Fields, int64_t *);
__global__ void k_step_close(const double *, int, const int *, const long long *, long long *);
__global__ void k_soa_to_aos(Fields, const int64_t *, int, double *, int64_t);
__global__ void k_ext_count(Grid, ExtGrid, const int *, int, int *);
__global__ void k_ext_fill(Grid, ExtGrid, const int *, const int *, const float4 *, int, float4 *);
__global__ void k_neigh_export(const uint32_t *, const int *, const int64_t *, const uint8_t *, int, int,
                               const int64_t *, int32_t *, int8_t *, int64_t);
__global__ void k_set_i64(int64_t *, int64_t);
__global__ void k_zero_ll(long long *, int);
__global__ void k_prerun_check(Fields, int, int *);
__global__ void k_counts_by_orig(const int *, const int64_t *, const uint8_t *, int, int64_t *);
__global__ void k_find_orig(const int64_t *, int, int64_t, int *);
__global__ void k_predict(Fields, int, double, int, double *, int);
__global__ void k_correct(Fields, int, double, int, const double *, int);
__global__ void k_rad_exchange(Fields, int, double, double, double, double, int *);
__global__ void k_dt_reduce(Fields, const uint8_t *, int, long long *);
__global__ void k_gravity(Fields, const uint8_t *, int, int);
__global__ void k_sts_stage(const double *, const double *, const double *, const double *, const double *, int,
                            double, double, double, double, double *);
__global__ void k_axpy_copy(double *, const double *, const double *, double, int);
template <int DIM, int KT>
__global__ void k_diff_weights(DiffOpArgs, double *, double *);
template <int DIM, int KT>
__global__ void k_diff_rweights(DiffOpArgs, const double *, double *);
template <int DIM, int KT>
__global__ void k_diff_gcontract(DiffOpArgs, const double *, const double *, double *);
__global__ void k_diff_spmv(DiffOpArgs, const double *, const double *, double *);
template <int NS, int UPD>
__global__ void k_diff_adj_spmv(DiffOpArgs, const double *, const double *, LZSet, double *, double *);
__global__ void k_stage_adj(int, const double *, const double *, double, double, double, double, double *, double *,
                            double *);
__global__ void k_adj_lz(int, const double *, const double *, const double *, double, double, const double4 *,
                         double2 *);
__global__ void k_gk_accum(const double *, int, double, double, double *);
__global__ void k_zrho(int, const double *, const double4 *, double *);
__global__ void k_diff_pack(int, const double4 *, const double4 *, double4 *, double *);
__global__ void k_rev_count(const uint8_t *, const uint32_t *, const int *, int, int, int *);
__global__ void k_rev_chunk(const int *, const uint32_t *, int, int, int *, uint32_t *);
__global__ void k_pair_enum(const uint8_t *, const uint32_t *, const int *, int, int, const int *, uint32_t *,
                            uint32_t *);
__global__ void k_to_sorted(const double *, const int64_t *, int, double *);
__global__ void k_from_sorted(const double *, const int64_t *, int, double *);
__global__ void k_lincomb(int, double *, double, const double *, double, const double *, double, const double *,
                          double, const double *, double, const double *);
__global__ void k_batch_merge(int, int, int64_t, int, const int *, const uint32_t *, const double *, const double4 *,
                              const int64_t *, int, int *, uint32_t *, double *, float *, double *, int64_t *, int *);
template <bool FIRST, typename WT>
__global__ void k_batch_stage(BatchStage);
__global__ void k_batch_in(const double *, const int64_t *, const double *, int, double *, double *);
template <int DIM, int KT, int MODE>
__global__ void k_nl_grad(DiffOpArgs, const double *, double4 *, double4 *);
__global__ void k_fld_coef(DiffOpArgs, const double *, const double4 *, double *, double *, double *);
template <int DIM, int KT, int OP>
__global__ void k_nl_apply(DiffOpArgs, const double *, const double4 *, const double *, double *);
template <int DIM, int KT, int OP>
__global__ void k_nl_nu(DiffOpArgs, const double *, const double4 *, double4 *);
__global__ void k_nl_mu(DiffOpArgs, const double4 *, const double4 *, const double4 *, double4 *, double *);
template <int DIM, int KT, int MODE>
__global__ void k_nl_gT(DiffOpArgs, const double4 *, const double4 *, double *, int);
template <int DIM, int KT>
__global__ void k_fld_adj(DiffOpArgs, const double *, const double *, const double *, double *, double *);
__global__ void k_fld_pv(DiffOpArgs, const double *, const double *, const double *, const double4 *, const double *,
                         double *, double4 *);
}  // namespace sph

namespace {

constexpr int kBlock = 256;
// Per-step counters in d_stripes (long long slots, kStripes per value): the
// search's pair counts and list lengths, k_step_close's record (energy bits,
// loop 2's error bits, h_max), k_post1's h_max, the error words (ints),
// k_prep's v / B flags; loop 1's per-sweep regions follow at kStepLL.
// k_prep zeroes [0, kZeroLL) as it starts a step, so the step's first
// search, first sweep, k_post1 and loop 2 find their counters cleared
// without a fill of their own.
constexpr int kSearchLL = 0;
constexpr int kCloseRec = 2 * kStripes;
constexpr int kHmaxLL = 4 * kStripes;
constexpr int kErrLL = 5 * kStripes;      // ints: 0 search, 1 loop 1, 2 loop 2
constexpr int kZeroLL = kErrLL + 4;
constexpr int kFlagSlots = 6 * kStripes;
constexpr int kStepLL = 9 * kStripes;     // the counters before the sweep regions
enum { kErrSearch = 0, kErrLoop1 = 1, kErrLoop2 = 2 };
constexpr int kAux = 16;  // auxiliary per-particle arrays permuted with the fields

// statevars slots (0-based), const.f90:14-66
enum { S_DIM = 0, S_EQS = 1, S_DDW = 2, S_ICS = 3, S_ADDEN = 4, S_ARTTS = 5, S_COORD = 6, S_DISO = 7,
       S_DCOND = 8, S_MUZERO = 9, S_HFAC = 12, S_XMIN = 16, S_BORD = 22, S_PAD = 23, S_REALPN = 24,
       S_FIXEDPN = 25, S_GAMMA = 26, S_PROCESS = 30, S_AU = 31, S_EQONHYDRO = 33, S_EQONDIFF = 34,
       S_EQONFLD = 35, S_MOLMASS = 46 };

inline unsigned grid_for(int64_t n) { return (unsigned)std::max<int64_t>(1, (n + kBlock - 1) / kBlock); }

inline double ord2d(long long v) {
  long long b = v >= 0 ? v : (v ^ 0x7FFFFFFFFFFFFFFFLL);
  double d;
  memcpy(&d, &b, 8);
  return d;
}
inline long long d2ord_h(double d) {
  long long b;
  memcpy(&b, &d, 8);
  return b >= 0 ? b : (b ^ 0x7FFFFFFFFFFFFFFFLL);
}

}  // namespace

struct sph_ctx {
  int device = 0;
  hipStream_t stream = nullptr;
  int kt = 1;
  sph_options opt{};
  // SPH_NEIGH_TRAVERSAL=lane at creation: the search takes the per-lane
  // traversal and builds no extended grid (the lists are the same; tests
  // compare the two)
  bool lane_search = false;
  // SPH_LOOP2_GENERIC=1 at creation: loop 2 takes the kernel with every term
  // under its run-time flag, never a compile-time equation set (tests compare
  // the two)
  bool loop2_generic = false;
  // the search's extended grid (ghost cells of the imaged axes): cell counts,
  // cell starts, records; valid for the positions, order and grid it was
  // built from
  int *ext_count = nullptr, *ecs = nullptr;
  float4 *erec = nullptr;
  void *ext_tmp = nullptr;
  size_t ext_cells_cap = 0, ext_rec_cap = 0, ext_tmp_bytes = 0;
  ExtGrid ext_g{};
  uint64_t ext_version = 0;  // pos_version it was built from (0: none)
  int64_t ext_total = 0;
  int ext_cur = -1, ext_n = -1;
  Grid ext_grid{};
  Phys frozen{};
  Box box{};
  int process = 0, ics = 0, do_eos = 1, eos_iso = 0, gravity = 0;
  int maxneib = 150;
  // particles
  int64_t n = 0, n_real = 0, n_fixed = 0, cap = 0, npad = 0;
  double *fbuf[2] = {nullptr, nullptr};  // (32 + kAux) * npad doubles each
  int cur = 0;
  int naux = 0;
  int64_t *orig[2] = {nullptr, nullptr};
  uint32_t *keys = nullptr, *keys2 = nullptr, *vals = nullptr, *vals2 = nullptr;
  void *sort_tmp = nullptr;
  size_t sort_tmp_bytes = 0;
  int *cstart = nullptr;   // first sorted index of every cell key
  int64_t ncell_cap = 0;
  double4 *pm = nullptr, *st[2] = {nullptr, nullptr};
  double *tr[2] = {nullptr, nullptr};  // T den of the loop-1 sweep state st[k]
  // mixed path: position records (31-bit fractions, `ff` their frame),
  // (m, T den) of sweep state st[k] and (m/den, T den) for loop 2
  uint3 *pfix = nullptr;
  FixFrame ff{};
  float2 *s1[2] = {nullptr, nullptr};
  float2 *s2 = nullptr;
  double4 *pd = nullptr;                // loop 2, diffusion mode: x, y, z, m/den
  double *trd = nullptr;                // loop 2, diffusion mode: T den
  double *h0 = nullptr, *eps = nullptr;
  float4 *pf = nullptr;   // FP32 search records (relative coordinates + band bits)
  uint8_t *real = nullptr;
  uint32_t *nb = nullptr;
  int nb_cap = 0;
  int last_maxcnt = 0;  // longest neighbour list of the last search
  long long step_flags = 0;       // k_prep's v / B flags of this step
  bool loop2_pending = false;     // loop 2's read-back still to be taken
  bool neigh_pending = false;     // the search's read-back still to be taken (loop 1's first sync brings it)
  int *cnt = nullptr;
  void *q[5] = {nullptr, nullptr, nullptr, nullptr, nullptr};
  double *energy_part = nullptr;
  int64_t energy_cap = 0;
  // small device scratch: [0..7] long long reductions, then ints
  long long *d_red = nullptr;
  int *d_int = nullptr;  // 0 err, 1 maxcnt, 2 sorted flag, 3 pos, 8.. sweep counters (max_sweeps+2)
  long long *d_stripes = nullptr;         // striped counters (kStripes slots per value), sweep regions
  long long *h_rec = nullptr;             // pinned mirror of d_stripes and the sweep regions
  int *h_int = nullptr;                   // pinned mirror of d_int
  long long *h_ll = nullptr;              // pinned scratch for small host<->device values (2048)
  double *d_staging = nullptr;
  int64_t staging_bytes = 0;
  double4 *final_st = nullptr;
  Grid grid{};
  // cell keys kept from an earlier step: valid while the positions they were
  // computed from are unchanged (pos_version) and the grid is the same
  uint64_t pos_version = 1;       // bumped by every write of positions (load, leapfrog, wrap)
  bool keep_q0 = false;           // the step writes the FP64 q0 pack for sph_diff_prepare
  uint64_t vb_version = 1;        // bumped by every write of v or B (load, leapfrog)
  uint64_t flags_version = 0;     // vb_version step_flags were computed from
  bool flags_pending = false;     // this step's k_prep computes them (read with the search)
  // the bounding box and h_max the next grid needs, when nothing but the
  // library's own step changed the positions (bbox_version) or h (hmax_version)
  uint64_t h_version = 1;         // bumped by every write of h other than the step's own (load, leapfrog)
  uint64_t bbox_version = 0, hmax_version = 0;
  double last_lo[3] = {0, 0, 0}, last_hi[3] = {0, 0, 0}, next_hmax = 0.0;
  uint64_t keys_version = 0;      // pos_version the keys and cell starts belong to
  // the position records k_prep last wrote (pm, real, pf): for which
  // positions, buffer, count and grid
  uint64_t prep_version = 0;
  int prep_cur = -1;
  int64_t prep_n = -1;
  Grid prep_grid{};
  Grid keys_grid{};
  int64_t keys_n = -1, keys_nkeys = -1;
  // per-sweep counters of loop 1 (still-active stripes, pair stripes, a
  // ran flag), one region per sweep of a step, and their pinned mirror: the
  // parts of d_stripes and h_rec from kStepLL on
  long long *d_sweep = nullptr, *h_sweep = nullptr;
  int sweep_hint = 4;             // sweeps the next step launches before it synchronises
  int sweeps_zeroed = 0;          // k_prep cleared the regions of sweeps 1..this for the next batch
  sph_stats stats{};
  bool have_step = false;
  // the resident store holds the derivatives of the last successful iterate
  // of this state (nothing loaded since): a leapfrog pass then starts at its
  // predictor, as every pass after main.f90:219's initial iterate does
  bool derivs_ready = false;
  hipEvent_t ev[11]{};            // the stats' phase and kernel times (8 unused; 9, 10: a batch of sweeps)
  bool timing = false;            // record them (sph_set_timing)
  // differentiable diffusion (sph_diff_*)
  bool diff_ready = false;
  Phys diff_phys{};
  int *roff = nullptr, *rtmp = nullptr;
  uint32_t *rent = nullptr;      // reverse entries (owner, image), list layout, rcnt per particle
  double4 *dpxh = nullptr;       // x, y, z, 1 / h of the frozen geometry
  double *dirho = nullptr;       // 1 / rho of the frozen geometry
  double *dW = nullptr, *ddW = nullptr;  // forward pair weights (list layout) and their row sums
  double *dR = nullptr, *dS = nullptr;   // reverse weights and gradient coefficients (reverse order)
  double *dZ = nullptr;
  double2 *dLZ[4] = {nullptr, nullptr, nullptr, nullptr};  // (lam, Zt) of adjoint stages, a ring
  int lz_pend = 0;               // stages in the ring whose S update is deferred (buffers 0..lz_pend-1)
  bool fw_tail = false;          // the forward kept its last checkpoint segment's stage states
  int *rcnt = nullptr;           // reverse-list lengths
  int rcap = 0;                  // reverse-list capacity (list layout, multiple of 8)
  bool w_valid = false;          // dW/dR hold the weights of w_K
  double w_K[6] = {0, 0, 0, 0, 0, 0};
  double *gkdev = nullptr;       // 12 tensor-gradient accumulators
  void *scan_tmp = nullptr;
  size_t scan_tmp_bytes = 0;
  double *dvec = nullptr;        // work vectors, diff_nvec x npad
  int diff_nvec = 0;
  double *gpart = nullptr;       // 12 partial sums per block
  double *ckpt = nullptr;        // checkpoints of T, (nsteps / every + 1) x npad
  int64_t ckpt_cap = 0;
  int fw_s = 0, fw_steps = 0, fw_every = 0;
  double fw_dt = 0.0, fw_K[6] = {0, 0, 0, 0, 0, 0};
  // operators that read grad T (2nw-+ 3, 2nw+- 4, flux-limited diffusion 5;
  // 0 = the Hessian operators of the assembled path)
  int diff_op = 0;
  double4 *nlF = nullptr, *nlU = nullptr, *nlUB = nullptr, *nlNu = nullptr, *nlMu = nullptr;
  double *nlD = nullptr, *nlDD = nullptr, *nlR = nullptr, *nlSD = nullptr;

  Fields fields() const {
    Fields F;
    for (int f = 0; f < F_TOTAL; f++) F.f[f] = fbuf[cur] + (size_t)f * npad;
    return F;
  }
  double *aux(int k) const { return fbuf[cur] + (size_t)(F_TOTAL + k) * npad; }
};

#define HIPCHK(x)                                                                  \
  do {                                                                             \
    hipError_t e_ = (x);                                                           \
    if (e_ != hipSuccess) {                                                        \
      fprintf(stderr, "sph_hip: %s failed: %s\n", #x, hipGetErrorString(e_));      \
      return SPH_ERR_HIP;                                                          \
    }                                                                              \
  } while (0)

#define RCHK(x)             \
  do {                      \
    int r_ = (x);           \
    if (r_ != SPH_OK) return r_; \
  } while (0)

static int dev_alloc(void **p, size_t bytes) {
  if (*p) { (void)hipFree(*p); *p = nullptr; }
  if (bytes == 0) bytes = 16;
  if (hipMalloc(p, bytes) != hipSuccess) { *p = nullptr; return SPH_ERR_NOMEM; }
  return SPH_OK;
}

template <typename T>
static int dalloc(T **p, size_t count) { return dev_alloc((void **)p, count * sizeof(T)); }

// --------------------------------------------------------------- kernel base
static double knorm(int kt, int dim) {
  const double pi = kPi;
  if (kt == KT_M4) { const double k[3] = {2. / 3., 10. / (7. * pi), 1. / pi}; return k[dim - 1]; }
  if (kt == KT_M6) { const double k[3] = {1. / 120., 7. / (478. * pi), 1. / (120. * pi)}; return k[dim - 1]; }
  const double k[3] = {1. / sqrt(pi), 1. / pi, 1. / pow(pi, 1.5)};
  return k[dim - 1];
}
static double fwc_of(int kt, int dim) {
  if (kt == KT_M4) { const double f[3] = {6., 6.3226, 6.66400}; return f[dim - 1]; }
  if (kt == KT_M6) { const double f[3] = {4., 9.31505 / 2.25, 9.6429 / 2.25}; return f[dim - 1]; }
  return 4.0;
}
static int maxneib_of(int kt, int dim) {
  static const int t[3][3] = {{35, 250, 150}, {50, 10000, 150}, {100, 5000, 50000}};
  return t[kt][dim - 1];
}
// CFL numbers cnarr (base_m6.f90:30-33, base_m4.f90:127-130, base_gauss.f90:222-225)
static void cn_of(int kt, int ddw, double *cnhydro, double *cnd2) {
  static const double t[3][4] = {{0.1, 0.96, 0.15, 0.2}, {0.3, 1.1, 0.2, 0.4}, {0.55, 1.25, 0.24, 0.50}};
  *cnhydro = t[kt][0];
  if (ddw == 401) *cnd2 = t[kt][3];                        // n2w
  else if (ddw == 403 || ddw == 404) *cnd2 = t[kt][1];     // 2nw
  else *cnd2 = t[kt][2];                                   // fab, fw, default
}

// getEqComponent, state.f90:117-149
static void eq_components(const double *s, Phys *P) {
  P->hydro = P->magneto = P->diff = P->fld = 0;
  switch ((int)s[S_EQS]) {
    case 300: P->hydro = 1; break;
    case 302: P->hydro = P->magneto = 1; break;
    case 301: P->diff = 1; break;
    case 306: P->fld = P->hydro = 1; break;
    case 303: P->hydro = P->magneto = P->diff = 1; break;
    case 304: P->hydro = P->diff = 1; break;
    case 307:
      P->diff = (int)s[S_EQONDIFF] == 1;
      P->hydro = (int)s[S_EQONHYDRO] == 1;
      P->fld = (int)s[S_EQONFLD] == 1;
      break;
    default: break;
  }
}

// ------------------------------------------------------------------ buffers
static int ensure_particles(sph_ctx *c, int64_t n) {
  if (n <= c->cap) return SPH_OK;
  int64_t cap = std::max<int64_t>(n, 1024);
  int64_t npad = (cap + 63) / 64 * 64;
  const size_t nf = F_TOTAL + kAux;
  for (int b = 0; b < 2; b++) {
    RCHK(dalloc(&c->fbuf[b], nf * npad));
    RCHK(dalloc(&c->orig[b], npad));
    RCHK(dalloc(&c->st[b], npad));
    RCHK(dalloc(&c->tr[b], npad));
    if (c->opt.precision == 1) RCHK(dalloc(&c->s1[b], npad));
  }
  if (c->opt.precision == 1) {
    RCHK(dalloc(&c->pfix, npad));
    RCHK(dalloc(&c->s2, npad));
  }
  RCHK(dalloc(&c->pd, npad));
  RCHK(dalloc(&c->trd, npad));
  RCHK(dalloc(&c->keys, npad));
  RCHK(dalloc(&c->keys2, npad));
  RCHK(dalloc(&c->vals, npad));
  RCHK(dalloc(&c->vals2, npad));
  RCHK(dalloc(&c->pm, npad));
  RCHK(dalloc(&c->h0, npad));
  RCHK(dalloc(&c->eps, npad));
  RCHK(dalloc(&c->pf, npad + 4));  // + 4 records read past the last cell by k_neigh
  HIPCHK(hipMemset(c->pf, 0, sizeof(float4) * (npad + 4)));
  RCHK(dalloc(&c->real, npad));
  RCHK(dalloc(&c->cnt, npad));
  size_t v4 = c->opt.precision == 1 ? sizeof(float4) : sizeof(double4);
  for (int k = 0; k < 5; k++) RCHK(dev_alloc(&c->q[k], v4 * npad));
  c->energy_cap = (npad + kBlock - 1) / kBlock;
  RCHK(dalloc(&c->energy_part, c->energy_cap));
  // radix-sort scratch for n keys
  size_t bytes = 0;
  (void)hipcub::DeviceRadixSort::SortPairs(nullptr, bytes, c->keys, c->keys2, c->vals, c->vals2, (int)npad, 0, 32,
                                     c->stream);
  RCHK(dev_alloc(&c->sort_tmp, bytes));
  c->sort_tmp_bytes = bytes;
  c->cap = cap;
  c->npad = npad;
  c->nb_cap = 0;
  if (c->nb) { (void)hipFree(c->nb); c->nb = nullptr; }
  return SPH_OK;
}

static int ensure_nb(sph_ctx *c, int cap) {
  if (cap <= c->nb_cap && c->nb) return SPH_OK;
  cap = (std::max(cap, 16) + 7) / 8 * 8;  // whole chunks of 8 (nb_slot)
  RCHK(dalloc(&c->nb, (size_t)cap * c->npad));
  c->nb_cap = cap;
  return SPH_OK;
}

static int ensure_cells(sph_ctx *c, int64_t nkeys) {
  if (nkeys + 1 <= c->ncell_cap) return SPH_OK;
  int64_t cap = std::max<int64_t>(nkeys + 1, 1024);
  RCHK(dalloc(&c->cstart, cap));
  c->ncell_cap = cap;
  return SPH_OK;
}

// A HIP call whose failure is tolerated (timings of events a failed step never
// recorded, best-effort clean-up) still leaves the thread's last error set,
// and hipcub reports that error as its own after its next launch: clear it.
static void tolerate(hipError_t e) {
  if (e != hipSuccess) (void)hipGetLastError();
}

// The events around a step's phases and kernels, recorded only with timing
// on: every record leaves the stream idle for a few microseconds between the
// two kernels it separates (about 40 us over a 128^3 step's ten of them).
static hipError_t mark(sph_ctx *c, int k) {
  return c->timing ? hipEventRecord(c->ev[k], c->stream) : hipSuccess;
}

static float elapsed(sph_ctx *c, hipEvent_t a, hipEvent_t b) {
  if (!c->timing) return 0.f;
  float ms = 0.f;
  hipError_t e = hipEventSynchronize(b);
  if (e == hipSuccess) e = hipEventElapsedTime(&ms, a, b);
  if (e != hipSuccess) {
    tolerate(e);
    return 0.f;
  }
  return ms;
}

static int map_err(int bits) {
  if (bits & EB_Q) return SPH_ERR_Q_NAN_OR_NEG;
  if (bits & EB_NEG_H) return SPH_ERR_NEG_H;
  if (bits & EB_TRUNC) return SPH_ERR_NEIGH_TRUNCATED;
  if (bits & EB_OUTSIDE) return SPH_ERR_UNSUPPORTED;
  return SPH_OK;
}

// The short waits inside a step poll the stream until its work is done:
// the runtime's blocking wait adds its wake-up to every host round trip of
// the step, and an event recorded to wait on adds the processing of its own
// marker (a few microseconds after the last copy).
static hipError_t wait_stream(sph_ctx *c) {
  hipError_t r;
  while ((r = hipStreamQuery(c->stream)) == hipErrorNotReady) {
  }
  return r;
}

static int read_ints(sph_ctx *c, int first, int count) {
  HIPCHK(hipMemcpyAsync(c->h_int + first, c->d_int + first, sizeof(int) * count, hipMemcpyDeviceToHost, c->stream));
  HIPCHK(wait_stream(c));
  return SPH_OK;
}

// ------------------------------------------------------- step: neighbours
// Sum / max / or of the kStripes slots of value v in a pinned copy.
static long long stripe_sum(const long long *h, int v) {
  long long t = 0;
  for (int k = 0; k < kStripes; k++) t += h[v * kStripes + k];
  return t;
}
static long long stripe_max(const long long *h, int v) {
  long long t = h[v * kStripes];
  for (int k = 1; k < kStripes; k++) t = std::max(t, h[v * kStripes + k]);
  return t;
}
static long long stripe_min(const long long *h, int v) {
  long long t = h[v * kStripes];
  for (int k = 1; k < kStripes; k++) t = std::min(t, h[v * kStripes + k]);
  return t;
}

// Stencil half-width of the neighbour search in cells (cells rmax/RS wide).
#ifndef SPH_NEIGH_RS
#define SPH_NEIGH_RS 3
#endif
static const int kNeighRS = SPH_NEIGH_RS;

// d_int slot k_wrap_reduce sets when it moved a particle back into the box
constexpr int kWrapFlag = 4;

// The mixed path's position-record frame: an imaged axis maps its box length
// L onto 2^(31-k), k the least with 2^k L beyond any listed separation along
// it (the stencil's reach), so wrap-around differences of the fractions stay
// exact; other axes map the grid span onto 2^31.
static FixFrame fix_frame(const sph_ctx *c, const Grid &g, int dim, const int rc[3], double rmax) {
  FixFrame f{};
  for (int a = 0; a < dim; a++) {
    if (g.periodic[a]) {
      const double L = c->box.emax[a] - c->box.emin[a];
      const double reach = std::max(rc[a] * g.cs[a], rmax) + g.cs[a];
      int k = 0;
      while (k < 24 && std::ldexp(L, k) <= reach) k++;
      const double m = std::ldexp(1.0, 31 - k);
      f.base[a] = c->box.emin[a];
      f.mul[a] = m / L;
      f.scale[a] = (float)(L / m);
      f.step[a] = (uint32_t)m;
    } else {
      const double span = g.nc[a] * g.cs[a];
      const double m = std::ldexp(1.0, 31);
      f.base[a] = g.lo[a];
      f.mul[a] = span > 0.0 ? m / span : 0.0;
      f.scale[a] = (float)(span / m);
      f.step[a] = 0;
    }
  }
  return f;
}

static bool same_grid(const Grid &a, const Grid &b) {
  for (int d = 0; d < 3; d++)
    if (a.lo[d] != b.lo[d] || a.inv_cs[d] != b.inv_cs[d] || a.cs[d] != b.cs[d] || a.nc[d] != b.nc[d] ||
        a.periodic[d] != b.periodic[d])
      return false;
  return a.mdim == b.mdim && a.ty == b.ty;
}

// loop 1's per-sweep counter regions: [0, kStripes) still active,
// [kStripes, 2 kStripes) pairs, [2 kStripes] ran; region k for sweep k
constexpr int kSweepRegion = 2 * kStripes + 8;
constexpr int kSweepRegions = 242;  // sweeps 1..240 (the cap on max_sweeps) and slack
constexpr int kRecLL = kStepLL + kSweepRegions * kSweepRegion;  // d_stripes and h_rec size

// The step's read-back, one copy into the pinned mirror h_rec: d_stripes'
// counters (the search's pair counts and list lengths, k_prep's flags, the
// error words, k_step_close's record) and the regions of sweeps 1..last_sweep
// (none for 0).  The search's part can wait for a later synchronisation (the
// step's last), as loop 1 and loop 2 write only their own words of it.
static hipError_t copy_record(sph_ctx *c, int last_sweep) {
  const size_t ll = kStepLL + (last_sweep > 0 ? (size_t)(last_sweep + 1) * kSweepRegion : 0);
  return hipMemcpyAsync(c->h_rec, c->d_stripes, sizeof(long long) * ll, hipMemcpyDeviceToHost, c->stream);
}

// The Newton sweeps cap, and the last sweep of a step's first batch (the
// sweeps launched before its first synchronisation).
static int sweep_cap(const sph_ctx *c) { return c->opt.max_sweeps > 0 ? c->opt.max_sweeps : 100; }
static int first_batch_last(const sph_ctx *c) { return std::min(sweep_cap(c), std::max(1, c->sweep_hint)); }
constexpr int kWhichInt = 6;        // d_int slot: parity of the last sweep that ran
constexpr int kExtTotalInt = 16;    // h_int slot: records of the extended grid
constexpr int kRedoNeigh = 1000; // internal: the lists overflowed, search again (not an SPH_ERR code)

// the error words of the step (ints in d_stripes), and the same words in a
// pinned copy of d_stripes starting at slot `base`
static int *step_err(sph_ctx *c) { return reinterpret_cast<int *>(c->d_stripes + kErrLL); }
static int copied_err(const long long *copy, int base, int word) {
  return reinterpret_cast<const int *>(copy + (kErrLL - base))[word];
}

static int finish_neighbours(sph_ctx *c, const long long *ll, int bits) {
  if (c->flags_pending) {  // k_prep's v / B flags of this step
    c->step_flags = 0;
    for (int k = 0; k < kStripes; k++) c->step_flags |= ll[kFlagSlots + k];
    c->flags_version = c->vb_version;
    c->flags_pending = false;
  }
  c->stats.n_pairs = stripe_sum(ll, 0);
  int maxcnt = (int)stripe_max(ll, 1);
  c->last_maxcnt = maxcnt;
  c->stats.ms_k_neigh = elapsed(c, c->ev[4], c->ev[7]);
  if (bits & EB_TRUNC) return SPH_ERR_NEIGH_TRUNCATED;
  if (!(bits & EB_OVERFLOW)) return SPH_OK;
  RCHK(ensure_nb(c, maxcnt + 8));
  return kRedoNeigh;
}

// The deferred read-back after a synchronisation the caller made.
static int take_neighbours_now(sph_ctx *c) {
  const long long *ll = c->h_rec;
  return finish_neighbours(c, ll, copied_err(ll, 0, kErrSearch));
}

// A deferred read-back, once the stream has passed its copies.
static int take_neighbours(sph_ctx *c) {
  if (!c->neigh_pending) return SPH_OK;
  c->neigh_pending = false;
  const long long *ll = c->h_rec;
  return finish_neighbours(c, ll, copied_err(ll, 0, kErrSearch));
}

// defer: launch the search without waiting for its read-back (a later
// copy_record brings it); the lists then hold at most nb_cap entries per
// particle, so what runs on them before take_neighbours stays in bounds.
static int build_neighbours(sph_ctx *c, bool defer = false) {
  const int n = (int)c->n;
  const int dim = c->frozen.dim;
  Fields F = c->fields();
  unsigned G = grid_for(n);
  const long long BIG = 0x7FFFFFFFFFFFFFFFLL, SMALL = -0x7FFFFFFFFFFFFFFFLL - 1;
  // 1. wrap escaped real particles, bounding box, max h (striped reductions).
  //    Positions unchanged since the last pass (so none escaped) and h last
  //    set by the previous step (whose k_post1 reduced h_max): nothing to do.
  double lo[3], hi[3], hmax;
  if (c->bbox_version == c->pos_version && c->hmax_version == c->h_version && c->n_real > 0) {
    for (int a = 0; a < 3; a++) { lo[a] = c->last_lo[a]; hi[a] = c->last_hi[a]; }
    hmax = c->next_hmax;
  } else {
    long long *init = c->h_ll;
    for (int v = 0; v < 7; v++)
      for (int k = 0; k < kStripes; k++) init[v * kStripes + k] = v < 3 ? BIG : SMALL;
    HIPCHK(hipMemcpyAsync(c->d_red, init, sizeof(long long) * 7 * kStripes, hipMemcpyHostToDevice, c->stream));
    HIPCHK(hipMemsetAsync(c->d_int, 0, sizeof(int) * 8, c->stream));
    // at most 1024 blocks (four per CU), each striding over the particles
    hipLaunchKernelGGL(k_wrap_reduce, dim3(std::min(G, 1024u)), dim3(kBlock), 0, c->stream, F, c->orig[c->cur], n,
                       c->n_real, c->box, dim, c->d_red, c->d_int);
    long long *red = c->h_ll + 512;
    HIPCHK(hipMemcpyAsync(red, c->d_red, sizeof(long long) * 7 * kStripes, hipMemcpyDeviceToHost, c->stream));
    RCHK(read_ints(c, 0, kWrapFlag + 1));
    if (c->h_int[0] & EB_OUTSIDE) return SPH_ERR_UNSUPPORTED;
    if (c->h_int[kWrapFlag]) c->pos_version++;  // a particle was wrapped back into the box
    for (int a = 0; a < 3; a++) { lo[a] = ord2d(stripe_min(red, a)); hi[a] = ord2d(stripe_max(red, 3 + a)); }
    hmax = c->n_real > 0 ? ord2d(stripe_max(red, 6)) : 0.0;
    for (int a = 0; a < 3; a++) { c->last_lo[a] = lo[a]; c->last_hi[a] = hi[a]; }
    c->bbox_version = c->pos_version;
  }
  if (!(hmax > 0.0) && c->n_real > 0) return SPH_ERR_Q_NAN_OR_NEG;
  double rmax = c->frozen.krad * hmax * (1.0 + 1e-9);
  if (!(rmax > 0.0)) rmax = 1.0;
  // 2. cell grid: cells at least rmax/RS wide; imaged axes tile
  //    [emin, emax] and need 2 RS + 1 cells, else RS drops to 1.  The search
  //    reads each (y, z) row as whole ranges along x, so thin rows follow the
  //    sphere closely at a few more row ranges, and they order the particles
  //    so that adjacent lanes gather adjacent neighbours in the loops.  Cells
  //    of a third of the radius measured fastest for the whole step (a half
  //    and a quarter were slower in the loops and in the search).
  const int axis_cap = dim == 3 ? 1024 : (dim == 2 ? 32768 : (1 << 30));
  Grid g{};
  double cs[3];
  int RS = c->opt.cell_div > 0 ? c->opt.cell_div : kNeighRS;
  for (int attempt = 0; attempt < 2; attempt++) {
    g = Grid{};
    g.mdim = dim;
    const double cw = rmax / RS;
    bool ok = true;
    for (int a = 0; a < 3; a++) {
      g.periodic[a] = 0;
      if (a >= dim) { g.lo[a] = lo[a]; g.nc[a] = 1; cs[a] = 0; continue; }
      if ((c->box.ghost_mask >> a) & 1) {
        double L = c->box.emax[a] - c->box.emin[a];
        int64_t nc = (int64_t)std::floor(L / cw);
        nc = std::max<int64_t>(1, std::min<int64_t>(nc, axis_cap));
        g.nc[a] = (int)nc;
        g.lo[a] = c->box.emin[a];
        cs[a] = L / (double)nc;
        g.periodic[a] = 1;
        if (RS > 1 && nc < 2 * RS + 1) ok = false;
      } else {
        double ext = hi[a] - lo[a];
        cs[a] = cw;
        if (ext / cs[a] + 1 > axis_cap) cs[a] = ext / (axis_cap - 1) * (1.0 + 1e-9);
        g.nc[a] = (int)std::min<int64_t>((int64_t)std::floor(ext / cs[a]) + 1, axis_cap);
        g.lo[a] = lo[a];
      }
      if (RS > 1 && cs[a] * RS < rmax) ok = false;  // capped axis: cells too small for the stencil
    }
    if (ok || RS == 1) break;
    RS = 1;
  }
  // bound the cell count (sparse sets): coarsen the largest axis
  while ((int64_t)g.nc[0] * g.nc[1] * g.nc[2] > (int64_t)1 << 26) {
    int a = g.nc[0] >= g.nc[1] ? (g.nc[0] >= g.nc[2] ? 0 : 2) : (g.nc[1] >= g.nc[2] ? 1 : 2);
    if (g.periodic[a]) {
      g.nc[a] = std::max(1, g.nc[a] / 2);
      cs[a] = (c->box.emax[a] - c->box.emin[a]) / g.nc[a];
      if (RS > 1 && g.nc[a] < 2 * RS + 1) RS = 1;
    } else {
      cs[a] *= 2;
      g.nc[a] = (int)std::floor((hi[a] - lo[a]) / cs[a]) + 1;
    }
  }
  for (int a = 0; a < 3; a++) {
    g.cs[a] = cs[a];
    g.inv_cs[a] = cs[a] > 0 ? 1.0 / cs[a] : 0.0;
  }
  // sort keys: row (y, z) times nx plus x; 3D rows in tiles of (y, z)
  int64_t nrow = 1;
  if (dim == 3) {
    g.ty = (g.nc[1] + kRowTile - 1) / kRowTile;
    nrow = (int64_t)g.ty * ((g.nc[2] + kRowTile - 1) / kRowTile) * kRowTile * kRowTile;
  } else if (dim == 2) {
    nrow = g.nc[1];
  }
  const int64_t nkeys = nrow * g.nc[0];
  c->grid = g;
  RCHK(ensure_cells(c, nkeys));
  int key_bits = 1;
  while (key_bits < 32 && ((int64_t)1 << key_bits) < nkeys) key_bits++;
  // 3. cell keys, sortedness, radix sort + reorder when needed; keys and
  //    cell starts of the same positions on the same grid are still valid
  c->stats.sorted = 0;
  const bool keys_kept = c->keys_version == c->pos_version && c->keys_n == n && c->keys_nkeys == nkeys &&
                         same_grid(g, c->keys_grid);
  if (!keys_kept) {
    // k_keys clears the order check's flag (no fill of its own)
    hipLaunchKernelGGL(k_keys, dim3(G), dim3(kBlock), 0, c->stream, F, n, g, c->keys, c->vals, c->d_int + 2);
    hipLaunchKernelGGL(k_check_sorted, dim3(std::min(G, 256u)), dim3(kBlock), 0, c->stream, c->keys, n,
                       c->d_int + 2);
    RCHK(read_ints(c, 2, 1));
    if (c->h_int[2]) {
      size_t bytes = c->sort_tmp_bytes;
      HIPCHK(hipcub::DeviceRadixSort::SortPairs(c->sort_tmp, bytes, c->keys, c->keys2, c->vals, c->vals2, n, 0,
                                                key_bits, c->stream));
      int nxt = 1 - c->cur;
      Fields src = c->fields();
      Fields dst;
      for (int f = 0; f < F_TOTAL; f++) dst.f[f] = c->fbuf[nxt] + (size_t)f * c->npad;
      hipLaunchKernelGGL(k_reorder, dim3(G), dim3(kBlock), 0, c->stream, src, dst, c->vals2, c->orig[c->cur],
                         c->orig[nxt], n, (int)F_TOTAL);
      // auxiliary arrays ride along in groups of up to 32
      for (int k0 = 0; k0 < c->naux; k0 += F_TOTAL) {
        Fields as{}, ad{};
        const int m = std::min((int)F_TOTAL, c->naux - k0);
        for (int f = 0; f < m; f++) {
          as.f[f] = c->fbuf[c->cur] + (size_t)(F_TOTAL + k0 + f) * c->npad;
          ad.f[f] = c->fbuf[nxt] + (size_t)(F_TOTAL + k0 + f) * c->npad;
        }
        hipLaunchKernelGGL(k_reorder, dim3(G), dim3(kBlock), 0, c->stream, as, ad, c->vals2, c->orig[c->cur],
                           (int64_t *)nullptr, n, m);
      }
      c->cur = nxt;
      std::swap(c->keys, c->keys2);
      F = c->fields();
      c->stats.sorted = 1;
    }
    hipLaunchKernelGGL(k_cell_start, dim3(grid_for(n + 1)), dim3(kBlock), 0, c->stream, c->keys, n, (uint32_t)nkeys,
                       c->cstart);
    c->keys_version = c->pos_version;
    c->keys_grid = g;
    c->keys_n = n;
    c->keys_nkeys = nkeys;
  }
  // the v / B flags are computed again only when v or B may have changed
  c->flags_pending = c->flags_version != c->vb_version;
  if (c->flags_pending) HIPCHK(hipMemsetAsync(c->d_stripes + kFlagSlots, 0, sizeof(long long) * kStripes, c->stream));
  // rows and cells within r_max (1 + 1e-6) of a cell; imaged axes at most
  // one wrap away
  int rc[3];
  for (int d = 0; d < 3; d++) {
    rc[d] = d < dim && cs[d] > 0 ? (int)std::ceil(rmax * (1.0 + 1e-6) / cs[d]) : 0;
    if (g.periodic[d]) rc[d] = std::min(rc[d], g.nc[d]);
  }
  const FixFrame ff = fix_frame(c, g, dim, rc, rmax);
  const bool pos_kept = keys_kept && c->prep_version == c->pos_version && c->prep_cur == c->cur &&
                        c->prep_n == n && same_grid(g, c->prep_grid) &&
                        std::memcmp(&ff, &c->ff, sizeof(ff)) == 0;
  c->ff = ff;
  // the start-state copy only for c1simple (a deferred search is
  // c1advanced's, whose first sweep reads the fields); the counters cleared
  // with the regions of the first batch of Newton sweeps when they follow
  // this search (c1advanced)
  const int zlast = defer ? first_batch_last(c) : 0;
  hipLaunchKernelGGL(k_prep, dim3(G), dim3(kBlock), 0, c->stream, F, c->orig[c->cur], n, c->n_real, c->box, g, dim,
                     c->pm, c->h0, c->pf, c->real, defer ? nullptr : c->st[0], c->tr[0], c->pfix, ff, c->s1[0],
                     c->eps,
                     c->flags_pending ? c->d_stripes + kFlagSlots : nullptr, c->d_stripes, kZeroLL,
                     c->d_sweep + kSweepRegion, zlast * kSweepRegion, pos_kept ? 0 : 1);
  c->sweeps_zeroed = zlast;
  c->prep_version = c->pos_version;
  c->prep_cur = c->cur;
  c->prep_n = n;
  c->prep_grid = g;
  // 4. the extended grid: rc ghost cells on each side of every axis (the
  //    periodic images on an imaged axis, empty on a walled one), so every
  //    stencil row of the search is one range of records and no row or cell
  //    needs a bounds test; rebuilt when the positions, their order, the
  //    grid or the stencil change
  bool use_ext = !c->lane_search;
  ExtGrid xg{};
  int64_t ncells = 1;
  for (int d = 0; d < 3; d++) {
    xg.gh[d] = rc[d];
    xg.ne[d] = g.nc[d] + 2 * xg.gh[d];
    xg.lenf[d] = g.periodic[d] ? (float)(c->box.emax[d] - c->box.emin[d]) : 0.f;
    ncells *= xg.ne[d];
  }
  // offsets are ints (every record has at most 27 copies) and the search
  // addresses the cell starts with 32-bit byte offsets
  if (ncells + 1 > INT32_MAX / 4 || (int64_t)n * 27 > INT32_MAX) use_ext = false;
  if (use_ext) {
    const bool ext_kept = c->ext_version == c->pos_version && c->ext_cur == c->cur && c->ext_n == n &&
                          same_grid(g, c->ext_grid) && std::memcmp(&xg, &c->ext_g, sizeof(xg)) == 0;
    if (!ext_kept) {
      const int E = (int)ncells;
      if ((size_t)E + 1 > c->ext_cells_cap) {
        if (c->ext_count) { (void)hipFree(c->ext_count); c->ext_count = nullptr; }
        if (c->ecs) { (void)hipFree(c->ecs); c->ecs = nullptr; }
        c->ext_cells_cap = 0;
        RCHK(dev_alloc((void **)&c->ext_count, sizeof(int) * ((size_t)E + 1)));
        RCHK(dev_alloc((void **)&c->ecs, sizeof(int) * ((size_t)E + 1)));
        c->ext_cells_cap = (size_t)E + 1;
      }
      size_t tb = 0;
      HIPCHK(hipcub::DeviceScan::ExclusiveSum(nullptr, tb, c->ext_count, c->ecs, E + 1, c->stream));
      if (tb > c->ext_tmp_bytes) {
        if (c->ext_tmp) { (void)hipFree(c->ext_tmp); c->ext_tmp = nullptr; }
        c->ext_tmp_bytes = 0;
        RCHK(dev_alloc(&c->ext_tmp, tb));
        c->ext_tmp_bytes = tb;
      }
      hipLaunchKernelGGL(k_ext_count, dim3(grid_for(E + 1)), dim3(kBlock), 0, c->stream, g, xg, c->cstart, E,
                         c->ext_count);
      tb = c->ext_tmp_bytes;
      HIPCHK(hipcub::DeviceScan::ExclusiveSum(c->ext_tmp, tb, c->ext_count, c->ecs, E + 1, c->stream));
      HIPCHK(hipMemcpyAsync(c->h_int + kExtTotalInt, c->ecs + E, sizeof(int), hipMemcpyDeviceToHost, c->stream));
      HIPCHK(wait_stream(c));
      int64_t total = c->h_int[kExtTotalInt];
      // 8 records of padding: the search reads whole batches past a row's
      // end.  The search addresses records with 32-bit byte offsets, so a
      // larger set (over 2^28 records) takes the per-lane traversal.
      if ((total + 8) * (int64_t)sizeof(float4) > (int64_t)UINT32_MAX) {
        total = -1;
      } else {
        if ((size_t)total + 8 > c->ext_rec_cap) {
          if (c->erec) { (void)hipFree(c->erec); c->erec = nullptr; }
          c->ext_rec_cap = 0;
          const size_t cap = std::min<size_t>((size_t)total + 8 + (size_t)total / 8, UINT32_MAX / sizeof(float4));
          RCHK(dev_alloc((void **)&c->erec, sizeof(float4) * cap));
          c->ext_rec_cap = cap;
        }
        hipLaunchKernelGGL(k_ext_fill, dim3(grid_for(E)), dim3(kBlock), 0, c->stream, g, xg, c->cstart, c->ecs,
                           c->pf, E, c->erec);
      }
      c->ext_total = total;
      c->ext_version = c->pos_version;
      c->ext_cur = c->cur;
      c->ext_n = n;
      c->ext_grid = g;
      c->ext_g = xg;
    }
    if (c->ext_total < 0) use_ext = false;
  }
  c->stats.ext_records = use_ext ? c->ext_total : 0;
  // 5. neighbour lists; grow the storage and redo if any list overflowed
  if (c->nb_cap == 0) RCHK(ensure_nb(c, std::min(c->maxneib, dim == 3 ? 160 : (dim == 2 ? 48 : 16))));
  for (int attempt = 0; attempt < 2; attempt++) {
    if (attempt > 0) {  // k_prep cleared the first search's counters
      HIPCHK(hipMemsetAsync(c->d_stripes + kSearchLL, 0, sizeof(long long) * 2 * kStripes, c->stream));
      HIPCHK(hipMemsetAsync(step_err(c) + kErrSearch, 0, sizeof(int), c->stream));
    }
    NeighArgs a;
    a.hs = c->h0; a.real = c->real; a.cstart = c->cstart;
    a.rx = rc[0]; a.ry = rc[1]; a.rz = rc[2];
    a.ecs = use_ext ? c->ecs : nullptr;
    a.erec = use_ext ? c->erec : nullptr;
    for (int d = 0; d < 3; d++) { a.ne[d] = xg.ne[d]; a.gh[d] = xg.gh[d]; }
    // a position in cell units is FP32: its rounding stays below 2^-23 of the
    // cell count along the axis
    for (int d = 0; d < 3; d++) a.mcell[d] = (float)(1e-4 + 1e-6 * g.nc[d]);
    a.g = g; a.b = c->box; a.krad = c->frozen.krad; a.n = n; a.npad = (int)c->npad; a.cap = c->nb_cap;
    // FP32 search records: coordinates relative to the grid origin lie within
    // the grid span S (imaged axes: the box), images within 2 S; each of the
    // roundings involved is at most 2^-24 S, six of them at most, so
    // 2^-21 S bounds the error of an FP32 separation component.
    double span = 0.0;
    for (int d = 0; d < dim; d++) span = std::max(span, g.nc[d] * cs[d]);
    a.delta = std::ldexp(span, -21);
    for (int d = 0; d < 3; d++)  // only imaged axes shift (elsewhere the shift is 0, keep 0 * lenf finite)
      a.lenf[d] = g.periodic[d] ? (float)(c->box.emax[d] - c->box.emin[d]) : 0.f;
    a.maxneib = c->maxneib; a.err = step_err(c) + kErrSearch; a.stripes = c->d_stripes + kSearchLL;
    // the search's own events (loop 1 reuses 5 and 6 before a deferred read-back is taken)
    HIPCHK(mark(c, 4));
#define NEIGH(D)                                                                                              \
  if (use_ext) hipLaunchKernelGGL((k_neigh<D, true>), dim3(G), dim3(kBlock), 0, c->stream, a, c->pm, c->pf, c->nb, c->cnt); \
  else hipLaunchKernelGGL((k_neigh<D, false>), dim3(G), dim3(kBlock), 0, c->stream, a, c->pm, c->pf, c->nb, c->cnt)
    if (dim == 1) { NEIGH(1); }
    else if (dim == 2) { NEIGH(2); }
    else { NEIGH(3); }
#undef NEIGH
    HIPCHK(mark(c, 7));
    // pair counts, list lengths, error words and k_prep's flags: deferred,
    // they come back with the step's next record copy
    if (defer) {
      c->neigh_pending = true;
      return SPH_OK;
    }
    HIPCHK(copy_record(c, 0));
    HIPCHK(wait_stream(c));
    const long long *ll = c->h_rec;
    const int fin = finish_neighbours(c, ll, copied_err(ll, 0, kErrSearch));
    if (fin != kRedoNeigh) return fin;
  }
  return SPH_ERR_HIP;
}

// ------------------------------------------------------------ step: loops
template <int DIM, int KT>
static void launch_loop1(sph_ctx *c, Loop1Args<double4> &a, int grad, bool simple, Fields F) {
  unsigned G = grid_for(a.n);
  bool mixed = c->opt.precision == 1;
#define L1(GR, PT)                                                                                   \
  if (simple) hipLaunchKernelGGL((k_loop1_simple<DIM, KT, GR, PT>), dim3(G), dim3(kBlock), 0, c->stream, a, F); \
  else hipLaunchKernelGGL((k_loop1<DIM, KT, GR, PT>), dim3(G), dim3(kBlock), 0, c->stream, a)
  if (mixed) {
    if (grad == 0) { L1(0, float); } else if (grad == 1) { L1(1, float); } else { L1(2, float); }
  } else {
    if (grad == 0) { L1(0, double); } else if (grad == 1) { L1(1, double); } else { L1(2, double); }
  }
#undef L1
}

template <int DIM>
static void launch_loop1_k(sph_ctx *c, Loop1Args<double4> &a, int grad, bool simple, Fields F) {
  if (c->kt == KT_M4) launch_loop1<DIM, KT_M4>(c, a, grad, simple, F);
  else if (c->kt == KT_M6) launch_loop1<DIM, KT_M6>(c, a, grad, simple, F);
  else launch_loop1<DIM, KT_GAUSS>(c, a, grad, simple, F);
}

// c1simple (circuit1.f90:227-305): one pass, grad T only with diffusion and
// the 2nw operators; the search is checked first (it writes h, rho, om in
// place), so this waits for the pass.
static int run_loop1_simple(sph_ctx *c, const Phys &P) {
  const int n = (int)c->n;
  Fields F = c->fields();
  Loop1Args<double4> a{};
  a.pm = c->pm; a.h0 = c->h0; a.real = c->real; a.nb = c->nb; a.cnt = c->cnt; a.eps = c->eps;
  a.dtdx = F.f[F_DTDX]; a.dtdy = F.f[F_DTDY]; a.dtdz = F.f[F_DTDZ]; a.err = step_err(c) + kErrLoop1;
  a.b = c->box; a.p = P; a.n = n; a.npad = (int)c->npad;
  const int grad = (P.diff && P.ddw == 403) ? 1 : ((P.diff && P.ddw == 404) ? 2 : 0);
  a.cur = c->st[0];
  a.next = c->st[1];
  HIPCHK(mark(c, 5));
  if (P.dim == 1) launch_loop1_k<1>(c, a, grad, true, F);
  else if (P.dim == 2) launch_loop1_k<2>(c, a, grad, true, F);
  else launch_loop1_k<3>(c, a, grad, true, F);
  HIPCHK(mark(c, 6));
  c->final_st = c->st[0];
  c->stats.sweeps = 0;
  HIPCHK(hipMemcpyAsync(c->h_ll, c->d_stripes + kErrLL, sizeof(long long) * 4, hipMemcpyDeviceToHost, c->stream));
  HIPCHK(wait_stream(c));
  RCHK(take_neighbours(c));
  c->stats.ms_k_loop1 = elapsed(c, c->ev[5], c->ev[6]);
  c->stats.loop1_pairs = c->stats.n_pairs;
  return map_err(copied_err(c->h_ll, kErrLL, kErrLoop1));
}

// Sweeps first..last of c1advanced (circuit1.f90:104-211) on the stream with
// no host round trip: each sweep runs only if the one before left a particle
// active (while maxval(eps) > tol, circuit1.f90:114) and no error stopped the
// search or loop 1 (k_loop1 decides on the device); their counters go to
// the sweep regions, zeroed here first (by k_prep for a step's first batch).
static void launch_sweeps(sph_ctx *c, const Phys &P, int first, int last) {
  const int n = (int)c->n;
  Fields F = c->fields();
  Loop1Args<double4> a{};
  a.pm = c->pm; a.h0 = c->h0; a.real = c->real; a.nb = c->nb; a.cnt = c->cnt; a.eps = c->eps;
  a.dtdx = F.f[F_DTDX]; a.dtdy = F.f[F_DTDY]; a.dtdz = F.f[F_DTDZ]; a.err = step_err(c) + kErrLoop1;
  a.b = c->box; a.p = P; a.n = n; a.npad = (int)c->npad;
  a.tol = c->opt.newton_tol > 0 ? c->opt.newton_tol : (c->opt.precision == 1 ? 1e-6 : 1e-8);
  a.serr = step_err(c) + kErrSearch;
  a.which = c->d_int + kWhichInt;
  const int grad = (P.diff || P.fld) ? (P.ddw == 404 ? 2 : 1) : 0;
  const int nreg = (last - first + 1) * kSweepRegion;
  if (!(first == 1 && last <= c->sweeps_zeroed))
    hipLaunchKernelGGL(k_zero_ll, dim3((nreg + 255) / 256), dim3(256), 0, c->stream,
                       c->d_sweep + (size_t)first * kSweepRegion, nreg);
  c->sweeps_zeroed = 0;
  tolerate(mark(c, 9));
  for (int it = first; it <= last; it++) {
    a.cur = c->st[(it - 1) & 1];
    a.next = c->st[it & 1];
    // the first sweep reads its start state from the fields (unchanged until
    // the tail's k_post1, which a redone search leaves idle)
    const bool s1st = it == 1;
    a.f_h = s1st ? F.f[F_H] : nullptr;
    a.f_den = s1st ? F.f[F_DEN] : nullptr;
    a.f_om = s1st ? F.f[F_OM] : nullptr;
    a.f_t = s1st ? F.f[F_T] : nullptr;
    a.trc = c->tr[(it - 1) & 1];
    a.trn = c->tr[it & 1];
    a.s1c = c->s1[(it - 1) & 1];
    a.s1n = c->s1[it & 1];
    a.pfix = c->pfix;
    a.ff = c->ff;
    a.stripes = c->d_sweep + (size_t)it * kSweepRegion;
    a.prev = it > 1 ? c->d_sweep + (size_t)(it - 1) * kSweepRegion : nullptr;
    a.sweep = it;
    if (P.dim == 1) launch_loop1_k<1>(c, a, grad, false, F);
    else if (P.dim == 2) launch_loop1_k<2>(c, a, grad, false, F);
    else launch_loop1_k<3>(c, a, grad, false, F);
  }
  tolerate(mark(c, 10));
}

template <int DIM, int KT, int MODE>
static void launch_loop2(sph_ctx *c, void *args) {
  unsigned G = grid_for(c->n);
  if constexpr (MODE == MODE_FAB) {
    if (c->opt.precision == 1)
      hipLaunchKernelGGL((k_loop2_fab<DIM, KT, float, float4>), dim3(G), dim3(kBlock), 0, c->stream,
                         *(Loop2Args<float4> *)args);
    else
      hipLaunchKernelGGL((k_loop2_fab<DIM, KT, double, double4>), dim3(G), dim3(kBlock), 0, c->stream,
                         *(Loop2Args<double4> *)args);
  } else if (c->opt.precision == 1) {
    hipLaunchKernelGGL((k_loop2<DIM, KT, float, float4, MODE>), dim3(G), dim3(kBlock), 0, c->stream,
                       *(Loop2Args<float4> *)args);
  } else {
    hipLaunchKernelGGL((k_loop2<DIM, KT, double, double4, MODE>), dim3(G), dim3(kBlock), 0, c->stream,
                       *(Loop2Args<double4> *)args);
  }
}
template <int DIM, int MODE>
static void launch_loop2_k(sph_ctx *c, void *args) {
  if (c->kt == KT_M4) launch_loop2<DIM, KT_M4, MODE>(c, args);
  else if (c->kt == KT_M6) launch_loop2<DIM, KT_M6, MODE>(c, args);
  else launch_loop2<DIM, KT_GAUSS, MODE>(c, args);
}
template <int MODE>
static void launch_loop2_d(sph_ctx *c, int dim, void *args) {
  if (dim == 1) launch_loop2_k<1, MODE>(c, args);
  else if (dim == 2) launch_loop2_k<2, MODE>(c, args);
  else launch_loop2_k<3, MODE>(c, args);
}
// modes built in 2D and 3D only
template <int MODE>
static void launch_loop2_d23(sph_ctx *c, int dim, void *args) {
  if (dim == 2) launch_loop2_k<2, MODE>(c, args);
  else launch_loop2_k<3, MODE>(c, args);
}

template <typename V4>
static int run_post_and_loop2(sph_ctx *c, const Phys &P) {
  const int n = (int)c->n;
  Fields F = c->fields();
  unsigned G = grid_for(n);
  // reference c1 for isothermal EOS reads particle 1's sound speed (eos.f90:43)
  double c1 = 0.0;
  if (c->do_eos && c->eos_iso && n > 0) {
    HIPCHK(hipMemsetAsync(c->d_int + 3, 0, sizeof(int), c->stream));
    hipLaunchKernelGGL(k_find_orig, dim3(G), dim3(kBlock), 0, c->stream, c->orig[c->cur], n, (int64_t)0,
                       c->d_int + 3);
    RCHK(read_ints(c, 3, 1));
    HIPCHK(hipMemcpyAsync(c->h_ll + 4, F.f[F_C] + c->h_int[3], sizeof(double), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    memcpy(&c1, c->h_ll + 4, sizeof(double));
  }
  // which j-side packs loop 2 may gather (must match k_loop2)
  const bool generic = !(P.diff && !P.hydro && !P.artts && !P.magneto && !P.fld &&
                         (P.ddw == 400 || P.ddw == 401 || P.ddw == 402));
  const bool want_v = P.adden || P.artts || P.hydro || P.magneto;
  const bool want_q2 = generic && ((P.diff && P.fld) || P.artts);
  const bool want_q3 = generic && P.diff && (P.fld || P.ddw == 403 || P.ddw == 404);
  const bool want_q4 = P.magneto || (P.diff && !P.fld && P.diso == 0);
  // k_prep's flags (v and B are what they were there): the diffusion-only
  // loop 2 reads v only where it is non-zero and B only where it varies
  const long long fl = c->step_flags;
  // q0 feeds the generic loop, and in FP64 also the adjoint and batched
  // operators sph_diff_prepare builds after its step; the diffusion loops
  // read pd alone
  V4 *q0 = !generic && (std::is_same_v<V4, float4> || !c->keep_q0) ? nullptr : (V4 *)c->q[0];
  V4 *q1 = want_v && (generic || (fl & 1)) ? (V4 *)c->q[1] : nullptr;
  V4 *q2 = want_q2 ? (V4 *)c->q[2] : nullptr;
  V4 *q3 = want_q3 ? (V4 *)c->q[3] : nullptr;
  V4 *q4 = want_q4 && (generic || (fl & 2)) ? (V4 *)c->q[4] : nullptr;
  long long *none = nullptr;
  PostSel sel{};
  sel.fin0 = P.adden ? c->st[0] : c->final_st;
  sel.fin1 = c->st[1];
  sel.which = P.adden ? c->d_int + kWhichInt : nullptr;
  sel.serr = step_err(c) + kErrSearch;
  sel.err1 = step_err(c) + kErrLoop1;
  hipLaunchKernelGGL(k_post1<V4>, dim3(G), dim3(kBlock), 0, c->stream, F, sel, c->real, n, P.adden ? 1 : 0,
                     c->do_eos, c->eos_iso, c1, P.gamma, q0, q1, q2, q3, q4,
                     generic || std::is_same_v<V4, float4> ? nullptr : c->pd,
                     generic ? nullptr : c->trd, generic || !std::is_same_v<V4, float4> ? nullptr : c->s2, none,
                     c->d_stripes + kHmaxLL);
  if (P.diff && P.fld && P.diso != 1) return SPH_ERR_FLD_ANISO;
  if (P.diff && !(P.ddw >= 400 && P.ddw <= 404)) return SPH_ERR_BAD_DDW;
  Loop2Args<V4> a{};
  a.pm = c->pm; a.real = c->real; a.nb = c->nb; a.cnt = c->cnt;
  a.q0 = q0; a.q1 = q1; a.q2 = q2; a.q3 = q3; a.q4 = q4;
  a.pd = generic ? nullptr : c->pd;
  a.trd = generic ? nullptr : c->trd;
  a.pfix = c->pfix;
  a.s2 = c->s2;
  a.ff = c->ff;
  a.v_nonzero = (fl & 1) ? 1 : 0;
  a.b_uniform = (fl & 2) ? 0 : 1;
  // loop 2's error word and k_post1's h_max stripes: cleared by this step's k_prep
  a.out = F; a.energy_part = c->energy_part; a.err = step_err(c) + kErrLoop2; a.b = c->box; a.p = P; a.n = n;
  a.npad = (int)c->npad;
  HIPCHK(mark(c, 5));
  // fab with one tensor for every pair and no velocity term: the loop
  // without run-time choices
  const bool fab = !generic && P.ddw == 400 && q1 == nullptr && q4 == nullptr;
  // the generic loop with its equation set fixed at compile time where one
  // is built for it (2D and 3D: hydro, magnetohydro, hydro with FLD)
  const int terms =
      c->loop2_generic ? -1 : (P.hydro ? 1 : 0) | (P.magneto ? 2 : 0) | (P.diff ? 4 : 0) | (P.fld ? 8 : 0);
  if (generic && P.dim >= 2 && terms == kSpecHydro) launch_loop2_d23<MODE_SPEC + kSpecHydro>(c, P.dim, &a);
  else if (generic && P.dim >= 2 && terms == kSpecMhd) launch_loop2_d23<MODE_SPEC + kSpecMhd>(c, P.dim, &a);
  else if (generic && P.dim >= 2 && terms == kSpecHydroFld)
    launch_loop2_d23<MODE_SPEC + kSpecHydroFld>(c, P.dim, &a);
  else if (generic) launch_loop2_d<MODE_GENERIC>(c, P.dim, &a);
  else if (fab) launch_loop2_d<MODE_FAB>(c, P.dim, &a);
  else launch_loop2_d<MODE_DIFF>(c, P.dim, &a);
  HIPCHK(mark(c, 6));
  // energy, error bits and h_max in one record of d_stripes, which the
  // caller's copy_record brings back with the step's closing sync
  hipLaunchKernelGGL(k_step_close, dim3(1), dim3(256), 0, c->stream, c->energy_part, (int)G,
                     step_err(c) + kErrLoop2, c->d_stripes + kHmaxLL, c->d_stripes + kCloseRec);
  c->loop2_pending = true;
  return SPH_OK;
}

// After the step's closing copy_record and sync: loop 2's energy sum (block
// partials summed in block order) and its error bits.
static int finish_loop2(sph_ctx *c) {
  if (!c->loop2_pending) return SPH_OK;
  c->loop2_pending = false;
  double e;
  std::memcpy(&e, c->h_rec + kCloseRec, sizeof e);
  c->stats.energy = e;
  // h_max after the step's own h update: the next step's grid needs no
  // bounding-box pass while positions and h change no other way
  const double hm = ord2d(c->h_rec[kCloseRec + 2]);
  if (hm > 0.0 && std::isfinite(hm)) {
    c->next_hmax = hm;
    c->hmax_version = c->h_version;
  }
  c->stats.ms_k_loop2 = elapsed(c, c->ev[5], c->ev[6]);
  return map_err((int)c->h_rec[kCloseRec + 1]);
}

static Phys call_phys(const sph_ctx *c, const double *state) {
  Phys P = c->frozen;
  // read at every call by the reference (circuit1.f90:96, circuit2.f90:86-90)
  P.hfac = state[S_HFAC];
  P.dcond = state[S_DCOND];
  P.diso = (int)state[S_DISO];
  P.muzero = state[S_MUZERO];
  P.gamma = state[S_GAMMA];
  return P;
}

// The tail of a step after loop 1: Newton bookkeeping, EOS and packs
// (k_post1), loop 2, mti gravity and the closing record's copy.
static int enqueue_tail(sph_ctx *c, const Phys &P) {
  HIPCHK(mark(c, 2));
  int rc = c->opt.precision == 1 ? run_post_and_loop2<float4>(c, P) : run_post_and_loop2<double4>(c, P);
  if (rc == SPH_OK && c->gravity)
    hipLaunchKernelGGL(k_gravity, dim3(grid_for(c->n)), dim3(kBlock), 0, c->stream, c->fields(), c->real, (int)c->n,
                       c->gravity);
  HIPCHK(mark(c, 3));
  return rc;
}

// Whether loop 2's packs or terms depend on k_prep's v / B flags (the
// "every velocity 0" and "one B for every particle" bits): the diffusion-only
// loops skip v and B by them; the general loop reads v for every pair unless
// no term needs it always (hydro, MHD, artificial conductivity), and B by
// them only for a diffusion tensor from B (k_loop2).  Otherwise a step need
// not wait for the flags before its tail: they come back at its end.
static bool loop2_reads_flags(const Phys &P) {
  const bool generic = !(P.diff && !P.hydro && !P.artts && !P.magneto && !P.fld &&
                         (P.ddw == 400 || P.ddw == 401 || P.ddw == 402));
  if (!generic) return true;
  const bool v_always = P.hydro || P.magneto || P.artts;
  const bool tensor_b = P.diff && !P.fld && P.diso == 0;
  return !v_always || tensor_b;
}

// c1advanced steps: the search (its read-back left pending), a batch of
// sweeps, the tail, then one synchronisation.  Only when k_prep's v / B
// flags must come back before the tail picks loop 2's packs (after a load or
// a leapfrog change of v or B, for a loop 2 that reads them) does the step
// wait after the sweeps.  Lists
// that overflowed leave loop 1 and k_post1 idle (device-side checks), so the
// step searches again with room and starts over; a batch that ends with
// particles still active is followed by more sweeps and the tail again.
static int run_advanced(sph_ctx *c, const Phys &P) {
  const int maxsw = sweep_cap(c);
  int first = 1, sweeps = 0, bits = 0;
  long long pairs = 0;
  float ms1 = 0.f;
  if (c->n_real <= 0) {  // no sweep, no active particle (circuit1.f90:114)
    if (c->neigh_pending) HIPCHK(copy_record(c, 0));
    HIPCHK(hipStreamSynchronize(c->stream));
    const int fin = take_neighbours(c);
    if (fin == kRedoNeigh) RCHK(build_neighbours(c));
    else RCHK(fin);
    c->stats.sweeps = 0;
    c->stats.loop1_pairs = 0;
    const int rc = enqueue_tail(c, P);
    HIPCHK(copy_record(c, 0));
    HIPCHK(wait_stream(c));
    const int rc2 = finish_loop2(c);
    return rc != SPH_OK ? rc : rc2;
  }
  for (;;) {
    const int last = std::min(maxsw, first + std::max(1, c->sweep_hint) - 1);
    launch_sweeps(c, P, first, last);
    if (c->neigh_pending && c->flags_pending && loop2_reads_flags(P)) {
      HIPCHK(copy_record(c, 0));
      HIPCHK(wait_stream(c));
      c->neigh_pending = false;
      const int fin = take_neighbours_now(c);
      if (fin == kRedoNeigh) {
        RCHK(build_neighbours(c));
        first = 1; sweeps = 0; pairs = 0; ms1 = 0.f;
        continue;
      }
      RCHK(fin);
    }
    const int rc = enqueue_tail(c, P);
    // the search's read-back if still pending, error words, the tail's record
    // and the sweep regions: one copy
    HIPCHK(copy_record(c, last));
    HIPCHK(wait_stream(c));
    ms1 += elapsed(c, c->ev[9], c->ev[10]);
    if (rc != SPH_OK) return rc;
    if (c->neigh_pending) {
      c->neigh_pending = false;
      const int fin = take_neighbours_now(c);
      if (fin == kRedoNeigh) {
        RCHK(build_neighbours(c));
        first = 1; sweeps = 0; pairs = 0; ms1 = 0.f;
        continue;
      }
      RCHK(fin);
    }
    for (int it = first; it <= last; it++) {
      const long long *r = c->h_sweep + (size_t)it * kSweepRegion;
      if (!r[2 * kStripes]) break;
      sweeps = it;
      pairs += stripe_sum(r, 1);
    }
    bits = copied_err(c->h_rec, 0, kErrLoop1);
    if (bits) break;  // circuit1.f90:202-204 stops inside the sweep
    if (sweeps == last && last < maxsw && stripe_sum(c->h_sweep + (size_t)last * kSweepRegion, 0) > 0) {
      first = last + 1;  // still active: more sweeps, then the tail again
      continue;
    }
    break;
  }
  c->stats.sweeps = sweeps;
  c->stats.loop1_pairs = pairs;
  c->stats.ms_k_loop1 = ms1;
  // one sweep more than this step took: a sweep past convergence exits at
  // once (its launch costs about 16 us at 2M particles), while a step that
  // needs more sweeps than launched runs its queued tail (EOS, packs,
  // loop 2) again after them
  c->sweep_hint = std::min(maxsw, std::max(2, sweeps + 1));
  const int rc2 = finish_loop2(c);
  const int r1 = map_err(bits);
  return r1 != SPH_OK ? r1 : rc2;
}

static int run_step(sph_ctx *c, const double *state) {
  if (c->n <= 0) return SPH_ERR_BAD_ARG;
  // a new step changes the geometry the differentiable operator was prepared on
  c->diff_ready = false;
  c->w_valid = false;
  c->fw_steps = 0;
  c->fw_tail = false;
  c->derivs_ready = false;
  Phys P = call_phys(c, state);
  c->stats.ms_k_neigh = c->stats.ms_k_loop1 = c->stats.ms_k_loop2 = 0.0;
  HIPCHK(mark(c, 0));
  // the search's read-back is taken with the step's closing synchronisation
  // (c1advanced); c1simple writes h, rho and om in place, so its search is
  // checked before it runs
  RCHK(build_neighbours(c, P.adden != 0));
  c->hmax_version = 0;  // h changes from here; this step's k_post1 reports the new h_max
  HIPCHK(mark(c, 1));
  int rc;
  if (P.adden) {
    rc = run_advanced(c, P);
  } else {
    int r1 = run_loop1_simple(c, P);
    c->neigh_pending = false;
    RCHK(r1);
    rc = enqueue_tail(c, P);
    HIPCHK(copy_record(c, 0));
    HIPCHK(wait_stream(c));
    const int rc2 = finish_loop2(c);
    if (rc == SPH_OK) rc = rc2;
  }
  c->neigh_pending = false;
  float t01 = 0, t12 = 0, t23 = 0;
  if (c->timing) {
    tolerate(hipEventElapsedTime(&t01, c->ev[0], c->ev[1]));
    tolerate(hipEventElapsedTime(&t12, c->ev[1], c->ev[2]));
    tolerate(hipEventElapsedTime(&t23, c->ev[2], c->ev[3]));
  }
  c->stats.ms_neigh = t01;
  c->stats.ms_loop1 = t12;
  c->stats.ms_loop2 = t23;
  c->stats.ms_total = t01 + t12 + t23;
  c->have_step = true;
  c->derivs_ready = rc == SPH_OK;
  return rc;
}

// ----------------------------------------------------------------- loading
static int load_impl(sph_ctx *c, int64_t n, const double *store, int64_t ld, int on_device) {
  if (n <= 0 || ld < F_TOTAL || !store) return SPH_ERR_BAD_ARG;
  if (n > ((int64_t)1 << 27) - 1) return SPH_ERR_UNSUPPORTED;
  if (n < c->n_real) return SPH_ERR_BAD_ARG;
  RCHK(ensure_particles(c, n));
  c->n = n;
  c->n_fixed = n - c->n_real;
  const double *src = store;
  int64_t bytes = n * ld * (int64_t)sizeof(double);
  if (!on_device) {
    if (bytes > c->staging_bytes) {
      RCHK(dev_alloc((void **)&c->d_staging, bytes));
      c->staging_bytes = bytes;
    }
    HIPCHK(hipMemcpyAsync(c->d_staging, store, bytes, hipMemcpyHostToDevice, c->stream));
    src = c->d_staging;
  }
  c->cur = 0;
  c->pos_version++;
  c->vb_version++;
  c->h_version++;
  // the particles (and their order) are new: lists, the prepared operator,
  // its weights and the recorded forward belong to the old ones
  c->have_step = false;
  c->derivs_ready = false;
  c->diff_ready = false;
  c->w_valid = false;
  c->fw_steps = 0;
  c->fw_tail = false;
  hipLaunchKernelGGL(k_aos_to_soa, dim3(grid_for(n)), dim3(kBlock), 0, c->stream, src, ld, (int)n, c->fields(),
                     c->orig[0]);
  HIPCHK(hipStreamSynchronize(c->stream));
  return SPH_OK;
}

static int save_impl(sph_ctx *c, double *store, int64_t ld, int on_device) {
  if (!store || ld < F_TOTAL || c->n <= 0) return SPH_ERR_BAD_ARG;
  int64_t bytes = c->n * ld * (int64_t)sizeof(double);
  double *dst = store;
  if (!on_device) {
    if (bytes > c->staging_bytes) {
      RCHK(dev_alloc((void **)&c->d_staging, bytes));
      c->staging_bytes = bytes;
    }
    dst = c->d_staging;
    if (ld > F_TOTAL) HIPCHK(hipMemcpyAsync(dst, store, bytes, hipMemcpyHostToDevice, c->stream));
  }
  hipLaunchKernelGGL(k_soa_to_aos, dim3(grid_for(c->n)), dim3(kBlock), 0, c->stream, c->fields(), c->orig[c->cur],
                     (int)c->n, dst, ld);
  if (!on_device) HIPCHK(hipMemcpyAsync(store, dst, bytes, hipMemcpyDeviceToHost, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  return SPH_OK;
}

// ================================================================== C-ABI
extern "C" {

int sph_default_options(sph_options *o) {
  if (!o) return SPH_ERR_BAD_ARG;
  memset(o, 0, sizeof(*o));
  o->precision = 0;
  o->tensor_mode = 0;
  o->kappa_perp = 0.0;
  o->max_sweeps = 100;
  o->newton_tol = 0.0;
  o->neigh_cap = 0;
  o->device = 0;
  o->cell_div = 0;
  return SPH_OK;
}

int sph_ctx_create(const double state[49], int kernel_id, const sph_options *opt, sph_ctx **out) {
  if (!state || !out || kernel_id < 0 || kernel_id > 2) return SPH_ERR_BAD_ARG;
  *out = nullptr;
  int dim = (int)state[S_DIM];
  if (dim < 1 || dim > 3) return SPH_ERR_BAD_ARG;
  if ((int)state[S_COORD] == 2) return SPH_ERR_UNSUPPORTED;  // cylindrical, kernel.f90:135-148
  sph_ctx *c = new sph_ctx();
  if (opt) c->opt = *opt;
  else sph_default_options(&c->opt);
  {
    const char *t = std::getenv("SPH_NEIGH_TRAVERSAL");
    c->lane_search = t && std::strcmp(t, "lane") == 0;
    const char *g = std::getenv("SPH_LOOP2_GENERIC");
    c->loop2_generic = g && std::strcmp(g, "1") == 0;
  }
  if (c->opt.precision < 0 || c->opt.precision > 1 || c->opt.tensor_mode < 0 || c->opt.tensor_mode > 1 ||
      c->opt.cell_div < 0 || c->opt.cell_div > 4 || c->opt.neigh_cap < 0 || c->opt.neigh_cap > (1 << 20)) {
    delete c;
    return SPH_ERR_BAD_ARG;
  }
  c->device = c->opt.device;
  if (hipSetDevice(c->device) != hipSuccess || hipStreamCreate(&c->stream) != hipSuccess) {
    delete c;
    return SPH_ERR_HIP;
  }
  for (auto &e : c->ev) (void)hipEventCreate(&e);
  c->kt = kernel_id;
  Phys &P = c->frozen;
  P.dim = dim;
  P.kt = kernel_id;
  P.ddw = (int)state[S_DDW];
  P.adden = (int)state[S_ADDEN] == 1;
  P.artts = (int)state[S_ARTTS] == 1;
  P.au = state[S_AU];
  P.tensor_mode = c->opt.tensor_mode;
  P.kappa_perp = c->opt.kappa_perp;
  eq_components(state, &P);
  P.wcv = knorm(kernel_id, dim);
  P.fwc = fwc_of(kernel_id, dim);
  P.krad = krad_of(kernel_id);
  c->maxneib = c->opt.neigh_cap > 0 ? c->opt.neigh_cap : maxneib_of(kernel_id, dim);
  // process / legacy-setup branch of iterate (iterator.f90:60-178)
  c->process = (int)state[S_PROCESS];
  c->ics = (int)state[S_ICS];
  int wrap = 0, ghost = 0;
  c->do_eos = 1;
  c->eos_iso = 0;
  if (c->process == 601) { wrap = ghost = 7; }
  else if (c->process == 602) { }
  else if (c->process == 600) {
    switch (c->ics) {
      case 102: ghost = 2; break;
      case 103: case 104: case 111: c->do_eos = 0; break;
      case 105: wrap = ghost = 7; c->eos_iso = 1; break;
      case 106: wrap = ghost = 6; break;
      case 108: wrap = ghost = 7; break;
      case 101: wrap = ghost = 1; c->gravity = 1; break;  // mti
      case 110: wrap = ghost = 1; c->gravity = 2; break;  // mtilowres
      default: delete c; return SPH_ERR_UNSUPPORTED;
    }
  } else {
    delete c;
    return SPH_ERR_BAD_ARG;
  }
  // box and ghost bands, BC.f90:46-68,210-223
  Box &b = c->box;
  double bs = state[S_BORD], pad = state[S_PAD];
  for (int a = 0; a < 3; a++) {
    b.emin[a] = state[S_XMIN + 2 * a];
    b.emax[a] = state[S_XMIN + 2 * a + 1];
    b.boxmax[a] = b.emax[a] - bs;
    b.boxmin[a] = b.emin[a] + bs;
    b.pmin[a] = b.emin[a];
    b.pmax[a] = b.emax[a];
    if (pad == 0) { b.pmin[a] = b.emin[a] + pad / 10; b.pmax[a] = b.emax[a] - pad / 10; }
  }
  int dmask = (1 << dim) - 1;
  b.wrap_mask = wrap & dmask;
  b.ghost_mask = ghost & dmask;
  c->n_real = (int64_t)state[S_REALPN];
  c->n_fixed = (int64_t)state[S_FIXEDPN];
  if (c->n_real < 0) { delete c; return SPH_ERR_BAD_ARG; }
  if (dev_alloc((void **)&c->d_red, 16 * kStripes * sizeof(long long)) != SPH_OK ||
      dev_alloc((void **)&c->d_int, 256 * sizeof(int)) != SPH_OK ||
      dev_alloc((void **)&c->d_stripes, kRecLL * sizeof(long long)) != SPH_OK ||
      hipHostMalloc((void **)&c->h_int, 256 * sizeof(int)) != hipSuccess ||
      hipHostMalloc((void **)&c->h_ll, 2048 * sizeof(long long)) != hipSuccess ||
      hipHostMalloc((void **)&c->h_rec, kRecLL * sizeof(long long)) != hipSuccess) {
    sph_ctx_destroy(c);
    return SPH_ERR_NOMEM;
  }
  c->d_sweep = c->d_stripes + kStepLL;
  c->h_sweep = c->h_rec + kStepLL;
  if (c->opt.max_sweeps > 240) c->opt.max_sweeps = 240;
  *out = c;
  return SPH_OK;
}

int sph_ctx_destroy(sph_ctx *c) {
  if (!c) return SPH_ERR_BAD_ARG;
  (void)hipSetDevice(c->device);
  if (c->stream) (void)hipStreamSynchronize(c->stream);
  void *ptrs[] = {c->fbuf[0], c->fbuf[1], c->orig[0], c->orig[1], c->keys, c->keys2, c->vals, c->vals2,
                  c->sort_tmp, c->cstart, c->pm, c->st[0], c->st[1], c->tr[0], c->tr[1], c->s1[0], c->s1[1], c->pfix, c->s2, c->pd, c->trd,
                  c->h0, c->eps, c->pf,
                  c->real, c->nb, c->cnt, c->q[0], c->q[1], c->q[2], c->q[3], c->q[4], c->energy_part,
                  c->d_red, c->d_int, c->d_stripes, c->d_staging, c->roff, c->rtmp, c->rent, c->scan_tmp,
                  c->dvec, c->gpart, c->ckpt, c->rcnt, c->dpxh, c->dirho, c->dW, c->ddW, c->dR, c->dS, c->dZ, c->dLZ[0], c->dLZ[1], c->dLZ[2], c->dLZ[3],
                  c->gkdev, c->nlF, c->nlU, c->nlUB, c->nlNu, c->nlMu, c->nlD, c->nlDD, c->nlR, c->nlSD,
                  c->ext_count, c->ecs, c->erec, c->ext_tmp};
  for (void *p : ptrs)
    if (p) (void)hipFree(p);
  if (c->h_int) (void)hipHostFree(c->h_int);
  if (c->h_ll) (void)hipHostFree(c->h_ll);
  if (c->h_rec) (void)hipHostFree(c->h_rec);
  for (auto &e : c->ev)
    if (e) (void)hipEventDestroy(e);
  if (c->stream) (void)hipStreamDestroy(c->stream);
  delete c;
  (void)hipGetLastError();  // clean-up failures are not the next call's
  return SPH_OK;
}

static int check_counts(sph_ctx *c, int64_t n, const double *state) {
  if (!state) return SPH_ERR_BAD_ARG;
  int64_t nr = (int64_t)state[S_REALPN];
  if (nr != c->n_real || n < nr) return SPH_ERR_BAD_ARG;
  return SPH_OK;
}

int sph_load(sph_ctx *c, int64_t n, const double *store, int64_t ld, int on_device) {
  if (!c) return SPH_ERR_BAD_ARG;
  (void)hipSetDevice(c->device);
  return load_impl(c, n, store, ld, on_device);
}

int sph_save(sph_ctx *c, double *store, int64_t ld, int on_device) {
  if (!c) return SPH_ERR_BAD_ARG;
  (void)hipSetDevice(c->device);
  return save_impl(c, store, ld, on_device);
}

int sph_step(sph_ctx *c, const double state[49], double *maxconsenrg) {
  if (!c || !state) return SPH_ERR_BAD_ARG;
  (void)hipSetDevice(c->device);
  int rc = run_step(c, state);
  if (maxconsenrg) *maxconsenrg = 0.0;  // circuit2.f90:101
  return rc;
}

int sph_iterate(sph_ctx *c, int64_t n, double *store, int64_t ld, const double state[49], double *maxconsenrg) {
  if (!c) return SPH_ERR_BAD_ARG;
  RCHK(check_counts(c, n, state));
  RCHK(sph_load(c, n, store, ld, 0));
  int rc = sph_step(c, state, maxconsenrg);
  int rs = save_impl(c, store, ld, 0);
  return rc != SPH_OK ? rc : rs;
}

int sph_iterate_device(sph_ctx *c, int64_t n, double *d_store, int64_t ld, const double state[49],
                       double *maxconsenrg) {
  if (!c) return SPH_ERR_BAD_ARG;
  RCHK(check_counts(c, n, state));
  RCHK(sph_load(c, n, d_store, ld, 1));
  int rc = sph_step(c, state, maxconsenrg);
  int rs = save_impl(c, d_store, ld, 1);
  return rc != SPH_OK ? rc : rs;
}

int sph_neighbours(sph_ctx *c, int64_t *off, int32_t *j, int8_t *code, int64_t cap_pairs, int64_t *n_pairs) {
  if (!c || !off || !c->have_step) return SPH_ERR_BAD_ARG;
  (void)hipSetDevice(c->device);
  const int n = (int)c->n;
  int64_t nr = c->n_real;
  int64_t *d_cnt = nullptr;
  RCHK(dalloc(&d_cnt, nr + 1));
  HIPCHK(hipMemsetAsync(d_cnt, 0, sizeof(int64_t) * (nr + 1), c->stream));
  hipLaunchKernelGGL(k_counts_by_orig, dim3(grid_for(n)), dim3(kBlock), 0, c->stream, c->cnt, c->orig[c->cur],
                     c->real, n, d_cnt);
  std::vector<int64_t> h(nr + 1);
  HIPCHK(hipMemcpyAsync(h.data(), d_cnt, sizeof(int64_t) * nr, hipMemcpyDeviceToHost, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  off[0] = 0;
  for (int64_t i = 0; i < nr; i++) off[i + 1] = off[i] + h[i];
  int64_t total = off[nr];
  if (n_pairs) *n_pairs = total;
  int rc = SPH_OK;
  if (j && code) {
    if (total > cap_pairs) rc = SPH_ERR_BAD_ARG;
    else {
      int32_t *dj = nullptr;
      int8_t *dc = nullptr;
      if (dalloc(&dj, total + 1) != SPH_OK || dalloc(&dc, total + 1) != SPH_OK) rc = SPH_ERR_NOMEM;
      else {
        HIPCHK(hipMemcpyAsync(d_cnt, off, sizeof(int64_t) * nr, hipMemcpyHostToDevice, c->stream));
        hipLaunchKernelGGL(k_neigh_export, dim3(grid_for(n)), dim3(kBlock), 0, c->stream, c->nb, c->cnt,
                           c->orig[c->cur], c->real, n, (int)c->npad, d_cnt, dj, dc, total);
        HIPCHK(hipMemcpyAsync(j, dj, sizeof(int32_t) * total, hipMemcpyDeviceToHost, c->stream));
        HIPCHK(hipMemcpyAsync(code, dc, total, hipMemcpyDeviceToHost, c->stream));
        HIPCHK(hipStreamSynchronize(c->stream));
      }
      if (dj) (void)hipFree(dj);
      if (dc) (void)hipFree(dc);
    }
  }
  (void)hipFree(d_cnt);
  return rc;
}

}  // extern "C"

// ------------------------------------------- allocation-free list export
// Scratch of sph_neighbours_device_noalloc, in device memory the library
// never allocates: the list lengths in original order (int64, at most
// kNoAllocMaxReal + 1 of them) and the prefix sum's temporary storage.  Each
// GPU the library runs on holds its own copy (about 33 MiB) until the process
// ends, and every call on that GPU shares it.
constexpr int64_t kNoAllocMaxReal = (int64_t)1 << 22;
constexpr size_t kNoAllocCountBytes = ((size_t)(kNoAllocMaxReal + 1) * sizeof(int64_t) + 255) / 256 * 256;
constexpr size_t kNoAllocTempBytes = (size_t)1 << 20;
static __device__ __align__(256) unsigned char scratch[kNoAllocCountBytes + kNoAllocTempBytes];
constexpr int kExportLL = 1600;  // h_ll slot of the pair-count read-back

// The device address of this device's `scratch`, looked up once per device
// (the device must be current).  Host code uses only this address: the
// array's own name is a host-side stand-in.
static std::mutex g_scratch_mu;
static unsigned char *g_scratch[64] = {};
static int scratch_on_device(int device, unsigned char **out) {
  if (device < 0 || device >= 64) return SPH_ERR_BAD_ARG;
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch[device]) {
    void *p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p) {
      fprintf(stderr, "sph_hip: hipGetSymbolAddress(scratch) failed on device %d: %s\n", device,
              hipGetErrorString(e));
      return SPH_ERR_HIP;
    }
    g_scratch[device] = static_cast<unsigned char *>(p);
  }
  *out = g_scratch[device];
  return SPH_OK;
}

extern "C" {

int sph_neighbours_device_noalloc(sph_ctx *c, int64_t *d_off, int32_t *d_j, int8_t *d_code, int64_t cap_pairs,
                                  int64_t *n_pairs) {
  if (!c || !d_off || !c->have_step || cap_pairs < 0 || (cap_pairs > 0 && (!d_j || !d_code)))
    return SPH_ERR_BAD_ARG;
  HIPCHK(hipSetDevice(c->device));
  const int64_t nr = c->n_real;
  if (nr > kNoAllocMaxReal) {
    fprintf(stderr, "sph_neighbours_device_noalloc: %lld real particles, at most %lld\n", (long long)nr,
            (long long)kNoAllocMaxReal);
    return SPH_ERR_BAD_ARG;
  }
  unsigned char *base = nullptr;
  RCHK(scratch_on_device(c->device, &base));
  int64_t *counts = reinterpret_cast<int64_t *>(base);
  void *temp = base + kNoAllocCountBytes;
  size_t temp_bytes = 0;
  HIPCHK(hipcub::DeviceScan::ExclusiveSum(nullptr, temp_bytes, counts, d_off, (int)(nr + 1), c->stream));
  if (temp_bytes > kNoAllocTempBytes) {
    fprintf(stderr, "sph_neighbours_device_noalloc: the prefix sum needs %zu bytes of scratch, %zu held\n",
            temp_bytes, kNoAllocTempBytes);
    return SPH_ERR_BAD_ARG;
  }
  const int n = (int)c->n;
  // every real particle writes its own slot; the slot after them (the
  // total) is the only one the last call's contents could leave stale
  hipLaunchKernelGGL(k_counts_by_orig, dim3(grid_for(n)), dim3(kBlock), 0, c->stream, c->cnt, c->orig[c->cur],
                     c->real, n, counts);
  hipLaunchKernelGGL(k_set_i64, dim3(1), dim3(64), 0, c->stream, counts + nr, (int64_t)0);
  HIPCHK(hipcub::DeviceScan::ExclusiveSum(temp, temp_bytes, counts, d_off, (int)(nr + 1), c->stream));
  if (cap_pairs > 0)
    hipLaunchKernelGGL(k_neigh_export, dim3(grid_for(n)), dim3(kBlock), 0, c->stream, c->nb, c->cnt,
                       c->orig[c->cur], c->real, n, (int)c->npad, d_off, d_j, d_code, cap_pairs);
  HIPCHK(hipMemcpyAsync(c->h_ll + kExportLL, d_off + nr, sizeof(int64_t), hipMemcpyDeviceToHost, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  const int64_t total = (int64_t)c->h_ll[kExportLL];
  if (n_pairs) *n_pairs = total;
  return (cap_pairs > 0 && total > cap_pairs) ? SPH_ERR_BAD_ARG : SPH_OK;
}

// timestep.calc, timestep.f90:40-166 (hydro, MHD, diffusion and FLD limits)
static int calc_dt(sph_ctx *c, const Phys &P, double *dt) {
  const long long BIG = 0x7FFFFFFFFFFFFFFFLL, SMALL = -0x7FFFFFFFFFFFFFFFLL - 1;
  const bool is_min[8] = {true, false, true, true, false, false, true, false};
  long long *init = c->h_ll;
  for (int v = 0; v < 8; v++)
    for (int k = 0; k < kStripes; k++) init[v * kStripes + k] = is_min[v] ? BIG : SMALL;
  HIPCHK(hipMemcpyAsync(c->d_red, init, sizeof(long long) * 8 * kStripes, hipMemcpyHostToDevice, c->stream));
  hipLaunchKernelGGL(k_dt_reduce, dim3(grid_for(c->n)), dim3(kBlock), 0, c->stream, c->fields(), c->real, (int)c->n,
                     c->d_red);
  long long *r = c->h_ll + 512;
  HIPCHK(hipMemcpyAsync(r, c->d_red, sizeof(long long) * 8 * kStripes, hipMemcpyDeviceToHost, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  double v[8];
  for (int k = 0; k < 8; k++) v[k] = ord2d(is_min[k] ? stripe_min(r, k) : stripe_max(r, k));
  double minh = v[0], maxc = v[1], minden = v[2], minc = v[3], maxkappa = v[4], maxb = v[5];
  double minkappa = v[6], maxlambda = v[7];
  double cnhydro, cnd2;
  cn_of(c->kt, P.ddw, &cnhydro, &cnd2);
  double d = HUGE_VAL;
  if (P.hydro) d = std::min(d, cnhydro * minh / maxc);
  if (P.magneto) d = std::min(d, .1 * minh / maxc / maxb);
  if (P.diff) d = std::min(d, cnd2 * minden * minc * minh * minh / maxkappa);
  if (P.fld) d = std::min(d, cnd2 * minh * minh * minden * minkappa / kLightspeed / maxlambda);
  *dt = d;
  return SPH_OK;
}

// checkVarsReady (preruncheck.f90:13-60) on the resident store; *what gets
// the failed test: 32 mu0 <= 0 with MHD, 64 isotropic conductivity <= 0,
// else the per-particle bits of k_prerun_check.
static int prerun_check(sph_ctx *c, const double *state, int *what) {
  *what = 0;
  const int eqs = (int)state[S_EQS];
  if (state[S_MUZERO] <= 0.0 && (eqs == 302 || eqs == 303)) *what = 32;
  else if ((int)state[S_DISO] > 0 && state[S_DCOND] <= 0.0) *what = 64;
  if (*what) return SPH_ERR_PRECONDITION;
  if (c->n <= 0) return SPH_ERR_BAD_ARG;
  HIPCHK(hipMemsetAsync(c->d_int + 5, 0, sizeof(int), c->stream));
  hipLaunchKernelGGL(k_prerun_check, dim3(grid_for(c->n)), dim3(kBlock), 0, c->stream, c->fields(), (int)c->n,
                     c->d_int + 5);
  RCHK(read_ints(c, 5, 1));
  const int b = c->h_int[5];
  // the reference stops at the first failing test, in this order
  for (int bit : {1, 2, 4, 8, 16})
    if (b & bit) { *what = bit; return SPH_ERR_PRECONDITION; }
  return SPH_OK;
}

int sph_prerun_check(sph_ctx *c, const double state[49], int *what) {
  if (!c || !state) return SPH_ERR_BAD_ARG;
  (void)hipSetDevice(c->device);
  int w = 0;
  const int rc = prerun_check(c, state, &w);
  if (what) *what = w;
  return rc;
}

int sph_leapfrog(sph_ctx *c, const double state[49], double dtcap, double *dt_used) {
  if (!c || !state || c->n <= 0) return SPH_ERR_BAD_ARG;
  (void)hipSetDevice(c->device);
  Phys P = call_phys(c, state);
  // main.f90:216-219: the pre-run checks and the initial iterate, once;
  // later passes take the derivatives the previous pass's iterate left
  // (main.f90:323-329)
  if (!c->derivs_ready) {
    int what = 0;
    RCHK(prerun_check(c, state, &what));
    RCHK(run_step(c, state));
  }
  double dt;
  RCHK(calc_dt(c, P, &dt));
  dt = std::min(dt, dtcap);
  if (dt_used) *dt_used = dt;
  c->naux = 9;
  hipLaunchKernelGGL(k_predict, dim3(grid_for(c->n)), dim3(kBlock), 0, c->stream, c->fields(), (int)c->n, dt, 1,
                     c->aux(0), (int)c->npad);
  c->pos_version++;
  c->vb_version++;
  c->h_version++;
  int rc = run_step(c, state);
  if (rc == SPH_OK)
    hipLaunchKernelGGL(k_correct, dim3(grid_for(c->n)), dim3(kBlock), 0, c->stream, c->fields(), (int)c->n, dt, 1,
                       c->aux(0), (int)c->npad);
  c->pos_version++;
  c->vb_version++;
  c->h_version++;
  c->naux = 0;
  HIPCHK(hipStreamSynchronize(c->stream));
  return rc;
}

// update_radenergy(n, store, dt), rad_exchange.f90:9-66, on every loaded
// particle (rows 1..n = real + fixed, as main.f90:367-369 passes n).
int sph_rad_exchange(sph_ctx *c, const double state[49], double dt) {
  if (!c || !state || c->n <= 0 || !(dt > 0.0)) return SPH_ERR_BAD_ARG;
  (void)hipSetDevice(c->device);
  const double a = 4. * kSigmaB / kLightspeed;                 // rad_exchange.f90:25
  const double cv1 = (state[S_GAMMA] - 1.) * state[S_MOLMASS] / kRg;  // :26
  HIPCHK(hipMemsetAsync(c->d_int, 0, sizeof(int), c->stream));
  hipLaunchKernelGGL(k_rad_exchange, dim3(grid_for(c->n)), dim3(kBlock), 0, c->stream, c->fields(), (int)c->n, dt, a,
                     cv1, kSigmaB, c->d_int);
  RCHK(read_ints(c, 0, 1));
  return (c->h_int[0] & EB_RAD) ? SPH_ERR_RAD_NEGATIVE : SPH_OK;
}

// sts_integrator.f90:10-51
static double sts_b(int j) { return j <= 2 ? 1. / 3. : (j * j + j - 2.) / (2. * j * (j + 1.)); }
static double sts_a(int j) { return 1 - sts_b(j); }
static double sts_mu(int j) { return (2. * j - 1.) / j * sts_b(j) / sts_b(j - 1); }
static double sts_nu(int j) { return -(j - 1.) / j * sts_b(j) / sts_b(j - 2); }
static double sts_w(int s) { return 4. / (s * s + s - 2.); }
static double sts_mu_hat(int j, int s) { return j == 1 ? sts_b(1) * sts_w(s) : sts_mu(j) * sts_w(s); }
static double sts_gm_hat(int j, int s) { return -sts_a(j - 1) * sts_mu_hat(j, s); }

int sph_sts_coeffs(int j, int s, double out[4]) {
  if (!out || j < 1 || s < 2) return SPH_ERR_BAD_ARG;
  out[0] = j >= 2 ? sts_mu(j) : 0.0;
  out[1] = j >= 2 ? sts_nu(j) : 0.0;
  out[2] = sts_mu_hat(j, s);
  out[3] = j >= 2 ? sts_gm_hat(j, s) : 0.0;
  return SPH_OK;
}

// sts_integrate, sts_integrator.f90:87-124: aux arrays 0 Y0, 1 Yjm1, 2 Yjm2,
// 3 DY0, 4 DYjm1, 5 Yj (permuted with the particles at every re-sort).
int sph_sts(sph_ctx *c, const double state[49], int s, double dt) {
  if (!c || !state || c->n <= 0 || s < 2) return SPH_ERR_BAD_ARG;
  (void)hipSetDevice(c->device);
  const int n = (int)c->n;
  unsigned G = grid_for(n);
  c->naux = 6;
  auto T = [&]() { return c->fields().f[F_T]; };
  auto DDT = [&]() { return c->fields().f[F_DDT]; };
  HIPCHK(hipMemcpyAsync(c->aux(0), T(), sizeof(double) * n, hipMemcpyDeviceToDevice, c->stream));
  int rc = run_step(c, state);
  if (rc != SPH_OK) { c->naux = 0; return rc; }
  HIPCHK(hipMemcpyAsync(c->aux(3), DDT(), sizeof(double) * n, hipMemcpyDeviceToDevice, c->stream));
  HIPCHK(hipMemcpyAsync(c->aux(2), c->aux(0), sizeof(double) * n, hipMemcpyDeviceToDevice, c->stream));
  hipLaunchKernelGGL(k_axpy_copy, dim3(G), dim3(kBlock), 0, c->stream, c->aux(1), c->aux(0), c->aux(3),
                     sts_mu_hat(1, s) * dt, n);
  for (int j = 2; j <= s; j++) {
    HIPCHK(hipMemcpyAsync(T(), c->aux(1), sizeof(double) * n, hipMemcpyDeviceToDevice, c->stream));
    rc = run_step(c, state);
    if (rc != SPH_OK) { c->naux = 0; return rc; }
    HIPCHK(hipMemcpyAsync(c->aux(4), DDT(), sizeof(double) * n, hipMemcpyDeviceToDevice, c->stream));
    double mu = sts_mu(j), nu = sts_nu(j);
    hipLaunchKernelGGL(k_sts_stage, dim3(G), dim3(kBlock), 0, c->stream, c->aux(0), c->aux(1), c->aux(2),
                       c->aux(3), c->aux(4), n, mu, nu, sts_mu_hat(j, s) * dt, sts_gm_hat(j, s) * dt, c->aux(5));
    HIPCHK(hipMemcpyAsync(c->aux(2), c->aux(1), sizeof(double) * n, hipMemcpyDeviceToDevice, c->stream));
    HIPCHK(hipMemcpyAsync(c->aux(1), c->aux(5), sizeof(double) * n, hipMemcpyDeviceToDevice, c->stream));
  }
  HIPCHK(hipMemcpyAsync(T(), c->aux(5), sizeof(double) * n, hipMemcpyDeviceToDevice, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  c->naux = 0;
  return SPH_OK;
}

int sph_set_timing(sph_ctx *c, int on) {
  if (!c || (on != 0 && on != 1)) return SPH_ERR_BAD_ARG;
  c->timing = on != 0;
  return SPH_OK;
}

int sph_get_stats(sph_ctx *c, sph_stats *s) {
  if (!c || !s) return SPH_ERR_BAD_ARG;
  *s = c->stats;
  s->n_real = c->n_real;
  s->n_fixed = c->n_fixed;
  return SPH_OK;
}

const char *sph_error_string(int code) {
  switch (code) {
    case SPH_OK: return "ok";
    case SPH_ERR_NEIGH_TRUNCATED: return "# <!> KDTree neibsearch result was truncated.";
    case SPH_ERR_NEG_H: return "Negative smoothing length in c1 advanced";
    case SPH_ERR_Q_NAN_OR_NEG: return "q is nan or negative";
    case SPH_ERR_FLD_ANISO: return "There is no anisotropic for FLD";
    case SPH_ERR_BAD_DDW: return "# <!> second deriv id not found";
    case SPH_ERR_PRECONDITION: return "pre-run check failed";
    case SPH_ERR_BAD_ARG: return "invalid argument";
    case SPH_ERR_HIP: return "HIP runtime error";
    case SPH_ERR_NOMEM: return "device allocation failed";
    case SPH_ERR_UNSUPPORTED: return "unsupported configuration";
    case SPH_ERR_RAD_NEGATIVE: return "radiation or thermal energy is negative after exchange";
    default: return "unknown error";
  }
}

}  // extern "C"

// ------------------------------------------------- differentiable diffusion
// Set-up, operator applications and the checkpointed RKL2 forward/backward of
// the STS super-steps (sts_integrator.f90:87-124) with a global tensor K.
static DiffOpArgs diff_args(sph_ctx *c, const double K[6]) {
  DiffOpArgs a{};
  a.pm = c->pm; a.q0 = (const double4 *)c->q[0]; a.real = c->real; a.nb = c->nb; a.cnt = c->cnt;
  a.rcnt = c->rcnt; a.rent = c->rent; a.pxh = c->dpxh; a.b = c->box;
  for (int k = 0; k < 6; k++) a.K[k] = K[k];
  a.wcv = c->diff_phys.wcv; a.fwc = c->diff_phys.fwc; a.ddw = c->diff_phys.ddw;
  a.n = (int)c->n; a.npad = (int)c->npad;
  a.kap = c->fields().f[F_KAPPA];
  return a;
}

// ---- operators that read grad T (frozen geometry; see sph_adjoint.hip)
template <int DIM, int KT>
static void nl_grad_k(sph_ctx *c, const DiffOpArgs &a, const double *Y, double4 *U, double4 *F) {
  const unsigned G = grid_for(c->n);
  if (c->diff_op == 4) hipLaunchKernelGGL((k_nl_grad<DIM, KT, 2>), dim3(G), dim3(kBlock), 0, c->stream, a, Y, U, F);
  else hipLaunchKernelGGL((k_nl_grad<DIM, KT, 1>), dim3(G), dim3(kBlock), 0, c->stream, a, Y, U, F);
}
template <int DIM, int KT>
static void nl_apply_k(sph_ctx *c, const DiffOpArgs &a, const double *Y, double *out) {
  const unsigned G = grid_for(c->n);
  nl_grad_k<DIM, KT>(c, a, Y, c->nlU, nullptr);
  if (c->diff_op == 5) {
    hipLaunchKernelGGL(k_fld_coef, dim3(G), dim3(kBlock), 0, c->stream, a, Y, c->nlU, c->nlD, c->nlDD, c->nlR);
    hipLaunchKernelGGL((k_nl_apply<DIM, KT, 5>), dim3(G), dim3(kBlock), 0, c->stream, a, Y, c->nlU, c->nlD, out);
  } else if (c->diff_op == 3) {
    hipLaunchKernelGGL((k_nl_apply<DIM, KT, 3>), dim3(G), dim3(kBlock), 0, c->stream, a, Y, c->nlU, c->nlD, out);
  } else {
    hipLaunchKernelGGL((k_nl_apply<DIM, KT, 4>), dim3(G), dim3(kBlock), 0, c->stream, a, Y, c->nlU, c->nlD, out);
  }
}
// out = A^T lam (2nw: A linear; FLD: the Jacobian of f at Ya, transposed);
// for 2nw the tensor gradients lam^T dA/dK Ya -> gkdev[0..5], Yb -> [6..11]
template <int DIM, int KT>
static void nl_adjoint_k(sph_ctx *c, const DiffOpArgs &a, const double *lam, const double *Ya, const double *Yb,
                         double *out) {
  const unsigned G = grid_for(c->n);
  if (c->diff_op == 5) {
    nl_grad_k<DIM, KT>(c, a, Ya, c->nlU, nullptr);
    hipLaunchKernelGGL(k_fld_coef, dim3(G), dim3(kBlock), 0, c->stream, a, Ya, c->nlU, c->nlD, c->nlDD, c->nlR);
    hipLaunchKernelGGL((k_fld_adj<DIM, KT>), dim3(G), dim3(kBlock), 0, c->stream, a, lam, Ya, c->nlD, out, c->nlSD);
    hipLaunchKernelGGL(k_fld_pv, dim3(G), dim3(kBlock), 0, c->stream, a, c->nlSD, c->nlDD, c->nlR, c->nlU, Ya, out,
                       c->nlMu);
    hipLaunchKernelGGL((k_nl_gT<DIM, KT, 1>), dim3(G), dim3(kBlock), 0, c->stream, a, c->nlMu, c->nlF, out, 1);
    return;
  }
  nl_grad_k<DIM, KT>(c, a, Ya, c->nlU, nullptr);
  if (Yb) nl_grad_k<DIM, KT>(c, a, Yb, c->nlUB, nullptr);
  if (c->diff_op == 3)
    hipLaunchKernelGGL((k_nl_nu<DIM, KT, 3>), dim3(G), dim3(kBlock), 0, c->stream, a, lam, c->nlF, c->nlNu);
  else
    hipLaunchKernelGGL((k_nl_nu<DIM, KT, 4>), dim3(G), dim3(kBlock), 0, c->stream, a, lam, c->nlF, c->nlNu);
  hipLaunchKernelGGL(k_nl_mu, dim3(G), dim3(kBlock), 0, c->stream, a, c->nlNu, c->nlU, Yb ? c->nlUB : nullptr,
                     c->nlMu, c->gpart);
  hipLaunchKernelGGL(k_gk_accum, dim3(1), dim3(256), 0, c->stream, c->gpart, (int)G, 1.0, 0.0, c->gkdev);
  hipLaunchKernelGGL(k_gk_accum, dim3(1), dim3(256), 0, c->stream, c->gpart, (int)G, 0.0, 1.0, c->gkdev + 6);
  if (c->diff_op == 4)
    hipLaunchKernelGGL((k_nl_gT<DIM, KT, 2>), dim3(G), dim3(kBlock), 0, c->stream, a, c->nlMu, c->nlF, out, 0);
  else
    hipLaunchKernelGGL((k_nl_gT<DIM, KT, 1>), dim3(G), dim3(kBlock), 0, c->stream, a, c->nlMu, c->nlF, out, 0);
}
#define NL_DISPATCH(FN, ...)                                                                       \
  do {                                                                                            \
    const int d_ = c->diff_phys.dim, k_ = c->kt;                                                  \
    if (d_ == 1) { if (k_ == KT_M4) FN<1, KT_M4>(__VA_ARGS__); else if (k_ == KT_M6) FN<1, KT_M6>(__VA_ARGS__); else FN<1, KT_GAUSS>(__VA_ARGS__); } \
    else if (d_ == 2) { if (k_ == KT_M4) FN<2, KT_M4>(__VA_ARGS__); else if (k_ == KT_M6) FN<2, KT_M6>(__VA_ARGS__); else FN<2, KT_GAUSS>(__VA_ARGS__); } \
    else { if (k_ == KT_M4) FN<3, KT_M4>(__VA_ARGS__); else if (k_ == KT_M6) FN<3, KT_M6>(__VA_ARGS__); else FN<3, KT_GAUSS>(__VA_ARGS__); } \
  } while (0)

template <int DIM>
static void launch_weights_d(sph_ctx *c, const DiffOpArgs &a) {
  unsigned G = grid_for(c->n);
#define WTS(KT)                                                                                                  \
  do {                                                                                                           \
    hipLaunchKernelGGL((k_diff_weights<DIM, KT>), dim3(G), dim3(kBlock), 0, c->stream, a, c->dW, c->ddW);       \
    hipLaunchKernelGGL((k_diff_rweights<DIM, KT>), dim3(G), dim3(kBlock), 0, c->stream, a, c->dirho, c->dR);    \
  } while (0)
  if (c->kt == KT_M4) WTS(KT_M4);
  else if (c->kt == KT_M6) WTS(KT_M6);
  else WTS(KT_GAUSS);
#undef WTS
}
// Pair weights of A(K), assembled once per K (list order and reverse order).
static void ensure_weights(sph_ctx *c, const double K[6]) {
  if (c->w_valid && std::equal(K, K + 6, c->w_K)) return;
  DiffOpArgs a = diff_args(c, K);
  if (c->diff_phys.dim == 1) launch_weights_d<1>(c, a);
  else if (c->diff_phys.dim == 2) launch_weights_d<2>(c, a);
  else launch_weights_d<3>(c, a);
  std::copy(K, K + 6, c->w_K);
  c->w_valid = true;
}
// out = A(K) Y (sorted vectors).
static void diff_apply(sph_ctx *c, const double K[6], const double *Y, double *out) {
  ensure_weights(c, K);
  DiffOpArgs a = diff_args(c, K);
  const unsigned G = grid_for(c->n);
  hipLaunchKernelGGL(k_zrho, dim3(G), dim3(kBlock), 0, c->stream, (int)c->n, Y, a.q0, c->dZ);
  hipLaunchKernelGGL(k_diff_spmv, dim3(G), dim3(kBlock), 0, c->stream, a, c->dW, c->dZ, out);
}
// One adjoint stage: out = A^T lam, and the stage's gradient coefficients
// S += lam_a (Zt_b - Zt_a), Zt = (wa Ya + wb Yb) rho (Yb may be null).  S
// updates are deferred and done four stages at a time (one pass over S per
// four stages); diff_flush settles the deferred ones.
static void diff_adj_stage(sph_ctx *c, const double K[6], const double *lam, const double *Ya, const double *Yb,
                           double wa, double wb, double *out) {
  ensure_weights(c, K);
  DiffOpArgs a = diff_args(c, K);
  const unsigned G = grid_for(c->n);
  // this stage's record goes first in the set; the deferred ones follow
  LZSet set{};
  set.p[0] = c->dLZ[c->lz_pend];
  for (int q = 0; q < c->lz_pend; q++) set.p[q + 1] = c->dLZ[q];
  hipLaunchKernelGGL(k_adj_lz, dim3(G), dim3(kBlock), 0, c->stream, (int)c->n, lam, Ya, Yb, wa, wb, a.q0,
                     c->dLZ[c->lz_pend]);
  if (c->lz_pend == 3) {
    hipLaunchKernelGGL((k_diff_adj_spmv<4, 1>), dim3(G), dim3(kBlock), 0, c->stream, a, c->dR, c->ddW, set, c->dS,
                       out);
    c->lz_pend = 0;
  } else {
    hipLaunchKernelGGL((k_diff_adj_spmv<1, 0>), dim3(G), dim3(kBlock), 0, c->stream, a, c->dR, c->ddW, set, c->dS,
                       out);
    c->lz_pend++;
  }
}
static void diff_flush(sph_ctx *c, const double K[6]) {
  if (c->lz_pend == 0) return;
  DiffOpArgs a = diff_args(c, K);
  const unsigned G = grid_for(c->n);
  LZSet set{};
  for (int q = 0; q < c->lz_pend; q++) set.p[q] = c->dLZ[q];
  double *none = nullptr;
  if (c->lz_pend == 1)
    hipLaunchKernelGGL((k_diff_adj_spmv<1, 1>), dim3(G), dim3(kBlock), 0, c->stream, a, c->dR, c->ddW, set, c->dS, none);
  else if (c->lz_pend == 2)
    hipLaunchKernelGGL((k_diff_adj_spmv<2, 1>), dim3(G), dim3(kBlock), 0, c->stream, a, c->dR, c->ddW, set, c->dS, none);
  else
    hipLaunchKernelGGL((k_diff_adj_spmv<3, 1>), dim3(G), dim3(kBlock), 0, c->stream, a, c->dR, c->ddW, set, c->dS, none);
  c->lz_pend = 0;
}
template <int DIM>
static void launch_gcontract_d(sph_ctx *c, const DiffOpArgs &a) {
  unsigned G = grid_for(c->n);
  if (c->kt == KT_M4)
    hipLaunchKernelGGL((k_diff_gcontract<DIM, KT_M4>), dim3(G), dim3(kBlock), 0, c->stream, a, c->dirho, c->dS, c->gpart);
  else if (c->kt == KT_M6)
    hipLaunchKernelGGL((k_diff_gcontract<DIM, KT_M6>), dim3(G), dim3(kBlock), 0, c->stream, a, c->dirho, c->dS, c->gpart);
  else
    hipLaunchKernelGGL((k_diff_gcontract<DIM, KT_GAUSS>), dim3(G), dim3(kBlock), 0, c->stream, a, c->dirho, c->dS,
                       c->gpart);
}
// gk[0..5] += sum over pairs of the accumulated coefficients times G (fixed
// reduction order), then the coefficients are cleared.
static int diff_gcontract(sph_ctx *c, const double K[6], double *gk) {
  diff_flush(c, K);
  DiffOpArgs a = diff_args(c, K);
  if (c->diff_phys.dim == 1) launch_gcontract_d<1>(c, a);
  else if (c->diff_phys.dim == 2) launch_gcontract_d<2>(c, a);
  else launch_gcontract_d<3>(c, a);
  hipLaunchKernelGGL(k_gk_accum, dim3(1), dim3(256), 0, c->stream, c->gpart, (int)grid_for(c->n), 1.0, 0.0, gk);
  HIPCHK(hipMemsetAsync(c->dS, 0, sizeof(double) * (size_t)c->npad * c->rcap, c->stream));
  return SPH_OK;
}

static void lincomb(sph_ctx *c, double *y, double c0, const double *x0, double c1 = 0, const double *x1 = nullptr,
                    double c2 = 0, const double *x2 = nullptr, double c3 = 0, const double *x3 = nullptr,
                    double c4 = 0, const double *x4 = nullptr) {
  hipLaunchKernelGGL(k_lincomb, dim3(grid_for(c->n)), dim3(kBlock), 0, c->stream, (int)c->n, y, c0, x0, c1, x1, c2,
                     x2, c3, x3, c4, x4);
}

static int ensure_dvec(sph_ctx *c, int nvec) {
  if (nvec <= c->diff_nvec) return SPH_OK;
  RCHK(dalloc(&c->dvec, (size_t)nvec * c->npad));
  HIPCHK(hipMemset(c->dvec, 0, sizeof(double) * nvec * c->npad));
  c->diff_nvec = nvec;
  return SPH_OK;
}
static double *dv(sph_ctx *c, int k) { return c->dvec + (size_t)k * c->npad; }

// One RKL2 super-step from Y[0] (sorted); stage states into Y[0..s]; work vector W.
static void sts_forward_step(sph_ctx *c, const double K[6], int s, double dt, double *const *Y, double *DY0,
                             double *DYm) {
  diff_apply(c, K, Y[0], DY0);
  lincomb(c, Y[1], 1.0, Y[0], sts_mu_hat(1, s) * dt, DY0);
  for (int j = 2; j <= s; j++) {
    diff_apply(c, K, Y[j - 1], DYm);
    double mu = sts_mu(j), nu = sts_nu(j);
    lincomb(c, Y[j], mu, Y[j - 1], nu, Y[j - 2], 1. - mu - nu, Y[0], sts_mu_hat(j, s) * dt, DYm,
            sts_gm_hat(j, s) * dt, DY0);
  }
}

// Adjoint of one super-step: lam[s] given (in L[s]); returns dL/dY0 in L[0];
// the tensor-gradient coefficients accumulate per pair (dS).
static int sts_backward_step(sph_ctx *c, const double K[6], int s, double dt, double *const *Y, double *const *L,
                             double *U) {
  for (int j = s; j >= 2; j--) {
    double mh = sts_mu_hat(j, s) * dt, gh = sts_gm_hat(j, s) * dt, mu = sts_mu(j), nu = sts_nu(j);
    diff_adj_stage(c, K, L[j], Y[j - 1], Y[0], mh, gh, U);
    hipLaunchKernelGGL(k_stage_adj, dim3(grid_for(c->n)), dim3(kBlock), 0, c->stream, (int)c->n, L[j], U, mu, nu,
                       mh, gh, L[j - 1], L[j - 2], L[0]);
  }
  double m1 = sts_mu_hat(1, s) * dt;
  diff_adj_stage(c, K, L[1], Y[0], nullptr, m1, 0.0, U);
  lincomb(c, L[0], 1.0, L[0], 1.0, L[1], m1, U);
  return SPH_OK;
}

extern "C" {

int sph_diff_prepare(sph_ctx *c, const double state[49]) {
  if (!c || !state || c->n <= 0) return SPH_ERR_BAD_ARG;
  (void)hipSetDevice(c->device);
  if (c->opt.precision != 0 || c->n_fixed != 0) return SPH_ERR_UNSUPPORTED;
  Phys P = call_phys(c, state);
  const bool hess = P.ddw == 400 || P.ddw == 401 || P.ddw == 402;
  int op;
  if (!P.diff) return SPH_ERR_UNSUPPORTED;
  if (P.fld) {
    // FLD: isotropic (circuit2.f90:189), one of the Hessian operators
    if (!hess || P.diso != 1) return SPH_ERR_UNSUPPORTED;
    op = 5;
  } else if (P.ddw == 403 || P.ddw == 404) {
    // the 2nw operators with one tensor K for every particle
    op = P.ddw == 403 ? 3 : 4;
  } else if (hess) {
    op = 0;
  } else {
    return SPH_ERR_UNSUPPORTED;
  }
  c->keep_q0 = true;
  const int step_rc = run_step(c, state);  // neighbours, Newton h, EOS, loop 2: h, rho, q0 now frozen
  c->keep_q0 = false;
  RCHK(step_rc);
  c->diff_phys = P;
  c->diff_op = op;
  const int n = (int)c->n;
  unsigned G = grid_for(n);
  const int64_t npairs = c->stats.n_pairs;
  if (npairs >= ((int64_t)1 << 31)) return SPH_ERR_UNSUPPORTED;  // 32-bit pair counts
  RCHK(dalloc(&c->roff, c->npad + 1));
  RCHK(dalloc(&c->rtmp, c->npad + 1));
  RCHK(dalloc(&c->rcnt, c->npad));
  RCHK(dalloc(&c->dpxh, c->npad));
  RCHK(dalloc(&c->dirho, c->npad));
  if (op == 0) RCHK(dalloc(&c->dW, (size_t)c->npad * c->nb_cap));
  RCHK(dalloc(&c->ddW, c->npad));
  RCHK(dalloc(&c->dZ, c->npad));
  for (int q = 0; q < 4; q++) RCHK(dalloc(&c->dLZ[q], c->npad));
  c->lz_pend = 0;
  RCHK(dalloc(&c->gkdev, 12));
  c->w_valid = false;
  hipLaunchKernelGGL(k_diff_pack, dim3(G), dim3(kBlock), 0, c->stream, n, c->pm, (const double4 *)c->q[0], c->dpxh,
                     c->dirho);
  // reverse lists: stable radix sort of the owner-major pairs on the neighbour
  // index, then into the list layout (capacity = longest reverse list)
  int *foff = nullptr;
  uint32_t *keys = nullptr, *keys2 = nullptr, *vals = nullptr, *flat = nullptr;
  auto release = [&]() {
    void *t[] = {foff, keys, keys2, vals, flat};
    for (void *q : t)
      if (q) (void)hipFree(q);
  };
  auto tmp_bytes = [&](size_t need) -> int {
    if (need > c->scan_tmp_bytes) {
      RCHK(dev_alloc(&c->scan_tmp, need));
      c->scan_tmp_bytes = need;
    }
    return SPH_OK;
  };
  int rc = SPH_OK;
  do {
    const size_t np = (size_t)std::max<int64_t>(npairs, 1);
    if ((rc = dalloc(&foff, c->npad + 1)) || (rc = dalloc(&keys, np)) || (rc = dalloc(&keys2, np)) ||
        (rc = dalloc(&vals, np)) || (rc = dalloc(&flat, np)))
      break;
    // reverse segment offsets: how often each particle is someone's neighbour
    HIPCHK(hipMemsetAsync(c->rtmp, 0, sizeof(int) * (n + 1), c->stream));
    hipLaunchKernelGGL(k_rev_count, dim3(G), dim3(kBlock), 0, c->stream, c->real, c->nb, c->cnt, n, (int)c->npad,
                       c->rtmp);
    size_t b1 = 0, b2 = 0, b3 = 0;
    HIPCHK(hipcub::DeviceScan::ExclusiveSum(nullptr, b1, c->rtmp, c->roff, n + 1, c->stream));
    HIPCHK(hipcub::DeviceScan::ExclusiveSum(nullptr, b2, c->cnt, foff, n, c->stream));
    HIPCHK(hipcub::DeviceRadixSort::SortPairs(nullptr, b3, keys, keys2, vals, flat, (int)npairs, 0, 32, c->stream));
    if ((rc = tmp_bytes(std::max(b1, std::max(b2, b3))))) break;
    size_t bytes = c->scan_tmp_bytes;
    HIPCHK(hipcub::DeviceScan::ExclusiveSum(c->scan_tmp, bytes, c->rtmp, c->roff, n + 1, c->stream));
    // the longest reverse list sets the capacity of the list-layout arrays
    std::vector<int> hoff(n + 1);
    HIPCHK(hipMemcpyAsync(hoff.data(), c->roff, sizeof(int) * (n + 1), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    int rmax = 0;
    for (int k = 0; k < n; k++) rmax = std::max(rmax, hoff[k + 1] - hoff[k]);
    c->rcap = std::max(8, (rmax + 7) / 8 * 8);
    const size_t rslots = (size_t)c->npad * c->rcap;
    if ((rc = dalloc(&c->rent, rslots))) break;
    if (op == 0) {
      if ((rc = dalloc(&c->dR, rslots)) || (rc = dalloc(&c->dS, rslots))) break;
      HIPCHK(hipMemsetAsync(c->dS, 0, sizeof(double) * rslots, c->stream));
    }
    // owner-major enumeration of the pairs, stable sort on the neighbour index
    bytes = c->scan_tmp_bytes;
    HIPCHK(hipcub::DeviceScan::ExclusiveSum(c->scan_tmp, bytes, c->cnt, foff, n, c->stream));
    hipLaunchKernelGGL(k_pair_enum, dim3(G), dim3(kBlock), 0, c->stream, c->real, c->nb, c->cnt, n, (int)c->npad,
                       foff, keys, vals);
    int bits = 1;
    while (bits < 32 && ((int64_t)1 << bits) < (int64_t)n) bits++;
    bytes = c->scan_tmp_bytes;
    if (npairs > 0)
      HIPCHK(hipcub::DeviceRadixSort::SortPairs(c->scan_tmp, bytes, keys, keys2, vals, flat, (int)npairs, 0, bits,
                                                c->stream));
    hipLaunchKernelGGL(k_rev_chunk, dim3(G), dim3(kBlock), 0, c->stream, c->roff, flat, n, (int)c->npad, c->rcnt,
                       c->rent);
    HIPCHK(hipStreamSynchronize(c->stream));
  } while (false);
  release();
  RCHK(rc);
  if (c->gpart) (void)hipFree(c->gpart);
  c->gpart = nullptr;
  RCHK(dalloc(&c->gpart, (size_t)G * 12));
  if (op != 0) {
    RCHK(dalloc(&c->nlF, c->npad));
    RCHK(dalloc(&c->nlU, c->npad));
    RCHK(dalloc(&c->nlUB, c->npad));
    RCHK(dalloc(&c->nlNu, c->npad));
    RCHK(dalloc(&c->nlMu, c->npad));
    RCHK(dalloc(&c->nlD, c->npad));
    RCHK(dalloc(&c->nlDD, c->npad));
    RCHK(dalloc(&c->nlR, c->npad));
    RCHK(dalloc(&c->nlSD, c->npad));
    const double K0[6] = {0, 0, 0, 0, 0, 0};
    DiffOpArgs a = diff_args(c, K0);
    NL_DISPATCH(nl_grad_k, c, a, nullptr, nullptr, c->nlF);  // F_s = sum_b m_b nw(r_sb, h_s)
  }
  HIPCHK(hipStreamSynchronize(c->stream));
  c->fw_tail = false;
  c->diff_ready = true;
  return SPH_OK;
}

int sph_diff_apply(sph_ctx *c, const double K[6], const double *d_T, double *d_out) {
  if (!c || !c->diff_ready || !d_T || !d_out || (!K && c->diff_op != 5)) return SPH_ERR_BAD_ARG;
  (void)hipSetDevice(c->device);
  RCHK(ensure_dvec(c, 2));
  const int n = (int)c->n;
  unsigned G = grid_for(n);
  hipLaunchKernelGGL(k_to_sorted, dim3(G), dim3(kBlock), 0, c->stream, d_T, c->orig[c->cur], n, dv(c, 0));
  if (c->diff_op != 0) {
    const double K0[6] = {0, 0, 0, 0, 0, 0};
    DiffOpArgs a = diff_args(c, K ? K : K0);
    NL_DISPATCH(nl_apply_k, c, a, dv(c, 0), dv(c, 1));
  } else {
    diff_apply(c, K, dv(c, 0), dv(c, 1));
  }
  hipLaunchKernelGGL(k_from_sorted, dim3(G), dim3(kBlock), 0, c->stream, dv(c, 1), c->orig[c->cur], n, d_out);
  HIPCHK(hipStreamSynchronize(c->stream));
  return SPH_OK;
}

int sph_diff_apply_adjoint(sph_ctx *c, const double K[6], const double *d_lam, const double *d_Ya,
                           const double *d_Yb, double *d_out, double gA[6], double gB[6]) {
  if (!c || !c->diff_ready || !d_lam || !d_Ya || !d_out || !gA || !gB) return SPH_ERR_BAD_ARG;
  if (c->diff_op != 5 && (!K || !d_Yb)) return SPH_ERR_BAD_ARG;
  (void)hipSetDevice(c->device);
  RCHK(ensure_dvec(c, 4));
  const int n = (int)c->n;
  unsigned G = grid_for(n);
  const int64_t *o = c->orig[c->cur];
  hipLaunchKernelGGL(k_to_sorted, dim3(G), dim3(kBlock), 0, c->stream, d_lam, o, n, dv(c, 0));
  hipLaunchKernelGGL(k_to_sorted, dim3(G), dim3(kBlock), 0, c->stream, d_Ya, o, n, dv(c, 1));
  if (d_Yb) hipLaunchKernelGGL(k_to_sorted, dim3(G), dim3(kBlock), 0, c->stream, d_Yb, o, n, dv(c, 2));
  HIPCHK(hipMemsetAsync(c->gkdev, 0, sizeof(double) * 12, c->stream));
  if (c->diff_op != 0) {
    const double K0[6] = {0, 0, 0, 0, 0, 0};
    DiffOpArgs a = diff_args(c, K ? K : K0);
    NL_DISPATCH(nl_adjoint_k, c, a, dv(c, 0), dv(c, 1), c->diff_op == 5 ? nullptr : dv(c, 2), dv(c, 3));
  } else {
    diff_adj_stage(c, K, dv(c, 0), dv(c, 1), nullptr, 1.0, 0.0, dv(c, 3));
    RCHK(diff_gcontract(c, K, c->gkdev));
    diff_adj_stage(c, K, dv(c, 0), dv(c, 2), nullptr, 1.0, 0.0, dv(c, 3));
    RCHK(diff_gcontract(c, K, c->gkdev + 6));
  }
  hipLaunchKernelGGL(k_from_sorted, dim3(G), dim3(kBlock), 0, c->stream, dv(c, 3), o, n, d_out);
  double g12[12];
  HIPCHK(hipMemcpyAsync(g12, c->gkdev, sizeof(double) * 12, hipMemcpyDeviceToHost, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  for (int k = 0; k < 6; k++) { gA[k] = g12[k]; gB[k] = g12[6 + k]; }
  return SPH_OK;
}

// Vectors: 0..s stage states, s+1 DY0, s+2 DYm, s+3 U, s+4 .. 2s+4 adjoints L[0..s];
// segment recompute states after that: every x (s+1).
int sph_diff_forward(sph_ctx *c, const double K[6], int s, double dt, int nsteps, int every, const double *d_T0,
                     double *d_T) {
  if (!c || !c->diff_ready || !K || !d_T0 || !d_T || s < 2 || nsteps < 1 || every < 1) return SPH_ERR_BAD_ARG;
  if (c->diff_op != 0) return SPH_ERR_UNSUPPORTED;  // stepped stage by stage by the caller (autograd.py)
  (void)hipSetDevice(c->device);
  const int n = (int)c->n;
  unsigned G = grid_for(n);
  int nck = nsteps / every + 1;
  if ((int64_t)nck * c->npad > c->ckpt_cap) {
    RCHK(dalloc(&c->ckpt, (size_t)nck * c->npad));
    c->ckpt_cap = (int64_t)nck * c->npad;
  }
  RCHK(ensure_dvec(c, 2 * s + 5 + every * (s + 1)));
  std::vector<double *> Y(s + 1), Yk(s + 1);
  for (int j = 0; j <= s; j++) Y[j] = dv(c, j);
  // the last checkpoint segment runs in the backward pass's segment buffers,
  // so the backward pass need not recompute it
  const int tail0 = ((nsteps + every - 1) / every - 1) * every, base = 2 * s + 5;
  hipLaunchKernelGGL(k_to_sorted, dim3(G), dim3(kBlock), 0, c->stream, d_T0, c->orig[c->cur], n, Y[0]);
  for (int st = 0; st < nsteps; st++) {
    if (st % every == 0)
      HIPCHK(hipMemcpyAsync(c->ckpt + (size_t)(st / every) * c->npad, Y[0], sizeof(double) * n,
                            hipMemcpyDeviceToDevice, c->stream));
    if (st < tail0) {
      sts_forward_step(c, K, s, dt, Y.data(), dv(c, s + 1), dv(c, s + 2));
      HIPCHK(hipMemcpyAsync(Y[0], Y[s], sizeof(double) * n, hipMemcpyDeviceToDevice, c->stream));
      continue;
    }
    const int k = st - tail0;
    for (int j = 0; j <= s; j++) Yk[j] = dv(c, base + k * (s + 1) + j);
    HIPCHK(hipMemcpyAsync(Yk[0], k == 0 ? Y[0] : dv(c, base + (k - 1) * (s + 1) + s), sizeof(double) * n,
                          hipMemcpyDeviceToDevice, c->stream));
    sts_forward_step(c, K, s, dt, Yk.data(), dv(c, s + 1), dv(c, s + 2));
    if (st == nsteps - 1)
      HIPCHK(hipMemcpyAsync(Y[0], Yk[s], sizeof(double) * n, hipMemcpyDeviceToDevice, c->stream));
  }
  hipLaunchKernelGGL(k_from_sorted, dim3(G), dim3(kBlock), 0, c->stream, Y[0], c->orig[c->cur], n, d_T);
  HIPCHK(hipStreamSynchronize(c->stream));
  c->fw_s = s; c->fw_steps = nsteps; c->fw_every = every; c->fw_dt = dt;
  for (int k = 0; k < 6; k++) c->fw_K[k] = K[k];
  c->fw_tail = true;
  return SPH_OK;
}

int sph_diff_backward(sph_ctx *c, const double *d_gT, double *d_gT0, double gK[6]) {
  if (!c || !c->diff_ready || c->fw_steps < 1 || !d_gT || !d_gT0 || !gK) return SPH_ERR_BAD_ARG;
  (void)hipSetDevice(c->device);
  const int n = (int)c->n, s = c->fw_s, every = c->fw_every, nsteps = c->fw_steps;
  const double dt = c->fw_dt;
  const double *K = c->fw_K;
  unsigned G = grid_for(n);
  std::vector<double *> L(s + 1);
  for (int j = 0; j <= s; j++) L[j] = dv(c, s + 4 + j);
  double *U = dv(c, s + 3);
  const int base = 2 * s + 5;
  auto seg = [&](int k, int j) { return dv(c, base + k * (s + 1) + j); };
  HIPCHK(hipMemsetAsync(c->dS, 0, sizeof(double) * (size_t)c->npad * c->rcap, c->stream));
  HIPCHK(hipMemsetAsync(c->gkdev, 0, sizeof(double) * 12, c->stream));
  double *lam = dv(c, s + 1);  // running dL/dT (sorted)
  hipLaunchKernelGGL(k_to_sorted, dim3(G), dim3(kBlock), 0, c->stream, d_gT, c->orig[c->cur], n, lam);
  int nseg = (nsteps + every - 1) / every;
  for (int sg = nseg - 1; sg >= 0; sg--) {
    int st0 = sg * every, st1 = std::min(nsteps, st0 + every);
    // recompute the segment's stage states from its checkpoint (the last
    // segment's are still there from the forward pass)
    const bool have = sg == nseg - 1 && c->fw_tail;
    if (!have)
      HIPCHK(hipMemcpyAsync(seg(0, 0), c->ckpt + (size_t)sg * c->npad, sizeof(double) * n, hipMemcpyDeviceToDevice,
                            c->stream));
    for (int st = st0; st < st1 && !have; st++) {
      int k = st - st0;
      std::vector<double *> Y(s + 1);
      for (int j = 0; j <= s; j++) Y[j] = seg(k, j);
      sts_forward_step(c, K, s, dt, Y.data(), dv(c, s + 2), U);
      if (st + 1 < st1)
        HIPCHK(hipMemcpyAsync(seg(k + 1, 0), Y[s], sizeof(double) * n, hipMemcpyDeviceToDevice, c->stream));
    }
    for (int st = st1 - 1; st >= st0; st--) {
      int k = st - st0;
      std::vector<double *> Y(s + 1);
      for (int j = 0; j <= s; j++) Y[j] = seg(k, j);
      HIPCHK(hipMemsetAsync(L[0], 0, sizeof(double) * (size_t)c->npad * s, c->stream));
      HIPCHK(hipMemcpyAsync(L[s], lam, sizeof(double) * n, hipMemcpyDeviceToDevice, c->stream));
      RCHK(sts_backward_step(c, K, s, dt, Y.data(), L.data(), U));
      HIPCHK(hipMemcpyAsync(lam, L[0], sizeof(double) * n, hipMemcpyDeviceToDevice, c->stream));
    }
  }
  hipLaunchKernelGGL(k_from_sorted, dim3(G), dim3(kBlock), 0, c->stream, lam, c->orig[c->cur], n, d_gT0);
  RCHK(diff_gcontract(c, K, c->gkdev));
  c->fw_tail = false;  // segment buffers now hold the first segment
  HIPCHK(hipMemcpyAsync(gK, c->gkdev, sizeof(double) * 6, hipMemcpyDeviceToHost, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  return SPH_OK;
}

}  // extern "C"

// ------------------------------------------------------- batched problems
// Independent prepared diffusion problems merged into one operator: one
// launch per RKL2 stage covers every particle of every problem, a super-step
// is captured once as a hipGraph and replayed, and the stage kernel's block
// size is picked by timing the candidates on the merged operator.
struct sph_batch {
  int device = 0;
  hipStream_t stream = nullptr;
  int nprob = 0;
  int64_t n = 0, npad = 0, n_pairs = 0;
  int cap = 0;
  std::vector<int64_t> off;            // nprob + 1 offsets of the problems' rows
  int *cnt = nullptr, *pid = nullptr;  // list lengths, problem of each row
  uint32_t *nb = nullptr;              // merged neighbour indices (list layout)
  double *W = nullptr, *rho = nullptr; // merged pair weights, densities
  float *W32 = nullptr;                // merged pair weights in FP32 (precision 1)
  int precision = 0;
  double *dt = nullptr;                // per-problem time steps
  int64_t *orig = nullptr;             // merged row -> position in the concatenated vectors
  double *vec = nullptr;               // Y0, three stage states, DY0, three Z (npad each)
  int block = 256;
  float tune_ms[4] = {0, 0, 0, 0};
  hipGraph_t graph = nullptr;
  hipGraphExec_t gexec = nullptr;
  int graph_s = 0, graph_block = 0;
};

namespace {
constexpr int kBatchBlocks[4] = {64, 128, 256, 512};
constexpr int kBatchVecs = 8;

double *bvec(sph_batch *b, int k) { return b->vec + (size_t)k * b->npad; }

// The launches of one RKL2 super-step of s stages (sts_integrator.f90:87-124)
// from Y_0 (buffer 0, its Z in buffer 5) back into buffers 0 and 5; stage
// states rotate over buffers 1-3, their Z over 6-7.
void batch_super_step(sph_batch *b, int s, int block) {
  double *Y0 = bvec(b, 0), *DY0 = bvec(b, 4), *Z0 = bvec(b, 5);
  double *R[3] = {bvec(b, 1), bvec(b, 2), bvec(b, 3)};
  double *ZR[2] = {bvec(b, 6), bvec(b, 7)};
  auto Y = [&](int j) -> double * { return (j == 0 || j == s) ? Y0 : R[j % 3]; };
  auto Z = [&](int j) -> double * { return (j == 0 || j == s) ? Z0 : ZR[j & 1]; };
  const unsigned G = (unsigned)((b->n + block - 1) / block);
  BatchStage a{};
  a.cnt = b->cnt; a.nb = b->nb; a.W = b->W; a.W32 = b->W32; a.rho = b->rho; a.pid = b->pid; a.dt = b->dt;
  a.Y0 = Y0; a.DY0 = DY0; a.n = (int)b->n; a.npad = (int)b->npad;
  a.Yjm1 = Y0; a.Yjm2 = Y0; a.Zjm1 = Z0; a.Yj = Y(1); a.Zj = Z(1);
  a.mhat = sts_mu_hat(1, s);
  const bool f32 = b->precision == 1;
  if (f32) hipLaunchKernelGGL((k_batch_stage<true, float>), dim3(G), dim3(block), 0, b->stream, a);
  else hipLaunchKernelGGL((k_batch_stage<true, double>), dim3(G), dim3(block), 0, b->stream, a);
  for (int j = 2; j <= s; j++) {
    a.Yjm1 = Y(j - 1); a.Yjm2 = Y(j - 2); a.Zjm1 = Z(j - 1); a.Yj = Y(j); a.Zj = Z(j);
    a.mu = sts_mu(j); a.nu = sts_nu(j); a.c0 = 1. - a.mu - a.nu;
    a.mhat = sts_mu_hat(j, s); a.ghat = sts_gm_hat(j, s);
    if (f32) hipLaunchKernelGGL((k_batch_stage<false, float>), dim3(G), dim3(block), 0, b->stream, a);
    else hipLaunchKernelGGL((k_batch_stage<false, double>), dim3(G), dim3(block), 0, b->stream, a);
  }
}

void batch_free(sph_batch *b) {
  if (b->gexec) (void)hipGraphExecDestroy(b->gexec);
  if (b->graph) (void)hipGraphDestroy(b->graph);
  void *p[] = {b->cnt, b->pid, b->nb, b->W, b->W32, b->rho, b->dt, b->orig, b->vec};
  for (void *q : p)
    if (q) (void)hipFree(q);
  if (b->stream) (void)hipStreamDestroy(b->stream);
  delete b;
}

// Time one 4-stage super-step per candidate block size (on the zeroed work
// vectors: the stage kernels' work does not depend on the values) and keep
// the fastest.
int batch_autotune(sph_batch *b) {
  hipEvent_t e0, e1;
  HIPCHK(hipEventCreate(&e0));
  HIPCHK(hipEventCreate(&e1));
  float best = 1e30f;
  for (int k = 0; k < 4; k++) {
    const int blk = kBatchBlocks[k];
    batch_super_step(b, 4, blk);  // warm-up
    float t = 1e30f;
    for (int r = 0; r < 3; r++) {
      tolerate(hipEventRecord(e0, b->stream));
      batch_super_step(b, 4, blk);
      tolerate(hipEventRecord(e1, b->stream));
      tolerate(hipEventSynchronize(e1));
      float ms = 0.f;
      tolerate(hipEventElapsedTime(&ms, e0, e1));
      t = std::min(t, ms);
    }
    b->tune_ms[k] = t;
    if (t < best) { best = t; b->block = blk; }
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  HIPCHK(hipGetLastError());
  HIPCHK(hipMemsetAsync(b->vec, 0, sizeof(double) * kBatchVecs * b->npad, b->stream));
  HIPCHK(hipStreamSynchronize(b->stream));
  return SPH_OK;
}
}  // namespace

extern "C" {

int sph_batch_create(sph_ctx *const *ctxs, int nprob, const double *K, int precision, sph_batch **out) {
  if (!ctxs || nprob < 1 || !K || !out || (precision != 0 && precision != 1)) return SPH_ERR_BAD_ARG;
  *out = nullptr;
  int64_t n = 0, np = 0;
  int maxc = 0;
  for (int p = 0; p < nprob; p++) {
    const sph_ctx *c = ctxs[p];
    if (!c || !c->diff_ready || c->device != ctxs[0]->device) return SPH_ERR_BAD_ARG;
    if (c->diff_op != 0) return SPH_ERR_UNSUPPORTED;  // the assembled Hessian operators only
    n += c->n;
    np += c->stats.n_pairs;
    maxc = std::max(maxc, c->last_maxcnt);
  }
  if (n >= ((int64_t)1 << 31) - 64) return SPH_ERR_UNSUPPORTED;  // 32-bit merged row indices
  sph_batch *b = new sph_batch();
  b->device = ctxs[0]->device;
  (void)hipSetDevice(b->device);
  b->nprob = nprob;
  b->precision = precision;
  b->n = n;
  b->n_pairs = np;
  b->npad = (n + 63) / 64 * 64;
  b->cap = std::max(8, (maxc + 7) / 8 * 8);
  b->off.assign(nprob + 1, 0);
  int rc = SPH_OK;
  auto fail = [&](int r) { batch_free(b); return r; };
  if (hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking) != hipSuccess) return fail(SPH_ERR_HIP);
  const size_t slots = (size_t)b->npad * b->cap;
  if ((rc = dalloc(&b->cnt, b->npad)) || (rc = dalloc(&b->pid, b->npad)) || (rc = dalloc(&b->nb, slots)) ||
      (rc = precision == 1 ? dalloc(&b->W32, slots) : dalloc(&b->W, slots)) || (rc = dalloc(&b->rho, b->npad)) ||
      (rc = dalloc(&b->dt, nprob)) ||
      (rc = dalloc(&b->orig, b->npad)) || (rc = dalloc(&b->vec, kBatchVecs * (size_t)b->npad)))
    return fail(rc);
  if (hipMemset(b->cnt, 0, sizeof(int) * b->npad) != hipSuccess ||
      hipMemset(b->dt, 0, sizeof(double) * nprob) != hipSuccess ||
      hipMemset(b->vec, 0, sizeof(double) * kBatchVecs * b->npad) != hipSuccess || hipDeviceSynchronize() != hipSuccess)
    return fail(SPH_ERR_HIP);
  // each problem's weights for its tensor, then into the merged layout
  for (int p = 0; p < nprob; p++) {
    sph_ctx *c = ctxs[p];
    ensure_weights(c, K + 6 * p);
    hipLaunchKernelGGL(k_batch_merge, dim3(grid_for(c->n)), dim3(kBlock), 0, c->stream, (int)c->n, (int)c->npad,
                       b->off[p], p, c->cnt, c->nb, c->dW, (const double4 *)c->q[0], c->orig[c->cur],
                       (int)b->npad, b->cnt, b->nb, b->W, b->W32, b->rho, b->orig, b->pid);
    if (hipStreamSynchronize(c->stream) != hipSuccess) return fail(SPH_ERR_HIP);
    b->off[p + 1] = b->off[p] + c->n;
  }
  if ((rc = batch_autotune(b))) return fail(rc);
  *out = b;
  return SPH_OK;
}

int sph_batch_destroy(sph_batch *b) {
  if (!b) return SPH_ERR_BAD_ARG;
  (void)hipSetDevice(b->device);
  (void)hipStreamSynchronize(b->stream);
  batch_free(b);
  return SPH_OK;
}

int sph_batch_sts(sph_batch *b, int s, const double *dt, int nsteps, const double *d_T0, double *d_T,
                  int use_graph) {
  if (!b || !dt || !d_T0 || !d_T || s < 2 || nsteps < 0) return SPH_ERR_BAD_ARG;
  (void)hipSetDevice(b->device);
  const unsigned G = grid_for(b->n);
  HIPCHK(hipMemcpyAsync(b->dt, dt, sizeof(double) * b->nprob, hipMemcpyHostToDevice, b->stream));
  hipLaunchKernelGGL(k_batch_in, dim3(G), dim3(kBlock), 0, b->stream, d_T0, b->orig, b->rho, (int)b->n, bvec(b, 0),
                     bvec(b, 5));
  if (use_graph && nsteps > 0) {
    if (!b->gexec || b->graph_s != s || b->graph_block != b->block) {
      if (b->gexec) { (void)hipGraphExecDestroy(b->gexec); b->gexec = nullptr; }
      if (b->graph) { (void)hipGraphDestroy(b->graph); b->graph = nullptr; }
      HIPCHK(hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal));
      batch_super_step(b, s, b->block);
      HIPCHK(hipStreamEndCapture(b->stream, &b->graph));
      HIPCHK(hipGraphInstantiate(&b->gexec, b->graph, nullptr, nullptr, 0));
      b->graph_s = s;
      b->graph_block = b->block;
    }
    for (int st = 0; st < nsteps; st++) HIPCHK(hipGraphLaunch(b->gexec, b->stream));
  } else {
    for (int st = 0; st < nsteps; st++) batch_super_step(b, s, b->block);
  }
  HIPCHK(hipGetLastError());
  hipLaunchKernelGGL(k_from_sorted, dim3(G), dim3(kBlock), 0, b->stream, bvec(b, 0), b->orig, (int)b->n, d_T);
  HIPCHK(hipStreamSynchronize(b->stream));
  return SPH_OK;
}

int sph_batch_get_info(sph_batch *b, sph_batch_info *info) {
  if (!b || !info) return SPH_ERR_BAD_ARG;
  info->n_problems = b->nprob;
  info->n_total = b->n;
  info->n_pairs = b->n_pairs;
  info->list_cap = b->cap;
  info->block = b->block;
  for (int k = 0; k < 4; k++) {
    info->tune_block[k] = kBatchBlocks[k];
    info->tune_ms[k] = b->tune_ms[k];
  }
  return SPH_OK;
}

int sph_batch_offsets(sph_batch *b, int64_t *off) {
  if (!b || !off) return SPH_ERR_BAD_ARG;
  for (int p = 0; p <= b->nprob; p++) off[p] = b->off[p];
  return SPH_OK;
}

}  // extern "C"
