// Device kernels of the two-loop hot path for gfx950.
//
// Data layout (HBM, "sorted order" = particles ordered by the Morton key of
// their cell, so a wavefront's 64 particles and their neighbours form compact
// 3D blobs):
//   fields     32 FP64 SoA arrays, one per store row (const.f90:69-103)
//   pm         double4 {x, y, z, m}: the 32-byte gather every pair makes
//   st (x2)    loop-1 Newton state {h, den, om, T}, ping-pong per sweep
//   q0..q4     loop-2 j-side packs (double4 for FP64, float4 for mixed)
//   nb         neighbour entries (sorted index << 5 | periodic image code) in
//              chunks of 8 per particle (nb_slot): one lane owns a 32-byte
//              sector per chunk; a wavefront reads a chunk with two
//              coalesced 16-byte loads per lane
//
// Every kernel maps one particle per lane (64-lane wavefronts, 256 per block)
// and remaps blockIdx so each XCD works on a contiguous range of the Morton
// curve (its particles' neighbours then stay in that XCD's L2).  Counters are
// reduced per block and added to one of 64 stripes, never all to one address.
#include <hip/hip_runtime.h>

#include <type_traits>

#include "sph_kernels.h"

namespace sph {

__device__ __forceinline__ int xcd_block(int b, int nb) {
  // blocks are dispatched round-robin over 8 XCDs; give XCD x the contiguous
  // logical range [x*q + min(x,r), ...) (a bijection on [0, nb))
  int x = b & 7, i = b >> 3, q = nb >> 3, r = nb & 7;
  return x * q + (x < r ? x : r) + i;
}

__device__ __forceinline__ int gtid() { return xcd_block(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x; }

// Order-preserving map of doubles onto int64 (for atomicMin/atomicMax).
__device__ __forceinline__ long long d2ord(double v) {
  long long b = __double_as_longlong(v);
  return b >= 0 ? b : (b ^ 0x7FFFFFFFFFFFFFFFLL);
}

template <typename T>
__device__ __forceinline__ T wave_sum(T v) {
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}
__device__ __forceinline__ long long wave_min(long long v) {
  for (int o = 32; o > 0; o >>= 1) { long long w = __shfl_xor(v, o, 64); v = w < v ? w : v; }
  return v;
}
__device__ __forceinline__ long long wave_max(long long v) {
  for (int o = 32; o > 0; o >>= 1) { long long w = __shfl_xor(v, o, 64); v = w > v ? w : v; }
  return v;
}
__device__ __forceinline__ int wave_or(int v) {
  for (int o = 32; o > 0; o >>= 1) v |= __shfl_xor(v, o, 64);
  return v;
}

// Block-wide reductions (256 threads = 4 wavefronts); every thread must call.
enum { R_SUM = 0, R_MIN = 1, R_MAX = 2, R_OR = 3 };
template <int OP>
__device__ __forceinline__ long long block_reduce(long long v, long long *lds) {
  if (OP == R_SUM) v = wave_sum(v);
  else if (OP == R_MIN) v = wave_min(v);
  else if (OP == R_MAX) v = wave_max(v);
  else v = wave_or((int)v);
  __syncthreads();
  if ((threadIdx.x & 63) == 0) lds[threadIdx.x >> 6] = v;
  __syncthreads();
  long long t = lds[0];
  for (int w = 1; w < (int)(blockDim.x >> 6); w++) {
    long long u = lds[w];
    if (OP == R_SUM) t += u;
    else if (OP == R_MIN) t = u < t ? u : t;
    else if (OP == R_MAX) t = u > t ? u : t;
    else t |= u;
  }
  return t;
}
__device__ __forceinline__ int stripe() { return blockIdx.x & (kStripes - 1); }

// Morton interleave of up to 10 (3D), 15 (2D) or 30 (1D) bits per axis.
__device__ __forceinline__ uint32_t spread3(uint32_t v) {
  v &= 0x3FFu;
  v = (v | (v << 16)) & 0x030000FFu;
  v = (v | (v << 8)) & 0x0300F00Fu;
  v = (v | (v << 4)) & 0x030C30C3u;
  v = (v | (v << 2)) & 0x09249249u;
  return v;
}
__device__ __forceinline__ uint32_t compact3(uint32_t v) {
  v &= 0x09249249u;
  v = (v | (v >> 2)) & 0x030C30C3u;
  v = (v | (v >> 4)) & 0x0300F00Fu;
  v = (v | (v >> 8)) & 0x030000FFu;
  v = (v | (v >> 16)) & 0x3FFu;
  return v;
}
__device__ __forceinline__ uint32_t spread2(uint32_t v) {
  v &= 0x7FFFu;
  v = (v | (v << 8)) & 0x00FF00FFu;
  v = (v | (v << 4)) & 0x0F0F0F0Fu;
  v = (v | (v << 2)) & 0x33333333u;
  v = (v | (v << 1)) & 0x55555555u;
  return v;
}
__device__ __forceinline__ uint32_t compact2(uint32_t v) {
  v &= 0x55555555u;
  v = (v | (v >> 1)) & 0x33333333u;
  v = (v | (v >> 2)) & 0x0F0F0F0Fu;
  v = (v | (v >> 4)) & 0x00FF00FFu;
  v = (v | (v >> 8)) & 0x7FFFu;
  return v;
}
__device__ __forceinline__ uint32_t morton_key(const int c[3], int mdim) {
  if (mdim == 3) return spread3(c[0]) | (spread3(c[1]) << 1) | (spread3(c[2]) << 2);
  if (mdim == 2) return spread2(c[0]) | (spread2(c[1]) << 1);
  return (uint32_t)c[0];
}
// Sort key of a cell: its row (y, z) times nx plus x, so the cells of a row
// are consecutive keys; in 3D the rows follow each other in square tiles of
// the (y, z) plane, which keeps the neighbours of a block of particles close
// in memory.
__device__ __forceinline__ uint32_t row_of(const Grid &g, int cy, int cz) {
  constexpr int L = kRowTileLog2, T = kRowTile;
  // 24-bit products: tiles and rows stay below 2^24 under the axis caps
  if (g.mdim == 3)
    return (__umul24((uint32_t)(cz >> L), (uint32_t)g.ty) + (uint32_t)(cy >> L)) * (uint32_t)(T * T) +
           (uint32_t)(((cz & (T - 1)) << L) | (cy & (T - 1)));
  return (uint32_t)cy;
}

__device__ __forceinline__ void morton_cell(uint32_t k, int mdim, int c[3]) {
  if (mdim == 3) { c[0] = compact3(k); c[1] = compact3(k >> 1); c[2] = compact3(k >> 2); }
  else if (mdim == 2) { c[0] = compact2(k); c[1] = compact2(k >> 1); c[2] = 0; }
  else { c[0] = (int)k; c[1] = 0; c[2] = 0; }
}

__device__ __forceinline__ void cell_of(const Grid &g, double x, double y, double z, int c[3]) {
  double p[3] = {x, y, z};
#pragma unroll
  for (int a = 0; a < 3; a++) {
    double t = floor((p[a] - g.lo[a]) * g.inv_cs[a]);
    // clamped to the grid; a NaN coordinate takes cell 0
    c[a] = !(t >= 0.0) ? 0 : (t >= (double)g.nc[a] ? g.nc[a] - 1 : (int)t);
  }
}

// ---------------------------------------------------------------- set-up
// reflectRealParticlesPeriodically (BC.f90:102-194) for real particles, then
// the bounding box of all particles and the largest h of the real ones.
// red[v * kStripes + stripe]: v = 0..2 min, 3..5 max (ordered int64), 6 max h.
__global__ __launch_bounds__(256) void k_wrap_reduce(Fields F, const int64_t *orig, int n, int64_t n_real, Box b,
                                                     int dim, long long *red, int *err) {
  __shared__ long long lds7[7 * 4];  // per-wave partials of the 256-thread block
  long long mn[3], mx[3], hm;
  const long long BIG = 0x7FFFFFFFFFFFFFFFLL, SMALL = -0x7FFFFFFFFFFFFFFFLL - 1;
  for (int a = 0; a < 3; a++) { mn[a] = BIG; mx[a] = SMALL; }
  hm = SMALL;
  bool outside = false, wrapped = false;
  // grid-stride: a few blocks' worth of particles per block keeps the
  // blocks' closing atomics few
  for (int s = gtid(); s < n; s += (int)(gridDim.x * blockDim.x)) {
    bool real = orig[s] < n_real;
    double x[3] = {F.f[F_RX][s], F.f[F_RY][s], F.f[F_RZ][s]};
    if (real && b.wrap_mask) {
      double dxmin[3], dxmax[3];
      for (int a = 0; a < 3; a++) { dxmin[a] = b.emin[a] - x[a]; dxmax[a] = x[a] - b.emax[a]; }
      for (int a = 0; a < dim; a++) {
        if (!((b.wrap_mask >> a) & 1)) continue;
        double nx = x[a];
        if (dxmax[a] > 0) nx = b.emin[a] + dxmax[a];
        if (dxmin[a] > 0) nx = b.emax[a] - dxmin[a];
        if (nx != x[a]) { x[a] = nx; F.f[F_RX + a][s] = nx; wrapped = true; }
      }
    }
    for (int a = 0; a < dim; a++)
      if (((b.ghost_mask >> a) & 1) && !(x[a] >= b.emin[a] && x[a] <= b.emax[a])) outside = true;
    for (int a = 0; a < 3; a++) {
      const long long o = d2ord(x[a]);
      mn[a] = o < mn[a] ? o : mn[a];
      mx[a] = o > mx[a] ? o : mx[a];
    }
    if (real) {
      const long long o = d2ord(F.f[F_H][s]);
      hm = o > hm ? o : hm;
    }
  }
  if (outside) atomicOr(err, EB_OUTSIDE);
  if (wrapped) atomicOr(err + 4, 1);  // positions changed: cell keys must be recomputed
  // the seven block reductions in one pass: per wave, then one barrier, then
  // thread v combines value v over the block's waves
  long long w7[7] = {wave_min(mn[0]), wave_min(mn[1]), wave_min(mn[2]), wave_max(mx[0]), wave_max(mx[1]),
                     wave_max(mx[2]), wave_max(hm)};
  const int nw = (int)(blockDim.x >> 6);
  if ((threadIdx.x & 63) == 0)
    for (int v = 0; v < 7; v++) lds7[v * 4 + (threadIdx.x >> 6)] = w7[v];
  __syncthreads();
  if (threadIdx.x < 7) {
    const int v = threadIdx.x;
    long long t = lds7[v * 4];
    for (int w = 1; w < nw; w++) {
      const long long u = lds7[v * 4 + w];
      t = v < 3 ? (u < t ? u : t) : (u > t ? u : t);
    }
    if (v < 3) atomicMin(&red[v * kStripes + stripe()], t);
    else atomicMax(&red[v * kStripes + stripe()], t);
  }
}

// sorted_flag: the order check's flag, cleared here for k_check_sorted
// (which follows on the stream)
__global__ void k_keys(Fields F, int n, Grid g, uint32_t *keys, uint32_t *vals, int *sorted_flag) {
  int s = gtid();
  if (blockIdx.x == 0 && threadIdx.x == 0) *sorted_flag = 0;
  if (s >= n) return;
  int c[3];
  cell_of(g, F.f[F_RX][s], F.f[F_RY][s], F.f[F_RZ][s], c);
  keys[s] = row_of(g, c[1], c[2]) * (uint32_t)g.nc[0] + (uint32_t)c[0];
  vals[s] = (uint32_t)s;
}

// Grid-stride over few blocks, one atomic per block that found a break
// (after particles move nearly every wave finds one, and atomics on one word
// serialise: one per wave took 130 us for a million particles).
__global__ __launch_bounds__(256) void k_check_sorted(const uint32_t *keys, int n, int *flag) {
  bool bad = false;
  for (int s = gtid(); s + 1 < n; s += gridDim.x * blockDim.x) bad = bad || keys[s] > keys[s + 1];
  if (__syncthreads_or(bad) && threadIdx.x == 0) atomicOr(flag, 1);
}

// Gather the first nf fields into the new order: dst[s] = src[perm[s]], and
// the original indices with them when orig_dst is given.
__global__ void k_reorder(Fields src, Fields dst, const uint32_t *perm, const int64_t *orig_src,
                          int64_t *orig_dst, int n, int nf) {
  int s = gtid();
  if (s >= n) return;
  uint32_t p = perm[s];
  // 16 columns' loads issued before their stores (a store may alias a later
  // column's load for the compiler, which would otherwise wait on each load)
  constexpr int kGroup = 16;
#pragma unroll
  for (int f0 = 0; f0 < F_TOTAL; f0 += kGroup) {
    double v[kGroup];
#pragma unroll
    for (int k = 0; k < kGroup; k++) v[k] = f0 + k < nf ? src.f[f0 + k][p] : 0.0;
#pragma unroll
    for (int k = 0; k < kGroup; k++)
      if (f0 + k < nf) dst.f[f0 + k][s] = v[k];
  }
  if (orig_dst) orig_dst[s] = orig_src[p];
}

// First sorted index of every cell key (cstart[K] = number of particles with
// a key below K, for every K in [0, nkeys]); index s fills the keys between
// its predecessor's and its own.  Short ranges are filled by their own lane;
// long ones (empty rows: the tiles of the key order pad each axis to a
// multiple of their side) by the whole wave, one range at a time.
__global__ __launch_bounds__(256) void k_cell_start(const uint32_t *keys, int n, uint32_t nkeys, int *cstart) {
  const int s = gtid();
  const int lane = threadIdx.x & 63;
  int64_t lo = 0, hi = -1;  // keys [lo, hi] take the value s
  if (s <= n) {
    lo = (s == 0 ? -1 : (int64_t)keys[s - 1]) + 1;
    hi = s == n ? (int64_t)nkeys : (int64_t)keys[s];
  }
  constexpr int64_t kShort = 8;
  const bool lng = hi - lo >= kShort;
  if (!lng)
    for (int64_t k = lo; k <= hi; k++) cstart[k] = s;
  unsigned long long m = __ballot(lng);
  while (m) {
    const int l = __ffsll((long long)m) - 1;
    m &= m - 1;
    const int64_t l0 = __shfl(lo, l), l1 = __shfl(hi, l);
    const int v = __shfl(s, l);
    for (int64_t k = l0 + lane; k <= l1; k += 64) cstart[k] = v;
  }
}

// Positions/masses, search h, the FP32 search record (coordinates relative
// to the grid origin + ghost-band bits), ghost-band codes (BC.f90:196-309: strict
// padding test for a particle's own band, the looser one a chained ghost
// gets, BC.f90:343-353), h at entry and T den for loop 1 (circuit1.f90:105-109),
// and c1simple's start state when st is given (c1advanced's first sweep reads
// h, den, om and T from the fields and takes every real particle as active,
// so eps needs no start value).
// pos = 0: the position records (pm, real, pf, pfix, band codes) of the
// previous prep still hold (same positions, order and grid) and are not
// rewritten.
__device__ __forceinline__ void prep_one(Fields F, const int64_t *orig, int s, int64_t n_real, Box b, Grid g,
                                         int dim, double4 *pm, double *h0, float4 *pf, uint8_t *real, double4 *st,
                                         double *tr, uint3 *pfix, const FixFrame &ff, float2 *s1, double *eps,
                                         int pos) {
  double h = F.f[F_H][s];
  h0[s] = h;
  // the start state for c1simple (c1advanced's first sweep reads the fields)
  if (st) st[s] = make_double4(h, F.f[F_DEN][s], F.f[F_OM][s], F.f[F_T][s]);
  (void)eps;
  // T den for grad T: FP64 in tr, or with m in FP32 on the mixed path
  if (s1) s1[s] = make_float2((float)F.f[F_M][s], (float)(F.f[F_T][s] * F.f[F_DEN][s]));
  else tr[s] = F.f[F_T][s] * F.f[F_DEN][s];
  if (!pos) return;
  const bool r = orig[s] < n_real;
  double x = F.f[F_RX][s], y = F.f[F_RY][s], z = F.f[F_RZ][s];
  pm[s] = make_double4(x, y, z, F.f[F_M][s]);
  real[s] = r ? 1 : 0;
  if (pfix) {  // the mixed path's position record: 31-bit fractions (the wrap keeps differences)
    const double xs[3] = {x, y, z};
    uint32_t u[3];
#pragma unroll
    for (int a = 0; a < 3; a++) u[a] = (uint32_t)(long long)floor((xs[a] - ff.base[a]) * ff.mul[a]);
    pfix[s] = make_uint3(u[0], u[1], u[2]);
  }
  int ty = (int)F.f[F_TYPE][s];
  uint32_t code = 0;
  if (ty == 1 || ty == 2 || ty == 21) {
    double xs[3] = {x, y, z};
    for (int a = 0; a < dim; a++) {
      if (!((b.ghost_mask >> a) & 1)) continue;
      double v = xs[a];
      uint32_t st_ = 0, lo = 0;
      if (v < b.boxmin[a]) {
        if (v > b.pmin[a]) st_ = 1;
        lo = 1;
      } else if (v > b.boxmax[a]) {
        if (v < b.pmax[a]) st_ = 2;
        lo = 2;
      }
      code |= (st_ << (2 * a)) | (lo << (6 + 2 * a));
    }
  }
  pf[s] = make_float4((float)(x - g.lo[0]), (float)(y - g.lo[1]), (float)(z - g.lo[2]), __uint_as_float(code));
}

// flags: OR of bit 0 = some velocity is non-zero, bit 1 = some B differs from
// particle 0's (loop 2 skips terms that are exactly zero otherwise; neither
// changes before loop 2).
// zero[0, nzero): the step's counters, cleared by the first block before the
// kernels that count into them run (they all follow k_prep on its stream);
// zero2[0, nzero2): the counter regions of the step's first Newton sweeps,
// cleared by the whole grid.
__global__ void k_prep(Fields F, const int64_t *orig, int n, int64_t n_real, Box b, Grid g, int dim, double4 *pm,
                       double *h0, float4 *pf, uint8_t *real, double4 *st, double *tr, uint3 *pfix, FixFrame ff,
                       float2 *s1, double *eps, long long *flags, long long *zero, int nzero, long long *zero2,
                       int nzero2, int pos) {
  __shared__ long long lds[4];
  int s = gtid();
  if (blockIdx.x == 0)
    for (int k = threadIdx.x; k < nzero; k += blockDim.x) zero[k] = 0;
  for (int k = s; k < nzero2; k += gridDim.x * blockDim.x) zero2[k] = 0;
  int fl = 0;
  if (s < n) {
    if (flags) {  // null: the caller still holds the flags of unchanged v and B
      const double vx = F.f[F_VX][s], vy = F.f[F_VY][s], vz = F.f[F_VZ][s];
      if (vx != 0.0 || vy != 0.0 || vz != 0.0) fl |= 1;
      if (!(F.f[F_BX][s] == F.f[F_BX][0] && F.f[F_BY][s] == F.f[F_BY][0] && F.f[F_BZ][s] == F.f[F_BZ][0])) fl |= 2;
    }
    prep_one(F, orig, s, n_real, b, g, dim, pm, h0, pf, real, st, tr, pfix, ff, s1, eps, pos);
  }
  if (flags) {
    long long t = block_reduce<R_OR>(fl, lds);
    if (threadIdx.x == 0 && t) atomicOr((unsigned long long *)&flags[stripe()], (unsigned long long)t);
  }
}

// Does the reference hold a ghost of j shifted by (sx, sy, sz)?  The first
// shifted axis uses j's own (strict) band, later axes the chained test
// (BC.f90:342-355).  For fixed shifts that is (band & m) == v, with m, v set
// here; no shift: m = v = 0.
__device__ __forceinline__ void image_mask(int sx, int sy, int sz, uint32_t &m, uint32_t &v) {
  const int sh[3] = {sx, sy, sz};
  bool first = true;
  m = 0;
  v = 0;
#pragma unroll
  for (int a = 0; a < 3; a++) {
    if (sh[a] == 0) continue;
    const uint32_t want = sh[a] > 0 ? 1u : 2u;
    const int pos = first ? 2 * a : 6 + 2 * a;
    m |= 3u << pos;
    v |= want << pos;
    first = false;
  }
}

// Squared separation exactly as kdtree2.f90:1673-1677 rounds it:
// sd = ((0 + dx^2) + dy^2) + dz^2 with dx = x_j - x_i, no fused multiply-add.
__device__ __forceinline__ double kd_sqdist(double xj, double yj, double zj, double xi, double yi, double zi) {
#pragma clang fp contract(off)
  double d = xj - xi;
  double sd = d * d;
  d = yj - yi;
  sd = sd + d * d;
  d = zj - zi;
  sd = sd + d * d;
  return sd;
}

// findneighboursKDT (neighboursearch.f90:74-194) as a cell-list search:
// particles are sorted so that the cells of one (y, z) row lie contiguously
// along x.  Every real i visits the rows within r_max in y and z (rows wrap
// with an image shift along imaged axes) and, in each, the contiguous range
// of records of the cells its sphere can reach along x at that row's distance
// (wrapped parts carry their shift), and tests them with the kd-tree's FP64
// inclusion rule sd <= (krad*h_i)^2; self (unshifted) is dropped.
//
// Candidates are 16-byte records (pf: FP32 coordinates relative to the grid
// origin, and the band bits).  The FP32 squared distance decides every
// candidate farther than the error band around the radius:
// |sd32 - |d|^2| <= 2 sqrt3 r delta + 3 delta^2 + 2^-20 sd with delta =
// a.delta (the FP32 coordinate error bound, 2^-21 x the grid span).
// Candidates inside the band, and any lane whose radius is not a finite
// number, take the exact FP64 rule on pm, so the sets are those of the FP64
// test.
//
// Two traversals give the same lists, entry for entry:
// - extended grid (the default): the periodic images live in ghost cells
//   around the real grid (built by k_ext_count / k_ext_fill whenever the
//   particles or the grid change), so each stencil row is one contiguous
//   range of records along x whose coordinates already carry the image shift,
//   a missing image is a NaN x and the list entry is the record's fourth
//   word.  Pre-shifting rounds q + s L once where the per-lane traversal
//   rounds p - s L: the same count of 2^-24 S roundings, so the FP32 bound
//   holds.
// - per lane (SPH_NEIGH_TRAVERSAL=lane, or no extended grid): each lane splits
//   a row's x range into the part left of the box, inside it and right of it,
//   and applies the shift and the image test to every record it reads.
// pending list entries per lane in k_neigh (a power of two, whole chunks)
constexpr int kNbRing = 16;
// records the extended-grid scan reads and tests together
#ifndef SPH_SCAN_BATCH
#define SPH_SCAN_BATCH 4
#endif
constexpr int kScanBatch = SPH_SCAN_BATCH;
static_assert(kScanBatch == 4 || kScanBatch == 8, "the ring holds 7 pending entries plus one batch");
// The rings of a block's 256 lanes are interleaved slot by slot: slot k of
// lane l is word 256 k + l, so when the 64 lanes of a wavefront write their
// next slots they touch every bank once per 32 lanes, whatever slot each is
// at (rings of one lane each, 64 bytes apart, put up to 32 lanes on a bank).
// A lane's cursor counts its entries in slot rows of 1024 bytes: the next
// slot's byte offset is (cursor & 0x3C00) | 4 threadIdx.x.
constexpr int kRingRow = 1024;  // bytes per slot row: 256 lanes x 4
constexpr int kRingShift = 10;
static_assert(kNbRing * kRingRow == 0x4000 && kRingRow == 256 * 4, "ring_slot masks 16 slot rows");
__device__ __forceinline__ uint32_t ring_slot(int cur, uint32_t lane_off) {
  uint32_t r;
  asm("v_and_or_b32 %0, %1, %2, %3" : "=v"(r) : "v"(cur), "s"(0x3C00), "v"(lane_off));
  return r;
}

#ifdef SPH_NEIGH_WAVES
#define SPH_NEIGH_OCC __attribute__((amdgpu_waves_per_eu(SPH_NEIGH_WAVES)))
#else
#define SPH_NEIGH_OCC
#endif
template <int DIM, bool EXT>
__global__ __launch_bounds__(256) SPH_NEIGH_OCC void k_neigh(NeighArgs a, const double4 *__restrict__ pm,
                                               const float4 *__restrict__ pf, uint32_t *__restrict__ nb,
                                               int *__restrict__ cnt) {
  __shared__ long long lds[4];
  __shared__ __align__(16) uint32_t nb_ring[256 * kNbRing];
  const int s = gtid();
  const bool live = s < a.n && a.real[s];
  const int sl = s < a.n ? s : 0;  // lanes past n read particle 0 and search nothing
  const double4 pi = pm[sl];
  const float4 pif = pf[sl];
  const double rr = a.krad * a.hs[sl];
  const double r2 = rr * rr;
  const double r2m = r2 * (1.0 + 1e-6);  // pruning bound, never tighter than the exact test
  // FP32 decision thresholds; a radius that is not a finite number sends
  // every candidate to the exact test
  const double mabs = 2.0 * 1.7320508075688772 * 1.01 * rr * a.delta + 3.0 * a.delta * a.delta;
  const bool finite_r = rr == rr && rr < HUGE_VAL;
  const float hi_thr = finite_r ? (float)(r2 * (1.0 + 1e-5) + mabs) : HUGE_VALF;
  const float lo_thr = finite_r ? (float)(r2 * (1.0 - 1e-5) - mabs) : -HUGE_VALF;
  // squared radius of the row culling (grown by 2e-5)
  const float r2g = finite_r ? (float)r2m * 1.00002f : HUGE_VALF;
  // entries [done, count) wait in the lane's ring and leave as whole 32-byte
  // chunks (see the traversals for when); cur = count x kRingRow
  const uint32_t lane_off = (uint32_t)threadIdx.x * 4;
  char *const ring_bytes = reinterpret_cast<char *>(nb_ring);
  int cur = 0;
  int done = 0;
  auto flush_chunk = [&]() {
    if (done < a.cap) {  // cap is whole chunks: the chunk fits
      // slots (done mod 16) .. + 7 of this lane, 256 words apart
      const uint32_t *src = nb_ring + (done & (kNbRing - 1)) * 256 + threadIdx.x;
      uint4 *dst = reinterpret_cast<uint4 *>(nb + nb_slot(done, s, a.npad));
      const uint4 lo = make_uint4(src[0], src[256], src[512], src[768]);
      const uint4 hi = make_uint4(src[1024], src[1280], src[1536], src[1792]);
      dst[0] = lo;
      dst[1] = hi;
    }
    done += 8;
  };
  int ci[3];
  cell_of(a.g, pi.x, pi.y, pi.z, ci);
  const int nx = a.g.nc[0];

  // ------------------------------------------------ per-lane traversal
  auto lane_rows = [&]() {
    // every candidate of one contiguous record range [t0, t1) of one image
    // shift: FP32 decision, the exact FP64 rule inside the band.  IMG is
    // false for the unshifted records, which always exist and never hold the
    // lane's own record (the caller splits the range around it); IMG true
    // tests the records' band bits for the image's existence.  One code path
    // for full and partial batches keeps the wave's lanes together.  Entries
    // leave the ring after each stencil row (the same point for every lane,
    // so the lanes' chunk stores share instructions) and when it is full.
    auto scan = [&](auto img, int t0, int t1, int sx, int sy, int sz) {
      constexpr bool IMG = decltype(img)::value;
      const uint32_t code = (uint32_t)((sx + 1) + 3 * (sy + 1) + 9 * (sz + 1));
      // the image shift in the FP32 frame, and the lane position in it
      const float ox = pif.x - (float)sx * a.lenf[0], oy = pif.y - (float)sy * a.lenf[1],
                  oz = pif.z - (float)sz * a.lenf[2];
      // image existence as (band & bm) == bv
      uint32_t bm = 0, bv = 0;
      if (IMG) image_mask(sx, sy, sz, bm, bv);
      for (int t = t0; t < t1; t += 4) {
        // records t..t+3 (pf holds 4 records of padding past n; the ones
        // past t1 belong to other cells and are masked)
        const int rem = t1 - t;
        const float4 *p = pf + t;
        float4 q[4];
#pragma unroll
        for (int u = 0; u < 4; u++) q[u] = p[u];
        const uint32_t e0 = nb_pack((uint32_t)t, code);
#pragma unroll
        for (int u = 0; u < 4; u++) {
          bool ok = u < rem;
          if (IMG) ok = ok && (__float_as_uint(q[u].w) & bm) == bv;
          const float fx = q[u].x - ox, fy = q[u].y - oy, fz = q[u].z - oz;
          const float sd32 = fx * fx + fy * fy + fz * fz;
          bool hit = ok && sd32 <= hi_thr;
          if (hit && !(sd32 < lo_thr)) {
            const double4 pj = pm[t + u];
            double xj = pj.x, yj = pj.y, zj = pj.z;
            if (IMG) {
              xj = image_coord(xj, sx, a.b.emin[0], a.b.emax[0]);
              yj = image_coord(yj, sy, a.b.emin[1], a.b.emax[1]);
              zj = image_coord(zj, sz, a.b.emin[2], a.b.emax[2]);
            }
            hit = kd_sqdist(xj, yj, zj, pi.x, pi.y, pi.z) <= r2;
          }
          // every lane writes the next free ring slot and keeps it on a hit
          // (at most kNbRing - 1 entries are pending here, so that slot
          // holds none of them)
          *reinterpret_cast<uint32_t *>(ring_bytes + ring_slot(cur, lane_off)) =
              e0 + ((uint32_t)u << kCodeBits);
          cur += hit ? kRingRow : 0;
        }
        // at most 11 pending before the next batch's 4 candidates
        if ((cur >> kRingShift) - done > kNbRing - 5) flush_chunk();
      }
    };
    using Img = std::true_type;
    using Own = std::false_type;
    // row geometry in FP32 and in cell units, widened by margins that only
    // add rows and cells: positions shrink towards the row by mcell, the
    // squared radius grows by 2e-5
    const float pcx = (float)((pi.x - a.g.lo[0]) * a.g.inv_cs[0]);
    const float pcy = (float)((pi.y - a.g.lo[1]) * a.g.inv_cs[1]);
    const float pcz = (float)((pi.z - a.g.lo[2]) * a.g.inv_cs[2]);
    const float csy = (float)a.g.cs[1], csz = (float)a.g.cs[2], icx = (float)a.g.inv_cs[0];
    const int rz = DIM == 3 ? a.rz : 0, ry = DIM >= 2 ? a.ry : 0;
    for (int dz = -rz; dz <= rz; dz++) {
      // the rows (cy, cz) with their image shifts; rows outside a non-imaged
      // axis, or farther than the radius, are skipped
      int cz = ci[2] + dz, sz = 0;
      bool usez = true;
      float gz2 = 0.f;
      if (DIM == 3) {
        const float t = (float)cz;
        gz2 = fmaxf(fmaxf(t - pcz, pcz - (t + 1.f)) - a.mcell[2], 0.f) * csz;
        gz2 *= gz2;
        if (a.g.periodic[2]) {
          if (cz < 0) { cz += a.g.nc[2]; sz = -1; }
          else if (cz >= a.g.nc[2]) { cz -= a.g.nc[2]; sz = 1; }
        } else if (cz < 0 || cz >= a.g.nc[2]) usez = false;
      }
      for (int dy = -ry; dy <= ry; dy++) {
        int cy = ci[1] + dy, sy = 0;
        bool use = usez;
        float gyz = gz2;
        if (DIM >= 2) {
          const float t = (float)cy;
          const float gy = fmaxf(fmaxf(t - pcy, pcy - (t + 1.f)) - a.mcell[1], 0.f) * csy;
          gyz += gy * gy;
          if (a.g.periodic[1]) {
            if (cy < 0) { cy += a.g.nc[1]; sy = -1; }
            else if (cy >= a.g.nc[1]) { cy -= a.g.nc[1]; sy = 1; }
          } else if (cy < 0 || cy >= a.g.nc[1]) use = false;
        }
        use = use && gyz <= r2g;
        if (use) {
          // the sphere's x extent at this row: cells [xl, xh] before
          // wrapping, at most one wrap away (a radius that is not finite
          // takes the whole stencil); the hardware square root's ulp is far
          // inside the margin mcell
          const float hxc =
              fminf(__builtin_amdgcn_sqrtf(fmaxf(r2g - gyz, 0.f)) * icx + a.mcell[0], (float)(a.rx + 1));
          int xl = (int)floorf(pcx - hxc);
          int xh = (int)floorf(pcx + hxc);
          xl = max(xl, ci[0] - a.rx);
          xh = min(xh, ci[0] + a.rx);
          if (a.g.periodic[0]) {  // at most one wrap
            xl = max(xl, -nx);
            xh = min(xh, 2 * nx - 1);
          }
          // first key of the row (rows and nx stay below 2^24 under the
          // axis caps: 1024 cells per axis in 3D, 32768 in 2D)
          const uint32_t row = __umul24(row_of(a.g, cy, cz), (uint32_t)nx);
          // every range bound of the row is loaded before the first scan, so
          // the loads overlap: the part left of the box (images shifted by
          // -L) or else the part right of it (+L) into w0, w1, and the part
          // inside into t0, t1
          const bool wl = a.g.periodic[0] && xl < 0, wr = a.g.periodic[0] && xh >= nx;
          int w0 = 0, w1 = 0;
          if (wl) {
            w0 = a.cstart[row + xl + nx];
            w1 = a.cstart[row + min(xh, -1) + nx + 1];
          } else if (wr) {
            w0 = a.cstart[row + max(xl, nx) - nx];
            w1 = a.cstart[row + xh - nx + 1];
          }
          const int l = max(xl, 0), h = min(xh, nx - 1);
          int t0 = 0, t1 = 0;
          if (l <= h) {
            t0 = a.cstart[row + l];
            t1 = a.cstart[row + h + 1];
          }
          if (wl) scan(Img{}, w0, w1, -1, sy, sz);
          if (l <= h) {
            if ((sy | sz) != 0) {
              scan(Img{}, t0, t1, 0, sy, sz);
            } else if (dy == 0 && dz == 0) {
              // the lane's own row: the records on either side of its own
              const int m = min(max(s, t0), t1);
              scan(Own{}, t0, m, 0, 0, 0);
              scan(Own{}, m == s ? m + 1 : m, t1, 0, 0, 0);
            } else {
              scan(Own{}, t0, t1, 0, 0, 0);
            }
          }
          if (wr) {
            if (wl) {  // both sides (a box under two stencils wide): load the right part now
              w0 = a.cstart[row + max(xl, nx) - nx];
              w1 = a.cstart[row + xh - nx + 1];
            }
            scan(Img{}, w0, w1, 1, sy, sz);
          }
        }
        while ((cur >> kRingShift) - done >= 8) flush_chunk();
      }
    }
  };

  // ------------------------------------------------ extended-grid traversal
  auto ext_rows = [&]() {
    const uint32_t own_e = nb_pack((uint32_t)s, kNoImage);
    // 32-bit byte offsets from the arrays' bases (the host keeps both arrays
    // under 4 GiB), so every load takes one offset register
    const char *const erec_b = reinterpret_cast<const char *>(a.erec);
    const char *const ecs_b = reinterpret_cast<const char *>(a.ecs);
    // the cursor at which 8 entries are pending; a list far past the
    // reference's cap (at most 2^20) stops early, since the step fails on it
    // anyway, which keeps the cursor in range
    int flush_at = 8 * kRingRow;
    const int stop_at = (a.maxneib + 8) * kRingRow;
    // every record of [t0, t1), kScanBatch at a time (a row's records,
    // usually fewer, are read together; erec holds 8 records of padding past
    // its end): FP32 decisions first, the exact FP64 rule for those inside
    // the band behind one branch, then the ring writes.  OWN (the lane's own
    // row) drops its own record.  No branch per candidate.
    auto scan = [&](auto own, int t0, int t1) {
      constexpr bool OWN = decltype(own)::value;
      uint32_t off = (uint32_t)t0 * (uint32_t)sizeof(float4);
      for (int rem = t1 - t0; rem > 0; rem -= kScanBatch, off += kScanBatch * (uint32_t)sizeof(float4)) {
        const float4 *p = reinterpret_cast<const float4 *>(erec_b + off);
        float4 q[kScanBatch];
#pragma unroll
        for (int u = 0; u < kScanBatch; u++) q[u] = p[u];
        // hit: inside the band's lower edge; near: inside its upper edge
        bool hit[kScanBatch], near[kScanBatch];
        bool any_band = false;
#pragma unroll
        for (int u = 0; u < kScanBatch; u++) {
          const float fx = q[u].x - pif.x, fy = q[u].y - pif.y, fz = q[u].z - pif.z;
          const float sd32 = fx * fx + fy * fy + fz * fz;
          bool ok = u < rem;
          if (OWN) ok = ok && __float_as_uint(q[u].w) != own_e;
          hit[u] = ok && sd32 < lo_thr;
          near[u] = ok && sd32 <= hi_thr;
          any_band = any_band || (hit[u] != near[u]);
        }
        if (any_band) {
#pragma unroll
          for (int u = 0; u < kScanBatch; u++) {
            if (hit[u] != near[u]) {
              const uint32_t e = __float_as_uint(q[u].w);
              const double4 pj = pm[nb_index(e)];
              double xj = pj.x, yj = pj.y, zj = pj.z;
              const uint32_t code = nb_code(e);
              if (code != kNoImage) image_shift(xj, yj, zj, code, a.b);
              hit[u] = kd_sqdist(xj, yj, zj, pi.x, pi.y, pi.z) <= r2;
            }
          }
        }
        // every lane writes the next free ring slot and keeps it on a hit
        // (at most 7 + u entries are pending here: the slot is free)
#pragma unroll
        for (int u = 0; u < kScanBatch; u++) {
          *reinterpret_cast<uint32_t *>(ring_bytes + ring_slot(cur, lane_off)) = __float_as_uint(q[u].w);
          cur += hit[u] ? kRingRow : 0;
        }
        if (cur >= flush_at) {  // at most 7 pending after it
          flush_chunk();
          flush_at += 8 * kRingRow;
          if (cur >= stop_at) break;
        }
      }
    };
    // Row geometry in FP32 cell units, widened by margins that only add rows
    // and cells (as in the per-lane traversal).  x counts extended cells, in
    // which a lane inside the grid sits at x >= gh0 >= rx, so truncation
    // floors every bound the clamps keep.  On y and z the gap between the
    // lane and the cells at offset d is max(d - (f + m), -d - (1 - f + m), 0)
    // cells, f the lane's offset inside its own cell and m the margin.
    // Walled axes have empty ghost cells, so no row or cell needs a bounds
    // test.
    const float pcx = (float)((pi.x - a.g.lo[0]) * a.g.inv_cs[0] + (double)a.gh[0]);
    const float pcy = (float)((pi.y - a.g.lo[1]) * a.g.inv_cs[1]);
    const float pcz = (float)((pi.z - a.g.lo[2]) * a.g.inv_cs[2]);
    const float fy = pcy - (float)ci[1], fz = pcz - (float)ci[2];
    const float ylo = fy + a.mcell[1], yhi = (1.f - fy) + a.mcell[1];
    const float zlo = fz + a.mcell[2], zhi = (1.f - fz) + a.mcell[2];
    const float csy = (float)a.g.cs[1], csz = (float)a.g.cs[2], icx = (float)a.g.inv_cs[0];
    const float hcap = (float)(a.rx + 1);
    const int rz = DIM == 3 ? a.rz : 0, ry = DIM >= 2 ? a.ry : 0;
    const int xlo = ci[0] + a.gh[0] - a.rx, xhi = ci[0] + a.gh[0] + a.rx;
    auto gap = [](float d, float lo, float hi, float cs) { return fmaxf(fmaxf(d - lo, -d - hi), 0.f) * cs; };
    // byte offset in ecs of cell 0 of row (dy, dz), stepped along the rows
    const uint32_t row_bytes = 4u * (uint32_t)a.ne[0];
    uint32_t rowb = (uint32_t)((ci[2] - rz + a.gh[2]) * a.ne[1] + ci[1] - ry + a.gh[1]) * row_bytes;
    const uint32_t plane_step = (uint32_t)(a.ne[1] - 2 * ry) * row_bytes;
    float gz2 = 0.f;
    if (DIM == 3) {
      const float gz = gap((float)-rz, zlo, zhi, csz);
      gz2 = gz * gz;
    }
    // the record range of row (dy, dz); a row farther than the radius is empty
    auto row_range = [&](int dy, int &t0, int &t1) {
      float gyz = gz2;
      if (DIM >= 2) {
        const float gy = gap((float)dy, ylo, yhi, csy);
        gyz = fmaf(gy, gy, gz2);
      }
      const float hxc = fminf(__builtin_amdgcn_sqrtf(fmaxf(r2g - gyz, 0.f)) * icx + a.mcell[0], hcap);
      // both bounds inside the stencil's cells [xlo, xhi] (xh one below xl
      // for an empty range), whatever the lane's position, so both loads
      // stay in the row
      const int xl = min(max((int)(pcx - hxc), xlo), xhi);
      int xh = min(max((int)(pcx + hxc), xlo - 1), xhi);
      if (!(gyz <= r2g)) xh = xl - 1;
      t0 = *reinterpret_cast<const int *>(ecs_b + (rowb + 4u * (uint32_t)xl));
      t1 = *reinterpret_cast<const int *>(ecs_b + (rowb + 4u * (uint32_t)xh + 4u));
    };
    // rows in (dz, dy) order; the next row's range is read while this row's
    // records are tested
    int dy = -ry, dz = -rz, t0n, t1n;
    row_range(dy, t0n, t1n);
    for (;;) {
      const int t0 = t0n, t1 = t1n;
      const bool own_row = dy == 0 && dz == 0;
      const bool last = dy == ry && dz == rz;
      if (!last) {
        if (++dy > ry) {
          dy = -ry;
          ++dz;
          rowb += plane_step;
          const float gz = gap((float)dz, zlo, zhi, csz);
          gz2 = gz * gz;
        } else {
          rowb += row_bytes;
        }
        row_range(dy, t0n, t1n);
      }
      if (own_row) scan(std::true_type{}, t0, t1);
      else scan(std::false_type{}, t0, t1);
      if (last || cur >= stop_at) break;
    }
  };

  if (live) {
    if constexpr (EXT) ext_rows();
    else lane_rows();
  }
  int count = 0;
  if (live) {
    count = cur >> kRingShift;
    if (count > done) flush_chunk();  // slots past count in this chunk are never read
    cnt[s] = min(count, a.cap);  // an overflowed list is searched again; what reads it first stays in bounds
    if (count + 1 > a.maxneib) atomicOr(a.err, EB_TRUNC);
    if (count > a.cap) atomicOr(a.err, EB_OVERFLOW);
  } else if (s < a.n) {
    cnt[s] = 0;
  }
  long long tot = block_reduce<R_SUM>(count, lds);
  if (threadIdx.x == 0 && tot) atomicAdd((unsigned long long *)&a.stripes[stripe()], (unsigned long long)tot);
  long long mx = block_reduce<R_MAX>(count, lds);
  if (threadIdx.x == 0 && mx) atomicMax(&a.stripes[kStripes + stripe()], mx);
}

// Separation a - b of two mixed-path position records (pfix), b at its
// image: per axis the 32-bit wrap-around difference of the fractions, less
// the image's mapped box length, times the axis scale in FP32.  Fractions are
// within one unit (at most L / 2^31) of the positions, so the separation is
// within two units before its FP32 conversion and product (each 2^-24).
template <int DIM>
__device__ __forceinline__ void fix_sep(const uint3 &qa, const uint3 &qb, uint32_t code, const FixFrame &f,
                                        float &r0, float &r1, float &r2) {
  uint32_t dx = qa.x - qb.x, dy = qa.y - qb.y, dz = qa.z - qb.z;
  if (code != kNoImage) {
    const int sx = (int)(code % 3) - 1, sy = (int)((code / 3) % 3) - 1, sz = (int)(code / 9) - 1;
    dx -= (uint32_t)sx * f.step[0];
    dy -= (uint32_t)sy * f.step[1];
    dz -= (uint32_t)sz * f.step[2];
  }
  r0 = (float)(int)dx * f.scale[0];
  r1 = DIM >= 2 ? (float)(int)dy * f.scale[1] : 0.f;
  r2 = DIM == 3 ? (float)(int)dz * f.scale[2] : 0.f;
}

// ---------------------------------------------------------------- loop 1
// One Jacobi sweep of c1advanced (circuit1.f90:114-211) for the particles
// still active (eps > tol); frozen ones carry their state to the next buffer.
// GRAD: 0 no grad T, 1 standard (:178-179), 2 the 2nw_sd form (:174-176).
// PT is the pair arithmetic type; the Newton update is always FP64.
template <int DIM, int KT, int GRAD, typename PT>
__global__ __launch_bounds__(256) void k_loop1(Loop1Args<double4> a) {
  __shared__ long long lds[4];
  __shared__ int skip;
  // while maxval(eps) > tol (circuit1.f90:114), decided here: the first
  // wave reads the previous sweep's still-active stripes (set before this
  // launch began) and the error words, so the whole block agrees
  if (threadIdx.x < 64) {
    const long long v = a.prev ? a.prev[threadIdx.x] : 1;
    const unsigned long long any = __ballot(v != 0);
    if (threadIdx.x == 0) skip = any == 0 || *a.serr != 0 || *a.err != 0;
  }
  __syncthreads();
  if (skip) return;
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    *a.which = a.sweep & 1;
    a.stripes[2 * kStripes] = 1;
  }
  int s = gtid();
  bool inb = s < a.n;
  // the first sweep starts from the fields themselves (h, den, om, T as the
  // step found them) with every real particle active; later sweeps from the
  // previous sweep's state and residuals
  const bool first = a.f_h != nullptr;
  double4 c = make_double4(0, 0, 0, 0);
  if (inb) c = first ? make_double4(a.f_h[s], a.f_den[s], a.f_om[s], a.f_t[s]) : a.cur[s];
  bool act = inb && a.real[s] && (first || a.eps[s] > a.tol);
  bool still = false;
  int nn = 0;
  // the FP32 path gathers 20 bytes per pair: the 12-byte position record
  // (31-bit fractions) and the FP32 m and T den
  constexpr bool PACK = GRAD == 1 && std::is_same_v<PT, float>;
  if (inb && !act) {
    a.next[s] = c;
    if constexpr (PACK) a.s1n[s] = a.s1c[s];
    else if constexpr (GRAD == 1) a.trn[s] = a.trc[s];
  }
  if (act) {
    using K = KernelFn<KT, PT>;
    const double ha = c.x, rhoa = c.y, oma = c.z, ta = c.w;
    const double4 pa = a.pm[s];
    const double ma = pa.w;
    const PT inv_ha = (PT)(1.0 / ha);
    PT dsum = 0, osum = 0;
    PT g0 = 0, g1 = 0, g2 = 0, h0x = 0, h0y = 0, h0z = 0;
    bool qbad = !(ha > 0.0);
    nn = a.cnt[s];
    // j side: x, y, z, m, and what grad T needs of the previous sweep (GRAD 1:
    // T_b rho_b alone; GRAD 2: h_b, rho_b, om_b, T_b)
    struct J { double4 pb, sb; uint3 pc; float2 sc; double tr; uint32_t code; };
    // the FP32 path's own position record
    uint3 qa{};
    if constexpr (PACK) qa = a.pfix[s];
    auto load = [&](uint32_t ent) {
      J d;
      const uint32_t j = nb_index(ent);
      d.code = nb_code(ent);
      if constexpr (PACK) {
        d.pc = a.pfix[j];
        d.sc = a.s1c[j];
      } else {
        d.pb = a.pm[j];
        if constexpr (GRAD == 1) d.tr = a.trc[j];
      }
      if constexpr (GRAD == 2) d.sb = a.cur[j];
      return d;
    };
    // four gathers in flight (with particles in row order, eight measured
    // slower in the FP32 path: 0.55 against 0.51 ms at 128^3)
    for_neighbours<4>(a.nb, s, a.npad, nn, load, [&](const J &d) {
      PT r0, r1, r2;
      if constexpr (PACK) {
        fix_sep<DIM>(qa, d.pc, d.code, a.ff, r0, r1, r2);
      } else {
        double xb = d.pb.x, yb = d.pb.y, zb = d.pb.z;
        if (d.code != kNoImage) image_shift(xb, yb, zb, d.code, a.b);
        r0 = (PT)(pa.x - xb);
        r1 = (PT)(pa.y - yb);
        r2 = (PT)(pa.z - zb);
      }
      PT rsq = r0 * r0 + r1 * r1 + r2 * r2;
      // FP64, and FP32 with grad T: 1/|r| from one reciprocal square root
      // (FP64: refined), |r| from it (FP64: 0 at r = 0, as the square root
      // gives; NaN stays NaN)
      constexpr bool RSQ = GRAD != 0 || std::is_same_v<PT, double>;
      PT inv_dr_rsq = 0;
      if constexpr (RSQ && std::is_same_v<PT, float>) inv_dr_rsq = __builtin_amdgcn_rsqf(rsq);
      if constexpr (RSQ && std::is_same_v<PT, double>) inv_dr_rsq = rsq_nr(rsq);
      PT dr = RSQ ? rsq * inv_dr_rsq : sqrt_pair(rsq);
      if constexpr (RSQ && std::is_same_v<PT, double>) dr = rsq == 0.0 ? 0.0 : dr;
      PT q = dr * inv_ha;
      PT f, df;
      K::f_df(q, f, df);
      PT mb;
      if constexpr (PACK) mb = d.sc.x;
      else mb = (PT)d.pb.w;
      dsum += mb * f;
      osum += mb * (PT(DIM) * f + q * df);
      if constexpr (GRAD != 0) {
        const double4 sb = d.sb;
        // r = 0 gives NaN here, as 1/|r| times the kernels' derivative (0
        // at q = 0) does
        const PT inv_dr = inv_dr_rsq;
        if constexpr (GRAD == 1) {
          PT trb;
          if constexpr (PACK) trb = d.sc.y;
          else trb = (PT)d.tr;
          PT w = mb * (trb - (PT)ta * (PT)rhoa) * df * inv_dr;
          g0 += w * r0; g1 += w * r1; g2 += w * r2;
        } else {
          PT inv_hb, i_ob_rb2;  // 1/h_b and 1/(om_b rho_b^2)
          if constexpr (std::is_same_v<PT, double>) {
            // FP64: both from one division of their product
            const double r = 1.0 / (sb.x * (sb.z * sb.y * sb.y));
            inv_hb = (sb.z * sb.y * sb.y) * r;
            i_ob_rb2 = sb.x * r;
          } else {
            inv_hb = (PT)(1.0 / sb.x);
            i_ob_rb2 = PT(1) / ((PT)sb.z * (PT)sb.y * (PT)sb.y);
          }
          PT qb = dr * inv_hb;
          PT dfb = K::df(qb);
          PT wa = mb * df * inv_dr;
          PT wb = mb * (PT)sb.w * (PT)rhoa * i_ob_rb2 * dfb * inv_dr * ipow<DIM + 1>(inv_hb);
          g0 += wa * r0; g1 += wa * r1; g2 += wa * r2;
          h0x += wb * r0; h0y += wb * r1; h0z += wb * r2;
        }
      }
      qbad |= !(q >= PT(0));
    });
    const double wcv = a.p.wcv;
    const double ih = 1.0 / ha;
    const double cW = wcv * ipow<DIM>(ih);
    const double cH = -wcv * ipow<DIM + 1>(ih);
    double f0, df0;
    KernelFn<KT, double>::f_df(0.0, f0, df0);
    double den = cW * (double)dsum + ma * (cW * f0);
    double osm = cH * (double)osum + ma * (cH * (DIM * f0));
    osm = 1. - osm * (-ha / (DIM * den));
    double dfdh = -DIM * den * osm / ha;
    double fh = ma * ipow<DIM>(a.p.hfac / ha) - den;
    double hn = ha - fh / dfdh;
    if (hn <= 0.) atomicOr(a.err, EB_NEG_H);
    if (qbad) atomicOr(a.err, EB_Q);
    // h at entry: the first sweep starts from it
    double ep = fabs(hn - ha) / (first ? ha : a.h0[s]);
    a.eps[s] = ep;
    still = ep > a.tol;
    a.next[s] = make_double4(hn, den, osm, ta);
    if constexpr (PACK) a.s1n[s] = make_float2((float)ma, (float)(ta * den));
    else if constexpr (GRAD == 1) a.trn[s] = ta * den;
    double gx = 0, gy = 0, gz = 0;
    if constexpr (GRAD == 1) {
      double cN = wcv * ipow<DIM + 1>(ih) / rhoa / oma;
      gx = cN * (double)g0; gy = cN * (double)g1; gz = cN * (double)g2;
    } else if constexpr (GRAD == 2) {
      double cA = wcv * ipow<DIM + 1>(ih) * ta / (oma * rhoa);
      gx = cA * (double)g0 + wcv * (double)h0x;
      gy = cA * (double)g1 + wcv * (double)h0y;
      gz = cA * (double)g2 + wcv * (double)h0z;
    }
    a.dtdx[s] = gx; a.dtdy[s] = gy; a.dtdz[s] = gz;
  }
  long long act_n = block_reduce<R_SUM>(still ? 1 : 0, lds);
  if (threadIdx.x == 0 && act_n) atomicAdd((unsigned long long *)&a.stripes[stripe()], (unsigned long long)act_n);
  long long pairs = block_reduce<R_SUM>(nn, lds);
  if (threadIdx.x == 0 && pairs) atomicAdd((unsigned long long *)&a.stripes[kStripes + stripe()], (unsigned long long)pairs);
}

// c1simple (circuit1.f90:227-305): direct summation, h = hfac (m/rho)^(1/dim),
// grad T only for 2nw_ds (:282-284) / 2nw_sd (:285-289), om = 1 (:62,301).
template <int DIM, int KT, int GRAD, typename PT>
__global__ __launch_bounds__(256) void k_loop1_simple(Loop1Args<double4> a, Fields F) {
  int s = gtid();
  if (s >= a.n) return;
  F.f[F_OM][s] = 1.0;
  if (!a.real[s]) return;
  using K = KernelFn<KT, PT>;
  const double4 c = a.cur[s];
  const double ha = c.x, rhoa = c.y, ta = c.w;
  const double4 pa = a.pm[s];
  const PT inv_ha = (PT)(1.0 / ha);
  PT dsum = 0, g0 = 0, g1 = 0, g2 = 0, h0x = 0, h0y = 0, h0z = 0;
  bool qbad = !(ha > 0.0);
  const int nn = a.cnt[s];
  struct J { double4 pb, sb; uint32_t code; };
  auto load = [&](uint32_t ent) {
    J d;
    const uint32_t j = nb_index(ent);
    d.code = nb_code(ent);
    d.pb = a.pm[j];
    if constexpr (GRAD != 0) d.sb = a.cur[j];
    return d;
  };
  for_neighbours<4>(a.nb, s, a.npad, nn, load, [&](const J &d) {
    double xb = d.pb.x, yb = d.pb.y, zb = d.pb.z;
    if (d.code != kNoImage) image_shift(xb, yb, zb, d.code, a.b);
    PT r0 = (PT)(pa.x - xb), r1 = (PT)(pa.y - yb), r2 = (PT)(pa.z - zb);
    PT dr = sqrt(r0 * r0 + r1 * r1 + r2 * r2);
    PT q = dr * inv_ha;
    PT mb = (PT)d.pb.w;
    if constexpr (GRAD == 0) {
      dsum += mb * K::f(q);
    } else {
      PT f, df;
      K::f_df(q, f, df);
      dsum += mb * f;
      const double4 sb = d.sb;
      PT inv_dr = PT(1) / dr;
      if constexpr (GRAD == 1) {  // 2nw_ds: m_b/rho_a (T_b - T_a) nabla W_a
        PT w = mb * ((PT)sb.w - (PT)ta) * df * inv_dr;
        g0 += w * r0; g1 += w * r1; g2 += w * r2;
      } else {  // 2nw_sd: m_b (T_a/rho_a^2 nabla W_a + T_b/rho_b^2 nabla W_b)
        PT inv_hb = (PT)(1.0 / sb.x);
        PT dfb = K::df(dr * inv_hb);
        PT wa = mb * df * inv_dr;
        PT wb = mb * (PT)sb.w / ((PT)sb.y * (PT)sb.y) * dfb * inv_dr * ipow<DIM + 1>(inv_hb);
        g0 += wa * r0; g1 += wa * r1; g2 += wa * r2;
        h0x += wb * r0; h0y += wb * r1; h0z += wb * r2;
      }
    }
    qbad |= !(q >= PT(0));
  });
  if (qbad) atomicOr(a.err, EB_Q);
  const double wcv = a.p.wcv, ih = 1.0 / ha;
  double f0 = KernelFn<KT, double>::f(0.0);
  double den = wcv * ipow<DIM>(ih) * (double)dsum + pa.w * (wcv * ipow<DIM>(ih) * f0);
  F.f[F_DEN][s] = den;
  F.f[F_H][s] = a.p.hfac * pow(pa.w / den, 1. / DIM);
  double gx = 0, gy = 0, gz = 0;
  if constexpr (GRAD == 1) {
    double cN = wcv * ipow<DIM + 1>(ih) / rhoa;
    gx = cN * (double)g0; gy = cN * (double)g1; gz = cN * (double)g2;
  } else if constexpr (GRAD == 2) {
    double cA = wcv * ipow<DIM + 1>(ih) * ta / (rhoa * rhoa);
    gx = cA * (double)g0 + wcv * (double)h0x;
    gy = cA * (double)g1 + wcv * (double)h0y;
    gz = cA * (double)g2 + wcv * (double)h0z;
  }
  F.f[F_DTDX][s] = gx; F.f[F_DTDY][s] = gy; F.f[F_DTDZ][s] = gz;
}

// After loop 1: write back h/den/om of real particles (circuit1.f90:213-215),
// the EOS (eos.f90:21-45) and the loop-2 j-side packs; flags[stripe] bit 1 =
// some velocity is non-zero, bit 2 = some B differs from particle 0's.
template <typename V4>
__global__ __launch_bounds__(256) void k_post1(Fields F, PostSel sel, const uint8_t *real, int n,
                                               int advanced, int do_eos, int iso, double c1, double gamma, V4 *q0,
                                               V4 *q1, V4 *q2, V4 *q3, V4 *q4, double4 *pd, double *trd,
                                               float2 *s2, long long *flags, long long *hmax_red) {
  __shared__ long long lds[4];
  using T = decltype(V4().x);
  if (*sel.serr != 0 || *sel.err1 != 0) return;  // set before this launch: uniform
  const double4 *fin = (sel.which && *sel.which) ? sel.fin1 : sel.fin0;
  int s = gtid();
  int fl = 0;
  // h and den as written here (no reload of the fields just stored)
  double h = 0.0;
  if (s < n) {
    bool r = real[s];
    double den, om;
    if (r && advanced) {
      const double4 v = fin[s];
      F.f[F_H][s] = v.x; F.f[F_DEN][s] = v.y; F.f[F_OM][s] = v.z;
      h = v.x; den = v.y; om = v.z;
    } else {
      h = F.f[F_H][s]; den = F.f[F_DEN][s]; om = F.f[F_OM][s];
    }
    if (r && do_eos) {
      if (iso) {
        F.f[F_P][s] = den * c1 * c1;
      } else {
        double p = (gamma - 1.) * den * F.f[F_U][s];
        F.f[F_P][s] = p;
        F.f[F_C][s] = sqrt(gamma * p / den);
      }
    }
    // v and B only where a pack or the flags need them
    double vx = 0, vy = 0, vz = 0, bx = 0, by = 0, bz = 0;
    if (flags || q1) { vx = F.f[F_VX][s]; vy = F.f[F_VY][s]; vz = F.f[F_VZ][s]; }
    if (flags || q4) { bx = F.f[F_BX][s]; by = F.f[F_BY][s]; bz = F.f[F_BZ][s]; }
    if (flags) {
      if (vx != 0.0 || vy != 0.0 || vz != 0.0) fl |= 1;
      if (!(bx == F.f[F_BX][0] && by == F.f[F_BY][0] && bz == F.f[F_BZ][0])) fl |= 2;
    }
    if (q0) q0[s] = V4{(T)(1.0 / h), (T)den, (T)om, (T)F.f[F_T][s]};
    if (q1) q1[s] = V4{(T)vx, (T)vy, (T)vz, (T)F.f[F_P][s]};
    if (q2) q2[s] = V4{(T)F.f[F_C][s], (T)F.f[F_U][s], (T)F.f[F_KAPPA][s], (T)0};
    if (q3) q3[s] = V4{(T)F.f[F_DTDX][s], (T)F.f[F_DTDY][s], (T)F.f[F_DTDZ][s], (T)0};
    if (q4) q4[s] = V4{(T)bx, (T)by, (T)bz, (T)0};
    if constexpr (std::is_same_v<V4, float4>) {  // FP32 path: m/den and T den in FP32 beside pfix
      if (s2) s2[s] = make_float2((float)(F.f[F_M][s] / den), (float)(F.f[F_T][s] * den));
    } else if (pd) {
      pd[s] = make_double4(F.f[F_RX][s], F.f[F_RY][s], F.f[F_RZ][s], F.f[F_M][s] / den);
      trd[s] = F.f[F_T][s] * den;
    }
  }
  if (flags) {
    long long t = block_reduce<R_OR>(fl, lds);
    if (threadIdx.x == 0 && t) atomicOr((unsigned long long *)&flags[stripe()], (unsigned long long)t);
  }
  // the largest h of the real particles after this step (ordered int64), for
  // the next step's grid when nothing else changes h or the positions
  if (hmax_red) {
    long long hm = -0x7FFFFFFFFFFFFFFFLL - 1;
    if (s < n && real[s]) hm = d2ord(h);
    const long long t = block_reduce<R_MAX>(hm, lds);
    if (threadIdx.x == 0) atomicMax(&hmax_red[stripe()], t);
  }
}

// ---------------------------------------------------------------- loop 2
template <typename T>
struct Sym3 {
  T xx, yy, zz, xy, xz, yz;
};

// Diffusion tensor of one particle from its unit field direction:
// reference modes (circuit2.f90:194-206, 370-374) or the full tensor.
template <typename T>
__device__ __forceinline__ Sym3<T> cond_tensor(const Phys &p, T ux, T uy, T uz) {
  Sym3<T> k;
  if (p.diso == 1) {
    k.xx = k.yy = k.zz = (T)p.dcond; k.xy = k.xz = k.yz = 0;
  } else if (p.tensor_mode == 1) {
    T kp = (T)p.dcond, kq = (T)p.kappa_perp, d = kp - kq;
    k.xx = d * ux * ux + kq; k.yy = d * uy * uy + kq; k.zz = d * uz * uz + kq;
    k.xy = d * ux * uy; k.xz = d * ux * uz; k.yz = d * uy * uz;
  } else {
    T kp = (T)p.dcond;
    k.xx = kp * ux * ux; k.yy = kp * uy * uy; k.zz = kp * uz * uz; k.xy = k.xz = k.yz = 0;
  }
  return k;
}

template <typename T>
__device__ __forceinline__ Sym3<T> tensor_of_b(const Phys &p, T bx, T by, T bz) {
  T b2 = bx * bx + by * by + bz * bz;
  T ux = 0, uy = 0, uz = 0;
  if (b2 != T(0)) { T r = T(1) / sqrt(b2); ux = bx * r; uy = by * r; uz = bz * r; }
  return cond_tensor<T>(p, ux, uy, uz);
}

// sum_lm kt_lm r_l r_m / r^2 (S) and trace (Tr) over the dim x dim block
template <int DIM, typename T>
__device__ __forceinline__ void tensor_contract(const Sym3<T> &kt, T rx, T ry, T rz, T inv_r2, T &S, T &Tr) {
  S = kt.xx * rx * rx;
  Tr = kt.xx;
  if (DIM >= 2) { S += kt.yy * ry * ry + T(2) * kt.xy * rx * ry; Tr += kt.yy; }
  if (DIM == 3) { S += kt.zz * rz * rz + T(2) * (kt.xz * rx * rz + kt.yz * ry * rz); Tr += kt.zz; }
  S *= inv_r2;
}


// c2 (circuit2.f90:58-505) for one real particle per lane over the loop-1
// neighbour list, post-loop-1 h/den/om and post-EOS p/c.
// The FP64 diffusion-mode loop is latency-bound: at 176 VGPRs it runs 2 waves
// per SIMD, capped to 168 it runs 3 (a few registers spill) and measured 17 %
// faster; the other variants are fastest uncapped.
template <int DIM, int KT, typename PT, typename V4, int MODE>
__global__ __launch_bounds__(256)
__attribute__((amdgpu_waves_per_eu((MODE == MODE_DIFF && std::is_same_v<PT, double>) ? 3 : 1)))
void k_loop2(Loop2Args<V4> a) {
  int s = gtid();
  double e_part = 0.0;
  if (s < a.n && a.real[s]) {
    using K = KernelFn<KT, PT>;
    const Phys &P = a.p;
    // GEN: every term under its flag (MODE_GENERIC), or the equation set
    // fixed at compile time (MODE_SPEC + mask)
    constexpr bool SPEC = MODE >= MODE_SPEC;
    constexpr bool GEN = MODE == MODE_GENERIC || SPEC;
    constexpr int TM = SPEC ? MODE - MODE_SPEC : 0;
    const bool hydro = SPEC ? (TM & 1) != 0 : (MODE == MODE_GENERIC && P.hydro);
    const bool magneto = SPEC ? (TM & 2) != 0 : (MODE == MODE_GENERIC && P.magneto);
    const bool diff = SPEC ? (TM & 4) != 0 : (MODE == MODE_DIFF || P.diff);
    const bool fld = SPEC ? (TM & 8) != 0 : (MODE == MODE_GENERIC && P.fld);
    const bool artts = GEN && P.artts;
    // optional j-side packs; a pack the host did not build is never read
    const bool need_v = (P.adden || artts || hydro || magneto) && (a.v_nonzero || artts || hydro || magneto) &&
                        a.q1 != nullptr;
    const bool need_q2 = GEN && ((diff && fld) || artts) && a.q2 != nullptr;
    const bool need_q3 = GEN && diff && (fld || P.ddw == 403 || P.ddw == 404) && a.q3 != nullptr;
    const bool tensor_b = diff && !fld && P.diso == 0 && !a.b_uniform && a.q4 != nullptr;
    const bool need_q4 = (magneto || tensor_b) && a.q4 != nullptr;
    const bool hess = diff && (P.ddw == 400 || P.ddw == 401 || P.ddw == 402);
    const bool need_dfb = GEN && (hydro || artts || magneto || P.ddw == 403);
    const double4 pa = a.pm[s];
    // i-side values straight from the FP64 fields (coalesced, read once)
    const double *const *Fi = a.out.f;
    const double ha_d = Fi[F_H][s];
    // den, om, v and B in FP64 as read, for the energy term after the loop
    const double rho_d = Fi[F_DEN][s], om_d = Fi[F_OM][s];
    const double vx_d = Fi[F_VX][s], vy_d = Fi[F_VY][s], vz_d = Fi[F_VZ][s];
    const double bx_d = Fi[F_BX][s], by_d = Fi[F_BY][s], bz_d = Fi[F_BZ][s];
    const PT inv_ha = (PT)(1.0 / ha_d), rhoa = (PT)rho_d, oma = (PT)om_d, pa_ = (PT)Fi[F_P][s];
    const PT vax = (PT)vx_d, vay = (PT)vy_d, vaz = (PT)vz_d, ca = (PT)Fi[F_C][s];
    const PT ua = (PT)Fi[F_U][s], ta = (PT)Fi[F_T][s], kappaa = (PT)Fi[F_KAPPA][s];
    const PT gax = (PT)Fi[F_DTDX][s], gay = (PT)Fi[F_DTDY][s], gaz = (PT)Fi[F_DTDZ][s];
    const PT bax = (PT)bx_d, bay = (PT)by_d, baz = (PT)bz_d;
    // i-side constants
    const PT wcv = (PT)P.wcv;
    const PT ihd1 = ipow<DIM + 1>(inv_ha);  // 1/h_a^(dim+1)
    const PT inv_rhoa = rcp_pair(rhoa);      // 1/rho_a once per particle
    // i-side reciprocals the pair terms multiply by: 1/(om_a rho_a) and
    // 1/(rho_a^2 om_a), once per particle
    const PT i_oa_ra = PT(1) / oma / rhoa;
    const PT i_ra2_oa = i_oa_ra / rhoa;
    const PT ub2 = bax * bax + bay * bay + baz * baz;
    Sym3<PT> kta = {0, 0, 0, 0, 0, 0}, ktb_fixed = {0, 0, 0, 0, 0, 0};
    const PT mu0inv = magneto ? (PT)(1.0 / P.muzero) : PT(0);
    if (diff) {
      if (fld) {
        PT gn = sqrt(gax * gax + gay * gay + gaz * gaz);
        PT fr = gn / (kappaa * rhoa * ta * rhoa);
        PT lam = (PT(2) + fr) / (PT(6) + PT(3) * fr + fr * fr);
        PT d = (PT)kLightspeed * lam / kappaa / rhoa;
        kta.xx = kta.yy = kta.zz = d;
        a.out.f[F_VFLD][s] = (double)(d / ((ta * rhoa) / gn));
      } else {
        kta = tensor_of_b<PT>(P, bax, bay, baz);
        ktb_fixed = kta;  // isotropic, or every particle has the same B
      }
    }
    // tensor and operator constants of the i side
    PT kdax = kta.xx * gax + kta.xy * gay + kta.xz * gaz;
    PT kday = kta.xy * gax + kta.yy * gay + kta.yz * gaz;
    PT kdaz = kta.xz * gax + kta.yz * gay + kta.zz * gaz;
    PT dvx = 0, dvy = 0, dvz = 0, du = 0, ddt = 0, dbx = 0, dby = 0, dbz = 0, drho = 0;
    bool qbad = !(inv_ha > PT(0));
    const int nn = a.cnt[s];
    // batched: generic mode gathers positions/mass and {1/h, den, om, T};
    // MODE_DIFF gathers only what the Hessian term reads, {x, y, z, m/den}
    // and T den (the same values the term forms from m, den, T).  The
    // optional packs (v, c/u/kappa, grad T, B) are gathered in the body.
    // the FP32 path's record holds m/den and T den as FP32 in its fourth slot
    constexpr bool PACK = std::is_same_v<V4, float4>;
    struct JG { double4 pb; V4 q0, q1, q2, q3, q4; uint32_t j, code; };
    struct JD { double4 pb; uint3 pc; float2 sc; double tr; uint32_t j, code; };
    using J = std::conditional_t<MODE == MODE_DIFF, JD, JG>;
    // MODE_DIFF's FP32 path: the own position record
    constexpr bool CELL = MODE == MODE_DIFF && PACK;
    uint3 qa{};
    if constexpr (CELL) qa = a.pfix[s];
    auto load = [&](uint32_t ent) {
      J d;
      d.j = nb_index(ent);
      d.code = nb_code(ent);
      if constexpr (MODE == MODE_DIFF) {
        if constexpr (PACK) {
          d.pc = a.pfix[d.j];
          d.sc = a.s2[d.j];
        } else {
          d.pb = a.pd[d.j];
          d.tr = a.trd[d.j];
        }
      } else {
        // every pack this step reads, gathered together with the position
        // (one round trip per neighbour)
        d.pb = a.pm[d.j];
        d.q0 = a.q0[d.j];
        if (need_v) d.q1 = a.q1[d.j];
        if (need_q2) d.q2 = a.q2[d.j];
        if (need_q3) d.q3 = a.q3[d.j];
        if (need_q4) d.q4 = a.q4[d.j];
      }
      return d;
    };
    // neighbours whose gathers are in flight together: two for the MHD and
    // FLD sets (1M-particle MHD shock tube: 5.23 -> 5.08 ms per step, FLD
    // unchanged; no occupancy lost), one where a second would cost waves
    constexpr int NBATCH = MODE == MODE_DIFF ? 4 : (SPEC && TM != kSpecHydro ? 2 : 1);
    for_neighbours<NBATCH>(a.nb, s, a.npad, nn, load, [&](const J &d) {
      const uint32_t code = d.code;
      PT rx, ry, rz;
      if constexpr (CELL) {
        fix_sep<DIM>(qa, d.pc, code, a.ff, rx, ry, rz);
      } else {
        const double4 pb = d.pb;
        double xb = pb.x, yb = pb.y, zb = pb.z;
        if (code != kNoImage) image_shift(xb, yb, zb, code, a.b);
        rx = (PT)(pa.x - xb);
        ry = (PT)(pa.y - yb);
        rz = (PT)(pa.z - zb);
      }
      // j-side scalars; in MODE_DIFF mb_rhob = m_b / rho_b and trb = T_b rho_b
      PT mb = 0, inv_hb = 0, rhob = 0, omb = 0, tb = 0, mb_rhob = 0, trb = 0;
      if constexpr (MODE == MODE_DIFF) {
        if constexpr (PACK) {
          mb_rhob = d.sc.x;
          trb = d.sc.y;
        } else {
          mb_rhob = (PT)d.pb.w;
          trb = (PT)d.tr;
        }
        if (need_v && P.adden) mb = (PT)a.pm[d.j].w;  // d rho/dt of moving particles only
      } else {
        const V4 B0 = d.q0;
        mb = (PT)d.pb.w;
        inv_hb = (PT)B0.x; rhob = (PT)B0.y; omb = (PT)B0.z; tb = (PT)B0.w;
      }
      const PT r2 = rx * rx + ry * ry + rz * rz;
      PT inv_r2, dr, inv_dr;
      if constexpr (std::is_same_v<PT, double>) {
        // 1/|r| refined from the hardware estimate; |r| = 0 at r = 0, as the
        // square root gives (the pair terms are then NaN, as with the
        // division), and NaN stays NaN
        inv_dr = rsq_nr(r2);
        inv_r2 = inv_dr * inv_dr;
        dr = r2 == 0.0 ? 0.0 : r2 * inv_dr;
      } else {
        inv_r2 = rcp_pair(r2);
        dr = sqrt_pair(r2);
        inv_dr = dr * inv_r2;
      }
      const PT qa = dr * inv_ha;
      PT dfa, ddfa = 0, fa = 0;
      if (hess && P.ddw == 401) K::df_ddf(qa, dfa, ddfa);
      else if (hess && P.ddw == 402) K::f_df(qa, fa, dfa);
      else dfa = K::df(qa);
      qbad |= !(qa >= PT(0));
      // nabla W_a = wcv df r / (h^(dim+1) dr)   (kernel.f90:189-197)
      const PT ga = wcv * dfa * ihd1 * inv_dr;
      const PT nwax = ga * rx, nway = ga * ry, nwaz = ga * rz;
      PT gb = 0, dfb = 0;
      if (need_dfb) {
        const PT qb = dr * inv_hb;
        dfb = K::df(qb);
        gb = wcv * dfb * ipow<DIM + 1>(inv_hb) * inv_dr;
      }
      const PT nwbx = gb * rx, nwby = gb * ry, nwbz = gb * rz;
      PT vabx = 0, vaby = 0, vabz = 0, pbb = 0;
      if (need_v) {
        V4 t;
        if constexpr (GEN) t = d.q1;
        else t = a.q1[d.j];
        vabx = vax - (PT)t.x; vaby = vay - (PT)t.y; vabz = vaz - (PT)t.z; pbb = (PT)t.w;
      }
      const PT vnwa = vabx * nwax + vaby * nway + vabz * nwaz;
      if (P.adden) drho += mb * vnwa;
      // j-side 1/(rho_b^2 om_b), once per pair where a term divides by it
      // (FP64: a division, so rho_b om_b = 0 still gives inf)
      PT i_rb2_ob = 0;
      if (GEN && (artts || hydro || magneto || (diff && P.ddw == 403)))
        i_rb2_ob = rcp_pair(rhob * rhob * omb);
      PT cb = 0, ub = 0, kappab = 0;
      if (need_q2) {
        V4 t;
        if constexpr (GEN) t = d.q2;
        else t = a.q2[d.j];
        cb = (PT)t.x; ub = (PT)t.y; kappab = (PT)t.z;
      }
      if (artts) {
        // Fab = wcv df dr / h^(dim+2) / q = wcv df / h^(dim+1)   (kernel.f90:251-259)
        const PT faba = wcv * dfa * ihd1;
        const PT fabb = wcv * dfb * ipow<DIM + 1>(inv_hb);
        const PT dvr = (vabx * rx + vaby * ry + vabz * rz) * inv_dr;
        PT qva = 0, qvb = 0;
        if (dvr < PT(0)) {  // art_viscosity, circuit2.f90:558-575
          qva = PT(-0.5) * rhoa * (ca + PT(2) * fabs(dvr)) * dvr;
          qvb = PT(-0.5) * rhob * (cb + PT(2) * fabs(dvr)) * dvr;
        }
        const PT ta_ = qva * i_ra2_oa, tb_ = qvb * i_rb2_ob;
        dvx -= mb * (ta_ * nwax + tb_ * nwbx);
        dvy -= mb * (ta_ * nway + tb_ * nwby);
        dvz -= mb * (ta_ * nwaz + tb_ * nwbz);
        PT qc = 0;
        if (P.au != -1.0) {  // art_termcond, circuit2.f90:520-529
          PT vs = sqrt_pair(div_pair((PT)fabs(pa_ - pbb), PT(0.5) * (rhoa + rhob)));
          qc = (PT)P.au * (vs * (ua - ub) * PT(0.5) * (faba * i_oa_ra + fabb * (rhob * i_rb2_ob)));
        }
        du += mb * (dvr * qva * faba * i_oa_ra + qc);
      }
      if (hydro) {
        const PT ta_ = pa_ * i_ra2_oa, tb_ = pbb * i_rb2_ob;
        dvx -= mb * (ta_ * nwax + tb_ * nwbx);
        dvy -= mb * (ta_ * nway + tb_ * nwby);
        dvz -= mb * (ta_ * nwaz + tb_ * nwbz);
        du += mb * pa_ * vnwa * i_ra2_oa;
      }
      PT bbx = 0, bby = 0, bbz = 0;
      if (need_q4) {
        V4 t;
        if constexpr (GEN) t = d.q4;
        else t = a.q4[d.j];
        bbx = (PT)t.x; bby = (PT)t.y; bbz = (PT)t.z;
      }
      if (diff) {
        Sym3<PT> ktb = ktb_fixed;
        PT gbx = 0, gby = 0, gbz = 0;
        if (need_q3) {
          V4 t;
          if constexpr (GEN) t = d.q3;
          else t = a.q3[d.j];
          gbx = (PT)t.x; gby = (PT)t.y; gbz = (PT)t.z;
        }
        if (fld) {
          PT gn = sqrt_pair(gbx * gbx + gby * gby + gbz * gbz);
          PT fr = div_pair(gn, kappab * rhob * rhob * tb);
          PT lam = div_pair(PT(2) + fr, PT(6) + PT(3) * fr + fr * fr);
          PT dfl = div_pair(div_pair((PT)kLightspeed * lam, kappab), rhob);
          ktb.xx = ktb.yy = ktb.zz = dfl; ktb.xy = ktb.xz = ktb.yz = 0;
        } else if (tensor_b) {
          ktb = tensor_of_b<PT>(P, bbx, bby, bbz);
        }
        if (hess) {
          // sum_lm ((kta+ktb)/2)_lm H_lm over the dim x dim block
          PT S, Tr;
          if (!tensor_b && !fld) {
            tensor_contract<DIM, PT>(kta, rx, ry, rz, inv_r2, S, Tr);
          } else {
            Sym3<PT> kt;
            kt.xx = (kta.xx + ktb.xx) * PT(0.5); kt.yy = (kta.yy + ktb.yy) * PT(0.5);
            kt.zz = (kta.zz + ktb.zz) * PT(0.5); kt.xy = (kta.xy + ktb.xy) * PT(0.5);
            kt.xz = (kta.xz + ktb.xz) * PT(0.5); kt.yz = (kta.yz + ktb.yz) * PT(0.5);
            tensor_contract<DIM, PT>(kt, rx, ry, rz, inv_r2, S, Tr);
          }
          PT contr;
          if (P.ddw == 401) {
            // hessian_ddw_cart, kernel.f90:377-417
            PT cst = wcv * ihd1 * inv_ha;
            // df / q with 1/q = h_a / |r| (q = 0 gives NaN, as the division)
            PT dfq = dfa * inv_dr * (PT)ha_d;
            contr = cst * ((ddfa - dfq) * S + dfq * Tr);
          } else {
            // hessian_fab_cart / hessian_fw_cart, kernel.f90:460-539
            PT half = (P.ddw == 402) ? PT(0.5) * (PT)P.fwc * wcv * fa * ihd1 * inv_ha
                                     : -wcv * dfa * ihd1 * inv_dr;  // 0.5*Gab
            contr = half * (PT(DIM + 2) * S - Tr);
          }
          if constexpr (MODE == MODE_DIFF) {
            ddt += mb_rhob * inv_rhoa * (trb - ta * rhoa) * contr;
          }
          else ddt += div_pair(mb, rhob) * inv_rhoa * (tb * rhob - ta * rhoa) * contr;
        } else if (GEN) {
          PT kdbx = ktb.xx * gbx + ktb.xy * gby + ktb.xz * gbz;
          PT kdby = ktb.xy * gbx + ktb.yy * gby + ktb.yz * gbz;
          PT kdbz = ktb.xz * gbx + ktb.yz * gby + ktb.zz * gbz;
          if (P.ddw == 403) {
            const PT odda = i_ra2_oa, oddb = i_rb2_ob;
            ddt += mb * ((kdax * nwax + kday * nway + kdaz * nwaz) * odda +
                         (kdbx * nwbx + kdby * nwby + kdbz * nwbz) * oddb);
          } else {
            ddt += mb * i_ra2_oa * ((kdbx - kdax) * nwax + (kdby - kday) * nway + (kdbz - kdaz) * nwaz);
          }
        }
      }
      if (magneto) {
        // stress tensors (circuit2.f90:166-175, 433-441), force (:443-454), induction (:458-461)
        PT b2a = ub2, b2b = bbx * bbx + bby * bby + bbz * bbz;
        PT ma_x = ((PT(0.5) * b2a - bax * bax) * nwax - bax * bay * nway - bax * baz * nwaz) * mu0inv;
        PT ma_y = (-bay * bax * nwax + (PT(0.5) * b2a - bay * bay) * nway - bay * baz * nwaz) * mu0inv;
        PT ma_z = (-baz * bax * nwax - baz * bay * nway + (PT(0.5) * b2a - baz * baz) * nwaz) * mu0inv;
        PT mb_x = ((PT(0.5) * b2b - bbx * bbx) * nwbx - bbx * bby * nwby - bbx * bbz * nwbz) * mu0inv;
        PT mb_y = (-bby * bbx * nwbx + (PT(0.5) * b2b - bby * bby) * nwby - bby * bbz * nwbz) * mu0inv;
        PT mb_z = (-bbz * bbx * nwbx - bbz * bby * nwby + (PT(0.5) * b2b - bbz * bbz) * nwbz) * mu0inv;
        dvx -= mb * (ma_x * i_ra2_oa + mb_x * i_rb2_ob);
        dvy -= mb * (ma_y * i_ra2_oa + mb_y * i_rb2_ob);
        dvz -= mb * (ma_z * i_ra2_oa + mb_z * i_rb2_ob);
        PT bnwa = bax * nwax + bay * nway + baz * nwaz;
        PT f = i_oa_ra * mb;
        dbx -= f * (vabx * bnwa - bbx * vnwa);
        dby -= f * (vaby * bnwa - bby * vnwa);
        dbz -= f * (vabz * bnwa - bbz * vnwa);
      }
    });
    if (qbad) atomicOr(a.err, EB_Q);
    double dh = 0.0;
    if (P.adden) dh = -DIM * ha_d / (double)rhoa * (double)drho / (double)oma;
    a.out.f[F_AX][s] = (double)dvx; a.out.f[F_AY][s] = (double)dvy; a.out.f[F_AZ][s] = (double)dvz;
    a.out.f[F_DH][s] = dh;
    a.out.f[F_DDT][s] = (double)ddt;
    a.out.f[F_DBX][s] = (double)dbx; a.out.f[F_DBY][s] = (double)dby; a.out.f[F_DBZ][s] = (double)dbz;
    a.out.f[F_DU][s] = (double)du;
    // FP64 energy sum (circuit2.f90:490-492)
    const double vx = vx_d, vy = vy_d, vz = vz_d, bx = bx_d, by = by_d, bz = bz_d, rh = rho_d;
    e_part = pa.w * (vx * (double)dvx + vy * (double)dvy + vz * (double)dvz + (double)du +
                     (bx * (double)dbx + by * (double)dby + bz * (double)dbz) / rh -
                     0.5 * (bx * bx + by * by + bz * bz) / rh / rh * (double)drho / om_d);
  }
  __shared__ double red[4];
  double w = wave_sum(e_part);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = w;
  __syncthreads();
  if (threadIdx.x == 0) {
    double t = 0;
    for (int i = 0; i < (int)(blockDim.x >> 6); i++) t += red[i];
    a.energy_part[xcd_block(blockIdx.x, gridDim.x)] = t;
  }
}

// c2 (circuit2.f90:58-505) of a diffusion-only step with the fab Hessian
// operator (hessian_fab_cart, kernel.f90:460-539) and one conduction tensor
// for every pair: isotropic conduction, or every particle carrying the same B
// bit for bit, and no velocity term (adden off, or every v exactly 0).  That
// is k_loop2's MODE_DIFF with all of its run-time choices fixed, so none of
// their branches remain in the pair loop; the other outputs of c2 are the
// zeros MODE_DIFF writes.  Both precisions fold (dim + 2) and the
// off-diagonal 2 into the tensor and apply the particle's constant factors
// -wcv / (h^(dim+1) rho_a) after the sum, so no pair divides by rho_a; both
// take 1/|r| from one reciprocal square root (FP64: the hardware estimate
// refined by two Newton-Raphson steps).
template <int DIM, int KT, typename PT, typename V4>
__global__ __launch_bounds__(256) void k_loop2_fab(Loop2Args<V4> a) {
  int s = gtid();
  double e_part = 0.0;
  if (s < a.n && a.real[s]) {
    using K = KernelFn<KT, PT>;
    constexpr bool F32 = std::is_same_v<PT, float>;
    const Phys &P = a.p;
    const double4 pa = a.pm[s];
    const double *const *Fi = a.out.f;
    const double ha_d = Fi[F_H][s], rho_d = Fi[F_DEN][s], om_d = Fi[F_OM][s];
    const PT inv_ha = (PT)(1.0 / ha_d), rhoa = (PT)rho_d, oma = (PT)om_d, ta = (PT)Fi[F_T][s];
    const PT wcv = (PT)P.wcv;
    const PT ihd1 = ipow<DIM + 1>(inv_ha);
    // one B for every particle (every component equal to particle 0's):
    // particle 0's, through scalar loads
    double bxa, bya, bza;
    if (a.b_uniform) {
      bxa = Fi[F_BX][0]; bya = Fi[F_BY][0]; bza = Fi[F_BZ][0];
    } else {
      bxa = Fi[F_BX][s]; bya = Fi[F_BY][s]; bza = Fi[F_BZ][s];
    }
    const Sym3<PT> kt = tensor_of_b<PT>(P, (PT)bxa, (PT)bya, (PT)bza);
    const PT tra = ta * rhoa;
    // (dim + 2) k with the off-diagonal terms doubled, and the trace
    constexpr PT D2 = PT(DIM + 2);
    const PT cxx = D2 * kt.xx, cyy = D2 * kt.yy, czz = D2 * kt.zz;
    const PT cxy = PT(2) * D2 * kt.xy, cxz = PT(2) * D2 * kt.xz, cyz = PT(2) * D2 * kt.yz;
    PT tr = kt.xx;
    if (DIM >= 2) tr += kt.yy;
    if (DIM == 3) tr += kt.zz;
    PT ddt = 0;
    bool qbad = !(inv_ha > PT(0));
    const int nn = a.cnt[s];
    // FP64: (x, y, z, m/den) and T den; FP32: the 12-byte position record and
    // (m/den, T den) in FP32
    struct J { double4 pb; uint3 pc; float2 sc; double tr; uint32_t code; };
    uint3 qa{};
    if constexpr (F32) qa = a.pfix[s];
    auto load = [&](uint32_t ent) {
      J d;
      const uint32_t j = nb_index(ent);
      d.code = nb_code(ent);
      if constexpr (F32) {
        d.pc = a.pfix[j];
        d.sc = a.s2[j];
      } else {
        d.pb = a.pd[j];
        d.tr = a.trd[j];
      }
      return d;
    };
    for_neighbours<4>(a.nb, s, a.npad, nn, load, [&](const J &d) {
      PT rx, ry, rz;
      if constexpr (F32) {
        fix_sep<DIM>(qa, d.pc, d.code, a.ff, rx, ry, rz);
      } else {
        double xb = d.pb.x, yb = d.pb.y, zb = d.pb.z;
        if (d.code != kNoImage) image_shift(xb, yb, zb, d.code, a.b);
        rx = (PT)(pa.x - xb);
        ry = (PT)(pa.y - yb);
        rz = (PT)(pa.z - zb);
      }
      if constexpr (F32) {
        const float mb_rhob = d.sc.x, trb = d.sc.y;
        const float r2 = fmaf(rz, rz, fmaf(ry, ry, rx * rx));
        const float inv_dr = __builtin_amdgcn_rsqf(r2);
        const float qa = r2 * inv_dr * inv_ha;
        const float dfa = K::df(qa);
        qbad |= !(qa >= 0.f);
        // (dim + 2) sum_lm k_lm r_l r_m over the dim x dim block
        float sx = cxx * rx, sy = 0.f;
        if (DIM >= 2) { sx = fmaf(cxy, ry, sx); sy = cyy * ry; }
        if (DIM == 3) { sx = fmaf(cxz, rz, sx); sy = fmaf(cyz, rz, sy); }
        float S = rx * sx;
        if (DIM >= 2) S = fmaf(ry, sy, S);
        if (DIM == 3) S = fmaf(czz * rz, rz, S);
        const float x = fmaf(S, inv_dr * inv_dr, -tr);
        ddt = fmaf(mb_rhob * (trb - tra), dfa * inv_dr * x, ddt);
      } else {
        const double mb_rhob = d.pb.w, trb = d.tr;
        const double r2 = rx * rx + ry * ry + rz * rz;
        // 1/|r| refined from the hardware estimate; |r| = 0 at r = 0, as the
        // square root gives (1/|r| is then NaN, and so was the pair term
        // with a division), and NaN stays NaN
        const double inv_dr = rsq_nr(r2);
        const double inv_r2 = inv_dr * inv_dr;
        const double dr = r2 == 0.0 ? 0.0 : r2 * inv_dr;
        const double qa = dr * inv_ha;
        const double dfa = K::df(qa);
        qbad |= !(qa >= 0.0);
        // (dim + 2) sum_lm k_lm r_l r_m over the dim x dim block, as above
        double sx = cxx * rx, sy = 0.0;
        if (DIM >= 2) { sx = fma(cxy, ry, sx); sy = cyy * ry; }
        if (DIM == 3) { sx = fma(cxz, rz, sx); sy = fma(cyz, rz, sy); }
        double S = rx * sx;
        if (DIM >= 2) S = fma(ry, sy, S);
        if (DIM == 3) S = fma(czz * rz, rz, S);
        const double x = fma(S, inv_r2, -tr);
        ddt = fma(mb_rhob * (trb - tra), dfa * inv_dr * x, ddt);
      }
    });
    ddt *= -wcv * ihd1 * rcp_pair(rhoa);
    if (qbad) atomicOr(a.err, EB_Q);
    const PT zero = 0;
    double dh = 0.0;
    if (P.adden) dh = -DIM * ha_d / (double)rhoa * (double)zero / (double)oma;
    a.out.f[F_AX][s] = 0.0; a.out.f[F_AY][s] = 0.0; a.out.f[F_AZ][s] = 0.0;
    a.out.f[F_DH][s] = dh;
    a.out.f[F_DDT][s] = (double)ddt;
    a.out.f[F_DBX][s] = 0.0; a.out.f[F_DBY][s] = 0.0; a.out.f[F_DBZ][s] = 0.0;
    a.out.f[F_DU][s] = 0.0;
    // FP64 energy sum (circuit2.f90:490-492) with zero dv, du, dB, drho.
    // Every velocity exactly 0 (a signed zero among them changes no term
    // after the + 0 below) or one B for every particle: no per-particle loads
    double vx = 0.0, vy = 0.0, vz = 0.0, bx, by, bz;
    if (a.v_nonzero) { vx = Fi[F_VX][s]; vy = Fi[F_VY][s]; vz = Fi[F_VZ][s]; }
    if (a.b_uniform) {
      bx = Fi[F_BX][0]; by = Fi[F_BY][0]; bz = Fi[F_BZ][0];
    } else {
      bx = Fi[F_BX][s]; by = Fi[F_BY][s]; bz = Fi[F_BZ][s];
    }
    const double rh = rho_d;
    const double z0 = 0.0;
    e_part = pa.w * (vx * z0 + vy * z0 + vz * z0 + z0 + (bx * z0 + by * z0 + bz * z0) / rh -
                     0.5 * (bx * bx + by * by + bz * bz) / rh / rh * z0 / om_d);
  }
  __shared__ double red[4];
  double w = wave_sum(e_part);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = w;
  __syncthreads();
  if (threadIdx.x == 0) {
    double t = 0;
    for (int i = 0; i < (int)(blockDim.x >> 6); i++) t += red[i];
    a.energy_part[xcd_block(blockIdx.x, gridDim.x)] = t;
  }
}

// The end of a step in one block: loop 2's energy partials summed in a fixed
// order (thread t over partials t, t + 256, ..., then a fixed binary tree
// over the threads), its error bits and k_post1's largest h, as one small
// record the host takes with one copy.
__global__ __launch_bounds__(256) void k_step_close(const double *part, int np, const int *err,
                                                    const long long *hmax_red, long long *out) {
  __shared__ double sums[256];
  __shared__ long long hms[kStripes];
  const int t = threadIdx.x;
  double acc = 0.0;
  for (int k0 = t; k0 < np; k0 += 256 * 8) {  // eight loads in flight, added in order
    double v[8];
#pragma unroll
    for (int u = 0; u < 8; u++) v[u] = k0 + 256 * u < np ? part[k0 + 256 * u] : 0.0;
#pragma unroll
    for (int u = 0; u < 8; u++) acc += v[u];
  }
  sums[t] = acc;
  if (t < kStripes) hms[t] = hmax_red[t];
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if (t < w) {
      sums[t] += sums[t + w];
      if (w < kStripes) hms[t] = hms[t + w] > hms[t] ? hms[t + w] : hms[t];
    }
    __syncthreads();
  }
  if (t == 0) {
    out[0] = __double_as_longlong(sums[0]);
    out[1] = err[0];
    out[2] = hms[0];
  }
}

// ------------------------------------------------------------- transfers
// AoS store rows (any leading dimension) <-> SoA fields in sorted order.
__global__ void k_aos_to_soa(const double *aos, int64_t ld, int n, Fields F, int64_t *orig) {
  int s = gtid();
  if (s >= n) return;
  const double *row = aos + (size_t)s * ld;
#pragma unroll 4
  for (int f = 0; f < F_TOTAL; f++) F.f[f][s] = row[f];
  orig[s] = s;
}

__global__ void k_soa_to_aos(Fields F, const int64_t *orig, int n, double *aos, int64_t ld) {
  int s = gtid();
  if (s >= n) return;
  double *row = aos + (size_t)orig[s] * ld;
#pragma unroll 4
  for (int f = 0; f < F_TOTAL; f++) row[f] = F.f[f][s];
}

// Neighbour lists back in original numbering (for sph_neighbours); entries
// at or past `cap` (the output's length) are not written.
// ---------------------------------------------------------------- extended grid
// Extended cell k = (ez * ne1 + ey) * ne0 + ex is real cell (ex - gh0, ey - gh1,
// ez - gh2): on an imaged axis taken modulo the cell count, with the box
// shift of the wrap; past the end of a walled axis it holds nothing.
__device__ __forceinline__ bool ext_axis(int e, int gh, int nc, bool periodic, int &c, int &shift) {
  c = e - gh;
  shift = c < 0 ? -1 : (c >= nc ? 1 : 0);
  if (!periodic) return shift == 0;
  c -= shift * nc;
  return true;
}

// the cell key an extended cell copies, or false for an empty ghost cell
__device__ __forceinline__ bool ext_source(const Grid &g, const ExtGrid &x, int k, int sh[3], uint32_t &key) {
  const int ex = k % x.ne[0], r = k / x.ne[0];
  const int ey = r % x.ne[1], ez = r / x.ne[1];
  int cx, cy, cz;
  const bool ux = ext_axis(ex, x.gh[0], g.nc[0], g.periodic[0] != 0, cx, sh[0]);
  const bool uy = ext_axis(ey, x.gh[1], g.nc[1], g.periodic[1] != 0, cy, sh[1]);
  const bool uz = ext_axis(ez, x.gh[2], g.nc[2], g.periodic[2] != 0, cz, sh[2]);
  key = row_of(g, cy, cz) * (uint32_t)g.nc[0] + (uint32_t)cx;
  return ux && uy && uz;
}

// records per extended cell (count[ncells] = 0 closes the exclusive scan)
__global__ void k_ext_count(Grid g, ExtGrid x, const int *cstart, int ncells, int *count) {
  const int k = gtid();
  if (k > ncells) return;
  if (k == ncells) {
    count[k] = 0;
    return;
  }
  int sh[3];
  uint32_t key;
  count[k] = ext_source(g, x, k, sh, key) ? cstart[key + 1] - cstart[key] : 0;
}

// Copy every extended cell's records: the source records with the image
// shift added to the FP32 coordinates, a NaN x where the reference holds no
// such image (the band test of createPeriodicBorder, BC.f90:342-355), and the
// list entry (sorted index << 5 | image code) in the fourth word.
__global__ void k_ext_fill(Grid g, ExtGrid x, const int *cstart, const int *ecs, const float4 *pf, int ncells,
                           float4 *erec) {
  const int k = gtid();
  if (k >= ncells) return;
  int sh[3];
  uint32_t key;
  if (!ext_source(g, x, k, sh, key)) return;
  uint32_t bm = 0, bv = 0;
  image_mask(sh[0], sh[1], sh[2], bm, bv);
  const uint32_t code = (uint32_t)((sh[0] + 1) + 3 * (sh[1] + 1) + 9 * (sh[2] + 1));
  const float dx = (float)sh[0] * x.lenf[0], dy = (float)sh[1] * x.lenf[1], dz = (float)sh[2] * x.lenf[2];
  const int t0 = cstart[key], t1 = cstart[key + 1], o = ecs[k];
  for (int t = t0; t < t1; t++) {
    float4 q = pf[t];
    const bool there = (__float_as_uint(q.w) & bm) == bv;
    q.x = there ? q.x + dx : __int_as_float(0x7fc00000);
    q.y = q.y + dy;
    q.z = q.z + dz;
    q.w = __uint_as_float(nb_pack((uint32_t)t, code));
    erec[o + (t - t0)] = q;
  }
}

__global__ void k_neigh_export(const uint32_t *nb, const int *cnt, const int64_t *orig, const uint8_t *real,
                               int n, int npad, const int64_t *row_off_by_orig, int32_t *jout, int8_t *cout,
                               int64_t cap) {
  int s = gtid();
  if (s >= n || !real[s]) return;
  int64_t o = row_off_by_orig[orig[s]];
  const int m = (int)max<int64_t>(0, min<int64_t>(cnt[s], cap - o));
  for (int k = 0; k < m; k++) {
    uint32_t e = nb[nb_slot(k, s, npad)];
    jout[o + k] = (int32_t)orig[nb_index(e)];
    cout[o + k] = (int8_t)nb_code(e);
  }
}

// checkVarsReady's per-particle tests (preruncheck.f90:37-58) on every loaded
// particle whose type is not empty: bits 1 density, 2 sound speed, 4 h,
// 8 pressure, 16 internal energy <= 0 (a NaN passes, as in the reference).
__global__ void k_prerun_check(Fields F, int n, int *bits) {
  int s = gtid();
  if (s >= n || (int)F.f[F_TYPE][s] == 0) return;
  int b = 0;
  if (F.f[F_DEN][s] <= 0.0) b |= 1;
  if (F.f[F_C][s] <= 0.0) b |= 2;
  if (F.f[F_H][s] <= 0.0) b |= 4;
  if (F.f[F_P][s] <= 0.0) b |= 8;
  if (F.f[F_U][s] <= 0.0) b |= 16;
  if (b) atomicOr(bits, b);
}

// count long longs set to 0 (a kernel rather than a memset, so that a captured
// graph replays it).
__global__ void k_zero_ll(long long *p, int count) {
  int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t < count) p[t] = 0;
}

// One int64 set on the device (a kernel, not a memset, so a captured graph
// replays it).
__global__ void k_set_i64(int64_t *p, int64_t v) {
  if (threadIdx.x == 0 && blockIdx.x == 0) *p = v;
}

__global__ void k_counts_by_orig(const int *cnt, const int64_t *orig, const uint8_t *real, int n, int64_t *out) {
  int s = gtid();
  if (s >= n || !real[s]) return;
  out[orig[s]] = cnt[s];
}

__global__ void k_find_orig(const int64_t *orig, int n, int64_t target, int *pos) {
  int s = gtid();
  if (s < n && orig[s] == target) *pos = s;
}

// ----------------------------------------------------- step-driver updates
// These elementwise kernels are HBM-bound; they turn FP contraction off so
// every product and sum rounds as the reference's unfused Fortran does.
//
// Leapfrog predictor (main.f90:323-342) on rows 1..n of the reference, i.e.
// real and fixed particles (main.f90:181-184); saves the old derivatives in
// `sv` (ax, ay, az, du, dh, ddt, dbx, dby, dbz) for the corrector.
__global__ void k_predict(Fields F, int n, double dt, int do_t, double *sv, int npad) {
#pragma clang fp contract(off)
  int s = gtid();
  if (s >= n) return;
  const int src[9] = {F_AX, F_AY, F_AZ, F_DU, F_DH, F_DDT, F_DBX, F_DBY, F_DBZ};
  double o[9];
  for (int k = 0; k < 9; k++) { o[k] = F.f[src[k]][s]; sv[(size_t)k * npad + s] = o[k]; }
  for (int a = 0; a < 3; a++) {
    double v = F.f[F_VX + a][s];
    F.f[F_RX + a][s] = F.f[F_RX + a][s] + dt * v + 0.5 * dt * dt * o[a];
    F.f[F_VX + a][s] = v + dt * o[a];
    F.f[F_BX + a][s] = F.f[F_BX + a][s] + dt * o[6 + a];
  }
  F.f[F_U][s] = F.f[F_U][s] + dt * o[3];
  F.f[F_H][s] = F.f[F_H][s] + dt * o[4];
  if (do_t) F.f[F_T][s] = F.f[F_T][s] + dt * o[5];
}

// Leapfrog corrector (main.f90:348-362), same rows as the predictor.
__global__ void k_correct(Fields F, int n, double dt, int do_t, const double *sv, int npad) {
#pragma clang fp contract(off)
  int s = gtid();
  if (s >= n) return;
  double o[9];
  for (int k = 0; k < 9; k++) o[k] = sv[(size_t)k * npad + s];
  for (int a = 0; a < 3; a++) {
    F.f[F_VX + a][s] = F.f[F_VX + a][s] + 0.5 * dt * (F.f[F_AX + a][s] - o[a]);
    F.f[F_BX + a][s] = F.f[F_BX + a][s] + 0.5 * dt * (F.f[F_DBX + a][s] - o[6 + a]);
  }
  F.f[F_U][s] = F.f[F_U][s] + 0.5 * dt * (F.f[F_DU][s] - o[3]);
  F.f[F_H][s] = F.f[F_H][s] + 0.5 * dt * (F.f[F_DH][s] - o[4]);
  if (do_t) F.f[F_T][s] = F.f[F_T][s] + 0.5 * dt * (F.f[F_DDT][s] - o[5]);
}

// Per-real-particle minima/maxima for timestep.calc (timestep.f90:62-133):
// red[v * kStripes + stripe] (ordered int64): 0 min h, 1 max c, 2 min den,
// 3 min c, 4 max kappa, 5 max of the B components, 6 min kappa,
// 7 max flux limiter lambda(R), R = |grad T| / (kappa rho^2 T)
__global__ __launch_bounds__(256) void k_dt_reduce(Fields F, const uint8_t *real, int n, long long *red) {
#pragma clang fp contract(off)
  __shared__ long long lds[4];
  int s = gtid();
  const long long BIG = 0x7FFFFFFFFFFFFFFFLL, SMALL = -0x7FFFFFFFFFFFFFFFLL - 1;
  long long v[8] = {BIG, SMALL, BIG, BIG, SMALL, SMALL, BIG, SMALL};
  if (s < n && real[s]) {
    v[0] = d2ord(F.f[F_H][s]);
    v[1] = d2ord(F.f[F_C][s]);
    v[2] = d2ord(F.f[F_DEN][s]);
    v[3] = v[1];
    double ka = F.f[F_KAPPA][s];
    v[4] = d2ord(ka);
    double bm = fmax(fmax(F.f[F_BX][s], F.f[F_BY][s]), F.f[F_BZ][s]);
    v[5] = d2ord(bm);
    v[6] = v[4];
    double gx = F.f[F_DTDX][s], gy = F.f[F_DTDY][s], gz = F.f[F_DTDZ][s], da = F.f[F_DEN][s];
    double fr = sqrt(gx * gx + gy * gy + gz * gz) / (ka * da * da * F.f[F_T][s]);
    v[7] = d2ord((2. + fr) / (6. + 3 * fr + fr * fr));
  }
  const bool is_min[8] = {true, false, true, true, false, false, true, false};
  for (int k = 0; k < 8; k++) {
    long long t = is_min[k] ? block_reduce<R_MIN>(v[k], lds) : block_reduce<R_MAX>(v[k], lds);
    if (threadIdx.x == 0) {
      if (is_min[k]) atomicMin(&red[k * kStripes + stripe()], t);
      else atomicMax(&red[k * kStripes + stripe()], t);
    }
  }
}

// Legacy gravity of the mti set-ups after loop 2 (iterator.f90:143,153-154):
// mode 1 a_y -= 1, mode 2 a_y -= erf(8 y), on real particles.
__global__ void k_gravity(Fields F, const uint8_t *real, int n, int mode) {
  int s = gtid();
  if (s >= n || !real[s]) return;
  if (mode == 1) F.f[F_AY][s] = F.f[F_AY][s] - 1.;
  else F.f[F_AY][s] = F.f[F_AY][s] - erf(F.f[F_RY][s] * 8.);
}

// update_radenergy, rad_exchange.f90:9-66 with solve_internal_energy_implicit
// (:100-118): per particle, Newton on
//   f(u) = u/dt - u0/dt - du - ack (rho (etot - u)/a - (u cv1)^4),
// ack = 4 sigma kappa, from u = u0 while |u - u_old| > 1e-16, at most 10
// iterations; then u = u_new, T (radiation energy per mass) = etot - u_new.
// Powers are products in the reference's order; no FP contraction.
__global__ void k_rad_exchange(Fields F, int n, double dt, double a, double cv1, double sigmab, int *err) {
#pragma clang fp contract(off)
  int s = gtid();
  if (s >= n) return;
  const double ack = 4. * sigmab * F.f[F_KAPPA][s];
  const double rho = F.f[F_DEN][s], u0 = F.f[F_U][s], dudt = F.f[F_DU][s];
  const double etot = u0 + F.f[F_T][s];
  const double cv2 = cv1 * cv1, cv4 = cv2 * cv2;
  double unew = u0, uold = 2 * u0;
  for (int iter = 0; fabs(unew - uold) > 1e-16 && iter < 10; iter++) {
    uold = unew;
    const double x = unew * cv1, x2 = x * x;
    const double fu = unew / dt - u0 / dt - dudt - ack * (rho * (etot - unew) / a - x2 * x2);
    const double dfu = 1. / dt + ack * (rho / a + 4. * (unew * (unew * unew) * cv4));
    unew = unew - fu / dfu;
  }
  const double t = etot - unew;
  F.f[F_U][s] = unew;
  F.f[F_T][s] = t;
  if (!(t > 0.0) || !(unew > 0.0)) atomicOr(err, EB_RAD);
}

// RKL2 stage combination on T (sts_integrator.f90:111-115) over every
// particle slot, as the reference updates whole store rows.
__global__ void k_sts_stage(const double *y0, const double *yjm1, const double *yjm2, const double *dy0,
                            const double *dyjm1, int n, double mu, double nu, double muh_dt, double gmh_dt,
                            double *yj) {
#pragma clang fp contract(off)
  int s = gtid();
  if (s >= n) return;
  yj[s] = mu * yjm1[s] + nu * yjm2[s] + (1. - mu - nu) * y0[s] + muh_dt * dyjm1[s] + gmh_dt * dy0[s];
}

__global__ void k_axpy_copy(double *dst, const double *x, const double *y, double a, int n) {
#pragma clang fp contract(off)
  int s = gtid();
  if (s >= n) return;
  dst[s] = x[s] + a * y[s];
}

// ------------------------------------------------------------ instantiation
#define SPH_INST_DK(DIM, KT)                                                                 \
  template __global__ void k_loop1<DIM, KT, 0, double>(Loop1Args<double4>);                  \
  template __global__ void k_loop1<DIM, KT, 1, double>(Loop1Args<double4>);                  \
  template __global__ void k_loop1<DIM, KT, 2, double>(Loop1Args<double4>);                  \
  template __global__ void k_loop1<DIM, KT, 0, float>(Loop1Args<double4>);                   \
  template __global__ void k_loop1<DIM, KT, 1, float>(Loop1Args<double4>);                   \
  template __global__ void k_loop1<DIM, KT, 2, float>(Loop1Args<double4>);                   \
  template __global__ void k_loop1_simple<DIM, KT, 0, double>(Loop1Args<double4>, Fields);   \
  template __global__ void k_loop1_simple<DIM, KT, 1, double>(Loop1Args<double4>, Fields);   \
  template __global__ void k_loop1_simple<DIM, KT, 2, double>(Loop1Args<double4>, Fields);   \
  template __global__ void k_loop1_simple<DIM, KT, 0, float>(Loop1Args<double4>, Fields);    \
  template __global__ void k_loop1_simple<DIM, KT, 1, float>(Loop1Args<double4>, Fields);    \
  template __global__ void k_loop1_simple<DIM, KT, 2, float>(Loop1Args<double4>, Fields);    \
  template __global__ void k_loop2<DIM, KT, double, double4, MODE_GENERIC>(Loop2Args<double4>); \
  template __global__ void k_loop2<DIM, KT, float, float4, MODE_GENERIC>(Loop2Args<float4>);    \
  template __global__ void k_loop2<DIM, KT, double, double4, MODE_DIFF>(Loop2Args<double4>);    \
  template __global__ void k_loop2<DIM, KT, float, float4, MODE_DIFF>(Loop2Args<float4>);       \
  template __global__ void k_loop2_fab<DIM, KT, double, double4>(Loop2Args<double4>);         \
  template __global__ void k_loop2_fab<DIM, KT, float, float4>(Loop2Args<float4>);

SPH_INST_DK(1, 0) SPH_INST_DK(1, 1) SPH_INST_DK(1, 2)
SPH_INST_DK(2, 0) SPH_INST_DK(2, 1) SPH_INST_DK(2, 2)
SPH_INST_DK(3, 0) SPH_INST_DK(3, 1) SPH_INST_DK(3, 2)

#define SPH_INST_SPEC1(DIM, KT, M)                                                                    \
  template __global__ void k_loop2<DIM, KT, double, double4, MODE_SPEC + M>(Loop2Args<double4>); \
  template __global__ void k_loop2<DIM, KT, float, float4, MODE_SPEC + M>(Loop2Args<float4>);
#define SPH_INST_SPEC(DIM, KT) \
  SPH_INST_SPEC1(DIM, KT, kSpecHydro) SPH_INST_SPEC1(DIM, KT, kSpecMhd) SPH_INST_SPEC1(DIM, KT, kSpecHydroFld)
SPH_INST_SPEC(2, 0) SPH_INST_SPEC(2, 1) SPH_INST_SPEC(2, 2)
SPH_INST_SPEC(3, 0) SPH_INST_SPEC(3, 1) SPH_INST_SPEC(3, 2)

#define SPH_INST_NEIGH(DIM)                                                                              \
  template __global__ void k_neigh<DIM, true>(NeighArgs, const double4 *__restrict__, const float4 *__restrict__, \
                                              uint32_t *__restrict__, int *__restrict__);                       \
  template __global__ void k_neigh<DIM, false>(NeighArgs, const double4 *__restrict__, const float4 *__restrict__, \
                                               uint32_t *__restrict__, int *__restrict__);
SPH_INST_NEIGH(1)
SPH_INST_NEIGH(2)
SPH_INST_NEIGH(3)
template __global__ void k_post1<double4>(Fields, PostSel, const uint8_t *, int, int, int, int, double,
                                          double, double4 *, double4 *, double4 *, double4 *, double4 *, double4 *,
                                          double *, float2 *, long long *, long long *);
template __global__ void k_post1<float4>(Fields, PostSel, const uint8_t *, int, int, int, int, double,
                                         double, float4 *, float4 *, float4 *, float4 *, float4 *, double4 *,
                                         double *, float2 *, long long *, long long *);

}  // namespace sph
