// Device kernels of the two-loop hot path and their argument blocks.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "sph_device.h"
#include "sph_types.h"

namespace sph {

// store rows (0-based es_* - 1), const.f90:69-103
enum : int {
  F_RX = 0, F_RY, F_RZ, F_VX, F_VY, F_VZ, F_AX, F_AY, F_AZ, F_M, F_DEN, F_H, F_DH, F_P, F_U, F_DU, F_T,
  F_KAPPA, F_DTDX, F_DTDY, F_DTDZ, F_DDT, F_BX, F_BY, F_BZ, F_DBX, F_DBY, F_DBZ, F_C, F_OM, F_TYPE,
  F_VFLD, F_TOTAL
};

// Error bits raised by kernels (mapped to SPH_ERR_* by the host).
enum : int {
  EB_TRUNC = 1,      // neighbour count + self > maxneibnum
  EB_NEG_H = 2,      // Newton step gave h <= 0
  EB_Q = 4,          // q NaN or negative
  EB_OVERFLOW = 8,   // neighbour storage too small (host grows it and reruns)
  EB_OUTSIDE = 16,   // particle outside the box on an imaged axis
  EB_RAD = 32,       // energy <= 0 after the radiative exchange
};

struct Fields {
  double *f[F_TOTAL];
};

// Neighbour entries are stored in chunks of 8 per particle:
//   nb[(k >> 3) * npad * 8 + s * 8 + (k & 7)]
// so each lane writes (and later reads, as two 16-byte loads) its own
// 32-byte sector, and a wavefront reads 64 consecutive sectors.
__host__ __device__ inline size_t nb_slot(int k, int s, int npad) {
  return (size_t)(k >> 3) * (size_t)npad * 8 + (size_t)s * 8 + (size_t)(k & 7);
}

constexpr int kStripes = 64;

// Loop-2 kernel variants.  MODE_GENERIC: every term under run-time flags.
// MODE_DIFF: diffusion with a Hessian operator (fab/fw/n2w) and no
// hydro/AV/MHD/FLD terms, the density+force step of the diffusion workloads.
// MODE_FAB: MODE_DIFF with the fab operator, one tensor for every pair and
// no velocity term (k_loop2_fab).
enum { MODE_GENERIC = 0, MODE_DIFF = 1, MODE_FAB = 2 };
// MODE_SPEC + mask: MODE_GENERIC with its equation set fixed at compile time,
// mask = hydro | magneto << 1 | diff << 2 | fld << 3 (terms outside the set
// are dead code, so the kernel holds fewer registers); built in 2D and 3D for
// hydro (1), magnetohydro (3) and hydro with flux-limited diffusion (13)
constexpr int MODE_SPEC = 16;
constexpr int kSpecHydro = 1, kSpecMhd = 3, kSpecHydroFld = 13;  // counters are striped over 64 slots to avoid same-address atomics

// Positions, search records, lists and counts travel as separate
// __restrict__ kernel arguments.
struct NeighArgs {
  const double *hs;        // h used for the search radius (pre-loop-1 h)
  const uint8_t *real;     // 1 = real particle
  const int *cstart;       // first sorted index of every cell key (nkeys + 1 entries)
  int rx, ry, rz;          // cells within r_max (1 + 1e-6) of a cell along x, y, z
  float mcell[3];          // margin (cells) covering FP32 rounding of a position in cell units
  Grid g;
  Box b;
  float lenf[3];           // box lengths in FP32 (image shifts of the search records)
  double delta;            // bound on the FP32 coordinate error of a search record
  double krad;
  int n, npad, cap, maxneib;
  int *err;                // error bits
  long long *stripes;      // [kStripes] pair counts, [kStripes..2k) max counts
  // extended grid: the real cells plus gh[a] ghost cells on each side of every
  // axis (ne[a] = nc[a] + 2 gh[a] cells; empty on a walled axis, the periodic
  // images on an imaged one), ecs the first record of every
  // extended cell (row-major, x fastest; ne0 ne1 ne2 + 1 entries) and erec the
  // records (the per-lane traversal reads pf and cstart instead)
  const int *ecs;
  const float4 *erec;
  int ne[3], gh[3];
};

// The extended grid's geometry for its build kernels.
struct ExtGrid {
  int ne[3], gh[3];
  float lenf[3];  // box lengths in FP32 (0 on axes that are not imaged)
};

// Neighbour iteration with batched gathers: entries come 8 per coalesced
// chunk; for every NB of them `load(entry)` issues all gathers first (slots
// past the list end reload particle s, always a valid address), then
// `body(data)` runs on the valid ones.  NB gathers are then in flight at once.
template <int NB, typename Load, typename Body>
__device__ __forceinline__ void for_neighbours(const uint32_t *nb, int s, int npad, int nn, Load load, Body body) {
  const uint32_t self = nb_pack((uint32_t)s, kNoImage);
  for (int k0 = 0; k0 < nn; k0 += 8) {
    const uint4 *p = reinterpret_cast<const uint4 *>(nb + nb_slot(k0, s, npad));
    const uint4 lo = p[0], hi = p[1];
    const uint32_t e8[8] = {lo.x, lo.y, lo.z, lo.w, hi.x, hi.y, hi.z, hi.w};
    const int m = min(8, nn - k0);
#pragma unroll
    for (int h = 0; h < 8; h += NB) {
      if (h < m) {
        decltype(load(0u)) jd[NB];
#pragma unroll
        for (int u = 0; u < NB; u++) jd[u] = load(h + u < m ? e8[h + u] : self);
#pragma unroll
        for (int u = 0; u < NB; u++)
          if (h + u < m) body(jd[u]);
      }
    }
  }
}

// The mixed path's 12-byte position record (pfix): per axis the position as
// a 31-bit fraction, u = (x - base) mul.  An imaged axis maps its box length L
// onto 2^(31-k) with 2^k L above the stencil's reach, an axis without images
// maps the grid span onto 2^31.  A pair's separation along an axis is the
// 32-bit wrap-around difference u_a - u_b - s step (s the image shift, step
// the mapped box length), exact while the separation is under 2^k L, times
// scale in FP32.
struct FixFrame {
  double base[3], mul[3];  // positions -> fractions (k_prep)
  float scale[3];          // fraction units -> length
  uint32_t step[3];        // mapped box length of an imaged axis (0 otherwise)
};

template <typename V4>
struct Loop1Args {
  const double4 *pm;
  const double *h0;        // h at entry (ha0 of circuit1.f90:133)
  const uint8_t *real;
  const uint32_t *nb;
  const int *cnt;
  const V4 *cur;           // h, den, om, T of sweep k
  // sweep 1 only: the fields it starts from (h, den, om, T; null later)
  const double *f_h, *f_den, *f_om, *f_t;
  V4 *next;                // sweep k+1
  const double *trc;       // T den of sweep k (the only j-side value grad T needs)
  double *trn;             // T den of sweep k+1
  // mixed path, grad T: the position records, and (m, T den) in FP32 of
  // sweep k and of sweep k+1
  const uint3 *pfix;
  const float2 *s1c;
  float2 *s1n;
  FixFrame ff;
  double *eps;
  double *dtdx, *dtdy, *dtdz;
  long long *stripes;      // [kStripes] still-active counts, [kStripes..2k) evaluated pairs, [2k] ran
  int *err;
  Box b;
  Phys p;
  double tol;
  int n, npad;
  // sweeps decided on the device: the sweep runs only if the previous one
  // (its stripes at `prev`, null for the first) left a particle active and no
  // error stopped loop 1 or the search (`serr`); a sweep that runs records
  // its parity in *which, the buffer its results went to
  const long long *prev;
  const int *serr;
  int *which;
  int sweep;
};

// Where k_post1 finds loop 1's result: st[*which] of the two sweep buffers
// (fin0 alone when which is null); nothing is written when the search (serr)
// or loop 1 (err1) raised an error bit: the step then stops or searches again.
struct PostSel {
  const double4 *fin0, *fin1;
  const int *which;
  const int *serr, *err1;
};

template <typename V4>
struct Loop2Args {
  const double4 *pm;
  const uint8_t *real;
  const uint32_t *nb;
  const int *cnt;
  const V4 *q0;  // 1/h, den, om, T
  const V4 *q1;  // vx, vy, vz, p
  const V4 *q2;  // c, u, kappa, -
  const V4 *q3;  // dtdx, dtdy, dtdz, -
  const V4 *q4;  // bx, by, bz, -
  const double4 *pd;  // MODE_DIFF, FP64: x, y, z, m / den
  const double *trd;  // MODE_DIFF, FP64: T den
  const uint3 *pfix;    // MODE_DIFF, mixed: the position records
  const float2 *s2;     // MODE_DIFF, mixed: m / den, T den in FP32
  FixFrame ff;
  int v_nonzero;  // 0: every velocity is exactly 0 (v_ab = 0 for all pairs)
  int b_uniform;  // 1: every particle carries the same B bit for bit (kappa_b = kappa_a)
  Fields out;
  double *energy_part;   // one FP64 partial per block
  int *err;
  Box b;
  Phys p;
  int n, npad;
};

// Up to four adjoint stages' (lam, Zt) records (sph_adjoint.hip).
// One RKL2 stage over merged problems (k_batch_stage): list-layout operator
// (cnt, nb, W), densities, each row's problem (pid) and the per-problem time
// steps; stage states Y_{j-1}, Y_{j-2}, Y_0, Z_{j-1} = Y_{j-1} rho, the
// stage-1 rate DY0 and the outputs Y_j, Z_j; coefficients mu, nu,
// 1 - mu - nu, mu_hat, gamma_hat.
struct BatchStage {
  const int *cnt;
  const uint32_t *nb;
  const double *W, *rho;
  const float *W32;          // FP32 weights (precision 1) instead of W
  const int *pid;
  const double *dt;
  const double *Yjm1, *Yjm2, *Y0, *Zjm1;
  double *DY0, *Yj, *Zj;
  double mu, nu, c0, mhat, ghat;
  int n, npad;
};

struct LZSet {
  const double2 *p[4];
};

// The differentiable diffusion operator (sph_adjoint.hip).
struct DiffOpArgs {
  const double4 *pm;   // x, y, z, m (sorted)
  const double4 *q0;   // 1/h, den, om, - (sorted)
  const uint8_t *real;
  const uint32_t *nb;
  const int *cnt;
  const int *rcnt;     // reverse-list lengths
  const uint32_t *rent;  // reverse entries, nb_slot layout: owner a and the image of b in a's list
  const double4 *pxh;  // x, y, z, 1 / h (sorted)
  Box b;
  double K[6];         // xx, yy, zz, xy, xz, yz
  double wcv, fwc;
  int ddw, n, npad;
  const double *kap;   // opacity of the frozen state (sorted), flux-limited diffusion
};

}  // namespace sph
