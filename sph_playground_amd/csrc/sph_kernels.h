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
};

struct Fields {
  double *f[F_TOTAL];
};

struct NeighArgs {
  const double4 *pm;       // x, y, z, m (sorted)
  const double *hs;        // h used for the search radius (pre-loop-1 h)
  const uint16_t *band;    // ghost-band codes
  const uint8_t *real;     // 1 = real particle
  const int *cstart, *cend;
  Grid g;
  Box b;
  double krad;
  int n, npad, cap, maxneib;
  uint32_t *nb;            // [cap][npad]
  int *cnt;                // [n]
  int *err;                // error bits
  int *maxcnt;
  unsigned long long *pairs;
};

template <typename V4>
struct Loop1Args {
  const double4 *pm;
  const double *h0;        // h at entry (ha0 of circuit1.f90:133)
  const uint8_t *real;
  const uint32_t *nb;
  const int *cnt;
  const V4 *cur;           // h, den, om, T of sweep k
  V4 *next;                // sweep k+1
  double *eps;
  double *dtdx, *dtdy, *dtdz;
  const int *prev_active;  // active count after the previous sweep (skip if 0)
  int *active;             // active count after this sweep
  int *err;
  unsigned long long *pairs;
  Box b;
  Phys p;
  double tol;
  int n, npad;
};

template <typename V4>
struct Loop2Args {
  const double4 *pm;
  const uint8_t *real;
  const uint32_t *nb;
  const int *cnt;
  const V4 *q0;  // 1/h, den, om, p
  const V4 *q1;  // vx, vy, vz, c
  const V4 *q2;  // u, T, kappa, -
  const V4 *q3;  // dtdx, dtdy, dtdz, -
  const V4 *q4;  // bx, by, bz, -
  Fields out;
  double *energy_part;   // one FP64 partial per block
  int *err;
  Box b;
  Phys p;
  int n, npad;
};

}  // namespace sph
