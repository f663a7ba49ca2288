// Plain parameter blocks shared by host code and device kernels.
#pragma once
#include <stdint.h>

namespace sph {

// Periodic box and the reference's ghost-band geometry (BC.f90:46-68, 196-309).
struct Box {
  double emin[3], emax[3];   // ec_xmin..ec_zmax (zeroed for unused dims, place.f90:70-73)
  double boxmin[3], boxmax[3];  // emin + bordersize, emax - bordersize (BC.f90:63-64)
  double pmin[3], pmax[3];      // padding-adjusted edges (BC.f90:210-223)
  int wrap_mask;   // axes wrapped by reflectRealParticlesPeriodically (bit a)
  int ghost_mask;  // axes imaged by createPeriodicBorder (bit a)
};

// Cell grid over the (wrapped) particle set.
struct Grid {
  double lo[3];
  double inv_cs[3];
  double cs[3];      // cell widths (the box of cell k spans lo + [k, k+1) cs)
  int nc[3];
  int periodic[3];  // 1 if the axis is imaged (cells wrap with a shift)
  int mdim;         // dimension of the sort key (= dim)
  int ty;           // 3D: tiles of kRowTile rows along y in the sort key
};

// 3D sort key: (y, z) rows in tiles of kRowTile x kRowTile rows.
#ifndef SPH_ROW_TILE_LOG2
#define SPH_ROW_TILE_LOG2 3
#endif
constexpr int kRowTileLog2 = SPH_ROW_TILE_LOG2;
constexpr int kRowTile = 1 << kRowTileLog2;

// Per-call physics flags and constants (circuit1.f90:38-51, circuit2.f90:45-56,86-90).
struct Phys {
  int dim, kt, ddw;
  int adden, artts, diso;
  int hydro, magneto, diff, fld;
  int tensor_mode;   // 0: reference diag(b^2)/isotropic tensor, 1: full kappa_par bb + kappa_perp (I-bb)
  double hfac, dcond, kappa_perp, muzero, au, gamma;
  double wcv, fwc, krad;
};

}  // namespace sph
