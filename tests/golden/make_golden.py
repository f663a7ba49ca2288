"""Write the golden vectors under tests/golden/ from the reference Fortran build.

Runs the reference's own hot-path modules (compiled from
/root/reference/playground/src by oracle/ref.mk with amdflang, FP64,
-O2 -ffp-contract=off, no fast-math) through oracle/ref_harness.f90 on seeded
inputs and stores inputs + outputs as compressed .npz fixtures:

  state      (49,)   statevars
  store_in   (n,32)  store before the call (row i = Fortran column i+1)
  store_out  (n,32)  store after the call
  nb_off/nb_j/nb_code  neighbour lists of the last iterate: real index of j
                     (0-based) and image code (sx+1)+3(sy+1)+9(sz+1)
  meta       json: kernel, mode (0 iterate, 1 leapfrog steps, 2 STS), sts_s,
             dtcap, dt_used (of the last step), steps (leapfrog passes),
             walls (the ebc_* side of the fixed walls, 0 = none), generator call

Stores with fixed walls are built by problems.fixed_walls; before such a case
is written, the wall positions and types are checked bit for bit against the
reference's own createFixedBorders (BC.f90:477-640, harness mode 6) run on the
case's real particles.

Run from the repository root:  python tests/golden/make_golden.py
"""
import json
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

from oracle import refrun  # noqa: E402
from sph_playground_amd import consts as C  # noqa: E402
from sph_playground_amd import problems as P  # noqa: E402

OUT = os.path.dirname(os.path.abspath(__file__))


def _slab(**kw):
    kw.setdefault("hfac", 1.0)
    kw.setdefault("velocity_seed", 9)
    return P.random_slab3d(2000, **kw)


def cases():
    """(name, kernel, mode, sts_s, dtcap, builder, state overrides).  For
    mode 1, sts_s is the number of leapfrog passes (0 = 1)."""
    c = []
    c.append(("hydro2d_m6_32", 1, 0, 0, 0.0, lambda: P.hydro2d(32), {}))
    c.append(("hydro2d_m6_32_leapfrog", 1, 1, 0, 1e-3, lambda: P.hydro2d(32), {}))
    c.append(("hydro2d_simple_16", 1, 0, 0, 0.0, lambda: P.hydro2d(16, artts="no", adden="no"), {}))
    c.append(("hydro2d_m4_16", 0, 0, 0, 0.0, lambda: P.hydro2d(16, kernel=0, hfac=1.2), {}))
    c.append(("hydro2d_gauss_16", 2, 0, 0, 0.0, lambda: P.hydro2d(16, kernel=2, hfac=0.8), {}))
    c.append(("hydro1d_m6_64", 1, 0, 0, 0.0, lambda: P.hydro1d(64), {}))
    for ddw, tag in (("fab", "fab"), ("n2w", "n2w"), ("fw", "fw"), ("2nw-+", "2nwds"), ("2nw+-", "2nwsd")):
        for iso in ("yes", "no"):
            c.append(("diff2d_%s_%s_16" % (tag, "iso" if iso == "yes" else "aniso"), 1, 0, 0, 0.0,
                      (lambda ddw=ddw, iso=iso: P.diffusion2d(16, ddw=ddw, disotropic=iso, b=(1.0, 0.5, 0.0),
                                                              seed=5, jitter=0.05)), {}))
    for ddw, tag in (("2nw-+", "2nwds"), ("2nw+-", "2nwsd")):
        c.append(("diff2d_simple_%s_16" % tag, 1, 0, 0, 0.0,
                  (lambda ddw=ddw: P.diffusion2d(16, ddw=ddw, adden="no", seed=5, jitter=0.05)), {}))
    c.append(("diff2d_fab_sts3_16", 1, 2, 3, 2e-4,
              lambda: P.diffusion2d(16, ddw="fab", disotropic="no", b=(1.0, 0.5, 0.0), seed=5, jitter=0.05), {}))
    c.append(("diff3d_fab_aniso_8", 1, 0, 0, 0.0, lambda: P.diffusion3d(8, ddw="fab"), {}))
    c.append(("diff3d_fab_iso_8", 1, 0, 0, 0.0, lambda: P.diffusion3d(8, ddw="fab", disotropic="yes"), {}))
    c.append(("mhd3d_periodic_2000", 1, 0, 0, 0.0, lambda: _slab(seed=3), {}))
    c.append(("mhd3d_hydroshock_2000", 1, 0, 0, 0.0, lambda: _slab(seed=3, process="backcompatibility"),
              {C.ec_ics: C.ett_hydroshock}))
    c.append(("mhd3d_anisodiff_2000", 1, 0, 0, 0.0, lambda: _slab(seed=4, equations="mhd"),
              {C.ec_disotropic: 0.0}))
    c.append(("mhd3d_anisodiff_2nwsd_noau_2000", 1, 0, 0, 0.0, lambda: _slab(seed=4, equations="mhd", au=-1.0),
              {C.ec_disotropic: 0.0, C.ec_ddw: C.esd_2nw_sd}))
    c.append(("fld3d_fab_2000", 1, 0, 0, 0.0, lambda: _slab(seed=5, fld=True), {}))
    c.append(("fld3d_2nwds_2000", 1, 0, 0, 0.0, lambda: _slab(seed=5, fld=True), {C.ec_ddw: C.esd_2nw_ds}))
    c.append(("hyrad3d_2000", 1, 0, 0, 0.0, lambda: _slab(seed=5, equations="hyrad"), {}))
    # legacy mti set-up: periodic in x only, gravity a_y -= 1 after loop 2 (iterator.f90:134-143)
    c.append(("hydro2d_mti_16", 1, 0, 0, 0.0, lambda: P.hydro2d(16, au=1.0),
              {C.ec_process: C.epc_backcompatibility, C.ec_ics: C.ett_mti}))
    # one leapfrog step with the flux-limited-diffusion time step (timestep.f90:83-136)
    c.append(("fld3d_fab_3000_leapfrog", 1, 1, 0, 1.0,
              lambda: P.random_slab3d(3000, seed=7, hfac=0.85, velocity_seed=9, fld=True), {}))
    # radiative exchange alone (rad_exchange.f90:9-118), then after a hyrad leapfrog step
    c.append(("radexch3d_3000", 1, 5, 0, 1e-3,
              lambda: P.random_slab3d(3000, seed=7, hfac=0.85, equations="hyrad"), {}))
    c.append(("hyrad3d_3000_leapfrog_radexch", 1, 4, 0, 1.0,
              lambda: P.random_slab3d(3000, seed=7, hfac=0.85, velocity_seed=9, equations="hyrad"), {}))
    # several leapfrog passes: the initial iterate runs once (main.f90:219),
    # then each pass is calc, predictor, iterate, corrector (main.f90:241-362)
    c.append(("hydro2d_m6_32_leapfrog3", 1, 1, 3, 1e-3, lambda: P.hydro2d(32), {}))
    # fixed walls (createFixedBorders): the mti set-up's y walls (IC/IC.f90:170),
    # the fld_gauss set-up's walls on every side (place_fld_gauss.f90:22), the
    # hydroshock set-up's x walls (IC/IC.f90:253); fixed particles are
    # neighbours and are never updated (eos.f90:27-32, circuit1/2 real loops)
    c.append(("hydro2d_mti_walls_16", 1, 0, 0, 0.0, lambda: P.hydro2d_mti_walls(16), {}))
    c.append(("hydro2d_mti_walls_16_leapfrog3", 1, 1, 3, 1e-3,
              lambda: P.hydro2d_mti_walls(16, velocity_seed=9), {}))
    c.append(("fld2d_gauss_walls_20", 1, 0, 0, 0.0, lambda: P.fld2d_gauss_walls(20), {}))
    c.append(("mhd3d_hydroshock_walls_2000", 1, 0, 0, 0.0,
              lambda: P.shock_slab_walls(2000, seed=3, hfac=1.0, velocity_seed=9), {}))
    c.append(("fld3d_hydroshock_walls_2000", 1, 0, 0, 0.0,
              lambda: P.shock_slab_walls(2000, seed=5, hfac=1.0, velocity_seed=9, fld=True), {}))
    c.append(("mhd3d_hydroshock_walls_2000_leapfrog2", 1, 1, 2, 1e-3,
              lambda: P.shock_slab_walls(2000, seed=3, hfac=0.9, velocity_seed=9, lattice=(80, 5, 5)), {}))
    # many leapfrog passes: 50 of the 2D mti set-up with its walls; 10 of the
    # MHD shock tube with its x walls on a jittered lattice (2000 particles,
    # dt capped at 5e-4; in longer runs of this small set the flow passes the
    # reference's neighbour cap of 150 near the walls).  Their lists are not
    # stored, only the states
    c.append(("hydro2d_mti_walls_16_leapfrog50", 1, 101, 50, 1e-3,
              lambda: P.hydro2d_mti_walls(16, velocity_seed=9), {}))
    c.append(("mhd3d_shocktube_2000_leapfrog10", 1, 101, 10, 5e-4,
              lambda: P.shock_slab_walls(2000, seed=3, hfac=0.9, lattice=(80, 5, 5), equations="magnetohydro",
                                         velocity_seed=9), {}))
    return c


WALL_SIDE = {C.ett_mti: C.ebc_y, C.ett_fld_gauss: C.ebc_all, C.ett_hydroshock: C.ebc_x}


def check_walls(state, st, kernel):
    """The store's fixed particles are the reference createFixedBorders' walls
    of its real particles (positions and types bit for bit)."""
    nr = int(state[C.ec_realpn - 1])
    if st.shape[0] == nr:
        return 0
    side = WALL_SIDE[int(state[C.ec_ics - 1])]
    s0 = state.copy()
    s0[C.ec_fixedpn - 1] = 0
    r = refrun.run(s0, st[:nr], kernel=kernel, mode=6, sts_s=side)
    assert r["store"].shape[0] == st.shape[0], (r["store"].shape, st.shape)
    assert np.array_equal(r["store"][:, 0:3], st[:, 0:3])
    assert np.array_equal(r["store"][:, C.es_type - 1], st[:, C.es_type - 1])
    return side


def main():
    if not refrun.available(1):
        sys.exit("reference harness not built: make -C oracle ref")
    only_missing = "--only-missing" in sys.argv
    for name, kernel, mode, sts_s, dtcap, build, over in cases():
        if only_missing and os.path.exists(os.path.join(OUT, name + ".npz")):
            continue
        state, st = build()
        for k, v in over.items():
            state[k - 1] = v
        walls = check_walls(state, st, kernel)
        lists = mode < 100
        r = refrun.run(state, st, kernel=kernel, mode=mode, sts_s=sts_s, dtcap=dtcap, threads=1)
        nr = int(state[C.ec_realpn - 1])
        off = np.zeros(nr + 1, dtype=np.int64)
        neigh = r["neigh"] if r["neigh"] is not None else []   # mode 5 searches nothing
        if neigh:
            off[1:] = np.cumsum([len(a) for a in neigh])
        nb = np.concatenate(neigh) if off[-1] else np.zeros((0, 2), np.int32)
        meta = dict(kernel=kernel, mode=mode % 100, sts_s=sts_s, dtcap=dtcap, dt_used=r["dt"],
                    steps=max(1, sts_s) if mode % 100 in (1, 4) else 0, walls=walls, lists=lists,
                    reference_build="amdflang -fdefault-real-8 -fdefault-double-8 -O2 -fopenmp "
                                    "-ffp-contract=off, OMP_NUM_THREADS=1")
        np.savez_compressed(os.path.join(OUT, name + ".npz"), state=state, store_in=st,
                            store_out=r["store"], nb_off=off, nb_j=nb[:, 0].astype(np.int32),
                            nb_code=nb[:, 1].astype(np.int8), meta=json.dumps(meta))
        print("%-36s n=%5d pairs=%7d dt=%.3e" % (name, st.shape[0], off[-1], r["dt"]))


if __name__ == "__main__":
    main()
