"""The two traversals of the neighbour search give the same lists.

By default k_neigh reads each stencil row as one range of an extended cell
grid whose ghost cells hold the periodic images (pre-shifted FP32
coordinates, missing images as NaN, the list entry in the record);
SPH_NEIGH_TRAVERSAL=lane makes every lane split the row into the parts left
of, inside and right of the box and apply the shift and the image test
itself.  Both must produce the kd-tree's exact sets in the same order, so the
neighbour lists agree entry for entry and every output of the step agrees
bit for bit.  The cases cover 1D, 2D and 3D, lattices and random sets,
periodic and walled axes, fixed wall particles, radii that vary between
particles, every cell size, and boxes under two stencils wide.
"""
import os

import numpy as np
import pytest

from sph_playground_amd import consts as C
from sph_playground_amd import problems as P

pytestmark = pytest.mark.gpu


def _step(state, st, lane, **kw):
    from sph_playground_amd.engine import Engine
    old = os.environ.get("SPH_NEIGH_TRAVERSAL")
    if lane:
        os.environ["SPH_NEIGH_TRAVERSAL"] = "lane"
    else:
        os.environ.pop("SPH_NEIGH_TRAVERSAL", None)
    try:
        eng = Engine(state, **kw)
    finally:
        if old is None:
            os.environ.pop("SPH_NEIGH_TRAVERSAL", None)
        else:
            os.environ["SPH_NEIGH_TRAVERSAL"] = old
    try:
        out = st.copy()
        eng.iterate(out)
        off, j, code = eng.neighbours()
        return out, off, j, code, eng.stats()
    finally:
        eng.close()


def _cases():
    cs = []
    cs.append(("lattice3d_32", lambda: P.diffusion3d(32, ddw="fab"), dict(kernel="m6")))
    cs.append(("lattice3d_48_square", lambda: P.diffusion3d(48, ddw="n2w", kind="square"), dict(kernel="m6")))
    for d in (1, 2, 4):
        cs.append(("lattice3d_40_cells%d" % d, lambda: P.diffusion3d(40, ddw="fab"), dict(kernel="m6", cell_div=d)))
    cs.append(("random3d_periodic", lambda: P.random_slab3d(20000, seed=7, hfac=1.2),
               dict(kernel="m6", neigh_cap=400)))
    cs.append(("shock3d_walls", lambda: P.shock_slab_walls(20000, seed=5, hfac=1.0),
               dict(kernel="m6", neigh_cap=400)))
    cs.append(("hydro2d_64", lambda: P.hydro2d(64, seed=3, jitter=0.05), dict(kernel="m6")))
    cs.append(("hydro2d_12_small_box", lambda: P.hydro2d(12, seed=4, jitter=0.05), dict(kernel="m4")))
    cs.append(("hydro2d_8_tiny_box", lambda: P.hydro2d(8, seed=8, jitter=0.05, hfac=1.2), dict(kernel="m6")))
    cs.append(("mti2d_walls", lambda: P.hydro2d_mti_walls(32, seed=2), dict(kernel="m6")))
    cs.append(("fld2d_gauss_walls", lambda: P.fld2d_gauss_walls(24), dict(kernel="m6")))
    cs.append(("hydro1d", lambda: P.hydro1d(3000, seed=5, hfac=1.2), dict(kernel="m6")))
    cs.append(("hydro1d_dense", lambda: P.hydro1d(4000, seed=6, hfac=70.0, kernel=C.KERNEL_M4),
               dict(kernel="m4", neigh_cap=1000)))
    return cs


@pytest.mark.parametrize("name,make,kw", _cases(), ids=[c[0] for c in _cases()])
def test_extended_grid_and_per_lane_traversals_agree(gpu, name, make, kw):
    print("case", name, flush=True)
    state, st = make()
    out_a, off_a, j_a, code_a, s_a = _step(state, st, False, **kw)
    out_b, off_b, j_b, code_b, s_b = _step(state, st, True, **kw)
    n = st.shape[0]
    assert s_b["ext_records"] == 0
    assert s_a["ext_records"] >= n, s_a
    assert np.array_equal(off_a, off_b)
    assert np.array_equal(j_a, j_b)
    assert np.array_equal(code_a, code_b)
    assert np.array_equal(out_a, out_b, equal_nan=True)


@pytest.mark.parametrize("lane", [False, True], ids=["extended_grid", "per_lane"])
def test_lists_past_the_cap_report_truncation(gpu, lane):
    # 114 neighbours per particle against a cap of 16: the extended-grid
    # traversal stops a list 8 entries past the cap, and both traversals
    # report the reference's truncation (neighboursearch.f90:127-132)
    from sph_playground_amd._lib import SphError
    state, st = P.diffusion3d(32, ddw="fab")
    with pytest.raises(SphError) as e:
        _step(state, st, lane, kernel="m6", neigh_cap=16)
    assert e.value.code == 1
