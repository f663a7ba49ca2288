"""The C restatement (oracle/) against the reference's own Fortran engine.

Golden vectors in tests/golden/ were written by tests/golden/make_golden.py
from the reference hot-path modules compiled with amdflang (oracle/ref.mk).
The oracle must reproduce the neighbour SETS exactly (only exact FP64 ties on
the search radius may differ, where the kd-tree's box pruning rounds its lower
bound differently) and every written store row to 1e-12 relative.
"""
import numpy as np
import pytest

from oracle import oracle
from sph_playground_amd import consts as C
from tests import goldens as G

ITERATE_CASES = G.names(mode=0)


@pytest.mark.parametrize("name", ITERATE_CASES)
def test_oracle_matches_reference(name):
    g = G.load(name)
    nr = int(g["state"][C.ec_realpn - 1])
    o = oracle.iterate(g["state"], g["store_in"], kernel=g["meta"]["kernel"])
    bad = G.compare_neighbours(g, o["off"], o["j"], o["code"])
    assert not bad, "neighbour mismatches (not exact ties): %s" % bad[:5]
    errs = G.field_errors(g["store_out"], o["store"], nr)
    worst = max(errs.items(), key=lambda kv: kv[1])
    assert worst[1] <= 1e-12, errs


def test_golden_set_is_complete():
    # every --ddw operator, every kernel, MHD, FLD, hydroshock branch, 1D/2D/3D
    names = set(G.names())
    for tag in ("fab", "n2w", "fw", "2nwds", "2nwsd"):
        assert any(tag in n for n in names), tag
    for n in ("hydro2d_m4_16", "hydro2d_gauss_16", "hydro1d_m6_64", "mhd3d_hydroshock_2000",
              "fld3d_fab_2000", "hyrad3d_2000", "diff2d_fab_sts3_16", "hydro2d_m6_32_leapfrog",
              "hydro2d_mti_16", "fld3d_fab_3000_leapfrog", "radexch3d_3000",
              "hyrad3d_3000_leapfrog_radexch", "hydro2d_m6_32_leapfrog3", "hydro2d_mti_walls_16",
              "hydro2d_mti_walls_16_leapfrog3", "fld2d_gauss_walls_20", "mhd3d_hydroshock_walls_2000",
              "fld3d_hydroshock_walls_2000", "mhd3d_hydroshock_walls_2000_leapfrog2"):
        assert n in names


@pytest.mark.parametrize("name", [n for n in G.names() if G.load(n)["meta"].get("walls")])
def test_fixed_walls_match_reference_create_fixed_borders(name):
    """problems.fixed_walls restates createFixedBorders (BC.f90:477-640): on the
    golden's real particles it must give the reference-built walls (positions
    and types of every fixed row) bit for bit, and ec_fixedpn."""
    from sph_playground_amd import problems as P
    g = G.load(name)
    nr = int(g["state"][C.ec_realpn - 1])
    s0 = g["state"].copy()
    s0[C.ec_fixedpn - 1] = 0
    s1, st = P.fixed_walls(s0, g["store_in"][:nr], g["meta"]["walls"])
    assert st.shape[0] == g["store_in"].shape[0] > nr
    assert np.array_equal(st[:, 0:3], g["store_in"][:, 0:3])
    assert np.array_equal(st[:, C.es_type - 1], g["store_in"][:, C.es_type - 1])
    assert s1[C.ec_fixedpn - 1] == g["state"][C.ec_fixedpn - 1]
    assert (g["store_in"][nr:, C.es_type - 1] == C.ept_fixed).all()


def test_tie_rule_detects_real_mismatch():
    """Plant a mistake in the oracle's output (drop one neighbour): the
    comparison must flag it, since it is not an exact tie."""
    g = G.load("hydro2d_m6_32")
    o = oracle.iterate(g["state"], g["store_in"], kernel=1)
    off, j, code = o["off"].copy(), o["j"].copy(), o["code"].copy()
    # drop the first neighbour of particle 0
    j2 = np.delete(j, 0)
    c2 = np.delete(code, 0)
    off2 = off.copy()
    off2[1:] -= 1
    bad = G.compare_neighbours(g, off2, j2, c2)
    assert len(bad) == 1 and bad[0][0] == 0


def test_oracle_error_codes_mirror_reference_aborts():
    from sph_playground_amd import problems as P
    # gauss kernel with h far too small: the reference stops with
    # "Negative smoothing length in c1 advanced" (circuit1.f90:202-204)
    state, st = P.hydro2d(16, kernel=2, hfac=0.3)
    assert oracle.iterate(state, st, kernel=2, check=False)["rc"] == 2
    # random 3D set with hfac 1.2 overflows the M6 cap of 150 (neighboursearch.f90:127-132)
    state, st = P.random_slab3d(2000, seed=3, hfac=1.2)
    assert oracle.iterate(state, st, kernel=1, check=False)["rc"] == 1


def test_oracle_rad_exchange_matches_reference():
    """update_radenergy (rad_exchange.f90:9-118) alone: every row of the
    reference's output, 1e-12 relative; u + T is conserved exactly and the
    Newton residual of the implicit equation vanishes."""
    g = G.load("radexch3d_3000")
    st, rc = oracle.rad_exchange(g["state"], g["store_in"], g["meta"]["dtcap"])
    assert rc == 0
    n = st.shape[0]
    errs = G.field_errors(g["store_out"], st, n)
    assert max(errs.values()) <= 1e-12, errs
    u0 = g["store_in"][:, C.es_u - 1]
    t0 = g["store_in"][:, C.es_t - 1]
    u = st[:, C.es_u - 1]
    assert np.array_equal(u + st[:, C.es_t - 1], u0 + t0) or \
        np.abs(u + st[:, C.es_t - 1] - (u0 + t0)).max() <= 1e-15 * np.abs(u0 + t0).max()
    # residual of f(u) = (u - u0)/dt - du - ack (rho (etot - u)/a - (u cv1)^4)
    dt = g["meta"]["dtcap"]
    sigmab, c, rg = 5.67051e-5, 2.997924e10, 8.314e7
    a = 4 * sigmab / c
    cv1 = (g["state"][C.ec_gamma - 1] - 1) * g["state"][C.ec_molecularmass - 1] / rg
    ack = 4 * sigmab * g["store_in"][:, C.es_kappa - 1]
    rho = g["store_in"][:, C.es_den - 1]
    du = g["store_in"][:, C.es_du - 1]
    f = (u - u0) / dt - du - ack * (rho * (u0 + t0 - u) / a - (u * cv1) ** 4)
    scale = np.abs(u / dt) + np.abs(u0 / dt) + np.abs(du) + np.abs(ack * rho * (u0 + t0) / a)
    assert np.max(np.abs(f) / scale) <= 1e-13


def test_oracle_rad_exchange_flags_negative_energy():
    """A planted negative radiation energy (on the CPU, in the oracle's input)
    makes the exchange report the reference's stop (rad_exchange.f90:60-63)."""
    g = G.load("radexch3d_3000")
    st = g["store_in"].copy()
    st[0, C.es_t - 1] = -10.0 * (abs(st[0, C.es_u - 1]) + 1.0)
    _, rc = oracle.rad_exchange(g["state"], st, g["meta"]["dtcap"], check=False)
    assert rc == 11
