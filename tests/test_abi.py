"""The C-ABI library: it loads, exports every entry point include/sph_hip.h
declares, and its host-only entry points work without a GPU."""
import ctypes as ct
import os
import re

import numpy as np
import pytest

from sph_playground_amd import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    txt = open(os.path.join(ROOT, "include", "sph_hip.h")).read()
    return sorted(set(re.findall(r"^\s*(?:int|const char \*)\s*(sph_\w+)\s*\(", txt, re.M)))


def test_header_and_binding_agree():
    assert set(declared_symbols()) == set(_lib.EXPORTS)


def test_library_exports_every_declared_symbol():
    L = _lib.lib()
    for name in declared_symbols():
        assert hasattr(L, name), name


def test_default_options():
    o = _lib.Options()
    assert _lib.lib().sph_default_options(ct.byref(o)) == 0
    assert (o.precision, o.tensor_mode, o.max_sweeps, o.newton_tol) == (0, 0, 100, 0.0)


def test_error_strings_mirror_reference_messages():
    L = _lib.lib()
    assert L.sph_error_string(1) == b"# <!> KDTree neibsearch result was truncated."
    assert L.sph_error_string(2) == b"Negative smoothing length in c1 advanced"


def test_sts_coefficients_known_answers():
    # sts_integrator.f90:54-68 (the reference's own `test` routine)
    from sph_playground_amd.engine import sts_coeffs
    mu2, nu2, _, gm23 = sts_coeffs(2, 3)
    mu3, nu3, mh33, gm33 = sts_coeffs(3, 3)
    mh13 = sts_coeffs(1, 3)[2]
    mh23 = sts_coeffs(2, 3)[2]
    assert mu2 == pytest.approx(3 / 2, abs=1e-15)
    assert mu3 == pytest.approx(25 / 12, abs=1e-15)
    assert nu2 == pytest.approx(-1 / 2, abs=1e-15)
    assert nu3 == pytest.approx(-5 / 6, abs=1e-15)
    assert mh13 == pytest.approx(2 / 15, abs=1e-15)
    assert mh23 == pytest.approx(3 / 5, abs=1e-15)
    assert mh33 == pytest.approx(5 / 6, abs=1e-15)
    assert gm23 == pytest.approx(-2 / 5, abs=1e-15)
    assert gm33 == pytest.approx(-5 / 9, abs=1e-15)
    assert 1 - mu2 - nu2 == pytest.approx(0.0, abs=1e-15)
    assert 1 - mu3 - nu3 == pytest.approx(-1 / 4, abs=1e-15)


def test_sts_coefficients_match_oracle():
    from oracle import oracle
    from sph_playground_amd.engine import sts_coeffs
    for s in range(2, 12):
        for j in range(2, s + 1):
            mu, nu, mh, gm = sts_coeffs(j, s)
            assert mu == oracle.sts("mu", j) and nu == oracle.sts("nu", j)
            assert mh == oracle.sts("mu_hat", j, s) and gm == oracle.sts("gm_hat", j, s)


def test_context_creation_fails_loudly_without_gpu():
    from tests.conftest import gpu_available
    if gpu_available():
        pytest.skip("GPU present")
    from sph_playground_amd import problems as P
    from sph_playground_amd.engine import Engine
    state, _ = P.hydro2d(8)
    with pytest.raises(_lib.SphError) as e:
        Engine(state)
    assert e.value.code == 8  # SPH_ERR_HIP: no silent CPU path


def test_bad_arguments_rejected():
    from sph_playground_amd import problems as P
    L = _lib.lib()
    state, _ = P.hydro2d(8)
    h = ct.c_void_p()
    s = np.ascontiguousarray(state)
    assert L.sph_ctx_create(s.ctypes.data_as(ct.POINTER(ct.c_double)), 7, None, ct.byref(h)) == 7
    s2 = s.copy()
    s2[0] = 4.0  # dim 4
    assert L.sph_ctx_create(s2.ctypes.data_as(ct.POINTER(ct.c_double)), 1, None, ct.byref(h)) == 7
    s3 = s.copy()
    s3[6] = 2.0  # cylindrical (kernel.f90:135-137)
    assert L.sph_ctx_create(s3.ctypes.data_as(ct.POINTER(ct.c_double)), 1, None, ct.byref(h)) == 10
    # a neighbour cap outside [0, 2^20] is refused before any device call
    for cap in (-1, (1 << 20) + 1):
        o = _lib.Options()
        L.sph_default_options(ct.byref(o))
        o.neigh_cap = cap
        assert L.sph_ctx_create(s.ctypes.data_as(ct.POINTER(ct.c_double)), 1, ct.byref(o), ct.byref(h)) == 7
    # timing needs a context
    assert L.sph_set_timing(None, 1) == 7
