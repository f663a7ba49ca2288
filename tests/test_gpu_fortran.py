"""The Fortran drop-in (sph_playground_amd/fortran/iterator_hip.f90) and the
reference's own iterate() in one Fortran program on the same inputs
(oracle/ref_vs_hip.f90, built from the reference sources where they exist)."""
import os
import re
import subprocess
import tempfile

import numpy as np
import pytest

from oracle import refrun
from tests import goldens as G

pytestmark = pytest.mark.gpu

KNAMES = {0: "m4", 1: "m6", 2: "gauss"}


def _binary(kernel):
    return os.path.join(os.path.dirname(refrun.HERE), "oracle", "_ref", KNAMES[kernel], "ref_vs_hip")


@pytest.mark.parametrize("name", ["hydro2d_m6_32", "diff3d_fab_aniso_8", "mhd3d_periodic_2000",
                                  "fld3d_fab_2000", "hydro2d_m4_16", "hydro1d_m6_64"])
def test_fortran_dropin_matches_reference_iterate(gpu, name):
    g = G.load(name)
    k = g["meta"]["kernel"]
    exe = _binary(k)
    if not os.path.exists(exe):
        pytest.skip("ref_vs_hip not built (needs the reference sources at build time)")
    print("case", name, flush=True)
    with tempfile.TemporaryDirectory() as td:
        fin, fout = os.path.join(td, "in.bin"), os.path.join(td, "out.bin")
        refrun.write_input(fin, g["state"], g["store_in"])
        p = subprocess.run([exe, fin, fout, str(k)], capture_output=True, text=True, timeout=120)
        assert p.returncode == 0, p.stdout + p.stderr
        worst = float(re.search(r"worst\s+(\S+)", p.stdout).group(1))
        assert worst <= 1e-12, p.stdout
        raw = open(fout, "rb").read()
    n = np.frombuffer(raw[:4], dtype="<i4")[0]
    out = np.frombuffer(raw[4:], dtype="<f8").reshape(n, 32)
    nr = out.shape[0]
    errs = G.field_errors(g["store_out"], out, int(g["state"][24]))
    assert max(errs.values()) <= 1e-12, errs
