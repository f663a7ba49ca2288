"""State dump and restore in the reference's HDF5 wire format
(include/sph_dump.h; dumper.f90).

The reference's dumper needs the HDF5 Fortran module of the compiler it was
built with, which this image lacks (its modules are from gfortran, which is
not installed), so it is not run here.  The files are checked against what
dumper.f90 creates (PutData1D/PutData2D, :86-148) with the HDF5 tools' own
reader (h5dump) where it is installed: dataset names, FP64 type, dims, max
dims and chunking, and every value; and against the fortdriver's access
pattern store[i, es - 1] (core.py:322-324,453).
"""
import os
import re
import shutil
import subprocess

import numpy as np
import pytest

from sph_playground_amd import consts as C

dump_mod = pytest.importorskip("sph_playground_amd.dump")
if not os.path.exists(dump_mod.LIB_PATH):
    pytest.skip("libsph_dump.so not built (no HDF5 library)", allow_module_level=True)

H5DUMP = shutil.which("h5dump") or ("/opt/conda/bin/h5dump" if os.path.exists("/opt/conda/bin/h5dump") else None)


def test_library_exports_every_declared_symbol():
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    txt = open(os.path.join(root, "include", "sph_dump.h")).read()
    names = sorted(set(re.findall(r"^\s*int\s*(sph_\w+)\s*\(", txt, re.M)))
    assert names == ["sph_dump_clean", "sph_dump_read", "sph_dump_rows", "sph_dump_write"]
    L = dump_mod.lib()
    for name in names:
        assert hasattr(L, name), name


def _case(n_real=37, n_fixed=5, extra=3, seed=4):
    rng = np.random.default_rng(seed)
    state = rng.normal(size=C.ec_total)
    state[C.ec_realpn - 1] = n_real
    state[C.ec_fixedpn - 1] = n_fixed
    state[C.ec_lastprint - 1] = 7
    store = rng.normal(size=(n_real + n_fixed + extra, C.es_total))
    return state, store


def test_dump_restore_round_trip(tmp_path):
    state, store = _case()
    dump_mod.dump(str(tmp_path), state, store, t=0.375)
    names = sorted(os.listdir(tmp_path))
    # new_dump_full.h5 renamed to dump_full.h5, copied to dump_00007.h5
    assert names == ["dump_00007.h5", "dump_full.h5"]
    assert open(tmp_path / "dump_full.h5", "rb").read() == open(tmp_path / "dump_00007.h5", "rb").read()
    st, s = dump_mod.restore(str(tmp_path))
    want = state.copy()
    want[C.ec_time - 1] = 0.375  # setStateVal(ec_time, t)
    assert np.array_equal(st, want)
    n = 37 + 5  # realpn + fixedpn rows; the rows past them are not written
    assert s.shape == (n, C.es_total) and np.array_equal(s, store[:n])
    assert dump_mod.exists(str(tmp_path))


def test_dump_numbering_and_clean(tmp_path):
    state, store = _case(n_real=3, n_fixed=0, extra=0)
    for lp, name in [(0, "dump_00000.h5"), (123.9, "dump_00123.h5"), (99999, "dump_99999.h5"),
                     (100000, "dump_*****.h5"), (-1, "dump_*****.h5")]:
        state[C.ec_lastprint - 1] = lp
        dump_mod.dump(str(tmp_path), state, store, t=1.0)
        assert (tmp_path / name).exists(), name
    (tmp_path / "new_dump_full.h5").write_bytes(b"partial")
    dump_mod.clean(str(tmp_path))
    assert not (tmp_path / "new_dump_full.h5").exists()
    dump_mod.clean(str(tmp_path))  # nothing to clean


def test_dump_errors(tmp_path):
    state, store = _case(n_real=0, n_fixed=0)
    with pytest.raises(dump_mod.DumpError):
        dump_mod.dump(str(tmp_path), state, store, t=0.0)
    with pytest.raises(dump_mod.DumpError):
        dump_mod.restore(str(tmp_path / "missing"))
    state, store = _case()
    with pytest.raises(ValueError):
        dump_mod.dump(str(tmp_path), state, store[:10], t=0.0)


@pytest.mark.skipif(H5DUMP is None, reason="h5dump not installed")
def test_file_layout_read_by_hdf5_tools(tmp_path):
    n_real, n_fixed = 6, 2
    state, store = _case(n_real=n_real, n_fixed=n_fixed, extra=1)
    dump_mod.dump(str(tmp_path), state, store, t=2.5)
    f = str(tmp_path / "dump_full.h5")
    head = subprocess.run([H5DUMP, "-p", "-H", f], capture_output=True, text=True, check=True).stdout
    blocks = dict(re.findall(r'DATASET "(\w+)" \{(.*?)\n   \}', head, re.S))
    assert set(blocks) == {"state", "store"}
    n = n_real + n_fixed
    for name, dims, maxd, chunk in [("state", "( 49 )", "( H5S_UNLIMITED )", "( 49 )"),
                                    ("store", "( %d, 32 )" % n, "( H5S_UNLIMITED, H5S_UNLIMITED )",
                                     "( %d, 32 )" % n)]:
        b = blocks[name]
        assert "H5T_IEEE_F64LE" in b, b
        assert "SIMPLE { %s / %s }" % (dims, maxd) in b, b
        assert "CHUNKED %s" % chunk in b, b
    # every value, through the tools' reader (17 significant digits)
    for name, want in [("state", None), ("store", store[:n])]:
        out = subprocess.run([H5DUMP, "-m", "%.17g", "-d", name, "-y", "-w", "0", f],
                             capture_output=True, text=True, check=True).stdout
        body = out.split("DATA {", 1)[1].split("}", 1)[0]
        vals = np.array([float(x) for x in re.split(r"[,\s]+", body.strip()) if x])
        if name == "state":
            want = state.copy()
            want[C.ec_time - 1] = 2.5
        assert np.array_equal(vals, want.ravel()), name
    # the fortdriver's view: store[i, es - 1] is field es of particle i
    st, s = dump_mod.restore(str(tmp_path))
    assert s[3, C.es_den - 1] == store[3, C.es_den - 1]
