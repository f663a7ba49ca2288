"""ctypes wrapper of the C restatement (oracle/liboracle.so).

Test infrastructure only: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, and only as the checker.
"""
import ctypes as ct
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "liboracle.so")


class Ext(ct.Structure):
    _fields_ = [("kappa_perp", ct.c_double), ("tensor_mode", ct.c_int),
                ("max_sweeps", ct.c_int), ("tol", ct.c_double), ("neigh_cap", ct.c_int)]


class Result(ct.Structure):
    _fields_ = [("n_real", ct.c_int64), ("n_pairs", ct.c_int64),
                ("nb_off", ct.POINTER(ct.c_int64)), ("nb_j", ct.POINTER(ct.c_int32)),
                ("nb_code", ct.POINTER(ct.c_int8)), ("sweeps", ct.c_int),
                ("energy", ct.c_double), ("n_ghost", ct.c_int64)]


_lib = None


def build():
    subprocess.run(["make", "-s", "-C", HERE, "liboracle.so"], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            build()
        L = ct.CDLL(LIB)
        L.orc_iterate.argtypes = [ct.POINTER(ct.c_double), ct.c_int, ct.POINTER(Ext), ct.c_int64,
                                  ct.POINTER(ct.c_double), ct.POINTER(Result)]
        L.orc_iterate.restype = ct.c_int
        L.orc_free_result.argtypes = [ct.POINTER(Result)]
        for name in ("orc_kf", "orc_kdf", "orc_kddf"):
            getattr(L, name).argtypes = [ct.c_int, ct.c_double]
            getattr(L, name).restype = ct.c_double
        L.orc_knorm.argtypes = [ct.c_int, ct.c_int]
        L.orc_knorm.restype = ct.c_double
        for name in ("orc_sts_mu", "orc_sts_nu"):
            getattr(L, name).argtypes = [ct.c_int]
            getattr(L, name).restype = ct.c_double
        for name in ("orc_sts_mu_hat", "orc_sts_gm_hat"):
            getattr(L, name).argtypes = [ct.c_int, ct.c_int]
            getattr(L, name).restype = ct.c_double
        L.orc_rad_exchange.argtypes = [ct.POINTER(ct.c_double), ct.c_int64, ct.POINTER(ct.c_double), ct.c_double]
        L.orc_rad_exchange.restype = ct.c_int
        _lib = L
    return _lib


class OracleError(RuntimeError):
    def __init__(self, code):
        super().__init__("oracle error code %d" % code)
        self.code = code


def iterate(state, store, kernel=1, kappa_perp=0.0, tensor_mode=0, max_sweeps=0, tol=0.0,
            neigh_cap=0, check=True):
    """Run ``iterate`` on a copy of ``store``.  Returns dict(store, off, j, code,
    sweeps, energy, n_ghost, rc)."""
    st = np.ascontiguousarray(np.array(store, dtype=np.float64, copy=True))
    s = np.ascontiguousarray(np.asarray(state, dtype=np.float64))
    ext = Ext(kappa_perp, tensor_mode, max_sweeps, tol, neigh_cap)
    res = Result()
    L = lib()
    rc = L.orc_iterate(s.ctypes.data_as(ct.POINTER(ct.c_double)), kernel, ct.byref(ext),
                       st.shape[0], st.ctypes.data_as(ct.POINTER(ct.c_double)), ct.byref(res))
    out = dict(store=st, rc=rc, sweeps=res.sweeps, energy=res.energy, n_ghost=res.n_ghost)
    if res.nb_off:
        nr = res.n_real
        npairs = res.n_pairs
        out["off"] = np.ctypeslib.as_array(res.nb_off, shape=(nr + 1,)).copy()
        out["j"] = np.ctypeslib.as_array(res.nb_j, shape=(max(npairs, 1),))[:npairs].copy()
        out["code"] = np.ctypeslib.as_array(res.nb_code, shape=(max(npairs, 1),))[:npairs].copy()
    L.orc_free_result(ct.byref(res))
    if check and rc != 0:
        raise OracleError(rc)
    return out


def kf(kernel, q):
    return lib().orc_kf(kernel, q)


def kdf(kernel, q):
    return lib().orc_kdf(kernel, q)


def kddf(kernel, q):
    return lib().orc_kddf(kernel, q)


def knorm(kernel, dim):
    return lib().orc_knorm(kernel, dim)


def sts(name, *args):
    return getattr(lib(), "orc_sts_" + name)(*args)


def rad_exchange(state, store, dt, check=True):
    """update_radenergy on a copy of ``store`` (all rows).  Returns (store, rc)."""
    st = np.ascontiguousarray(np.array(store, dtype=np.float64, copy=True))
    s = np.ascontiguousarray(np.asarray(state, dtype=np.float64))
    rc = lib().orc_rad_exchange(s.ctypes.data_as(ct.POINTER(ct.c_double)), st.shape[0],
                                st.ctypes.data_as(ct.POINTER(ct.c_double)), dt)
    if check and rc != 0:
        raise OracleError(rc)
    return st, rc
