"""Host-side mirror of the reference's step-kernel interface, backed by HIP.

The reference exposes ``iterate(n, store, maxconsenrg)`` (iterator.f90:44)
with its configuration in module state (``statevars``, state.f90:36) and the
kernel picked at compile time (Makefile:14).  ``Engine`` is that module state
made explicit: it is created from a ``statevars`` vector and a kernel name,
freezes what the reference freezes at first call, and its ``iterate`` updates a
``store`` in place exactly as the reference does.  All arithmetic runs in
libsph_hip.so on the GPU; there is no CPU path.
"""
import ctypes as ct

import numpy as np

from . import consts as C
from ._lib import Options, Stats, check, lib

_PREC = {"fp64": 0, "mixed": 1}
_TENSOR = {"reference": 0, "full": 1}


def _dptr(a):
    return a.ctypes.data_as(ct.POINTER(ct.c_double))


class Engine:
    """The two-loop hot path for one problem set-up.

    Parameters mirror the reference's knobs: ``state`` (49 statevars),
    ``kernel`` ('m4' | 'm6' | 'gauss'), plus the device-path options:
    ``precision`` ('fp64' throughout | 'mixed'), ``tensor``
    ('reference' diffusion tensor | 'full' kappa_par bb + kappa_perp (I-bb)),
    ``kappa_perp``, ``max_sweeps`` (100), ``newton_tol`` (1e-8 fp64, 1e-6
    mixed), ``neigh_cap`` (reference maxneibnum), ``cell_div`` (search cells
    r_max / cell_div wide, 0 = 3; the lists do not depend on it).
    """

    def __init__(self, state, kernel="m6", precision="fp64", tensor="reference", kappa_perp=0.0,
                 max_sweeps=100, newton_tol=0.0, neigh_cap=0, device=0, cell_div=0):
        L = lib()
        self._state = np.ascontiguousarray(np.asarray(state, dtype=np.float64))
        if self._state.shape != (C.ec_total,):
            raise ValueError("state must have %d entries" % C.ec_total)
        kid = C.KERNELS[kernel] if isinstance(kernel, str) else int(kernel)
        opt = Options(_PREC[precision], _TENSOR[tensor], float(kappa_perp), int(max_sweeps),
                      float(newton_tol), int(neigh_cap), int(device), int(cell_div))
        h = ct.c_void_p()
        # the state's pointer, made once (the per-step calls pass it again)
        self._state_p = _dptr(self._state)
        check(L.sph_ctx_create(self._state_p, kid, ct.byref(opt), ct.byref(h)), "sph_ctx_create")
        self._h = h
        self.kernel = kid
        self.precision = precision
        self.device = int(device)
        self.n_real = int(self._state[C.ec_realpn - 1])
        self.n = 0
        # counts read by the autograd wrapper: calls that replace the resident
        # particles or their geometry, and recorded sph_diff_forward runs
        self._geometry_gen = 0
        self._diff_forward_gen = 0

    # -- the reference's entry point ------------------------------------
    def iterate(self, store, state=None):
        """``call iterate(n, store, maxconsenrg)`` on a host (n, 32) FP64 store
        (updated in place).  Returns maxconsenrg (always 0, circuit2.f90:101)."""
        st = self._check_store(store)
        s = self._state if state is None else np.ascontiguousarray(np.asarray(state, dtype=np.float64))
        e = ct.c_double(0.0)
        self._geometry_gen += 1
        check(lib().sph_iterate(self._h, st.shape[0], _dptr(st), st.shape[1], _dptr(s), ct.byref(e)),
              "sph_iterate")
        self.n = st.shape[0]
        if st is not store:
            store[...] = st
        return e.value

    # -- resident use ---------------------------------------------------
    def load(self, store):
        st = self._check_store(store)
        self._geometry_gen += 1
        check(lib().sph_load(self._h, st.shape[0], st.ctypes.data, st.shape[1], 0), "sph_load")
        self.n = st.shape[0]

    def load_device(self, ptr, n, ld=C.es_total):
        self._geometry_gen += 1
        check(lib().sph_load(self._h, n, ct.c_void_p(ptr), ld, 1), "sph_load")
        self.n = n

    def save(self, out=None):
        out = np.empty((self.n, C.es_total)) if out is None else out
        check(lib().sph_save(self._h, out.ctypes.data, out.shape[1], 0), "sph_save")
        return out

    def save_device(self, ptr, ld=C.es_total):
        check(lib().sph_save(self._h, ct.c_void_p(ptr), ld, 1), "sph_save")

    def step(self, state=None):
        sp = self._state_p if state is None else _dptr(np.ascontiguousarray(np.asarray(state, dtype=np.float64)))
        self._geometry_gen += 1
        check(lib().sph_step(self._h, sp, None), "sph_step")

    def leapfrog(self, dtcap, state=None):
        """One pass of the leapfrog loop (main.f90:241-362) on the resident
        state; returns dt.  The initial iterate (main.f90:219) runs first when
        the state was loaded since the last step, so N calls after ``load``
        are the reference's run of N steps."""
        sp = self._state_p if state is None else _dptr(np.ascontiguousarray(np.asarray(state, dtype=np.float64)))
        dt = ct.c_double(0.0)
        self._geometry_gen += 1
        check(lib().sph_leapfrog(self._h, sp, float(dtcap), ct.byref(dt)), "sph_leapfrog")
        return dt.value

    PRERUN_MESSAGES = {1: "density <= 0", 2: "csound <= 0", 4: "smoothing length <= 0", 8: "pressure <= 0",
                       16: "internal energy <= 0", 32: "mhd_magneticconstant <= 0",
                       64: "diff_conductivity <= 0"}

    def prerun_check(self, state=None):
        """checkVarsReady (preruncheck.f90:13-60) on the resident store: raises
        SphError (PRECONDITION) naming the reference's message."""
        s = self._state if state is None else np.ascontiguousarray(np.asarray(state, dtype=np.float64))
        what = ct.c_int(0)
        rc = lib().sph_prerun_check(self._h, _dptr(s), ct.byref(what))
        check(rc, "sph_prerun_check: %s" % self.PRERUN_MESSAGES.get(what.value, "") if rc else "")

    def sts(self, s_stages, dt, state=None):
        """One RKL2 super-step of T (sts_integrator.f90:87-124) on the resident state."""
        s = self._state if state is None else np.ascontiguousarray(np.asarray(state, dtype=np.float64))
        self._geometry_gen += 1
        check(lib().sph_sts(self._h, _dptr(s), int(s_stages), float(dt)), "sph_sts")

    def rad_exchange(self, dt, state=None):
        """Implicit gas/radiation energy exchange over dt on the resident state
        (update_radenergy, rad_exchange.f90:9-118)."""
        s = self._state if state is None else np.ascontiguousarray(np.asarray(state, dtype=np.float64))
        check(lib().sph_rad_exchange(self._h, _dptr(s), float(dt)), "sph_rad_exchange")

    def neighbours(self):
        """Neighbour lists of the last step: (off, j, code) in original numbering."""
        nr = self.n_real
        off = np.zeros(nr + 1, dtype=np.int64)
        npairs = ct.c_int64(0)
        L = lib()
        check(L.sph_neighbours(self._h, off.ctypes.data_as(ct.POINTER(ct.c_int64)), None, None, 0,
                               ct.byref(npairs)), "sph_neighbours")
        j = np.zeros(max(npairs.value, 1), dtype=np.int32)
        code = np.zeros(max(npairs.value, 1), dtype=np.int8)
        check(L.sph_neighbours(self._h, off.ctypes.data_as(ct.POINTER(ct.c_int64)),
                               j.ctypes.data_as(ct.POINTER(ct.c_int32)), code.ctypes.data_as(ct.POINTER(ct.c_int8)),
                               j.shape[0], ct.byref(npairs)), "sph_neighbours")
        return off, j[:npairs.value], code[:npairs.value]

    def set_timing(self, on=True):
        """Record device times of the phases and kernels of later steps (the
        stats' ms_* fields, 0 while off, the default); the events cost each
        step a few microseconds per kernel boundary."""
        check(lib().sph_set_timing(self._h, 1 if on else 0), "sph_set_timing")

    def stats(self):
        s = Stats()
        check(lib().sph_get_stats(self._h, ct.byref(s)), "sph_get_stats")
        return {k: getattr(s, k) for k, _ in Stats._fields_}

    def close(self):
        if getattr(self, "_h", None):
            lib().sph_ctx_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _check_store(self, store):
        if not isinstance(store, np.ndarray) or store.ndim != 2 or store.shape[1] < C.es_total:
            raise ValueError("store must be an (n, >=32) array")
        st = store if (store.dtype == np.float64 and store.flags.c_contiguous) else \
            np.ascontiguousarray(store, dtype=np.float64)
        if st.shape[0] < self.n_real:
            raise ValueError("store holds fewer particles than ec_realpn")
        return st


def sts_coeffs(j, s):
    """(mu_j, nu_j, mu_hat_j, gm_hat_j) of the RKL2 scheme, FP64 (sts_integrator.f90:10-51)."""
    out = np.zeros(4)
    check(lib().sph_sts_coeffs(int(j), int(s), _dptr(out)), "sph_sts_coeffs")
    return tuple(out)


def iterate(n, store, state, kernel="m6", **kw):
    """One-shot ``iterate(n, store, maxconsenrg)``: builds an Engine, runs it on
    ``store[:n]`` in place and returns maxconsenrg (0)."""
    eng = Engine(state, kernel=kernel, **kw)
    try:
        view = store[:n]
        return eng.iterate(view)
    finally:
        eng.close()
