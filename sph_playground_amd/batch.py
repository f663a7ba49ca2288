"""Many independent anisotropic-diffusion problems stepped together.

The reference runs one ``./execute`` per problem of a parameter study
(e.g. 000_anisopaper_v1_gls_kxx.py:44-118 loops over resolutions and field
angles); here each problem is prepared on its own ``Engine`` (neighbours,
Newton h and the EOS of its frozen geometry, ``sph_diff_prepare``) and the
prepared operators are merged into one: every RKL2 stage
(sts_integrator.f90:87-124) of every problem is one kernel launch, a
super-step is one replayed hipGraph (``sph_batch_*`` in libsph_hip.so).
"""
import ctypes as ct

import numpy as np

from ._lib import BatchInfo, check, lib

_D = ct.POINTER(ct.c_double)


_PREC = {"fp64": 0, "mixed": 1}


class Batch:
    """Merged operator of ``engines`` (each already prepared) with the
    per-problem tensors ``K`` ((nprob, 6): xx, yy, zz, xy, xz, yz).
    ``precision``: 'fp64' (each problem bitwise its unbatched super-steps) or
    'mixed' (pair weights stored in FP32; products, sums and states FP64)."""

    def __init__(self, engines, K, precision="fp64"):
        self.nprob = len(engines)
        self.precision = precision
        Kh = np.ascontiguousarray(np.asarray(K, dtype=np.float64).reshape(self.nprob, 6))
        hs = (ct.c_void_p * self.nprob)(*[e._h.value for e in engines])
        h = ct.c_void_p()
        check(lib().sph_batch_create(hs, self.nprob, Kh.ctypes.data_as(_D), _PREC[precision], ct.byref(h)),
              "sph_batch_create")
        self._h = h
        off = np.zeros(self.nprob + 1, dtype=np.int64)
        check(lib().sph_batch_offsets(self._h, off.ctypes.data_as(ct.POINTER(ct.c_int64))), "sph_batch_offsets")
        self.offsets = off
        self.n_total = int(off[-1])

    def info(self):
        i = BatchInfo()
        check(lib().sph_batch_get_info(self._h, ct.byref(i)), "sph_batch_get_info")
        return dict(n_problems=i.n_problems, n_total=i.n_total, n_pairs=i.n_pairs, list_cap=i.list_cap,
                    block=i.block, tune={int(b): float(t) for b, t in zip(i.tune_block, i.tune_ms)})

    def sts(self, T0, s, dt, nsteps, graph=True, out=None):
        """nsteps RKL2 super-steps of s stages; problem p uses dt[p].  T0: CUDA
        float64 tensor of n_total values (problem p at offsets[p]:offsets[p+1],
        original particle order).  Returns the final T (a new tensor or out)."""
        import torch
        if T0.dtype != torch.float64 or not T0.is_cuda or T0.numel() != self.n_total:
            raise ValueError("T0 must be a CUDA float64 tensor of %d values" % self.n_total)
        T0c = T0.contiguous()
        out = torch.empty_like(T0c) if out is None else out
        dth = np.ascontiguousarray(np.broadcast_to(np.asarray(dt, dtype=np.float64), (self.nprob,)))
        torch.cuda.synchronize()
        check(lib().sph_batch_sts(self._h, int(s), dth.ctypes.data_as(_D), int(nsteps), ct.c_void_p(T0c.data_ptr()),
                                  ct.c_void_p(out.data_ptr()), 1 if graph else 0), "sph_batch_sts")
        return out

    def split(self, T):
        """Per-problem views of a batch vector."""
        return [T[self.offsets[p]:self.offsets[p + 1]] for p in range(self.nprob)]

    def close(self):
        if getattr(self, "_h", None):
            lib().sph_batch_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
