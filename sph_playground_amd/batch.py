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
import math
import time
from concurrent.futures import ThreadPoolExecutor

import numpy as np

from ._lib import BatchInfo, check, lib

_D = ct.POINTER(ct.c_double)


_PREC = {"fp64": 0, "mixed": 1}

# problem-size buckets of the cell-size autotuner: N in [1024 4^k, 1024 4^(k+1))
BUCKET_BASE = 1024
CELL_DIVS = (1, 2, 3, 4)


def bucket_of(n):
    """Size bucket of an N-particle problem: floor(log4(N / 1024)), >= 0."""
    return max(0, int(math.floor(math.log(max(n, 1) / BUCKET_BASE, 4) + 1e-12)))


def tune_cells(problems, reps=3, cell_divs=CELL_DIVS, device=0):
    """The autotuner of the search cells: for every size bucket, the bucket's
    largest problem is stepped with search cells r_max / d wide for each d
    in ``cell_divs`` (device time of the whole step: search, Newton sweeps,
    loop 2, since the cell size also orders the particles the loops gather)
    and the fastest d is kept.  ``problems``: (state, store, kappa_perp)
    triples.  Returns ({bucket: d}, {bucket: {d: ms}})."""
    from .engine import Engine
    rep = {}
    for k, (state, st, kperp) in enumerate(problems):
        b = bucket_of(st.shape[0])
        if b not in rep or st.shape[0] > problems[rep[b]][1].shape[0]:
            rep[b] = k
    best, table = {}, {}
    for b, k in sorted(rep.items()):
        state, st, kperp = problems[k]
        table[b] = {}
        for d in cell_divs:
            eng = Engine(state, kernel="m6", tensor="full", kappa_perp=kperp, cell_div=d, device=device)
            eng.load(st)
            eng.set_timing(True)
            eng.step()
            ms = []
            for _ in range(reps):
                eng.step()
                ms.append(eng.stats()["ms_total"])
            eng.close()
            table[b][d] = float(np.median(ms))
        best[b] = min(table[b], key=table[b].get)
    return best, table


def prepare_problems(problems, cells=None, threads=8, device=0):
    """Engines for ``problems`` ((state, store, kappa_perp) triples), each
    loaded and prepared (sph_diff_prepare: search, Newton h, EOS, reverse
    lists), with search cells from ``cells`` ({bucket: d}, see tune_cells;
    missing buckets use the default).  Contexts are independent, so
    ``threads`` host threads prepare them concurrently, each on its own
    stream: the small problems' launches and host round trips overlap instead
    of queueing one problem after another.  Returns (engines, seconds)."""
    from . import autograd as AG
    from .engine import Engine
    cells = cells or {}

    def one(pr):
        state, st, kperp = pr
        eng = Engine(state, kernel="m6", tensor="full", kappa_perp=kperp,
                     cell_div=cells.get(bucket_of(st.shape[0]), 0), device=device)
        eng.load(st)
        AG.prepare(eng)
        return eng

    t0 = time.perf_counter()
    with ThreadPoolExecutor(max_workers=max(1, threads)) as ex:
        engines = list(ex.map(one, problems))
    return engines, time.perf_counter() - t0


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
