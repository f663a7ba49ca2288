"""The batched convergence study at its workload: 256 independent 2D
anisotropic-diffusion problems, N log-uniform in [32^2, 512^2] on periodic
square lattices, field angle U[0, pi), kappa_perp / kappa_par in
{1, 0.1, 0.01, 0.001}, sine T, M6, fab (problems.batch2d_specs(256)).

The search cells of each size bucket come from the autotuner
(batch.tune_cells), the problems are prepared concurrently
(batch.prepare_problems), then merged; RKL2 super-steps
(sts_integrator.f90:87-124) run one launch per stage, replayed as a hipGraph.
Checks: every problem's result is bitwise its own sph_diff_forward result
(graph replays twice); the smallest, median and largest problems follow the
RKL2 recurrence with the oracle's coefficients and the oracle's loop-2
dT/dt as each stage's rate, FP64 within 1e-12, and their neighbour sets are
the oracle's.
"""
import ctypes as ct

import numpy as np
import pytest

from oracle import oracle
from sph_playground_amd import consts as C
from sph_playground_amd import problems as P
from tests import goldens as G

pytestmark = pytest.mark.gpu

NPROB = 256
S, NSTEPS = 5, 2
TOL = 1e-12


def _dt(st, s):
    dx = float(st[0, C.es_h - 1])
    return 0.05 * dx * dx * (s * s + s - 2) / 4.0


def test_batch_study_256_problems(gpu):
    import torch
    from sph_playground_amd import batch as B
    from sph_playground_amd._lib import check, lib
    specs = P.batch2d_specs(NPROB)
    built = [P.batch2d_problem(sp) for sp in specs]
    triples = [(state, st, sp["ratio"]) for (state, st, _), sp in zip(built, specs)]
    Ks = np.stack([K for _, _, K in built])
    sizes = np.array([st.shape[0] for _, st, _ in built])
    best, table = B.tune_cells(triples)
    print("cell autotuner", best, table, flush=True)
    assert set(best) == {B.bucket_of(n) for n in sizes}
    engines, secs = B.prepare_problems(triples, best)
    print("prepared %d problems (%d particles) in %.2f s" % (NPROB, sizes.sum(), secs), flush=True)
    batch = B.Batch(engines, Ks)
    dts = [_dt(st, S) for _, st, _ in built]
    T0 = torch.from_numpy(np.concatenate([st[:, C.es_t - 1] for _, st, _ in built])).cuda()
    out = batch.sts(T0, S, dts, NSTEPS, graph=True)
    _ = (T0 * 2.0).sum().item()   # other GPU work between replays
    assert torch.equal(batch.sts(T0, S, dts, NSTEPS, graph=True), out)
    parts = batch.split(out)
    for p, eng in enumerate(engines):
        ref = torch.empty_like(parts[p])
        Kh = np.ascontiguousarray(Ks[p])
        x0 = batch.split(T0)[p].contiguous()
        torch.cuda.synchronize()
        check(lib().sph_diff_forward(eng._h, Kh.ctypes.data_as(ct.POINTER(ct.c_double)), S, float(dts[p]), NSTEPS,
                                     NSTEPS, ct.c_void_p(x0.data_ptr()), ct.c_void_p(ref.data_ptr())),
              "sph_diff_forward")
        assert torch.equal(parts[p], ref), p
    order = np.argsort(sizes, kind="stable")
    got = out.cpu().numpy()
    for p in (int(order[0]), int(order[NPROB // 2]), int(order[-1])):
        state, st, K = built[p]
        r = specs[p]["ratio"]
        print("oracle recurrence, problem", p, "N", st.shape[0], "cells r/%d" % best[B.bucket_of(st.shape[0])],
              flush=True)
        first = oracle.iterate(state, st, kernel=1, tensor_mode=1, kappa_perp=r)
        off, j, code = engines[p].neighbours()
        assert G.pair_sets(off, j, code) == G.pair_sets(first["off"], first["j"], first["code"])

        def rate(T):
            x = st.copy()
            x[:, C.es_t - 1] = T
            return oracle.iterate(state, x, kernel=1, tensor_mode=1, kappa_perp=r)["store"][:, C.es_ddt - 1]
        dt = dts[p]
        T = st[:, C.es_t - 1].copy()
        for _ in range(NSTEPS):
            Y0 = T
            DY0 = rate(Y0)
            Ym2, Ym1 = Y0, Y0 + oracle.sts("mu_hat", 1, S) * dt * DY0
            for jj in range(2, S + 1):
                mu, nu = oracle.sts("mu", jj), oracle.sts("nu", jj)
                mh, gh = oracle.sts("mu_hat", jj, S), oracle.sts("gm_hat", jj, S)
                Yj = mu * Ym1 + nu * Ym2 + (1 - mu - nu) * Y0 + mh * dt * rate(Ym1) + gh * dt * DY0
                Ym2, Ym1 = Ym1, Yj
            T = Ym1
        mine = got[batch.offsets[p]:batch.offsets[p + 1]]
        assert np.abs(mine - T).max() <= TOL * np.abs(T).max(), p
        assert np.abs(T - st[:, C.es_t - 1]).max() > 1e-6 * np.abs(T).max()
    batch.close()
    for e in engines:
        e.close()
