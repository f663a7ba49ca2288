"""Batched problems: many prepared diffusion problems merged into one
operator, one launch per RKL2 stage (sts_integrator.f90:87-124).

Checks: (1) every problem's result is bitwise the unbatched sph_diff_forward
result, for direct launches and for hipGraph replays (several replays, with
other GPU work and a changed time step between them, then a new stage count);
(2) against the oracle: the RKL2 recurrence with the oracle's STS
coefficients, each stage's rate being the oracle's loop-2 dT/dt of the
problem's geometry, FP64 within 1e-12 relative per problem.
"""
import ctypes as ct
import math

import numpy as np
import pytest

from oracle import oracle
from sph_playground_amd import consts as C
from sph_playground_amd import problems as P

pytestmark = pytest.mark.gpu

TOL = 1e-12
# sizes of different magnitude and one 3D problem among the 2D ones
SPECS = [dict(n_side=8, theta=0.3, ratio=1.0), dict(n_side=17, theta=2.1, ratio=0.01),
         dict(n_side=12, theta=1.2, ratio=0.1), dict(n_side=23, theta=0.7, ratio=0.001)]


def _problems():
    from sph_playground_amd import autograd as AG
    from sph_playground_amd.engine import Engine
    out = []
    for sp in SPECS:
        state, st, K = P.batch2d_problem(sp)
        out.append((state, st, K, sp["ratio"]))
    b = (1.0, 0.4, 0.3)
    state, st = P.diffusion3d(6, ddw="fab", b=b, h_scale=1.03)
    bn = np.asarray(b) / np.linalg.norm(b)
    Km = 0.05 * np.eye(3) + 0.95 * np.outer(bn, bn)
    out.append((state, st, np.array([Km[0, 0], Km[1, 1], Km[2, 2], Km[0, 1], Km[0, 2], Km[1, 2]]), 0.05))
    engines = []
    for state, st, K, r in out:
        eng = Engine(state, kernel="m6", tensor="full", kappa_perp=r)
        eng.load(st)
        AG.prepare(eng)
        engines.append(eng)
    return out, engines


def _dt(st, s):
    dx = float(st[0, C.es_h - 1])
    return 0.05 * dx * dx * (s * s + s - 2) / 4.0


def _forward(eng, K, s, dt, nsteps, T0):
    import torch
    from sph_playground_amd._lib import check, lib
    out = torch.empty_like(T0)
    Kh = np.ascontiguousarray(K, dtype=np.float64)
    torch.cuda.synchronize()
    check(lib().sph_diff_forward(eng._h, Kh.ctypes.data_as(ct.POINTER(ct.c_double)), s, float(dt), nsteps, nsteps,
                                 ct.c_void_p(T0.data_ptr()), ct.c_void_p(out.data_ptr())), "sph_diff_forward")
    return out


def test_batch_is_bitwise_each_problems_forward(gpu):
    import torch
    from sph_playground_amd.batch import Batch
    probs, engines = _problems()
    batch = Batch(engines, np.stack([K for _, _, K, _ in probs]))
    info = batch.info()
    print("batch", info, flush=True)
    assert info["n_total"] == sum(st.shape[0] for _, st, _, _ in probs)
    assert info["block"] in info["tune"]
    rng = np.random.default_rng(7)
    T0 = torch.from_numpy(np.concatenate([st[:, C.es_t - 1] for _, st, _, _ in probs])).cuda()
    for s, nsteps, scale in ((5, 3, 1.0), (5, 2, 0.5), (8, 2, 1.0)):
        dts = [scale * _dt(st, s) for _, st, _, _ in probs]
        T0 = T0 + torch.from_numpy(1e-3 * rng.standard_normal(T0.numel())).cuda()
        eager = batch.sts(T0, s, dts, nsteps, graph=False)
        graphs = [batch.sts(T0, s, dts, nsteps, graph=True)]
        _ = (T0 * 2.0).sum().item()  # other work between replays
        graphs.append(batch.sts(T0, s, dts, nsteps, graph=True))
        for g in graphs:
            assert torch.equal(g, eager)
        for p, ((_, st, K, _), eng) in enumerate(zip(probs, engines)):
            ref = _forward(eng, K, s, dts[p], nsteps, batch.split(T0)[p].contiguous())
            assert torch.equal(batch.split(eager)[p], ref), (s, p)
    batch.close()


def test_batch_matches_oracle_recurrence(gpu):
    import torch
    from sph_playground_amd.batch import Batch
    probs, engines = _problems()
    batch = Batch(engines, np.stack([K for _, _, K, _ in probs]))
    s, nsteps = 4, 2
    dts = [_dt(st, s) for _, st, _, _ in probs]
    T0 = np.concatenate([st[:, C.es_t - 1] for _, st, _, _ in probs])
    got = batch.sts(torch.from_numpy(T0).cuda(), s, dts, nsteps).cpu().numpy()
    for p, (state, st, K, r) in enumerate(probs):
        def rate(T):
            x = st.copy()
            x[:, C.es_t - 1] = T
            return oracle.iterate(state, x, kernel=1, tensor_mode=1, kappa_perp=r)["store"][:, C.es_ddt - 1]
        dt = dts[p]
        T = st[:, C.es_t - 1].copy()
        for _ in range(nsteps):
            Y0 = T
            DY0 = rate(Y0)
            Ym2, Ym1 = Y0, Y0 + oracle.sts("mu_hat", 1, s) * dt * DY0
            for j in range(2, s + 1):
                mu, nu = oracle.sts("mu", j), oracle.sts("nu", j)
                mh, gh = oracle.sts("mu_hat", j, s), oracle.sts("gm_hat", j, s)
                Yj = mu * Ym1 + nu * Ym2 + (1 - mu - nu) * Y0 + mh * dt * rate(Ym1) + gh * dt * DY0
                Ym2, Ym1 = Ym1, Yj
            T = Ym1
        mine = got[batch.offsets[p]:batch.offsets[p + 1]]
        assert np.abs(mine - T).max() <= TOL * np.abs(T).max(), p
        # the field decays: the super-steps did change it
        assert np.abs(T - st[:, C.es_t - 1]).max() > 1e-6 * np.abs(T).max()
    batch.close()


def test_batch_rejects_unprepared_and_bad_arguments(gpu):
    import torch
    from sph_playground_amd._lib import SphError
    from sph_playground_amd.batch import Batch
    from sph_playground_amd.engine import Engine
    state, st, K = P.batch2d_problem(SPECS[0])
    eng = Engine(state, kernel="m6")
    eng.load(st)
    with pytest.raises(SphError) as e:
        Batch([eng], K[None, :])
    assert e.value.code == 7  # BAD_ARG: not prepared
    probs, engines = _problems()
    batch = Batch(engines[:1], probs[0][2][None, :])
    with pytest.raises(SphError):
        batch.sts(torch.zeros(batch.n_total, dtype=torch.float64, device="cuda"), 1, [1e-5], 1)  # s < 2
    batch.close()


def test_batch_fp32_weights_close_to_fp64(gpu):
    # FP32-stored pair weights (6e-8 relative each), FP64 products, sums and
    # states: the change of T over the super-steps agrees with the FP64
    # batch to 1e-5 of its size (the sum over a smooth field cancels to
    # about h k of its terms, k the field's wavenumber); graph replays equal
    # direct launches bit for bit
    import torch
    from sph_playground_amd.batch import Batch
    probs, engines = _problems()
    K = np.stack([K for _, _, K, _ in probs])
    b64 = Batch(engines, K)
    b32 = Batch(engines, K, precision="mixed")
    s, nsteps = 6, 3
    dts = [_dt(st, s) for _, st, _, _ in probs]
    T0 = torch.from_numpy(np.concatenate([st[:, C.es_t - 1] for _, st, _, _ in probs])).cuda()
    ref = b64.sts(T0, s, dts, nsteps)
    got = b32.sts(T0, s, dts, nsteps, graph=False)
    assert torch.equal(b32.sts(T0, s, dts, nsteps, graph=True), got)
    for p in range(len(probs)):
        r, g, t0 = (x[p].cpu().numpy() for x in (b64.split(ref), b32.split(got), b64.split(T0)))
        change = np.abs(r - t0).max()
        assert change > 0
        assert np.abs(g - r).max() <= 1e-5 * change, (p, np.abs(g - r).max() / change)
    b64.close()
    b32.close()
