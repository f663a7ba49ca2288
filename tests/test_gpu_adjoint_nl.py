"""Adjoints of the operators that read loop 1's grad T: 2nw-+ and 2nw+- with
one conduction tensor K (linear in T and K), and flux-limited diffusion
(FLD: the conductivity follows grad E through the Levermore-Pomraning
limiter, circuit2.f90:177-190, so the operator is not linear in T).  The
geometry (h, rho, Omega) is frozen at sph_diff_prepare.

Checks: (1) the operator equals loop 2's dT/dt of the prepared step, within
the difference between frozen and stepped geometry (loop 1 forms grad T with
the previous Newton sweep's density, circuit1.f90:137-179); (2) <u, A v> =
<A^T u, v> (2nw) and <lam, J v> = <J^T lam, v> by central differences (FLD);
(3) the tensor gradient against <u, A(e_k) v>; (4) the gradient of a loss
after RKL2 super-steps (sts_integrator.f90:87-124) against central finite
differences and an FP64 reference gradient: torch autograd through dense
matrices assembled from the operator (2nw), and through a torch
restatement of the FLD operator held to ours first (FLD).
"""
import math

import numpy as np
import pytest

from sph_playground_amd import consts as C
from sph_playground_amd import problems as P

pytestmark = pytest.mark.gpu

KPERP = 0.1
B = (1.0, 0.4, 0.3)


def _K(b=B, kpar=1.0, kperp=KPERP):
    b = np.asarray(b, dtype=np.float64)
    b = b / np.linalg.norm(b)
    K = kperp * np.eye(3) + (kpar - kperp) * np.outer(b, b)
    return np.array([K[0, 0], K[1, 1], K[2, 2], K[0, 1], K[0, 2], K[1, 2]])


def _fld_problem(n_side=8):
    # 3D periodic lattice, diffusion with the FLD conductivity (manual set),
    # opacity 1, a positive smooth xi = T
    state, st = P.diffusion3d(n_side, ddw="fab", b=B, h_scale=1.03, disotropic="yes")
    state[C.ec_eqs - 1] = C.eeq_manual
    state[C.ec_eqondiff - 1] = 1
    state[C.ec_eqonfluxlim - 1] = 1
    x = st[:, 0:3]
    st[:, C.es_t - 1] = 1.5 + np.sin(2 * np.pi * x[:, 0]) * np.sin(2 * np.pi * x[:, 1]) * np.cos(2 * np.pi * x[:, 2])
    st[:, C.es_kappa - 1] = 1.0
    return state, st


def _setup(kind, n_side=8):
    from sph_playground_amd import autograd as AG
    from sph_playground_amd.engine import Engine
    if kind == "fld":
        state, st = _fld_problem(n_side)
        eng = Engine(state, kernel="m6")
    else:
        state, st = P.diffusion3d(n_side, ddw=kind, b=B, h_scale=1.03)
        eng = Engine(state, kernel="m6", tensor="full", kappa_perp=KPERP)
    eng.load(st)
    AG.prepare(eng)
    assert AG.operator_kind(eng) == ("fld" if kind == "fld" else "2nw")
    return state, st, eng, AG


@pytest.mark.parametrize("kind", ["2nw-+", "2nw+-", "fld"])
def test_operator_equals_loop2(gpu, kind):
    import torch
    state, st, eng, AG = _setup(kind, 10)
    after = eng.save()
    T = torch.from_numpy(st[:, C.es_t - 1].copy()).cuda()
    got = AG.apply(eng, None if kind == "fld" else _K(), T).cpu().numpy()
    ddt = after[:, C.es_ddt - 1]
    err = np.abs(got - ddt).max() / np.abs(ddt).max()
    print(kind, "frozen operator vs loop 2: %.2e" % err, flush=True)
    assert err <= 1e-6


@pytest.mark.parametrize("kind", ["2nw-+", "2nw+-"])
def test_2nw_transpose_and_tensor_gradient(gpu, kind):
    import torch
    state, st, eng, AG = _setup(kind)
    n = st.shape[0]
    rng = np.random.default_rng(3)
    u, v, w = (torch.from_numpy(rng.standard_normal(n)).cuda() for _ in range(3))
    K = _K()
    Av = AG.apply(eng, K, v)
    ATu, gA, gB = AG.apply_adjoint(eng, K, u, v, w)
    lhs, rhs = float(torch.dot(u, Av)), float(torch.dot(ATu, v))
    assert abs(lhs - rhs) <= 1e-12 * float(torch.dot(u.abs(), Av.abs())), (lhs, rhs)
    for k in range(6):
        ek = np.zeros(6)
        ek[k] = 1.0
        ra = float(torch.dot(u, AG.apply(eng, ek, v)))
        rb = float(torch.dot(u, AG.apply(eng, ek, w)))
        sc = float(torch.dot(u.abs(), AG.apply(eng, ek, v).abs())) + 1e-300
        assert abs(gA[k] - ra) <= 1e-12 * sc and abs(gB[k] - rb) <= 1e-12 * sc, (k, gA[k], ra, gB[k], rb)


def test_fld_jacobian_transpose_vs_central_differences(gpu):
    import torch
    state, st, eng, AG = _setup("fld")
    n = st.shape[0]
    rng = np.random.default_rng(4)
    T = torch.from_numpy(st[:, C.es_t - 1].copy()).cuda()
    lam = torch.from_numpy(rng.standard_normal(n)).cuda()
    JT, _, _ = AG.apply_adjoint(eng, None, lam, T, None)
    for seed in (5, 6):
        v = torch.from_numpy(np.random.default_rng(seed).standard_normal(n)).cuda()
        eps = 1e-6
        fd = float(torch.dot(lam, AG.apply(eng, None, T + eps * v) - AG.apply(eng, None, T - eps * v))) / (2 * eps)
        ad = float(torch.dot(JT, v))
        print("FLD <lam, J v>: adjoint %.12e, central differences %.12e" % (ad, fd), flush=True)
        assert abs(ad - fd) <= 1e-6 * abs(fd)


def _loss(AG, eng, T0, params, target, s, dt, nsteps, every, kind, need_grad=True):
    import torch
    T0v = T0.clone().requires_grad_(need_grad)
    if kind == "fld":
        T = AG.sts(T0v, None, eng, s, dt, nsteps, every)
        p = []
    else:
        p = [torch.tensor(x, dtype=torch.float64, requires_grad=need_grad) for x in params]
        T = AG.sts(T0v, AG.conduction_tensor(*p), eng, s, dt, nsteps, every)
    loss = 0.5 * torch.sum((T - target) ** 2)
    if not need_grad:
        return float(loss)
    loss.backward()
    return float(loss), [float(x.grad) for x in p], T0v.grad.clone()


@pytest.mark.parametrize("kind", ["2nw-+", "2nw+-", "fld"])
def test_staged_sts_gradient_vs_central_differences(gpu, kind):
    import torch
    state, st, eng, AG = _setup(kind)
    n = st.shape[0]
    T0 = torch.from_numpy(st[:, C.es_t - 1].copy()).cuda()
    target = torch.from_numpy(np.random.default_rng(5).standard_normal(n) * 0.1 + 1.0).cuda()
    params = [1.0, KPERP, 1.1, 0.4]
    s, nsteps, every = 3, 4, 2
    dt = 1e-15 if kind == "fld" else 2e-5
    _, g, gT0 = _loss(AG, eng, T0, params, target, s, dt, nsteps, every, kind)
    eps = 1e-4
    for k in range(len(g)):
        p1, p2 = list(params), list(params)
        p1[k] += eps
        p2[k] -= eps
        fd = (_loss(AG, eng, T0, p1, target, s, dt, nsteps, every, kind, False)
              - _loss(AG, eng, T0, p2, target, s, dt, nsteps, every, kind, False)) / (2 * eps)
        assert abs(g[k] - fd) <= 1e-6 * max(1.0, abs(fd)), (k, g[k], fd)
    d = torch.from_numpy(np.random.default_rng(6).standard_normal(n)).cuda()
    e = 1e-6
    fd = (_loss(AG, eng, T0 + e * d, params, target, s, dt, nsteps, every, kind, False)
          - _loss(AG, eng, T0 - e * d, params, target, s, dt, nsteps, every, kind, False)) / (2 * e)
    ad = float(torch.dot(gT0, d))
    print(kind, "dL/dT0 . d: adjoint %.10e, central differences %.10e" % (ad, fd), flush=True)
    assert abs(ad - fd) <= 1e-6 * max(1.0, abs(fd))


def _dense_sts(f, T, s, dt, nsteps):
    from sph_playground_amd.engine import sts_coeffs
    for _ in range(nsteps):
        Y0 = T
        DY0 = f(Y0)
        Ym2, Ym1 = Y0, Y0 + sts_coeffs(1, s)[2] * dt * DY0
        for j in range(2, s + 1):
            mu, nu, mh, gh = sts_coeffs(j, s)
            Yj = mu * Ym1 + nu * Ym2 + (1 - mu - nu) * Y0 + mh * dt * f(Ym1) + gh * dt * DY0
            Ym2, Ym1 = Ym1, Yj
        T = Ym1
    return T


@pytest.mark.parametrize("kind", ["2nw-+", "2nw+-"])
def test_2nw_sts_gradient_vs_dense_fp64_reference(gpu, kind):
    import torch
    state, st, eng, AG = _setup(kind, 6)
    n = st.shape[0]
    A = np.zeros((6, n, n))
    for k in range(6):
        ek = np.zeros(6)
        ek[k] = 1.0
        for c in range(n):
            e = torch.zeros(n, dtype=torch.float64, device="cuda")
            e[c] = 1.0
            A[k][:, c] = AG.apply(eng, ek, e).cpu().numpy()
    At = torch.from_numpy(A)
    T0 = torch.from_numpy(st[:, C.es_t - 1].copy())
    target = torch.from_numpy(np.random.default_rng(5).standard_normal(n) * 0.1)
    params = [1.0, KPERP, 1.1, 0.4]
    s, nsteps, every, dt = 4, 5, 2, 2e-5
    kp = [torch.tensor(p, dtype=torch.float64, requires_grad=True) for p in params]
    T0r = T0.clone().requires_grad_(True)
    Am = torch.einsum("k,kij->ij", AG.conduction_tensor(*kp), At)
    loss_ref = 0.5 * torch.sum((_dense_sts(lambda y: Am @ y, T0r, s, dt, nsteps) - target) ** 2)
    loss_ref.backward()
    loss, g, gT0 = _loss(AG, eng, T0.cuda(), params, target.cuda(), s, dt, nsteps, every, kind)
    assert abs(loss - float(loss_ref)) <= 1e-12 * abs(float(loss_ref))
    for k in range(4):
        ref = float(kp[k].grad)
        assert abs(g[k] - ref) <= 1e-9 * max(1.0, abs(ref)), (k, g[k], ref)
    assert np.abs(gT0.cpu().numpy() - T0r.grad.numpy()).max() <= 1e-11 * np.abs(T0r.grad.numpy()).max()


def _fld_torch(eng, state, st):
    """The FLD operator restated in torch (CPU, FP64) on the frozen geometry
    of the engine's prepared step and its neighbour lists: U = grad E (loop-1
    form), D from the limiter, f (fab Hessian trace)."""
    import torch
    out = eng.save()
    off, j, code = eng.neighbours()
    n = st.shape[0]
    i = np.repeat(np.arange(n), np.diff(off))
    L = np.array([state[C.ec_xmax - 1] - state[C.ec_xmin - 1], state[C.ec_ymax - 1] - state[C.ec_ymin - 1],
                  state[C.ec_zmax - 1] - state[C.ec_zmin - 1]])
    sh = np.stack([code % 3 - 1, (code // 3) % 3 - 1, code // 9 - 1], axis=1) * L
    r = out[i, 0:3] - (out[j, 0:3] + sh)
    dr = np.sqrt(np.sum(r * r, axis=1))
    h = out[:, C.es_h - 1]
    rho, om, m, kap = (out[:, getattr(C, f) - 1] for f in ("es_den", "es_om", "es_m", "es_kappa"))
    q = dr / h[i]
    t3, t2, t1 = (np.maximum(c - q, 0.0) for c in (3.0, 2.0, 1.0))
    df = -5 * t3 ** 4 + 30 * t2 ** 4 - 75 * t1 ** 4
    wcv = 1.0 / (120.0 * np.pi)
    nwc = wcv * df / h[i] ** 5 / q                   # nw(r, h_a) / r
    trh = -2.0 * wcv * df / h[i] ** 5 / q           # fab: trace of the Hessian = Gab
    T = lambda a: torch.from_numpy(np.ascontiguousarray(a))  # noqa: E731
    ti, tj = T(i), T(j)
    gvec = T(r * (m[j] / (rho[i] * om[i]) * nwc)[:, None])
    hw = T(m[j] / rho[j] * trh)
    rho_t, kap_t = T(rho), T(kap)
    c = 2.997924e10

    def f(Y):
        Z = Y * rho_t
        dZ = Z[tj] - Z[ti]
        U = torch.zeros(n, 3, dtype=torch.float64).index_add_(0, ti, gvec * dZ[:, None])
        R = torch.sqrt(torch.sum(U * U, dim=1)) / (kap_t * rho_t * Y * rho_t)
        D = c * ((2 + R) / (6 + 3 * R + R * R)) / kap_t / rho_t
        return torch.zeros(n, dtype=torch.float64).index_add_(0, ti, hw * 0.5 * (D[ti] + D[tj]) * dZ) / rho_t
    return f


def test_fld_sts_gradient_vs_torch_restatement(gpu):
    import torch
    state, st, eng, AG = _setup("fld", 6)
    f = _fld_torch(eng, state, st)
    T0 = torch.from_numpy(st[:, C.es_t - 1].copy())
    ours = AG.apply(eng, None, T0.cuda()).cpu()
    assert float((f(T0) - ours).abs().max()) <= 1e-12 * float(ours.abs().max())
    target = torch.from_numpy(np.random.default_rng(5).standard_normal(st.shape[0]) * 0.1 + 1.0)
    s, nsteps, every, dt = 4, 5, 2, 1e-15
    T0r = T0.clone().requires_grad_(True)
    loss_ref = 0.5 * torch.sum((_dense_sts(f, T0r, s, dt, nsteps) - target) ** 2)
    loss_ref.backward()
    loss, _, gT0 = _loss(AG, eng, T0.cuda(), None, target.cuda(), s, dt, nsteps, every, "fld")
    assert abs(loss - float(loss_ref)) <= 1e-12 * abs(float(loss_ref))
    err = np.abs(gT0.cpu().numpy() - T0r.grad.numpy()).max() / np.abs(T0r.grad.numpy()).max()
    print("FLD dL/dT0 vs torch autograd of the restatement: %.2e" % err, flush=True)
    assert err <= 1e-10
