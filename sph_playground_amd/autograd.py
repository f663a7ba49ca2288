"""Differentiable diffusion behind ``torch.autograd.Function``.

The forward runs ``nsteps`` RKL2 super-time-steps (sts_integrator.f90:87-124)
of loop 2's diffusion term on the GPU with the geometry (h, rho, Omega) of
the prepared step frozen; the backward is the hand-written adjoint.

  - Hessian operators (fab, fw, n2w) with a global conduction tensor K: the
    fused library path (``sph_diff_forward`` / ``sph_diff_backward``):
    transposed operator over reverse neighbour lists, the tensor gradient
    reduced in FP64, recomputation from checkpoints kept every ``every``
    steps.
  - Operators that read loop 1's grad T: 2nw-+ and 2nw+- with a global K
    (linear in T and K), and flux-limited diffusion (not linear in T: its
    conductivity follows grad E through the Levermore-Pomraning limiter).
    Stage by stage here, each stage one ``sph_diff_apply``; the backward
    recomputes each checkpoint segment's stage states and runs the adjoint
    recurrence with ``sph_diff_apply_adjoint`` (the transposed operator, or
    the transposed Jacobian at the stage state, and the tensor gradient).

PyTorch only carries the tensors and the chain rule from K to (kappa_par,
kappa_perp, theta, phi).
"""
import ctypes as ct

import numpy as np
import torch

from . import consts as C
from ._lib import check, lib


def conduction_tensor(kpar, kperp, theta, phi):
    """K = kappa_perp I + (kappa_par - kappa_perp) b b^T with
    b = (sin t cos p, sin t sin p, cos t), as (xx, yy, zz, xy, xz, yz)."""
    b = torch.stack([torch.sin(theta) * torch.cos(phi), torch.sin(theta) * torch.sin(phi), torch.cos(theta)])
    d = kpar - kperp
    return torch.stack([kperp + d * b[0] * b[0], kperp + d * b[1] * b[1], kperp + d * b[2] * b[2],
                        d * b[0] * b[1], d * b[0] * b[2], d * b[1] * b[2]])


def _vector(engine, t, name):
    """A device vector the C-ABI reads or writes: float64, on the engine's
    GPU, ``engine.n_real`` entries (the kernels index it without bounds)."""
    if not isinstance(t, torch.Tensor):
        raise ValueError("%s must be a torch tensor" % name)
    if t.dtype != torch.float64:
        raise ValueError("%s must be float64, got %s" % (name, t.dtype))
    if not t.is_cuda or t.device.index != engine.device:
        raise ValueError("%s must live on cuda:%d, got %s" % (name, engine.device, t.device))
    if t.numel() != engine.n_real:
        raise ValueError("%s must hold %d values (the real particles), got %d" % (name, engine.n_real, t.numel()))
    return t.detach().contiguous()


def _tensor6(K):
    Kh = np.ascontiguousarray(np.asarray(K.detach().cpu() if isinstance(K, torch.Tensor) else K,
                                         dtype=np.float64)).reshape(-1)
    if Kh.shape != (6,):
        raise ValueError("K must have 6 components (xx, yy, zz, xy, xz, yz)")
    return Kh


def _forward(engine, Kh, s, dt, nsteps, every, T0c, out):
    torch.cuda.synchronize()
    check(lib().sph_diff_forward(engine._h, Kh.ctypes.data_as(ct.POINTER(ct.c_double)), int(s), float(dt),
                                 int(nsteps), int(every), ct.c_void_p(T0c.data_ptr()), ct.c_void_p(out.data_ptr())),
          "sph_diff_forward")
    engine._diff_forward_gen += 1
    return engine._diff_forward_gen


class DiffusionSTS(torch.autograd.Function):
    """T_final = STS^nsteps(T0; K).  Inputs: T0 (n, float64, cuda), K (6,).

    The library keeps the trajectory of the context's most recent
    sph_diff_forward and differentiates that one.  Each forward records the
    engine's forward count and geometry count; a backward whose forward is no
    longer the most recent one (two sts() calls on one engine in one graph)
    runs its forward again from the saved T0 and K first.  A backward after
    the geometry changed (load, step, prepare) raises: its operator is gone."""

    @staticmethod
    def forward(ctx, T0, K, engine, s, dt, nsteps, every):
        T0c = _vector(engine, T0, "T0")
        Kh = _tensor6(K)
        out = torch.empty_like(T0c)
        ctx.gen = _forward(engine, Kh, s, dt, nsteps, every, T0c, out)
        ctx.geom = engine._geometry_gen
        ctx.engine = engine
        ctx.Kdev = K.device
        ctx.args = (Kh, s, dt, nsteps, every)
        ctx.save_for_backward(T0c)
        return out

    @staticmethod
    def backward(ctx, gT):
        engine = ctx.engine
        if engine._geometry_gen != ctx.geom:
            raise RuntimeError("sts backward: the engine's particles changed since this forward "
                               "(load / step / prepare); its operator no longer exists")
        if engine._diff_forward_gen != ctx.gen:
            (T0c,) = ctx.saved_tensors
            Kh, s, dt, nsteps, every = ctx.args
            ctx.gen = _forward(engine, Kh, s, dt, nsteps, every, T0c, torch.empty_like(T0c))
        g = _vector(engine, gT, "grad")
        gT0 = torch.empty_like(g)
        gK = np.zeros(6)
        torch.cuda.synchronize()
        check(lib().sph_diff_backward(engine._h, ct.c_void_p(g.data_ptr()), ct.c_void_p(gT0.data_ptr()),
                                      gK.ctypes.data_as(ct.POINTER(ct.c_double))), "sph_diff_backward")
        return gT0, torch.from_numpy(gK).to(ctx.Kdev), None, None, None, None, None


def operator_kind(engine):
    """'fld' (flux-limited diffusion), '2nw' (2nw-+ / 2nw+-) or 'hessian'
    (fab, fw, n2w) for the engine's equation set and second derivative."""
    st = engine._state
    fld = int(st[C.ec_eqs - 1]) == C.eeq_manual and int(st[C.ec_eqonfluxlim - 1]) == 1
    if fld:
        return "fld"
    if int(st[C.ec_ddw - 1]) in (C.esd_2nw_ds, C.esd_2nw_sd):
        return "2nw"
    return "hessian"


def _coeffs(s):
    from .engine import sts_coeffs
    return [sts_coeffs(j, s) for j in range(1, s + 1)]


def _super_step(engine, Kh, s, dt, Y, keep=None):
    """One RKL2 super-step (sts_integrator.f90:100-123) of dY/dt = f(Y);
    with ``keep`` (a list) the stage states Y_0..Y_s go into it."""
    co = _coeffs(s)
    DY0 = apply(engine, Kh, Y)
    Ym2, Ym1 = Y, Y + co[0][2] * dt * DY0
    if keep is not None:
        keep.extend([Y, Ym1])
    for j in range(2, s + 1):
        mu, nu, mh, gh = co[j - 1]
        Yj = mu * Ym1 + nu * Ym2 + (1 - mu - nu) * Y + mh * dt * apply(engine, Kh, Ym1) + gh * dt * DY0
        Ym2, Ym1 = Ym1, Yj
        if keep is not None:
            keep.append(Yj)
    return Ym1


def _super_step_adjoint(engine, Kh, s, dt, Ys, lam, gK):
    """dL/dY_0 of one super-step from dL/dY_s = lam (the recurrence run
    backwards; Ys its stage states); adds dL/dK into gK (2nw)."""
    co = _coeffs(s)
    L = [None] * (s + 1)
    L[s] = lam
    Ldy0 = torch.zeros_like(lam)

    def add(k, v):
        L[k] = v if L[k] is None else L[k] + v

    for j in range(s, 1, -1):
        mu, nu, mh, gh = co[j - 1]
        Lj = L[j]
        ATl, gA, _ = apply_adjoint(engine, Kh, Lj, Ys[j - 1], None)
        add(j - 1, mu * Lj + mh * dt * ATl)
        add(j - 2, nu * Lj)
        add(0, (1 - mu - nu) * Lj)
        Ldy0 = Ldy0 + gh * dt * Lj
        if gK is not None:
            gK += mh * dt * gA
    add(0, L[1])
    Ldy0 = Ldy0 + co[0][2] * dt * L[1]
    ATl, gA, _ = apply_adjoint(engine, Kh, Ldy0, Ys[0], None)
    if gK is not None:
        gK += gA
    return L[0] + ATl


class DiffusionSTSStaged(torch.autograd.Function):
    """T_final = STS^nsteps(T0; K) for the operators that read grad T (2nw-+,
    2nw+-, flux-limited diffusion), stage by stage.  K: (6,) for 2nw, None
    for FLD (its conductivity comes from the state).  Checkpoints of T every
    ``every`` steps; the backward recomputes one segment's stage states at a
    time."""

    @staticmethod
    def forward(ctx, T0, K, engine, s, dt, nsteps, every):
        Y = _vector(engine, T0, "T0").clone()
        Kh = _tensor6(K) if K is not None else None
        ckpts = []
        with torch.no_grad():
            for st in range(nsteps):
                if st % every == 0:
                    ckpts.append(Y.clone())
                Y = _super_step(engine, Kh, s, dt, Y)
        ctx.engine, ctx.Kh, ctx.args = engine, Kh, (s, dt, nsteps, every)
        ctx.geom = engine._geometry_gen
        ctx.Kdev = K.device if isinstance(K, torch.Tensor) else None
        ctx.ckpts = ckpts
        return Y

    @staticmethod
    def backward(ctx, gT):
        engine, Kh = ctx.engine, ctx.Kh
        s, dt, nsteps, every = ctx.args
        if engine._geometry_gen != ctx.geom:
            raise RuntimeError("sts backward: the engine's particles changed since this forward "
                               "(load / step / prepare); its operator no longer exists")
        lam = _vector(engine, gT, "grad").clone()
        gK = np.zeros(6) if Kh is not None else None
        with torch.no_grad():
            for seg in range(len(ctx.ckpts) - 1, -1, -1):
                first = seg * every
                last = min(first + every, nsteps)
                states, Y = [], ctx.ckpts[seg]
                for _ in range(first, last):
                    keep = []
                    Y = _super_step(engine, Kh, s, dt, Y, keep)
                    states.append(keep)
                for keep in reversed(states):
                    lam = _super_step_adjoint(engine, Kh, s, dt, keep, lam, gK)
        gKt = None
        if gK is not None and ctx.Kdev is not None:
            gKt = torch.from_numpy(gK).to(ctx.Kdev)
        return lam, gKt, None, None, None, None, None


def prepare(engine, state=None):
    """Freeze h/rho/Omega of the resident particles and build reverse lists."""
    s = engine._state if state is None else np.ascontiguousarray(state, dtype=np.float64)
    engine._geometry_gen += 1
    check(lib().sph_diff_prepare(engine._h, s.ctypes.data_as(ct.POINTER(ct.c_double))), "sph_diff_prepare")


def sts(T0, K, engine, s, dt, nsteps, every=5):
    """T after nsteps RKL2 super-steps of s stages from T0 (differentiable in
    T0 and, except for flux-limited diffusion, the tensor K)."""
    if operator_kind(engine) == "hessian":
        return DiffusionSTS.apply(T0, K, engine, s, dt, nsteps, every)
    return DiffusionSTSStaged.apply(T0, K, engine, s, dt, nsteps, every)


def apply(engine, K, T):
    """A(K) T (device tensors, original particle order); for flux-limited
    diffusion f(T) (K None)."""
    T = _vector(engine, T, "T")
    out = torch.empty_like(T)
    Kp = _tensor6(K).ctypes.data_as(ct.POINTER(ct.c_double)) if K is not None else None
    torch.cuda.synchronize()
    check(lib().sph_diff_apply(engine._h, Kp, ct.c_void_p(T.data_ptr()), ct.c_void_p(out.data_ptr())),
          "sph_diff_apply")
    return out


def apply_adjoint(engine, K, lam, Ya, Yb):
    """(A(K)^T lam, lam^T dA/dK Ya, lam^T dA/dK Yb); Yb may be None (gB 0).
    For flux-limited diffusion (K None): (J(Ya)^T lam, 0, 0), J the Jacobian
    of f at Ya."""
    lam = _vector(engine, lam, "lam")
    Ya = _vector(engine, Ya, "Ya")
    Yb = _vector(engine, Yb, "Yb") if Yb is not None else None
    out = torch.empty_like(lam)
    gA = np.zeros(6)
    gB = np.zeros(6)
    torch.cuda.synchronize()
    dp = ct.POINTER(ct.c_double)
    Kp = _tensor6(K).ctypes.data_as(dp) if K is not None else None
    Yb_ = Yb if Yb is not None else (Ya if K is not None else None)
    check(lib().sph_diff_apply_adjoint(engine._h, Kp, ct.c_void_p(lam.data_ptr()), ct.c_void_p(Ya.data_ptr()),
                                       ct.c_void_p(Yb_.data_ptr()) if Yb_ is not None else None,
                                       ct.c_void_p(out.data_ptr()), gA.ctypes.data_as(dp), gB.ctypes.data_as(dp)),
          "sph_diff_apply_adjoint")
    return out, gA, (gB if Yb is not None else np.zeros(6))
