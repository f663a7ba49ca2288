"""Differentiable anisotropic diffusion behind ``torch.autograd.Function``.

The forward runs ``nsteps`` RKL2 super-time-steps (sts_integrator.f90:87-124)
of the loop-2 diffusion operator with a global conduction tensor on the GPU
(``sph_diff_forward``); the backward is the hand-written adjoint
(``sph_diff_backward``): transposed operator over reverse neighbour lists,
the tensor gradient reduced in FP64, and recomputation from checkpoints kept
every ``every`` steps.  PyTorch only carries the tensors and the chain rule
from K to (kappa_par, kappa_perp, theta, phi).
"""
import ctypes as ct

import numpy as np
import torch

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


def prepare(engine, state=None):
    """Freeze h/rho/Omega of the resident particles and build reverse lists."""
    s = engine._state if state is None else np.ascontiguousarray(state, dtype=np.float64)
    engine._geometry_gen += 1
    check(lib().sph_diff_prepare(engine._h, s.ctypes.data_as(ct.POINTER(ct.c_double))), "sph_diff_prepare")


def sts(T0, K, engine, s, dt, nsteps, every=5):
    return DiffusionSTS.apply(T0, K, engine, s, dt, nsteps, every)


def apply(engine, K, T):
    """A(K) T (device tensors, original particle order)."""
    T = _vector(engine, T, "T")
    out = torch.empty_like(T)
    Kh = _tensor6(K)
    torch.cuda.synchronize()
    check(lib().sph_diff_apply(engine._h, Kh.ctypes.data_as(ct.POINTER(ct.c_double)), ct.c_void_p(T.data_ptr()),
                               ct.c_void_p(out.data_ptr())), "sph_diff_apply")
    return out


def apply_adjoint(engine, K, lam, Ya, Yb):
    """(A(K)^T lam, lam^T dA/dK Ya, lam^T dA/dK Yb)."""
    lam = _vector(engine, lam, "lam")
    Ya = _vector(engine, Ya, "Ya")
    Yb = _vector(engine, Yb, "Yb")
    out = torch.empty_like(lam)
    Kh = _tensor6(K)
    gA = np.zeros(6)
    gB = np.zeros(6)
    torch.cuda.synchronize()
    dp = ct.POINTER(ct.c_double)
    check(lib().sph_diff_apply_adjoint(engine._h, Kh.ctypes.data_as(dp), ct.c_void_p(lam.data_ptr()),
                                       ct.c_void_p(Ya.data_ptr()), ct.c_void_p(Yb.data_ptr()),
                                       ct.c_void_p(out.data_ptr()), gA.ctypes.data_as(dp), gB.ctypes.data_as(dp)),
          "sph_diff_apply_adjoint")
    return out, gA, gB
