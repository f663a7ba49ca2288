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


class DiffusionSTS(torch.autograd.Function):
    """T_final = STS^nsteps(T0; K).  Inputs: T0 (n, float64, cuda), K (6,)."""

    @staticmethod
    def forward(ctx, T0, K, engine, s, dt, nsteps, every):
        L = lib()
        T0c = T0.detach().contiguous()
        out = torch.empty_like(T0c)
        Kh = np.ascontiguousarray(K.detach().cpu().numpy(), dtype=np.float64)
        torch.cuda.synchronize()
        check(L.sph_diff_forward(engine._h, Kh.ctypes.data_as(ct.POINTER(ct.c_double)), int(s), float(dt),
                                 int(nsteps), int(every), ct.c_void_p(T0c.data_ptr()), ct.c_void_p(out.data_ptr())),
              "sph_diff_forward")
        ctx.engine = engine
        ctx.Kdev = K.device
        return out

    @staticmethod
    def backward(ctx, gT):
        L = lib()
        g = gT.detach().contiguous()
        gT0 = torch.empty_like(g)
        gK = np.zeros(6)
        torch.cuda.synchronize()
        check(L.sph_diff_backward(ctx.engine._h, ct.c_void_p(g.data_ptr()), ct.c_void_p(gT0.data_ptr()),
                                  gK.ctypes.data_as(ct.POINTER(ct.c_double))), "sph_diff_backward")
        return gT0, torch.from_numpy(gK).to(ctx.Kdev), None, None, None, None, None


def prepare(engine, state=None):
    """Freeze h/rho/Omega of the resident particles and build reverse lists."""
    s = engine._state if state is None else np.ascontiguousarray(state, dtype=np.float64)
    check(lib().sph_diff_prepare(engine._h, s.ctypes.data_as(ct.POINTER(ct.c_double))), "sph_diff_prepare")


def sts(T0, K, engine, s, dt, nsteps, every=5):
    return DiffusionSTS.apply(T0, K, engine, s, dt, nsteps, every)


def apply(engine, K, T):
    """A(K) T (device tensors, original particle order)."""
    out = torch.empty_like(T)
    Kh = np.ascontiguousarray(np.asarray(K, dtype=np.float64))
    torch.cuda.synchronize()
    check(lib().sph_diff_apply(engine._h, Kh.ctypes.data_as(ct.POINTER(ct.c_double)), ct.c_void_p(T.data_ptr()),
                               ct.c_void_p(out.data_ptr())), "sph_diff_apply")
    return out


def apply_adjoint(engine, K, lam, Ya, Yb):
    """(A(K)^T lam, lam^T dA/dK Ya, lam^T dA/dK Yb)."""
    out = torch.empty_like(lam)
    Kh = np.ascontiguousarray(np.asarray(K, dtype=np.float64))
    gA = np.zeros(6)
    gB = np.zeros(6)
    torch.cuda.synchronize()
    dp = ct.POINTER(ct.c_double)
    check(lib().sph_diff_apply_adjoint(engine._h, Kh.ctypes.data_as(dp), ct.c_void_p(lam.data_ptr()),
                                       ct.c_void_p(Ya.data_ptr()), ct.c_void_p(Yb.data_ptr()),
                                       ct.c_void_p(out.data_ptr()), gA.ctypes.data_as(dp), gB.ctypes.data_as(dp)),
          "sph_diff_apply_adjoint")
    return out, gA, gB
