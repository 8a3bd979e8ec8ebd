"""Adjoint (reverse-mode) residual and differentiable time stepping.

No counterpart in the reference, which has no gradients. `mcfd_residual_vjp`
is a hand-written backward kernel chain for the residual (Rusanov and HLLC
face fluxes and their ghost states in reverse mode, Roe by 8-way
forward-mode duals per face, then reconstruction, limiter, LSQ gradient or
quadratic fit and primitive conversion reversed as deterministic per-cell
gathers).
`ResidualFunction` exposes it to `torch.autograd`; `rk3_rollout` advances
SSP-RK3 steps with a fixed step size, checkpointing the state at every step
and recomputing the stages during the backward pass, so long transients need
one state per step of HBM.
"""

from __future__ import annotations

import ctypes as C

import numpy as np
import torch
from torch.utils.checkpoint import checkpoint

from . import native

native.EXPORTS.update(
    {
        "mcfd_mesh_set_adjoint": (C.c_int, [C.c_void_p, C.c_void_p]),
        "mcfd_mesh_set_parabolic_adjoint": (C.c_int, [C.c_void_p, C.c_void_p]),
        "mcfd_residual_vjp": (
            C.c_int,
            [C.c_void_p, C.POINTER(native.McfdSettings), C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p,
             C.c_void_p],
        ),
    }
)


def _p(t):
    return C.c_void_p(t.data_ptr())


def enable(dm) -> None:
    """Upload the adjoint structure of a DeviceMesh (idempotent)."""
    if getattr(dm, "_adjoint_ready", False):
        return
    lib = native.load()
    for name in ("mcfd_mesh_set_adjoint", "mcfd_residual_vjp", "mcfd_mesh_set_parabolic_adjoint"):
        restype, argtypes = native.EXPORTS[name]
        fn = getattr(lib, name)
        fn.restype, fn.argtypes = restype, argtypes
    rev = np.ascontiguousarray(dm.layout.grad_rev, dtype=np.int32)
    native.check(lib.mcfd_mesh_set_adjoint(dm.handle, C.c_void_p(rev.ctypes.data)))
    par = getattr(dm, "_parabolic", None)
    if par is not None:
        prev = np.ascontiguousarray(par["rev"], dtype=np.int32)
        native.check(lib.mcfd_mesh_set_parabolic_adjoint(dm.handle, C.c_void_p(prev.ctypes.data)))
    dm._adjoint_ready = True


def residual_vjp(dm, u: torch.Tensor, lam: torch.Tensor, out: torch.Tensor | None = None,
                 state_bar: torch.Tensor | None = None) -> torch.Tensor:
    """lam^T dR/dU at u (device order, (C,4)). With ``state_bar`` ((n_bc, 4)
    device tensor) also lam^T dR/d(exterior state) per boundary record
    (velocity inlets and far field; 0 for the other records)."""
    enable(dm)
    u = u.contiguous()
    lam = lam.contiguous()
    out = torch.empty_like(u) if out is None else out
    sbar = None if state_bar is None else C.c_void_p(state_bar.data_ptr())
    native.check(
        dm.lib.mcfd_residual_vjp(dm.handle, C.byref(dm.native_settings), _p(u), _p(lam), _p(out), sbar, dm.stream)
    )
    return out


class ResidualFunction(torch.autograd.Function):
    """R = residual(U) with the hand-written VJP as its backward."""

    @staticmethod
    def forward(ctx, u, dm):
        u = u.contiguous()
        r = dm.residual(u)
        ctx.save_for_backward(u)
        ctx.dm = dm
        return r

    @staticmethod
    def backward(ctx, grad_r):
        (u,) = ctx.saved_tensors
        return residual_vjp(ctx.dm, u, grad_r.contiguous()), None


def residual(dm, u):
    return ResidualFunction.apply(u, dm)


def rk3_step(dm, u, dt: float):
    """SSP-RK3 with the reference's Shu-Osher coefficients (solver.py:567-756)."""
    one = u + dt * residual(dm, u)
    two = 0.75 * u + 0.25 * (one + dt * residual(dm, one))
    return (1.0 / 3.0) * u + (2.0 / 3.0) * (two + dt * residual(dm, two))


def rk3_rollout(dm, u0, dt: float, steps: int, checkpointed: bool = True):
    """`steps` SSP-RK3 steps of fixed size; with ``checkpointed`` only the
    state at each step start is kept for the backward pass."""
    u = u0
    for _ in range(steps):
        if checkpointed and u.requires_grad:
            u = checkpoint(rk3_step, dm, u, dt, use_reentrant=False)
        else:
            u = rk3_step(dm, u, dt)
    return u


def kinetic_energy(u, areas):
    """sum_i A_i |m_i|^2 / (2 rho_i)."""
    return torch.sum(areas * (u[:, 1] ** 2 + u[:, 2] ** 2) / (2.0 * u[:, 0]))


def vortex_state(centers, strength, gamma=1.4, free_stream=(1.0, 1.0, 1.0, 1.0), center=(0.0, 0.0)):
    """Conservative isentropic-vortex state as a differentiable function of
    the vortex strength (same formula as meshgen.isentropic_vortex)."""
    rho_inf, u_inf, v_inf, p_inf = free_stream
    x = centers[:, 0] - center[0]
    y = centers[:, 1] - center[1]
    bump = torch.exp(0.5 * (1.0 - (x * x + y * y)))
    du = -strength / (2.0 * np.pi) * bump * y
    dv = strength / (2.0 * np.pi) * bump * x
    t = p_inf / rho_inf - (gamma - 1.0) * strength**2 / (8.0 * gamma * np.pi**2) * bump * bump
    entropy = p_inf / rho_inf**gamma
    rho = (t / entropy) ** (1.0 / (gamma - 1.0))
    p = rho * t
    u = u_inf + du
    v = v_inf + dv
    e = p / (gamma - 1.0) + 0.5 * rho * (u * u + v * v)
    return torch.stack((rho, rho * u, rho * v, e), dim=1)
