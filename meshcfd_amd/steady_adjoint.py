"""Steady solves and the steady adjoint on the device.

No counterpart in the reference, which has neither objectives nor
gradients. At a converged steady state R(U*, beta) = 0 an objective
J(U*, beta) has

    dJ/dbeta = dJ/dbeta|_U + lambda^T dR/dbeta,   (dR/dU)^T lambda = -(dJ/dU)^T.

The adjoint equation is solved by implicit pseudo-time stepping of
d(lambda)/d(tau) = (dR/dU)^T lambda + (dJ/dU)^T: with A = I - D dR/dU
(D = per-cell pseudo-time steps) each step solves A^T w = r_k with the
adjoint GMRES (`mcfd_adjoint_gmres`: exact transposed Jacobian through the
reverse kernels, left-preconditioned by the transposed LU-SGS of the
forward preconditioner) and sets lambda += D w.

Objectives are written in torch on the device state, so dJ/dU comes from
autograd; dR/dbeta for the inflow Mach number and angle comes from the
reverse kernels' boundary-state adjoint (lambda^T dR/d(far-field state))
times the free stream's derivative.
"""

from __future__ import annotations

import ctypes as C
import math

import numpy as np
import torch

from . import adjoint, native
from .geometry import BOUNDARY_CODES, SolverError
from .implicit import ImplicitStepper, McfdGmres
from .layout import FACE_FAR_FIELD


def _p(t):
    return C.c_void_p(t.data_ptr())


def residual_rms(dm, u=None) -> float:
    """sqrt(mean(R^2)) over all cells and components (`solver.py:945-957`)."""
    r = dm.residual(dm.u if u is None else u)
    s = dm.sum_squares(r)
    return math.sqrt(float(s.sum().item()) / (4.0 * dm.n))


def solve_steady(dm, tol: float | None = None, max_steps: int | None = None, raise_cfl: float = 0.0,
                 max_cfl: float = 1e8, progress=None):
    """Implicit-Euler pseudo-time steps (`DeviceMesh.step_implicit`, local or
    global steps as the case says) until the RMS residual is at most ``tol``
    (default: the case's residual tolerance). Returns the RMS history.

    The reference's Newton loop stops once max|U - U^n - dt R| is below 1e-9
    (`solver.py:1418-1425`), so at a given CFL the residual stalls near
    1e-9/dt. With ``raise_cfl`` > 1 the CFL number is multiplied by it
    whenever a step leaves the residual unchanged (up to ``max_cfl``), which
    turns the last steps into Newton steps on R = 0. ``progress(step, rms,
    cfl)`` is called after every step."""
    from .device import native_settings

    s = dm.settings
    tol = s.residual_tolerance if tol is None else tol
    max_steps = s.max_iterations if max_steps is None else max_steps
    history = [residual_rms(dm)]
    t = 0.0
    cfl0 = s.cfl
    try:
        for _ in range(max_steps):
            if history[-1] <= tol:
                break
            t = dm.step_implicit(t)
            history.append(residual_rms(dm))
            if progress is not None:
                progress(len(history) - 1, history[-1], s.cfl)
            if raise_cfl > 1.0 and history[-1] >= history[-2] and s.cfl < max_cfl:
                s.cfl = min(s.cfl * raise_cfl, max_cfl)
                dm.native_settings = native_settings(s)
    finally:
        if s.cfl != cfl0:
            s.cfl = cfl0
            dm.native_settings = native_settings(s)
    return history


class WallForce:
    """Pressure force on the wall faces projected on (cos a, sin a) with a
    given or the free-stream angle: sum_f p_owner (n_f . d) L_f; with
    ``lift`` the direction (-sin a, cos a). Faces of kind Wall (not
    symmetry)."""

    def __init__(self, dm, alpha_deg: float, lift: bool = False):
        g = dm.geometry
        walls = np.asarray(g.wall_faces, dtype=np.int64)
        walls = walls[g.face_kind_code[walls] == BOUNDARY_CODES["Wall"]] if g.face_kind_code is not None else walls
        new_id = np.empty(dm.n, dtype=np.int64)
        new_id[dm.layout.cell_order] = np.arange(dm.n)
        opts = dict(device=dm.device)
        self.cells = torch.as_tensor(new_id[g.owners[walls]], **opts)
        self.nl = torch.as_tensor(g.face_normals[walls] * g.face_lengths[walls, None], dtype=torch.float64, **opts)
        self.gamma = dm.settings.gamma
        self.alpha = math.radians(alpha_deg)
        self.lift = lift

    def direction(self, alpha):
        return (-math.sin(alpha), math.cos(alpha)) if self.lift else (math.cos(alpha), math.sin(alpha))

    def d_direction(self, alpha):
        return (-math.cos(alpha), -math.sin(alpha)) if self.lift else (-math.sin(alpha), math.cos(alpha))

    def __call__(self, u):
        w = u[self.cells]
        p = (self.gamma - 1.0) * (w[:, 3] - 0.5 * (w[:, 1] ** 2 + w[:, 2] ** 2) / w[:, 0])
        dx, dy = self.direction(self.alpha)
        return torch.sum(p * (self.nl[:, 0] * dx + self.nl[:, 1] * dy))

    def d_alpha(self, u) -> float:
        """Explicit derivative with respect to the angle (radians)."""
        w = u[self.cells]
        p = (self.gamma - 1.0) * (w[:, 3] - 0.5 * (w[:, 1] ** 2 + w[:, 2] ** 2) / w[:, 0])
        dx, dy = self.d_direction(self.alpha)
        return float(torch.sum(p * (self.nl[:, 0] * dx + self.nl[:, 1] * dy)).item())


class OutletTotalPressureLoss:
    """Length-weighted mean of (p0_ref - p0) over the pressure-outlet faces,
    p0 = p (1 + (gamma-1)/2 M^2)^(gamma/(gamma-1)) of the owner cells."""

    def __init__(self, dm, p0_ref: float):
        g = dm.geometry
        faces = np.asarray(g.pressure_outlet_faces, dtype=np.int64)
        if len(faces) == 0:
            raise SolverError("The case has no pressure outlet.")
        new_id = np.empty(dm.n, dtype=np.int64)
        new_id[dm.layout.cell_order] = np.arange(dm.n)
        self.cells = torch.as_tensor(new_id[g.owners[faces]], device=dm.device)
        self.length = torch.as_tensor(g.face_lengths[faces], dtype=torch.float64, device=dm.device)
        self.gamma = dm.settings.gamma
        self.p0_ref = p0_ref

    def __call__(self, u):
        gm = self.gamma
        w = u[self.cells]
        rho = w[:, 0]
        v2 = (w[:, 1] ** 2 + w[:, 2] ** 2) / rho**2
        p = (gm - 1.0) * (w[:, 3] - 0.5 * rho * v2)
        m2 = rho * v2 / (gm * p)
        p0 = p * (1.0 + 0.5 * (gm - 1.0) * m2) ** (gm / (gm - 1.0))
        return torch.sum((self.p0_ref - p0) * self.length) / torch.sum(self.length)


def objective_gradient(objective, u):
    """(J, dJ/dU) at the device state u (C,4)."""
    x = u.detach().clone().requires_grad_(True)
    j = objective(x)
    (g,) = torch.autograd.grad(j, x)
    return float(j.item()), g.contiguous()


class SteadyAdjoint:
    """Adjoint solves at the converged state of a DeviceMesh."""

    def __init__(self, dm, cfl: float = 1e3, krylov: int = 30, rtol: float = 1e-3, max_gmres: int = 300):
        self.dm = dm
        # the LU-SGS structure on the mesh (shared with the forward implicit steps)
        if getattr(dm, "_implicit", None) is None:
            dm._implicit = ImplicitStepper(dm)
        self.stepper = dm._implicit
        adjoint.enable(dm)
        self.lib = dm.lib
        self.cfl = cfl
        self.krylov = krylov
        self.rtol = rtol
        self.max_gmres = max_gmres
        self.settings = native.McfdSettings.from_buffer_copy(dm.native_settings)
        self.dt_settings = native.McfdSettings.from_buffer_copy(dm.native_settings)
        self.dt_settings.cfl = cfl
        opts = dict(dtype=torch.float64, device=dm.device)
        self.dt = torch.empty((dm.n,), **opts)
        self.dt_min = torch.empty((1,), **opts)
        self.w = torch.empty((dm.n, 4), **opts)
        self.r = torch.empty((dm.n, 4), **opts)
        self.gmres_iterations = 0

    def transpose_apply(self, u, v, out=None):
        """A^T v = v - (dR/dU)^T (D v) at u with the adjoint pseudo-time steps."""
        dm = self.dm
        out = torch.empty_like(v) if out is None else out
        self._time_steps(u)
        g = McfdGmres()
        g.state = _p(u)
        g.dt = _p(self.dt)
        g.dt_per_cell = 1
        native.check(self.lib.mcfd_adjoint_apply(dm.handle, C.byref(self.settings), C.byref(g), _p(v.contiguous()),
                                                 _p(out), dm.stream))
        return out

    def _time_steps(self, u):
        dm = self.dm
        native.check(dm.lib.mcfd_time_steps(dm.handle, C.byref(self.dt_settings), _p(u), _p(self.dt),
                                            _p(self.dt_min), dm.stream))

    def solve(self, dj_du, u=None, tol: float = 1e-10, max_iterations: int = 100, progress=None):
        """lambda with (dR/dU)^T lambda = -dJ/dU at u (default: the mesh state).
        Returns (lambda, history of ||(dR/dU)^T lambda + dJ/dU|| / ||dJ/dU||)."""
        dm = self.dm
        u = (dm.u if u is None else u).contiguous()
        g_norm = float(torch.linalg.norm(dj_du).item())
        lam = torch.zeros_like(dj_du)
        history = []
        if g_norm == 0.0:
            return lam, history
        self._time_steps(u)
        dm.check_status("adjoint time steps: ")
        self.stepper.assemble(u, self.dt, True)
        for _ in range(max_iterations):
            adjoint.residual_vjp(dm, u, lam, out=self.r)
            self.r.add_(dj_du)
            rel = float(torch.linalg.norm(self.r).item()) / g_norm
            history.append(rel)
            if progress is not None:
                progress(len(history) - 1, rel)
            if rel <= tol:
                break
            g = McfdGmres()
            g.state = _p(u)
            g.rhs = _p(self.r)
            g.x = _p(self.w)
            g.dt = _p(self.dt)
            g.dt_per_cell = 1
            g.krylov = self.krylov
            g.max_iterations = self.max_gmres
            g.tolerance = self.rtol
            native.check(self.lib.mcfd_adjoint_gmres(dm.handle, C.byref(self.settings), C.byref(g), dm.stream))
            self.gmres_iterations += g.iterations
            lam.add_(self.dt[:, None] * self.w)
        dm.check_status("adjoint: ")
        return lam, history

    def inflow_gradient(self, lam, u=None):
        """lambda^T dR/d(far-field state) summed over the far-field records
        (4 values: rho, u, v, p of the free stream)."""
        dm = self.dm
        u = (dm.u if u is None else u).contiguous()
        n_bc = len(dm.layout.bc_kind)
        sbar = torch.zeros((max(n_bc, 1), 4), dtype=torch.float64, device=dm.device)
        adjoint.residual_vjp(dm, u, lam, out=self.r, state_bar=sbar)
        far = torch.as_tensor(dm.layout.bc_kind == FACE_FAR_FIELD, device=dm.device)
        return sbar[:n_bc][far].sum(dim=0).cpu().numpy()


def free_stream_derivatives(mach: float, alpha_deg: float, gamma: float, density: float, pressure: float):
    """d(rho, u, v, p)/dM and d(rho, u, v, p)/d(alpha in radians) of
    `meshgen.free_stream` at fixed density and pressure."""
    c = math.sqrt(gamma * pressure / density)
    a = math.radians(alpha_deg)
    d_mach = np.array((0.0, c * math.cos(a), c * math.sin(a), 0.0))
    d_alpha = np.array((0.0, -mach * c * math.sin(a), mach * c * math.cos(a), 0.0))
    return d_mach, d_alpha
