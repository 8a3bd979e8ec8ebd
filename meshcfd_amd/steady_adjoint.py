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


def _csr(keys, n):
    """Rows of ``keys`` grouped by value: (order, ptr) with order[ptr[i]:ptr[i+1]]
    the positions whose key is i, in their original order."""
    order = np.argsort(keys, kind="stable")
    ptr = np.zeros(n + 1, dtype=np.int64)
    np.cumsum(np.bincount(keys, minlength=n), out=ptr[1:])
    return order, ptr


class ViscousWallTraction:
    """Viscous traction (tx, ty) on given wall faces as torch functions of the
    conservative state: the viscous residual's wall-face flux, restated from
    the reference's viscous gradients (`solver.py:3615-3714`: weighted least
    squares of (u, v, T) with one ghost entry per non-periodic boundary face)
    and its boundary-face branch (`:3927-4113`: no-slip / slip velocity
    correction, fixed temperature / heat flux, viscosity at the wall value,
    Stokes stress). The residual adds (tx, ty) L / A to the owner cell, so the
    fluid pushes on the wall with -(tx, ty) L.

    Ghost entries of the owner's other boundary faces follow the same branch
    table (walls, symmetry planes, velocity inlets; outlets and transmissive
    faces add a zero difference). Owners that also touch a far-field face are
    rejected: their ghost state is the characteristic far-field state."""

    WALL = BOUNDARY_CODES["Wall"]
    SYMMETRY = BOUNDARY_CODES["Symmetry"]
    INLET = BOUNDARY_CODES["Velocity Inlet"]
    FAR_FIELD = BOUNDARY_CODES["Far Field"]

    def __init__(self, g, settings, faces, new_id, device):
        s = settings
        self.s = s
        n_cells = len(g.areas)
        faces = np.asarray(faces, dtype=np.int64)
        owners = g.owners[faces]
        kind = g.face_kind_code
        t = lambda a, dt=torch.float64: torch.as_tensor(np.ascontiguousarray(a), dtype=dt, device=device)

        # interior / periodic LSQ entries of each owner
        src = np.asarray(g.gradient_sources, dtype=np.int64)
        e_order, e_ptr = _csr(src, n_cells)
        cnt = e_ptr[owners + 1] - e_ptr[owners]
        E = max(int(cnt.max(initial=0)), 1)
        slot = np.arange(E)
        ok = slot[None, :] < cnt[:, None]
        pos = np.where(ok, e_ptr[owners][:, None] + slot[None, :], 0)
        ent = e_order[pos] if len(e_order) else np.zeros_like(pos)
        dx = np.where(ok, np.asarray(g.gradient_dx)[ent] if len(src) else 0.0, 0.0)
        dy = np.where(ok, np.asarray(g.gradient_dy)[ent] if len(src) else 0.0, 0.0)
        nbr = np.where(ok, np.asarray(g.gradient_neighbours, dtype=np.int64)[ent] if len(src) else 0, owners[:, None])
        w = np.where(ok, 1.0 / np.maximum(dx * dx + dy * dy, 1e-300), 0.0)

        # ghost entries: the owner's non-periodic boundary faces
        bfaces = np.asarray(g.diffusive_boundary_faces, dtype=np.int64)
        b_order, b_ptr = _csr(g.owners[bfaces], n_cells)
        bcnt = b_ptr[owners + 1] - b_ptr[owners]
        G = max(int(bcnt.max(initial=0)), 1)
        gslot = np.arange(G)
        gok = gslot[None, :] < bcnt[:, None]
        gpos = np.where(gok, b_ptr[owners][:, None] + gslot[None, :], 0)
        gf = bfaces[b_order[gpos]] if len(bfaces) else np.zeros_like(gpos)
        gkind = np.where(gok, kind[gf], -1)
        if np.any(gkind == self.FAR_FIELD):
            raise SolverError("The viscous wall force needs wall cells without far-field faces.")
        gn = g.face_normals[gf]
        goff = g.face_centers[gf] - g.centers[owners][:, None, :]
        gdn = np.sum(goff * gn, axis=2)
        mirror = (gkind == self.WALL) | (gkind == self.SYMMETRY)
        gvec = np.where(mirror[..., None], 2.0 * gdn[..., None] * gn, 2.0 * goff)
        gw = np.where(gok, 1.0 / np.maximum(np.sum(gvec * gvec, axis=2), 1e-300), 0.0)
        inlet_state = np.zeros(gf.shape + (3,))
        if len(g.velocity_inlet_faces):
            vi = np.full(len(g.owners), -1, dtype=np.int64)
            vi[g.velocity_inlet_faces] = np.arange(len(g.velocity_inlet_faces))
            st = g.velocity_inlet_states
            idx = vi[gf]
            has = (gkind == self.INLET) & (idx >= 0)
            sv = st[np.maximum(idx, 0)]
            inlet_state = np.where(has[..., None], np.stack(
                (sv[..., 1], sv[..., 2], sv[..., 3] / (sv[..., 0] * s.gas_constant)), axis=-1), 0.0)

        # the normal-equation matrix depends on the geometry only
        xx = np.sum(w * dx * dx, axis=1) + np.sum(gw * gvec[..., 0] ** 2, axis=1)
        xy = np.sum(w * dx * dy, axis=1) + np.sum(gw * gvec[..., 0] * gvec[..., 1], axis=1)
        yy = np.sum(w * dy * dy, axis=1) + np.sum(gw * gvec[..., 1] ** 2, axis=1)
        det = xx * yy - xy * xy
        valid = det > 1e-14 * np.maximum(xx * yy, 1e-300)
        inv = np.where(valid, 1.0 / np.where(valid, det, 1.0), 0.0)

        n = g.face_normals[faces]
        off = g.face_centers[faces] - g.centers[owners]
        self.owner = t(new_id[owners], torch.int64)
        self.nbr = t(new_id[nbr], torch.int64)
        self.wx, self.wy = t(w * dx), t(w * dy)
        self.gwx, self.gwy = t(gw * gvec[..., 0]), t(gw * gvec[..., 1])
        self.gkind = t(gkind, torch.int64)
        self.gn, self.gdn = t(gn), t(gdn)
        gw_vel = g.wall_velocities[gf]
        self.g_wall_vel = t(gw_vel)
        self.g_no_slip = t(g.wall_no_slip[gf] & (gkind == self.WALL), torch.bool)
        self.g_fixed_t = t((g.wall_thermal[gf] == 1) & (gkind == self.WALL), torch.bool)
        self.g_thermal = t(np.where(gkind == self.WALL, g.wall_thermal_value[gf], 0.0))
        self.g_inlet = t(inlet_state)
        self.xx, self.xy, self.yy, self.inv = t(xx), t(xy), t(yy), t(inv)
        self.n, self.off = t(n), t(off)
        self.dn = t(np.sum(off * n, axis=1))
        if torch.any(self.dn <= 1e-14):
            raise SolverError("Boundary face has invalid cell-center distance.")
        self.wall_vel = t(g.wall_velocities[faces])
        self.no_slip = t(g.wall_no_slip[faces], torch.bool)
        self.fixed_t = t(g.wall_thermal[faces] == 1, torch.bool)
        self.thermal = t(g.wall_thermal_value[faces])

    def viscosity(self, temperature):
        s = self.s
        if s.viscosity_model == "Sutherland":
            tr = s.sutherland_reference_temperature
            return (s.sutherland_reference_viscosity * (temperature / tr) ** 1.5
                    * ((tr + s.sutherland_temperature) / (temperature + s.sutherland_temperature)))
        return torch.full_like(temperature, s.dynamic_viscosity)

    def conductivity(self, mu):
        s = self.s
        return mu * s.gamma * s.gas_constant / ((s.gamma - 1.0) * s.prandtl_number)

    def __call__(self, u):
        s = self.s
        rho = u[:, 0]
        vel = u[:, 1:3] / rho[:, None]
        p = (s.gamma - 1.0) * (u[:, 3] - 0.5 * (u[:, 1] ** 2 + u[:, 2] ** 2) / rho)
        values = torch.cat((vel, (p / (rho * s.gas_constant))[:, None]), dim=1)
        vo = values[self.owner]  # (F, 3)
        diff = values[self.nbr] - vo[:, None, :]
        rx = torch.sum(self.wx[..., None] * diff, dim=1)
        ry = torch.sum(self.wy[..., None] * diff, dim=1)
        # ghost differences by boundary kind
        uv = vo[:, None, :2]
        un = torch.sum(uv * self.gn, dim=2, keepdim=True)
        mirrored = -2.0 * un * self.gn
        d_uv = torch.where(self.g_no_slip[..., None], 2.0 * (self.g_wall_vel - uv), mirrored)
        wall_or_sym = (self.gkind == self.WALL) | (self.gkind == self.SYMMETRY)
        d_uv = torch.where(wall_or_sym[..., None], d_uv, torch.zeros_like(d_uv))
        mu_o = self.viscosity(vo[:, 2])
        d_t = torch.where(self.g_fixed_t, 2.0 * (self.g_thermal - vo[:, None, 2]),
                          2.0 * self.gdn * self.g_thermal / self.conductivity(mu_o)[:, None])
        d_t = torch.where(self.gkind == self.WALL, d_t, torch.zeros_like(d_t))
        gdiff = torch.cat((d_uv, d_t[..., None]), dim=2)
        inlet = (self.gkind == self.INLET)[..., None]
        gdiff = torch.where(inlet, 2.0 * (self.g_inlet - vo[:, None, :]), gdiff)
        rx = rx + torch.sum(self.gwx[..., None] * gdiff, dim=1)
        ry = ry + torch.sum(self.gwy[..., None] * gdiff, dim=1)
        gx = (rx * self.yy[:, None] - ry * self.xy[:, None]) * self.inv[:, None]
        gy = (ry * self.xx[:, None] - rx * self.xy[:, None]) * self.inv[:, None]
        fg = torch.stack((gx, gy), dim=2)  # (F, 3, 2)

        n, off, dn = self.n, self.off, self.dn
        tang = off - dn[:, None] * n
        # no-slip: normal derivative from the wall velocity
        nd_ns = (self.wall_vel - vo[:, :2] - torch.sum(fg[:, :2] * tang[:, None, :], dim=2)) / dn[:, None]
        fg_ns = fg[:, :2] + (nd_ns - torch.sum(fg[:, :2] * n[:, None, :], dim=2))[..., None] * n[:, None, :]
        bv_ns = self.wall_vel
        # slip: tangential velocity extrapolated to the face, u.n = 0 enforced
        un_o = torch.sum(vo[:, :2] * n, dim=1)
        bv_sl = vo[:, :2] + torch.sum(fg[:, :2] * off[:, None, :], dim=2)
        bv_sl = bv_sl - torch.sum(bv_sl * n, dim=1, keepdim=True) * n
        gnv = n[:, 0:1] * fg[:, 0] + n[:, 1:2] * fg[:, 1]
        nd_sl = -(un_o + torch.sum(gnv * tang, dim=1)) / dn
        desired = (nd_sl[:, None] * n)[:, :, None] * n[:, None, :]
        fg_sl = fg[:, :2] + desired - torch.sum(fg[:, :2] * n[:, None, :], dim=2)[..., None] * n[:, None, :]
        ns = self.no_slip[:, None]
        bv_uv = torch.where(ns, bv_ns, bv_sl)
        fg_uv = torch.where(ns[..., None], fg_ns, fg_sl)
        # temperature
        t_face = torch.where(self.fixed_t, self.thermal, vo[:, 2] + torch.sum(fg[:, 2] * off, dim=1))
        nd_t = torch.where(self.fixed_t,
                           (self.thermal - vo[:, 2] - torch.sum(fg[:, 2] * tang, dim=1)) / dn,
                           self.thermal / self.conductivity(mu_o))
        fg_t = fg[:, 2] + (nd_t - torch.sum(fg[:, 2] * n, dim=1))[:, None] * n
        mu = self.viscosity(t_face)
        du_dx, du_dy = fg_uv[:, 0, 0], fg_uv[:, 0, 1]
        dv_dx, dv_dy = fg_uv[:, 1, 0], fg_uv[:, 1, 1]
        div = du_dx + dv_dy
        txx = mu * (2.0 * du_dx - (2.0 / 3.0) * div)
        tyy = mu * (2.0 * dv_dy - (2.0 / 3.0) * div)
        txy = mu * (du_dy + dv_dx)
        del fg_t, bv_uv  # the energy flux does not enter the traction
        return torch.stack((txx * n[:, 0] + txy * n[:, 1], txy * n[:, 0] + tyy * n[:, 1]), dim=1)


class WallForce:
    """Force of the fluid on the wall faces projected on (cos a, sin a) with a
    given or the free-stream angle; with ``lift`` the direction
    (-sin a, cos a). Faces of kind Wall (not symmetry). The pressure part is
    sum_f p_owner (n_f . d) L_f; viscous cases add the viscous traction,
    -sum_f (t_f . d) L_f with t_f the viscous residual's wall flux
    (`ViscousWallTraction`)."""

    def __init__(self, dm, alpha_deg: float, lift: bool = False):
        g = dm.geometry
        walls = np.asarray(g.wall_faces, dtype=np.int64)
        walls = walls[g.face_kind_code[walls] == BOUNDARY_CODES["Wall"]] if g.face_kind_code is not None else walls
        new_id = np.empty(dm.n, dtype=np.int64)
        new_id[dm.layout.cell_order] = np.arange(dm.n)
        opts = dict(device=dm.device)
        self.cells = torch.as_tensor(new_id[g.owners[walls]], **opts)
        self.nl = torch.as_tensor(g.face_normals[walls] * g.face_lengths[walls, None], dtype=torch.float64, **opts)
        self.length = torch.as_tensor(g.face_lengths[walls], dtype=torch.float64, **opts)
        self.gamma = dm.settings.gamma
        self.alpha = math.radians(alpha_deg)
        self.lift = lift
        self.viscous = None
        if dm.settings.viscosity_model != "Inviscid" and len(walls):
            self.viscous = ViscousWallTraction(g, dm.settings, walls, new_id, dm.device)

    def direction(self, alpha):
        return (-math.sin(alpha), math.cos(alpha)) if self.lift else (math.cos(alpha), math.sin(alpha))

    def d_direction(self, alpha):
        return (-math.cos(alpha), -math.sin(alpha)) if self.lift else (-math.sin(alpha), math.cos(alpha))

    def components(self, u):
        """(pressure force, viscous force) on the walls, each (2,)."""
        w = u[self.cells]
        p = (self.gamma - 1.0) * (w[:, 3] - 0.5 * (w[:, 1] ** 2 + w[:, 2] ** 2) / w[:, 0])
        fp = torch.sum(p[:, None] * self.nl, dim=0)
        if self.viscous is None:
            return fp, torch.zeros_like(fp)
        return fp, -torch.sum(self.viscous(u) * self.length[:, None], dim=0)

    def _project(self, u, d):
        fp, fv = self.components(u)
        f = fp + fv
        return f[0] * d[0] + f[1] * d[1]

    def __call__(self, u):
        return self._project(u, self.direction(self.alpha))

    def d_alpha(self, u) -> float:
        """Explicit derivative with respect to the angle (radians)."""
        return float(self._project(u, self.d_direction(self.alpha)).item())


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
