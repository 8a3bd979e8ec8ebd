"""ORACLE (test infrastructure only): residual, time steps and explicit steps.

CPU restatement of the reference solver's hot path on a precomputed geometry
(any object with the `_Geometry` field names, `solver.py:122-176`):

* `residual`            <- `_residual` (`solver.py:2437-2718`), inviscid, with
  the boundary kinds periodic / wall+symmetry / velocity inlet / pressure
  outlet / far field, plus transmissive (no reference counterpart: exterior
  state = interior reconstructed state, processed after far field);
* `stable_time_steps`   <- `_stable_time_steps` (`solver.py:4132-4343`);
* `far_field_states`    <- `_far_field_boundary_states` (`solver.py:4350-4409`);
* `step`                <- the explicit branches of `solve_case`
  (`solver.py:437-812`) and `_explicit_rk4_step` (`solver.py:1128-1182`).

Summation orders follow the reference: interior faces through a CSR matvec
whose rows sum in ascending face order, then ``np.add.at`` for each boundary
group in turn.
"""

from __future__ import annotations

import math

import numpy as np
import scipy.sparse as sp

from . import kernels
from .euler import (
    conservative_to_primitive,
    numerical_flux,
    primitive_to_conservative,
    sound_speed,
)

LIMITER_CODES = {"Barth-Jespersen": 0, "Minmod": 1, "Van Leer": 2, "Venkatakrishnan": 3}


class OracleError(ValueError):
    pass


def divergence_operator(geometry):
    """CSR (C x F_int) with -L/A_owner and +L/A_neighbour (`solver.py:2157-2174`)."""
    interior = geometry.interior_faces
    io, ine = geometry.interior_owners, geometry.interior_neighbours
    ids = np.arange(len(interior))
    lengths = geometry.face_lengths[interior]
    data = np.concatenate((-lengths / geometry.areas[io], lengths / geometry.areas[ine]))
    rows = np.concatenate((io, ine))
    cols = np.concatenate((ids, ids))
    return sp.coo_matrix((data, (rows, cols)), shape=(len(geometry.areas), len(interior))).tocsr()


class Oracle:
    """Residual / dt / step evaluator bound to one geometry and settings."""

    def __init__(self, geometry, settings):
        self.g = geometry
        self.s = settings
        self.D = divergence_operator(geometry)
        self.stencil = None
        if settings.reconstruction == "MUSCL parabolic":
            from .parabolic import parabolic_stencil

            self.stencil = parabolic_stencil(
                len(geometry.areas), geometry.gradient_sources, geometry.gradient_neighbours,
                geometry.gradient_dx, geometry.gradient_dy, geometry.domain_scale,
            )

    # -- reconstruction ---------------------------------------------------
    def gradients_and_factors(self, primitive):
        g, s = self.g, self.s
        if s.reconstruction == "MUSCL linear" and len(g.gradient_sources):
            return kernels.linear_gradients_and_factors(
                primitive,
                g.gradient_sources,
                g.gradient_neighbours,
                g.gradient_coefficients_x,
                g.gradient_coefficients_y,
                g.limiter_cells,
                g.limiter_dx,
                g.limiter_dy,
                LIMITER_CODES[s.limiter],
            )
        if s.reconstruction not in ("First order", "MUSCL linear"):
            raise OracleError(f"Unsupported reconstruction '{s.reconstruction}'.")
        n = len(primitive)
        return np.zeros((n, 4, 2)), np.ones((n, 4))

    def parabolic(self, primitive):
        from .parabolic import parabolic_reconstruction

        return parabolic_reconstruction(primitive, self.g, self.stencil, LIMITER_CODES[self.s.limiter])

    # -- residual ---------------------------------------------------------
    def residual(self, conservative):
        g, s = self.g, self.s
        gamma = s.gamma
        primitive = conservative_to_primitive(conservative, gamma)
        curvature = None
        if s.reconstruction == "MUSCL parabolic":
            grads, curvature, factors = self.parabolic(primitive)
        else:
            grads, factors = self.gradients_and_factors(primitive)

        def rec(cells, offsets):
            if curvature is not None:
                from .parabolic import reconstruct as prec

                return prec(primitive, grads, factors, curvature, cells, offsets)
            return kernels.reconstruct(primitive, grads, factors, cells, offsets)

        deriv = np.zeros_like(conservative)
        if len(g.interior_faces):
            if s.flux == "HLLC" and curvature is None:
                flux = kernels.reconstruct_hllc_flux(
                    primitive, grads, factors, g.interior_owners, g.interior_neighbours,
                    g.interior_owner_offsets, g.interior_neighbour_offsets, g.interior_normals, gamma,
                )
            else:
                left = rec(g.interior_owners, g.interior_owner_offsets)
                right = rec(g.interior_neighbours, g.interior_neighbour_offsets)
                flux = numerical_flux(left, right, g.interior_normals, gamma, s.flux)
            deriv += self.D @ flux

        def boundary_offsets(faces, cells):
            return g.face_centers[faces] - g.centers[cells]

        if len(g.periodic_faces_a):
            fa, fb = g.periodic_faces_a, g.periodic_faces_b
            oa, ob = g.owners[fa], g.owners[fb]
            left = rec(oa, boundary_offsets(fa, oa))
            right = rec(ob, boundary_offsets(fb, ob))
            flux = numerical_flux(left, right, g.face_normals[fa], gamma, s.flux)
            integ = flux * g.face_lengths[fa, None]
            np.add.at(deriv, oa, -integ / g.areas[oa, None])
            np.add.at(deriv, ob, integ / g.areas[ob, None])

        def owner_side(faces, exterior_fn):
            owners = g.owners[faces]
            interior = rec(owners, boundary_offsets(faces, owners))
            exterior = exterior_fn(interior, faces)
            flux = numerical_flux(interior, exterior, g.face_normals[faces], gamma, s.flux)
            integ = flux * g.face_lengths[faces, None]
            np.add.at(deriv, owners, -integ / g.areas[owners, None])

        if len(g.wall_faces):
            def wall(interior, faces):
                right = interior.copy()
                rel = interior[:, 1:3] - g.wall_velocities[faces]
                n = g.face_normals[faces]
                un = rel[:, 0] * n[:, 0] + rel[:, 1] * n[:, 1]
                right[:, 1:3] -= 2.0 * un[:, None] * n
                return right
            owner_side(g.wall_faces, wall)
        if len(g.velocity_inlet_faces):
            owner_side(g.velocity_inlet_faces, lambda interior, faces: g.velocity_inlet_states)
        if len(g.pressure_outlet_faces):
            def outlet(interior, faces):
                ext = interior.copy()
                ext[:, 3] = g.pressure_outlet_pressures
                return ext
            owner_side(g.pressure_outlet_faces, outlet)
        if len(g.far_field_faces):
            owner_side(
                g.far_field_faces,
                lambda interior, faces: far_field_states(
                    interior, g.far_field_states, g.face_normals[faces], gamma
                ),
            )
        if len(g.pressure_inlet_faces):
            from .inflow import pressure_inlet_states

            owner_side(g.pressure_inlet_faces, lambda interior, faces: pressure_inlet_states(
                interior, g.face_normals[faces], g.pressure_inlet_params, g.pressure_inlet_patch, gamma,
                s.gas_constant, g.inflow_patch_names[0]))
        if len(g.mass_flow_inlet_faces):
            from .inflow import mass_flow_inlet_states

            owner_side(g.mass_flow_inlet_faces, lambda interior, faces: mass_flow_inlet_states(
                interior, g.face_normals[faces], g.face_lengths[faces], g.mass_flow_inlet_params,
                g.mass_flow_inlet_patch, gamma, s.gas_constant, g.inflow_patch_names[1]))
        if len(g.mass_flow_outlet_faces):
            from .inflow import mass_flow_outlet_states

            owner_side(g.mass_flow_outlet_faces, lambda interior, faces: mass_flow_outlet_states(
                interior, g.face_normals[faces], g.mass_flow_outlet_params))
        if len(getattr(g, "transmissive_faces", ())):
            owner_side(g.transmissive_faces, lambda interior, faces: interior.copy())
        if s.viscosity_model != "Inviscid":
            from .viscous import viscous_residual

            deriv += viscous_residual(primitive, g, s, self.D)
        if s.acceleration_x != 0.0 or s.acceleration_y != 0.0:
            deriv[:, 1] += primitive[:, 0] * s.acceleration_x
            deriv[:, 2] += primitive[:, 0] * s.acceleration_y
            deriv[:, 3] += primitive[:, 0] * (
                s.acceleration_x * primitive[:, 1] + s.acceleration_y * primitive[:, 2]
            )
        if not np.isfinite(deriv).all():
            raise OracleError("Spatial residual contains non-finite values.")
        return deriv

    # -- time step --------------------------------------------------------
    def stable_time_steps(self, primitive):
        g, s = self.g, self.s
        gamma = s.gamma
        n = g.face_normals
        un = primitive[g.owners, 1] * n[:, 0] + primitive[g.owners, 2] * n[:, 1]
        c = sound_speed(primitive, gamma)
        radius = np.abs(un) + c[g.owners]
        bfaces = np.concatenate((g.velocity_inlet_faces, g.pressure_outlet_faces, g.far_field_faces,
                                 g.pressure_inlet_faces, g.mass_flow_inlet_faces,
                                 g.mass_flow_outlet_faces)).astype(np.int64)
        if len(bfaces):
            owners = g.owners[bfaces]
            ext = primitive[owners].copy()
            ni = len(g.velocity_inlet_faces)
            no = ni + len(g.pressure_outlet_faces)
            if ni:
                ext[:ni] = g.velocity_inlet_states
            if len(g.pressure_outlet_faces):
                ext[ni:no, 3] = g.pressure_outlet_pressures
            nf = no + len(g.far_field_faces)
            if len(g.far_field_faces):
                ext[no:nf] = far_field_states(primitive[owners[no:nf]], g.far_field_states, n[bfaces[no:nf]], gamma)
            k = nf
            if len(g.pressure_inlet_faces):
                from .inflow import pressure_inlet_states

                m = len(g.pressure_inlet_faces)
                ext[k:k + m] = pressure_inlet_states(primitive[owners[k:k + m]], n[bfaces[k:k + m]],
                                                     g.pressure_inlet_params, g.pressure_inlet_patch, gamma,
                                                     s.gas_constant, g.inflow_patch_names[0])
                k += m
            if len(g.mass_flow_inlet_faces):
                from .inflow import mass_flow_inlet_states

                m = len(g.mass_flow_inlet_faces)
                ext[k:k + m] = mass_flow_inlet_states(primitive[owners[k:k + m]], n[bfaces[k:k + m]],
                                                      g.face_lengths[bfaces[k:k + m]], g.mass_flow_inlet_params,
                                                      g.mass_flow_inlet_patch, gamma, s.gas_constant,
                                                      g.inflow_patch_names[1])
                k += m
            if len(g.mass_flow_outlet_faces):
                from .inflow import mass_flow_outlet_states

                m = len(g.mass_flow_outlet_faces)
                ext[k:k + m] = mass_flow_outlet_states(primitive[owners[k:k + m]], n[bfaces[k:k + m]],
                                                       g.mass_flow_outlet_params)
            eun = ext[:, 1] * n[bfaces, 0] + ext[:, 2] * n[bfaces, 1]
            er = np.abs(eun) + sound_speed(ext, gamma)
            radius[bfaces] = np.maximum(radius[bfaces], er)
        total = np.zeros(len(primitive))
        np.add.at(total, g.owners, radius * g.face_lengths)
        inner = g.interior_faces
        if len(inner):
            nb = g.neighbours[inner]
            nun = primitive[nb, 1] * n[inner, 0] + primitive[nb, 2] * n[inner, 1]
            np.add.at(total, nb, (np.abs(nun) + c[nb]) * g.face_lengths[inner])
        if s.viscosity_model != "Inviscid":
            from .viscous import viscous_time_step_sum

            total += viscous_time_step_sum(primitive, g, s)
        steps = s.cfl * g.areas / total
        if not np.isfinite(steps).all() or np.any(steps <= 0.0):
            raise OracleError("Local stable time steps must be positive and finite.")
        return steps

    # -- explicit step ----------------------------------------------------
    def step(self, conservative, time):
        """One explicit step of `solve_case` (`solver.py:443-812`).

        Returns (new conservative, reported time step, per-cell or scalar
        step sizes used, stage-one derivative)."""
        s = self.s
        primitive = conservative_to_primitive(conservative, s.gamma)
        if s.time_step_mode == "Local (steady)":
            sizes = self.stable_time_steps(primitive)
            dt = float(np.mean(sizes))
        elif s.time_step_mode == "Global fixed":
            dt = min(s.fixed_time_step, s.end_time - time)
            sizes = dt
        else:
            dt = min(float(np.min(self.stable_time_steps(primitive))), s.end_time - time)
            sizes = dt
        if not math.isfinite(dt) or dt <= 0.0:
            raise OracleError("The stable time step is non-positive or non-finite.")

        def scale(rate, h):
            if np.ndim(h) == 0:
                return h * rate
            return np.asarray(h)[:, None] * rate

        start = conservative
        d0 = self.residual(start)
        ti = s.time_integration
        if ti == "Explicit RK4":
            new = self.rk4(start, sizes)
        elif ti == "Explicit Euler":
            new = start + scale(d0, sizes)
        elif ti in ("Explicit RK2", "Explicit RK3"):
            one = start + scale(d0, sizes)
            conservative_to_primitive(one, s.gamma)
            d1 = self.residual(one)
            if ti == "Explicit RK2":
                new = 0.5 * start + 0.5 * (one + scale(d1, sizes))
            else:
                two = 0.75 * start + 0.25 * (one + scale(d1, sizes))
                conservative_to_primitive(two, s.gamma)
                d2 = self.residual(two)
                new = (1.0 / 3.0) * start + (2.0 / 3.0) * (two + scale(d2, sizes))
        else:
            raise OracleError(f"Unsupported time integration '{ti}'.")
        conservative_to_primitive(new, s.gamma)
        return new, dt, sizes, d0

    def rk4(self, state, h):
        """SSPRK(5,4) (`solver.py:1128-1182`)."""
        def scale(rate, k):
            if np.ndim(h) == 0:
                return (k * h) * rate
            return (k * np.asarray(h))[:, None] * rate

        r = self.residual
        d = r(state)
        s1 = state + scale(d, 0.391752226571890)
        d = r(s1)
        s2 = 0.444370493651235 * state + 0.555629506348765 * s1 + scale(d, 0.368410593050371)
        d = r(s2)
        s3 = 0.620101851488403 * state + 0.379898148511597 * s2 + scale(d, 0.251891774271694)
        d3 = r(s3)
        s4 = 0.178079954393132 * state + 0.821920045606868 * s3 + scale(d3, 0.544974750228521)
        d4 = r(s4)
        return (
            0.517231671970585 * s2
            + 0.096059710526147 * s3
            + scale(d3, 0.063692468666290)
            + 0.386708617503269 * s4
            + scale(d4, 0.226007483236906)
        )

    def solve(self, primitive0, time=0.0):
        """The explicit part of `solve_case`'s loop (`solver.py:437-999`):
        end-time test with the 1e-13 relative slack, local-mode steady test on
        the RMS of an extra residual; returns (conservative, time, steps,
        completed)."""
        s = self.s
        u = primitive_to_conservative(primitive0, s.gamma)
        steps = 0
        completed = False
        local = s.time_step_mode == "Local (steady)"
        for step in range(1, s.max_iterations + 1):
            u, dt, _, _ = self.step(u, time)
            time += dt
            steps = step
            if local:
                r = self.residual(u)
                if float(np.sqrt(np.mean(r**2))) <= s.residual_tolerance:
                    completed = True
                    break
            elif time >= s.end_time * (1.0 - 1e-13):
                time = s.end_time
                completed = True
                break
        return u, time, steps, completed

    def run(self, primitive0, steps, time=0.0):
        """Advance ``steps`` explicit steps from a primitive state."""
        u = primitive_to_conservative(primitive0, self.s.gamma)
        for _ in range(steps):
            u, dt, _, _ = self.step(u, time)
            time += dt
        return u, time


def far_field_states(interior, far_field, normals, gamma):
    """Characteristic far-field ghost states (`solver.py:4350-4409`)."""
    nx, ny = normals[:, 0], normals[:, 1]
    un_i = interior[:, 1] * nx + interior[:, 2] * ny
    un_f = far_field[:, 1] * nx + far_field[:, 2] * ny
    c_i = sound_speed(interior, gamma)
    c_f = sound_speed(far_field, gamma)
    ext = interior.copy()
    sup_in = un_i <= -c_i
    sup_out = un_i >= c_i
    ext[sup_in] = far_field[sup_in]
    sub = ~(sup_in | sup_out)
    if not np.any(sub):
        return ext
    n = normals[sub]
    inside, outside = interior[sub], far_field[sub]
    r_out = un_i[sub] + 2.0 * c_i[sub] / (gamma - 1.0)
    r_in = un_f[sub] - 2.0 * c_f[sub] / (gamma - 1.0)
    vb = 0.5 * (r_out + r_in)
    cb = 0.25 * (gamma - 1.0) * (r_out - r_in)
    if not np.isfinite(cb).all() or np.any(cb <= 0.0):
        raise OracleError("Far-field characteristics produced invalid sound speed.")
    inflow = vb < 0.0
    src = np.where(inflow[:, None], outside, inside)
    entropy = src[:, 3] / src[:, 0] ** gamma
    rho = (cb**2 / (gamma * entropy)) ** (1.0 / (gamma - 1.0))
    p = entropy * rho**gamma
    tx, ty = -n[:, 1], n[:, 0]
    tsrc = np.where(inflow[:, None], outside[:, 1:3], inside[:, 1:3])
    vt = tsrc[:, 0] * tx + tsrc[:, 1] * ty
    vx = vb * n[:, 0] + vt * tx
    vy = vb * n[:, 1] + vt * ty
    states = np.column_stack((rho, vx, vy, p))
    if not np.isfinite(states).all() or np.any(states[:, 0] <= 0.0) or np.any(states[:, 3] <= 0.0):
        raise OracleError("Far-field characteristics produced an invalid state.")
    ext[sub] = states
    return ext
