"""ORACLE (test infrastructure only): laminar viscous terms.

Restates, for the laminar (no SST) case:

* `dynamic_viscosity`         <- `_dynamic_viscosity` (`solver.py:2721-2737`);
* `weighted_least_squares`    <- the numba kernel `weighted_least_squares_gradients`
  (`_kernels.py:163-207`), sequential in entry order;
* `viscous_gradients`         <- `_viscous_gradients` (`solver.py:3615-3714`):
  face-neighbour entries, then one ghost entry per non-periodic boundary face;
* `viscous_flux`              <- `_viscous_flux` (`solver.py:3717-3747`);
* `viscous_residual`          <- `_viscous_residual` (`solver.py:3750-4114`);
* `viscous_time_step_sum`     <- the diffusive part of `_stable_time_steps`
  (`solver.py:4246-4332`).

Boundary kinds handled: wall (no-slip or slip; fixed temperature or heat
flux), symmetry, velocity inlet and far field (inlet states), and the rest
(pressure outlet, transmissive) as zero-normal-gradient faces.
"""

from __future__ import annotations

import numpy as np

from .euler import sound_speed  # noqa: F401  (kept for symmetry with solver)


class ViscousError(ValueError):
    pass


def dynamic_viscosity(temperature, s):
    if s.viscosity_model == "Constant":
        return np.full_like(temperature, s.dynamic_viscosity, dtype=float)
    if s.viscosity_model == "Sutherland":
        tr = s.sutherland_reference_temperature
        mu = s.sutherland_reference_viscosity * (temperature / tr) ** 1.5
        mu *= (tr + s.sutherland_temperature) / (temperature + s.sutherland_temperature)
        if not np.isfinite(mu).all() or np.any(mu <= 0.0):
            raise ViscousError("Sutherland law produced an invalid viscosity.")
        return mu
    return np.zeros_like(temperature, dtype=float)


def conductivity_of(mu, s):
    return mu * s.gamma * s.gas_constant / ((s.gamma - 1.0) * s.prandtl_number)


def weighted_least_squares(n, sources, dx, dy, diff):
    comps = diff.shape[1]
    xx = np.zeros(n)
    xy = np.zeros(n)
    yy = np.zeros(n)
    rx = np.zeros((n, comps))
    ry = np.zeros((n, comps))
    for e in range(len(sources)):
        c = sources[e]
        w = 1.0 / max(dx[e] * dx[e] + dy[e] * dy[e], 1e-300)
        wx = w * dx[e]
        wy = w * dy[e]
        xx[c] += wx * dx[e]
        xy[c] += wx * dy[e]
        yy[c] += wy * dy[e]
        for k in range(comps):
            rx[c, k] += wx * diff[e, k]
            ry[c, k] += wy * diff[e, k]
    grads = np.zeros((n, comps, 2))
    for c in range(n):
        det = xx[c] * yy[c] - xy[c] * xy[c]
        if det > 1e-14 * max(xx[c] * yy[c], 1e-300):
            inv = 1.0 / det
            for k in range(comps):
                grads[c, k, 0] = (rx[c, k] * yy[c] - ry[c, k] * xy[c]) * inv
                grads[c, k, 1] = (ry[c, k] * xx[c] - rx[c, k] * xy[c]) * inv
    return grads


def inlet_states(primitive, g, s):
    """Boundary-face -> exterior primitive state for velocity inlets, pressure
    and mass-flow inlets (their exterior states from the owner cells'
    unreconstructed states) and far field (`_viscous_inlet_states`,
    `solver.py:2740-2785`)."""
    from .solver import far_field_states

    states = {int(f): st for f, st in zip(g.velocity_inlet_faces, g.velocity_inlet_states)}
    if len(getattr(g, "pressure_inlet_faces", ())):
        from .inflow import pressure_inlet_states

        faces = g.pressure_inlet_faces
        ext = pressure_inlet_states(primitive[g.owners[faces]], g.face_normals[faces], g.pressure_inlet_params,
                                    g.pressure_inlet_patch, s.gamma, s.gas_constant, g.inflow_patch_names[0])
        states.update({int(f): st for f, st in zip(faces, ext)})
    if len(getattr(g, "mass_flow_inlet_faces", ())):
        from .inflow import mass_flow_inlet_states

        faces = g.mass_flow_inlet_faces
        ext = mass_flow_inlet_states(primitive[g.owners[faces]], g.face_normals[faces], g.face_lengths[faces],
                                     g.mass_flow_inlet_params, g.mass_flow_inlet_patch, s.gamma, s.gas_constant,
                                     g.inflow_patch_names[1])
        states.update({int(f): st for f, st in zip(faces, ext)})
    if len(g.far_field_faces):
        faces = g.far_field_faces
        ext = far_field_states(primitive[g.owners[faces]], g.far_field_states, g.face_normals[faces], s.gamma)
        states.update({int(f): st for f, st in zip(faces, ext)})
    return states


WALL, SYMMETRY = 1, 2


def viscous_gradients(primitive, g, s):
    rho = primitive[:, 0]
    temperature = primitive[:, 3] / (rho * s.gas_constant)
    if not np.isfinite(temperature).all() or np.any(temperature <= 0.0):
        raise ViscousError("Viscous flow requires finite positive temperature.")
    values = np.column_stack((primitive[:, 1:3], temperature))
    src = list(g.gradient_sources)
    dxs = list(g.gradient_dx)
    dys = list(g.gradient_dy)
    diffs = list(values[g.gradient_neighbours] - values[g.gradient_sources])
    inlet = {f: np.array((st[1], st[2], st[3] / (st[0] * s.gas_constant))) for f, st in
             inlet_states(primitive, g, s).items()}
    for f in g.diffusive_boundary_faces:
        f = int(f)
        o = int(g.owners[f])
        n = g.face_normals[f]
        off = g.face_centers[f] - g.centers[o]
        kind = int(g.face_kind_code[f])
        dn = float(np.dot(off, n))
        goff = 2.0 * dn * n if kind in (WALL, SYMMETRY) else 2.0 * off
        ghost = values[o].copy()
        if kind == WALL:
            if g.wall_no_slip[f]:
                ghost[:2] = 2.0 * g.wall_velocities[f] - ghost[:2]
            else:
                ghost[:2] -= 2.0 * float(np.dot(ghost[:2], n)) * n
            if g.wall_thermal[f] == 1:
                ghost[2] = 2.0 * g.wall_thermal_value[f] - values[o, 2]
            else:
                mu = float(dynamic_viscosity(temperature[o:o + 1], s)[0])
                k = conductivity_of(mu, s)
                ghost[2] += 2.0 * dn * g.wall_thermal_value[f] / k
        elif kind == SYMMETRY:
            vel = values[o, :2]
            ghost[:2] -= 2.0 * float(np.dot(vel, n)) * n
        elif f in inlet:
            ghost = 2.0 * inlet[f] - ghost
        src.append(o)
        dxs.append(goff[0])
        dys.append(goff[1])
        diffs.append(ghost - values[o])
    return weighted_least_squares(len(primitive), np.asarray(src, dtype=int), np.asarray(dxs), np.asarray(dys),
                                  np.asarray(diffs).reshape(-1, 3))


def viscous_flux(grads, vel, mu, k, normals):
    du_dx, du_dy = grads[:, 0, 0], grads[:, 0, 1]
    dv_dx, dv_dy = grads[:, 1, 0], grads[:, 1, 1]
    div = du_dx + dv_dy
    txx = mu * (2.0 * du_dx - (2.0 / 3.0) * div)
    tyy = mu * (2.0 * dv_dy - (2.0 / 3.0) * div)
    txy = mu * (du_dy + dv_dx)
    tx = txx * normals[:, 0] + txy * normals[:, 1]
    ty = txy * normals[:, 0] + tyy * normals[:, 1]
    q = k * (grads[:, 2, 0] * normals[:, 0] + grads[:, 2, 1] * normals[:, 1])
    return np.column_stack((np.zeros(len(normals)), tx, ty, vel[:, 0] * tx + vel[:, 1] * ty + q))


def viscous_residual(primitive, g, s, D):
    """D: the interior divergence operator (CSR, -L/A owner, +L/A neighbour)."""
    rho = primitive[:, 0]
    temperature = primitive[:, 3] / (rho * s.gas_constant)
    values = np.column_stack((primitive[:, 1:3], temperature))
    mu = dynamic_viscosity(temperature, s)
    kc = conductivity_of(mu, s)
    grads = viscous_gradients(primitive, g, s)
    deriv = np.zeros((len(primitive), 4))
    inner = g.interior_faces
    if len(inner):
        o, nb = g.owners[inner], g.neighbours[inner]
        flux = viscous_flux(0.5 * (grads[o] + grads[nb]), 0.5 * (primitive[o, 1:3] + primitive[nb, 1:3]),
                            0.5 * (mu[o] + mu[nb]), 0.5 * (kc[o] + kc[nb]), g.interior_normals)
        deriv -= D @ flux
    fa, fb = g.periodic_faces_a, g.periodic_faces_b
    if len(fa):
        oa, ob = g.owners[fa], g.owners[fb]
        flux = viscous_flux(0.5 * (grads[oa] + grads[ob]), 0.5 * (primitive[oa, 1:3] + primitive[ob, 1:3]),
                            0.5 * (mu[oa] + mu[ob]), 0.5 * (kc[oa] + kc[ob]), g.face_normals[fa])
        integ = flux * g.face_lengths[fa, None]
        np.add.at(deriv, oa, integ / g.areas[oa, None])
        np.add.at(deriv, ob, -integ / g.areas[ob, None])
    inl = inlet_states(primitive, g, s)
    for f in g.diffusive_boundary_faces:
        f = int(f)
        o = int(g.owners[f])
        n = g.face_normals[f]
        off = g.face_centers[f] - g.centers[o]
        dn = float(np.dot(off, n))
        if dn <= 1e-14:
            raise ViscousError("Boundary face has invalid cell-center distance.")
        kind = int(g.face_kind_code[f])
        fg = grads[o].copy()
        bv = np.array((primitive[o, 1], primitive[o, 2], temperature[o]))
        if kind == WALL:
            if g.wall_no_slip[f]:
                wv = g.wall_velocities[f]
                for c in (0, 1):
                    nd = (wv[c] - bv[c] - float(np.dot(fg[c], off - dn * n))) / dn
                    fg[c] += (nd - float(np.dot(fg[c], n))) * n
                    bv[c] = wv[c]
            else:
                tang = off - dn * n
                un = float(np.dot(bv[:2], n))
                bv[:2] += fg[:2] @ off
                bv[:2] -= float(np.dot(bv[:2], n)) * n
                gnv = n[0] * fg[0] + n[1] * fg[1]
                nd = -(un + float(np.dot(gnv, tang))) / dn
                desired = np.outer(nd * n, n)
                vg = fg[:2]
                vg += desired - (vg @ n)[:, None] * n
            if g.wall_thermal[f] == 1:
                nd = (g.wall_thermal_value[f] - bv[2] - float(np.dot(fg[2], off - dn * n))) / dn
                bv[2] = g.wall_thermal_value[f]
            else:
                nd = g.wall_thermal_value[f] / kc[o]
                bv[2] += float(np.dot(fg[2], off))
            fg[2] += (nd - float(np.dot(fg[2], n))) * n
        elif kind == SYMMETRY:
            un = float(np.dot(bv[:2], n))
            dv = -un / dn * n
            for c in (0, 1):
                fg[c] += (dv[c] - float(np.dot(fg[c], n))) * n
            fg[2] -= float(np.dot(fg[2], n)) * n
            bv[:2] -= un * n
        else:
            st = inl.get(f)
            if st is not None:
                bv = np.array((st[1], st[2], st[3] / (st[0] * s.gas_constant)))
                for c in range(3):
                    nd = (bv[c] - values[o, c] - float(np.dot(fg[c], off - dn * n))) / dn
                    fg[c] += (nd - float(np.dot(fg[c], n))) * n
            else:
                fg -= np.outer(fg @ n, n)
                bv += fg @ off
        fmu = float(dynamic_viscosity(np.array([bv[2]]), s)[0])
        fk = conductivity_of(fmu, s)
        flux = viscous_flux(fg[None], bv[None, :2], np.array([fmu]), np.array([fk]), n[None])[0]
        if kind == WALL and g.wall_thermal[f] == 2:
            flux[3] = g.wall_thermal_value[f]
        deriv[o] += flux * g.face_lengths[f] / g.areas[o]
    return deriv


def viscous_time_step_sum(primitive, g, s):
    """2 * diffusion sum added to the spectral-radius sum (`solver.py:4246-4332`)."""
    rho = primitive[:, 0]
    temperature = primitive[:, 3] / (rho * s.gas_constant)
    mu = dynamic_viscosity(temperature, s)
    eff = np.maximum((4.0 / 3.0) * mu / rho, mu / (rho * min(s.prandtl_number, 0.9)))
    total = np.zeros(len(primitive))
    inner = g.interior_faces
    if len(inner):
        o, nb = g.owners[inner], g.neighbours[inner]
        dist = np.abs(np.sum((g.centers[nb] - g.centers[o]) * g.face_normals[inner], axis=1))
        contrib = np.maximum(eff[o], eff[nb]) * g.face_lengths[inner] / np.maximum(dist, 1e-14)
        np.add.at(total, o, contrib)
        np.add.at(total, nb, contrib)
    fa, fb = g.periodic_faces_a, g.periodic_faces_b
    if len(fa):
        oa, ob = g.owners[fa], g.owners[fb]
        dist = np.abs(np.sum((g.face_centers[fa] - g.centers[oa]) * g.face_normals[fa], axis=1)) + np.abs(
            np.sum((g.face_centers[fb] - g.centers[ob]) * g.face_normals[fb], axis=1)
        )
        contrib = np.maximum(eff[oa], eff[ob]) * g.face_lengths[fa] / np.maximum(dist, 1e-14)
        np.add.at(total, oa, contrib)
        np.add.at(total, ob, contrib)
    for f in g.diffusive_boundary_faces:
        o = g.owners[f]
        dist = abs(float(np.dot(g.face_centers[f] - g.centers[o], g.face_normals[f])))
        total[o] += eff[o] * g.face_lengths[f] / max(dist, 1e-14)
    return 2.0 * total
