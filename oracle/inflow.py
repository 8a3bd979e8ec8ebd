"""ORACLE (test infrastructure only): pressure-inlet and mass-flow boundaries.

* `pressure_inlet_states`  <- `_pressure_inlet_boundary_states` (`solver.py:4412-4496`):
  per face, 64-step bisection on the inflow Mach number so that the incoming
  state's Riemann invariant equals the interior's outgoing one;
* `mass_flow_inlet_states` <- `_mass_flow_inlet_boundary_states` (`solver.py:4499-4554`):
  per patch, 80-step bisection on one inflow speed so that the patch's mass
  flow (interior pressures, total temperature) equals the target;
* `mass_flow_outlet_states` <- `_mass_flow_outlet_boundary_states` (`solver.py:4557-4577`).

Patches are given as the per-face parameter rows and patch ordinals
`Geometry` stores; each patch is handled by one call of the restated
function, as the reference calls its function once per patch.
"""

from __future__ import annotations

import math

import numpy as np


class InflowError(ValueError):
    pass


def pressure_inlet_states(interior, normals, params, patch, gamma, gas_constant, names=None):
    """One call of the reference function per patch ordinal in ``patch``
    (faces of a patch contiguous, patches ascending)."""
    out = np.empty_like(interior)
    for k in np.unique(patch):
        sel = np.flatnonzero(patch == k)
        name = names[k] if names else "inlet"
        out[sel] = _pressure_inlet_patch(interior[sel], normals[sel], params[sel[0]], gamma, gas_constant, name)
    return out


def _pressure_inlet_patch(interior, normals, params, gamma, gas_constant, name):
    p0, t0 = float(params[0]), float(params[1])
    direction = np.array(params[2:4], dtype=float)
    direction_normal = normals @ direction
    sound_interior = np.sqrt(gamma * interior[:, 3] / interior[:, 0])
    normal_velocity = np.sum(interior[:, 1:3] * normals, axis=1)
    mach_interior = normal_velocity / sound_interior
    if np.any(mach_interior >= 1.0):
        raise InflowError(f"Pressure inlet '{name}' cannot prescribe a supersonic outflow.")
    total_sound = math.sqrt(gamma * gas_constant * t0)
    mach = np.empty(len(interior))
    for f in range(len(interior)):
        if mach_interior[f] <= -1.0:
            raise InflowError(f"Pressure inlet '{name}' needs Mach number to define a supersonic inflow.")
        outgoing = float(normal_velocity[f]) + 2.0 * float(sound_interior[f]) / (gamma - 1.0)
        dn = float(direction_normal[f])

        def invariant_at(m):
            ratio = 1.0 + 0.5 * (gamma - 1.0) * m**2
            return total_sound / math.sqrt(ratio) * (dn * m + 2.0 / (gamma - 1.0))

        if not invariant_at(1.0) <= outgoing <= invariant_at(0.0):
            raise InflowError(
                f"Pressure inlet '{name}' total conditions are incompatible with the interior characteristic state."
            )
        low, high = 0.0, 1.0
        for _ in range(64):
            middle = 0.5 * (low + high)
            if invariant_at(middle) > outgoing:
                low = middle
            else:
                high = middle
        mach[f] = 0.5 * (low + high)
    temperature = t0 / (1.0 + 0.5 * (gamma - 1.0) * mach**2)
    pressure = p0 * (temperature / t0) ** (gamma / (gamma - 1.0))
    density = pressure / (gas_constant * temperature)
    speed = mach * np.sqrt(gamma * gas_constant * temperature)
    states = np.column_stack((density, speed * direction[0], speed * direction[1], pressure))
    if not np.isfinite(states).all() or np.any(states[:, (0, 3)] <= 0.0):
        raise InflowError(f"Pressure inlet '{name}' produced an invalid state.")
    return states


def mass_flow_inlet_states(interior, normals, lengths, params, patch, gamma, gas_constant, names=None):
    """params per face (mass flow, total temperature, direction); one
    bisection per patch ordinal in ``patch``."""
    out = np.empty_like(interior)
    for k in np.unique(patch):
        sel = np.flatnonzero(patch == k)
        name = names[k] if names else "inlet"
        mdot, t0 = float(params[sel[0], 0]), float(params[sel[0], 1])
        direction = np.array(params[sel[0], 2:4], dtype=float)
        inward = -(normals[sel] @ direction)
        pressure = interior[sel, 3]
        if np.any(pressure <= 0.0) or not np.isfinite(pressure).all():
            raise InflowError(f"Mass-flow inlet '{name}' has invalid interior pressure.")
        cp = gamma * gas_constant / (gamma - 1.0)
        vmax = math.sqrt(2.0 * cp * t0) * (1.0 - 1e-12)
        w = lengths[sel]

        def flow_at(speed):
            temperature = t0 - speed**2 / (2.0 * cp)
            density = pressure / (gas_constant * temperature)
            return float(np.sum(density * speed * inward * w))

        low, high = 0.0, vmax
        if flow_at(high) < mdot:
            raise InflowError(f"Mass-flow inlet '{name}' cannot satisfy the configured flow rate.")
        for _ in range(80):
            middle = 0.5 * (low + high)
            if flow_at(middle) < mdot:
                low = middle
            else:
                high = middle
        speed = 0.5 * (low + high)
        temperature = t0 - speed**2 / (2.0 * cp)
        density = pressure / (gas_constant * temperature)
        out[sel] = np.column_stack((density, np.full(len(sel), speed * direction[0]),
                                    np.full(len(sel), speed * direction[1]), pressure))
    return out


def mass_flow_outlet_states(interior, normals, params, name="outlet"):
    """params per face (mass flow, summed face length of the patch)."""
    if np.any(interior[:, 0] <= 0.0) or not np.isfinite(interior[:, 0]).all():
        raise InflowError(f"Mass-flow outlet '{name}' has invalid interior density.")
    states = interior.copy()
    un = params[:, 0] / (params[:, 1] * interior[:, 0])
    tangent = np.column_stack((-normals[:, 1], normals[:, 0]))
    ut = np.sum(interior[:, 1:3] * tangent, axis=1)
    states[:, 1:3] = un[:, None] * normals + ut[:, None] * tangent
    return states
