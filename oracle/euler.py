"""ORACLE (test infrastructure only): ideal-gas states and numerical fluxes.

CPU restatement of the reference's `src/meshcfd/core/euler.py`, kept in the
reference's operation order so that a device implementation with the same
order (and no FMA contraction) reproduces it bit for bit. Only `tests/`,
`__graft_entry__.smoke()` and `bench.py`'s CPU-baseline leg may import this
package; the product path never does.
"""

from __future__ import annotations

import numpy as np


def primitive_to_conservative(primitive: np.ndarray, gamma: float) -> np.ndarray:
    """`euler.py:8-17`: E = p/(gamma-1) + (0.5*rho)*(u*u + v*v)."""
    rho, u, v, p = primitive.T
    energy = p / (gamma - 1.0) + 0.5 * rho * (u**2 + v**2)
    return np.column_stack((rho, rho * u, rho * v, energy))


def conservative_to_primitive(conservative: np.ndarray, gamma: float) -> np.ndarray:
    """`euler.py:20-40`; raises on non-finite or non-positive rho / p."""
    rho = conservative[:, 0]
    with np.errstate(divide="ignore", invalid="ignore", over="ignore"):
        u = conservative[:, 1] / rho
        v = conservative[:, 2] / rho
        p = (gamma - 1.0) * (
            conservative[:, 3] - 0.5 * (conservative[:, 1] ** 2 + conservative[:, 2] ** 2) / rho
        )
    primitive = np.column_stack((rho, u, v, p))
    if not np.isfinite(primitive).all() or np.any(rho <= 0.0) or np.any(p <= 0.0):
        raise ValueError("Euler state has non-finite or non-positive density or pressure.")
    return primitive


def physical_flux(primitive: np.ndarray, normals: np.ndarray, gamma: float):
    """`euler.py:43-57`: returns (F.n, conservative state)."""
    state = primitive_to_conservative(primitive, gamma)
    rho, u, v, p = primitive.T
    un = u * normals[:, 0] + v * normals[:, 1]
    flux = np.column_stack(
        (
            rho * un,
            rho * u * un + p * normals[:, 0],
            rho * v * un + p * normals[:, 1],
            (state[:, 3] + p) * un,
        )
    )
    return flux, state


def sound_speed(primitive: np.ndarray, gamma: float) -> np.ndarray:
    """`euler.py:76-77`."""
    return np.sqrt(gamma * primitive[:, 3] / primitive[:, 0])


def rusanov_flux(left, right, normals, gamma):
    """`euler.py:80-93`: 0.5*(F_L + F_R - s*(U_R - U_L)), s = max(|u_n|+c)."""
    fl, ul = physical_flux(left, normals, gamma)
    fr, ur = physical_flux(right, normals, gamma)
    unl = left[:, 1] * normals[:, 0] + left[:, 2] * normals[:, 1]
    unr = right[:, 1] * normals[:, 0] + right[:, 2] * normals[:, 1]
    s = np.maximum(np.abs(unl) + sound_speed(left, gamma), np.abs(unr) + sound_speed(right, gamma))
    return 0.5 * (fl + fr - s[:, None] * (ur - ul))


def hllc_flux(left, right, normals, gamma, fused_order: bool = False):
    """HLLC with Davis wave speeds (`euler.py:96-166`).

    ``fused_order`` selects the branch order of the compiled interior-face
    kernel (`_kernels.py:136-159`: S_L>=0, S_R<=0, S*>=0) instead of the
    vectorised one (S_L>=0, S*>=0, S_R>0); the two differ only when
    S_R <= 0 <= S*.
    """
    fl, ul = physical_flux(left, normals, gamma)
    fr, ur = physical_flux(right, normals, gamma)
    rl, rr = left[:, 0], right[:, 0]
    pl, pr = left[:, 3], right[:, 3]
    unl = left[:, 1] * normals[:, 0] + left[:, 2] * normals[:, 1]
    unr = right[:, 1] * normals[:, 0] + right[:, 2] * normals[:, 1]
    cl, cr = sound_speed(left, gamma), sound_speed(right, gamma)
    sl = np.minimum(unl - cl, unr - cr)
    sr = np.maximum(unl + cl, unr + cr)
    den = rl * (sl - unl) - rr * (sr - unr)
    den = np.where(np.abs(den) > 1e-14, den, np.copysign(1e-14, den))
    star = (pr - pl + rl * unl * (sl - unl) - rr * unr * (sr - unr)) / den

    def star_flux(prim, cons, flux, un, s):
        with np.errstate(divide="ignore", invalid="ignore", over="ignore"):
            rho = prim[:, 0]
            factor = rho * (s - un) / (s - star)
            vx = prim[:, 1] + (star - un) * normals[:, 0]
            vy = prim[:, 2] + (star - un) * normals[:, 1]
            energy = factor * (cons[:, 3] / rho + (star - un) * (star + prim[:, 3] / (rho * (s - un))))
            q = np.column_stack((factor, factor * vx, factor * vy, energy))
            return flux + s[:, None] * (q - cons)

    fsl = star_flux(left, ul, fl, unl, sl)
    fsr = star_flux(right, ur, fr, unr, sr)
    if fused_order:
        result = np.where(
            (sl >= 0.0)[:, None],
            fl,
            np.where((sr <= 0.0)[:, None], fr, np.where((star >= 0.0)[:, None], fsl, fsr)),
        )
        return result
    result = np.where(
        (sl >= 0.0)[:, None],
        fl,
        np.where((star >= 0.0)[:, None], fsl, np.where((sr > 0.0)[:, None], fsr, fr)),
    )
    if not np.isfinite(result).all():
        raise ValueError("HLLC flux produced a non-finite value.")
    return result


def roe_flux(left, right, normals, gamma):
    """Roe flux with Harten entropy fix delta = 0.1*a (`euler.py:169-262`).

    Eigenvector columns (wave order -, contact, shear, +) in the (n, t) frame,
    t = (-ny, nx), rotated back to (x, y).
    """
    fl, _ = physical_flux(left, normals, gamma)
    fr, _ = physical_flux(right, normals, gamma)
    nx, ny = normals[:, 0], normals[:, 1]
    sl = np.sqrt(left[:, 0])
    sr = np.sqrt(right[:, 0])
    den = sl + sr
    u = (sl * left[:, 1] + sr * right[:, 1]) / den
    v = (sl * left[:, 2] + sr * right[:, 2]) / den
    h = (
        sl * (left[:, 3] / left[:, 0] + 0.5 * (left[:, 1] ** 2 + left[:, 2] ** 2))
        + sr * (right[:, 3] / right[:, 0] + 0.5 * (right[:, 1] ** 2 + right[:, 2] ** 2))
    ) / den
    un = u * nx + v * ny
    ut = u * -ny + v * nx
    a2 = (gamma - 1.0) * (h - 0.5 * (u**2 + v**2))
    if np.any(a2 <= 0.0) or not np.isfinite(a2).all():
        raise ValueError("Roe average has a non-positive sound speed.")
    a = np.sqrt(a2)
    rho_avg = sl * sr
    unl = left[:, 1] * nx + left[:, 2] * ny
    unr = right[:, 1] * nx + right[:, 2] * ny
    utl = left[:, 1] * -ny + left[:, 2] * nx
    utr = right[:, 1] * -ny + right[:, 2] * nx
    drho = right[:, 0] - left[:, 0]
    dun = unr - unl
    dut = utr - utl
    dp = right[:, 3] - left[:, 3]
    alpha_m = (dp - rho_avg * a * dun) / (2.0 * a2)
    alpha_p = (dp + rho_avg * a * dun) / (2.0 * a2)
    alpha_c = drho - dp / a2
    alpha_s = rho_avg * dut
    lam = np.column_stack((un - a, un, un, un + a))
    fix = 0.1 * a[:, None]
    lam = np.abs(lam)
    lam = np.where(lam < fix, 0.5 * (lam**2 / fix + fix), lam)
    w = lam * np.column_stack((alpha_m, alpha_c, alpha_s, alpha_p))
    ones = np.ones_like(un)
    # R[i, j]: component i of wave j.
    r = np.empty((len(un), 4, 4))
    r[:, :, 0] = np.column_stack((ones, un - a, ut, h - un * a))
    r[:, :, 1] = np.column_stack((ones, un, ut, 0.5 * (un**2 + ut**2)))
    r[:, :, 2] = np.column_stack((0.0 * ones, 0.0 * ones, ones, ut))
    r[:, :, 3] = np.column_stack((ones, un + a, ut, h + un * a))
    local = np.einsum("nij,nj->ni", r, w)
    diss = np.column_stack(
        (local[:, 0], local[:, 1] * nx - local[:, 2] * ny, local[:, 1] * ny + local[:, 2] * nx, local[:, 3])
    )
    result = 0.5 * (fl + fr - diss)
    if not np.isfinite(result).all():
        raise ValueError("Roe flux produced a non-finite value.")
    return result


def numerical_flux(left, right, normals, gamma, scheme: str, fused_order: bool = False):
    """`euler.py:60-73` dispatch."""
    if scheme == "Rusanov":
        return rusanov_flux(left, right, normals, gamma)
    if scheme == "HLLC":
        return hllc_flux(left, right, normals, gamma, fused_order=fused_order)
    if scheme == "Roe":
        return roe_flux(left, right, normals, gamma)
    raise ValueError(f"Unsupported convective flux '{scheme}'.")
