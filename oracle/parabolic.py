"""ORACLE (test infrastructure only): MUSCL-parabolic reconstruction.

* `parabolic_stencil`        <- `_parabolic_stencil` (`solver.py:2241-2291`),
  a literal BFS with the reference's dedup key;
* `parabolic_reconstruction` <- `_parabolic_reconstruction` (`solver.py:2322-2415`):
  normal equations assembled and solved per call, SVD full-rank test,
  two-ring min/max, quadratic face increments, same limiter family;
* `reconstruct`              <- `_reconstruct` with curvature (`solver.py:2418-2434`).
"""

from __future__ import annotations

import math

import numpy as np

from .kernels import limiter_function, linear_gradients_and_factors


def parabolic_stencil(cell_count, sources, neighbours, dx, dy, domain_scale):
    adjacency = [[] for _ in range(cell_count)]
    for s, n, x, y in zip(sources, neighbours, dx, dy):
        adjacency[int(s)].append((int(n), float(x), float(y)))
    out_s, out_n, out_x, out_y = [], [], [], []
    for source in range(cell_count):
        frontier = [(source, 0.0, 0.0)]
        seen = {(source, 0, 0)}
        for _ in range(2):
            nxt = []
            for cur, ox, oy in frontier:
                for nb, ex, ey in adjacency[cur]:
                    px, py = ox + ex, oy + ey
                    key = (nb, round(px / domain_scale, 10), round(py / domain_scale, 10))
                    if key in seen:
                        continue
                    seen.add(key)
                    nxt.append((nb, px, py))
                    if math.hypot(px, py) > domain_scale * 1e-12:
                        out_s.append(source)
                        out_n.append(nb)
                        out_x.append(px)
                        out_y.append(py)
            frontier = nxt
    return (np.asarray(out_s, dtype=int), np.asarray(out_n, dtype=int), np.asarray(out_x), np.asarray(out_y))


def parabolic_reconstruction(primitive, g, stencil, limiter_code):
    """Returns (gradients, curvature, factors)."""
    gradients, factors = linear_gradients_and_factors(
        primitive, g.gradient_sources, g.gradient_neighbours, g.gradient_coefficients_x,
        g.gradient_coefficients_y, g.limiter_cells, g.limiter_dx, g.limiter_dy, limiter_code,
    )
    n = len(primitive)
    curvature = np.zeros((n, 4, 3))
    sources, nbrs, dx, dy = stencil
    if len(sources) == 0:
        return gradients, curvature, factors
    scales = np.zeros(n)
    np.maximum.at(scales, sources, np.maximum(np.abs(dx), np.abs(dy)))
    scales = np.maximum(scales, 1e-300)
    x = dx / scales[sources]
    y = dy / scales[sources]
    design = np.column_stack((x, y, 0.5 * x**2, x * y, 0.5 * y**2))
    w = 1.0 / np.maximum(x**2 + y**2, 1e-300)
    M = np.zeros((n, 5, 5))
    rhs = np.zeros((n, 5, 4))
    np.add.at(M, sources, w[:, None, None] * design[:, :, None] * design[:, None, :])
    diff = primitive[nbrs] - primitive[sources]
    np.add.at(rhs, sources, w[:, None, None] * design[:, :, None] * diff[:, None, :])
    full = np.linalg.matrix_rank(M) == 5
    if not np.any(full):
        return gradients, curvature, factors
    coef = np.linalg.solve(M[full], rhs[full])
    s = scales[full, None]
    gradients[full, :, 0] = coef[:, 0, :] / s
    gradients[full, :, 1] = coef[:, 1, :] / s
    curvature[full, :, 0] = coef[:, 2, :] / s**2
    curvature[full, :, 1] = coef[:, 3, :] / s**2
    curvature[full, :, 2] = coef[:, 4, :] / s**2
    lo = primitive.copy()
    hi = primitive.copy()
    np.minimum.at(lo, sources, primitive[nbrs])
    np.maximum.at(hi, sources, primitive[nbrs])
    cells = g.limiter_cells
    fx, fy = g.limiter_dx, g.limiter_dy
    inc = (
        gradients[cells, :, 0] * fx[:, None]
        + gradients[cells, :, 1] * fy[:, None]
        + 0.5 * curvature[cells, :, 0] * fx[:, None] ** 2
        + curvature[cells, :, 1] * (fx * fy)[:, None]
        + 0.5 * curvature[cells, :, 2] * fy[:, None] ** 2
    )
    val = primitive[cells]
    tol = 1e-14 * np.maximum(np.abs(val), 1.0)
    with np.errstate(divide="ignore", invalid="ignore"):
        ratio = np.where(inc > tol, (hi[cells] - val) / inc, np.where(inc < -tol, (lo[cells] - val) / inc, 1.0))
    ratio = np.maximum(ratio, 0.0)
    face = limiter_function(ratio, limiter_code)
    pf = np.ones((n, 4))
    np.minimum.at(pf, cells, face)
    factors[full] = pf[full]
    return gradients, curvature, factors


def reconstruct(primitive, gradients, factors, curvature, cell_ids, offsets):
    dx, dy = offsets[:, 0], offsets[:, 1]
    inc = gradients[cell_ids, :, 0] * dx[:, None] + gradients[cell_ids, :, 1] * dy[:, None]
    inc = inc + (
        0.5 * curvature[cell_ids, :, 0] * dx[:, None] ** 2
        + curvature[cell_ids, :, 1] * (dx * dy)[:, None]
        + 0.5 * curvature[cell_ids, :, 2] * dy[:, None] ** 2
    )
    return primitive[cell_ids] + factors[cell_ids] * inc
