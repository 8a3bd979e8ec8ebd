"""ORACLE (test infrastructure only): the reference's compiled kernels, restated.

`src/meshcfd/core/_kernels.py` holds numba loops; this module restates the
three on the residual path in NumPy with the same per-element arithmetic and
the same summation order (``np.add.at`` is unbuffered and sequential in entry
order, exactly like the numba loops):

* `limited_gradient_factors` (`_kernels.py:9-67`)
* `reconstruct_hllc_flux` (`_kernels.py:70-160`)
* `linear_gradients_and_factors` (`_kernels.py:239-273`)

`oracle/mcfd_oracle.c` holds single-threaded C stand-ins of the same loops,
used for CPU-baseline timing (the reference runs them compiled).
"""

from __future__ import annotations

import numpy as np

from .euler import hllc_flux


def limiter_function(ratio: np.ndarray, code: int) -> np.ndarray:
    """phi(r) for codes 0/1 (Barth-Jespersen / Minmod), 2 (Van Leer), 3
    ("Venkatakrishnan", a function of r only) — `_kernels.py:54-63`."""
    if code <= 1:
        return np.minimum(ratio, 1.0)
    if code == 2:
        with np.errstate(divide="ignore", invalid="ignore"):
            return np.where(ratio > 0.0, np.minimum(1.0, 2.0 * ratio / (1.0 + ratio)), 0.0)
    eps = 1e-12
    return np.minimum((ratio * ratio + 2.0 * ratio + eps) / (ratio * ratio + ratio + 2.0 + eps), 1.0)


def limited_gradient_factors(
    primitive, gradients, sources, neighbours, limiter_cells, limiter_dx, limiter_dy, code
):
    """Per-cell min/max over face neighbours, then phi = min over the cell's
    face-midpoint increments (`_kernels.py:9-67`)."""
    local_min = primitive.copy()
    local_max = primitive.copy()
    np.minimum.at(local_min, sources, primitive[neighbours])
    np.maximum.at(local_max, sources, primitive[neighbours])
    cells = limiter_cells
    value = primitive[cells]
    inc = gradients[cells, :, 0] * limiter_dx[:, None] + gradients[cells, :, 1] * limiter_dy[:, None]
    tol = 1e-14 * np.maximum(np.abs(value), 1.0)
    ratio = np.ones_like(inc)
    pos = inc > tol
    neg = inc < -tol
    with np.errstate(divide="ignore", invalid="ignore"):
        up = (local_max[cells] - value) / inc
        lo = (local_min[cells] - value) / inc
    ratio = np.where(pos, up, np.where(neg, lo, ratio))
    ratio = np.maximum(ratio, 0.0)
    face_factor = limiter_function(ratio, code)
    factors = np.ones_like(primitive)
    np.minimum.at(factors, cells, face_factor)
    return factors


def linear_gradients(primitive, sources, neighbours, coefficients_x, coefficients_y):
    """grad W_i = sum_e c_e (W_nb - W_i), accumulated in entry order
    (`_kernels.py:252-262`)."""
    diff = primitive[neighbours] - primitive[sources]
    grads = np.zeros((len(primitive), primitive.shape[1], 2))
    np.add.at(grads[:, :, 0], sources, coefficients_x[:, None] * diff)
    np.add.at(grads[:, :, 1], sources, coefficients_y[:, None] * diff)
    return grads


def linear_gradients_and_factors(
    primitive, sources, neighbours, coefficients_x, coefficients_y,
    limiter_cells, limiter_dx, limiter_dy, code,
):
    """`_kernels.py:239-273`."""
    grads = linear_gradients(primitive, sources, neighbours, coefficients_x, coefficients_y)
    factors = limited_gradient_factors(
        primitive, grads, sources, neighbours, limiter_cells, limiter_dx, limiter_dy, code
    )
    return grads, factors


def reconstruct(primitive, gradients, factors, cell_ids, offsets):
    """W_f = W_c + phi*(g_x*dx + g_y*dy) (`solver.py:2418-2434`, linear part;
    `_kernels.py:90-98` computes the same)."""
    inc = gradients[cell_ids, :, 0] * offsets[:, 0:1] + gradients[cell_ids, :, 1] * offsets[:, 1:2]
    return primitive[cell_ids] + factors[cell_ids] * inc


def reconstruct_hllc_flux(
    primitive, gradients, factors, owners, neighbours, owner_offsets, neighbour_offsets, normals, gamma
):
    """Fused reconstruction + HLLC for interior faces (`_kernels.py:70-160`)."""
    left = reconstruct(primitive, gradients, factors, owners, owner_offsets)
    right = reconstruct(primitive, gradients, factors, neighbours, neighbour_offsets)
    return hllc_flux(left, right, normals, gamma, fused_order=True)
