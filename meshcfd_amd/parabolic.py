"""MUSCL-parabolic stencils and precomputed fit coefficients.

The reference (`src/meshcfd/core/solver.py`) builds a two-ring stencil per cell
(`_parabolic_stencil`, `:2241-2291`), then on *every* residual evaluation
assembles the 5x5 weighted normal equations of the quadratic fit
(x, y, x^2/2, xy, y^2/2), tests full rank with an SVD and solves them
(`_parabolic_reconstruction`, `:2322-2372`). The matrix depends on geometry
only, so here it is factored once at setup: each stencil entry e of cell i
gets the 5-vector P_e = M_i^-1 w_e d_e, scaled so that

    grad_i  = sum_e P_e[0:2] (W_e - W_i),   curvature_i = sum_e P_e[2:5] (W_e - W_i),

and the device evaluates one weighted sum per cell. Results equal the
reference's solve up to rounding (the reference's own LAPACK solve order is
not reproducible either).
"""

from __future__ import annotations

import numpy as np

from .geometry import Geometry


def parabolic_stencil(g: Geometry):
    """Two-ring stencil in the reference's BFS order with its dedup key
    (neighbour, round(x/scale, 10), round(y/scale, 10)); the cell itself at
    zero offset is excluded (`solver.py:2241-2291`). Vectorised."""
    n = g.cell_count
    src = g.gradient_sources
    order = np.argsort(src, kind="stable")
    ptr = np.zeros(n + 1, dtype=np.int64)
    np.cumsum(np.bincount(src, minlength=n), out=ptr[1:])
    a_nb = g.gradient_neighbours[order]
    a_dx = g.gradient_dx[order]
    a_dy = g.gradient_dy[order]
    deg = np.diff(ptr)
    # ring 1 in adjacency order
    r1_src = np.repeat(np.arange(n), deg)
    r1_nb, r1_x, r1_y = a_nb, a_dx, a_dy
    r1_rank = np.arange(len(r1_src)) - ptr[r1_src]
    # ring 2: expand every ring-1 entry by its neighbour's adjacency
    d2 = deg[r1_nb]
    r2_src = np.repeat(r1_src, d2)
    r2_parent = np.repeat(np.arange(len(r1_src)), d2)
    offs = np.arange(d2.sum()) - np.repeat(np.cumsum(d2) - d2, d2)
    idx = np.repeat(ptr[r1_nb], d2) + offs
    r2_nb = a_nb[idx]
    r2_x = np.repeat(r1_x, d2) + a_dx[idx]
    r2_y = np.repeat(r1_y, d2) + a_dy[idx]
    scale = g.domain_scale
    srcs = np.concatenate((r1_src, r2_src))
    nbs = np.concatenate((r1_nb, r2_nb))
    xs = np.concatenate((r1_x, r2_x))
    ys = np.concatenate((r1_y, r2_y))
    # BFS order key: depth, then ring-1 rank (parent rank for ring 2), then position in parent adjacency
    depth = np.concatenate((np.zeros(len(r1_src), np.int64), np.ones(len(r2_src), np.int64)))
    rank1 = np.concatenate((r1_rank, r1_rank[r2_parent]))
    rank2 = np.concatenate((np.zeros(len(r1_src), np.int64), offs))
    bfs = np.lexsort((rank2, rank1, depth, srcs))
    srcs, nbs, xs, ys = srcs[bfs], nbs[bfs], xs[bfs], ys[bfs]
    kx = np.round(xs / scale, 10)
    ky = np.round(ys / scale, 10)
    # drop the (source, 0, 0) key and keep the first occurrence of every key
    self_key = (nbs == srcs) & (kx == 0.0) & (ky == 0.0)
    keep_idx = np.flatnonzero(~self_key)
    rec = np.rec.fromarrays((srcs[keep_idx], nbs[keep_idx], kx[keep_idx], ky[keep_idx]))
    _, first = np.unique(rec, return_index=True)
    first = np.sort(first)
    sel = keep_idx[first]
    sel = sel[np.hypot(xs[sel], ys[sel]) > scale * 1e-12]
    return srcs[sel], nbs[sel], xs[sel], ys[sel]


def fit_coefficients(n: int, sources, dx, dy):
    """Per entry (P0..P4) with gradient = P0:2 . dW and curvature = P2:5 . dW,
    and the per-cell full-rank flag (numpy matrix_rank == 5, `solver.py:2361`)."""
    scales = np.zeros(n)
    np.maximum.at(scales, sources, np.maximum(np.abs(dx), np.abs(dy)))
    scales = np.maximum(scales, 1e-300)
    x = dx / scales[sources]
    y = dy / scales[sources]
    design = np.column_stack((x, y, 0.5 * x**2, x * y, 0.5 * y**2))
    w = 1.0 / np.maximum(x**2 + y**2, 1e-300)
    M = np.zeros((n, 5, 5))
    np.add.at(M, sources, w[:, None, None] * design[:, :, None] * design[:, None, :])
    full = np.linalg.matrix_rank(M) == 5
    P = np.zeros((len(sources), 5))
    if np.any(full):
        Minv = np.zeros((n, 5, 5))
        Minv[full] = np.linalg.inv(M[full])
        P = np.einsum("eab,eb->ea", Minv[sources], w[:, None] * design)
        s = scales[sources]
        P[:, 0:2] /= s[:, None]
        P[:, 2:5] /= (s * s)[:, None]
        P[~full[sources]] = 0.0
    return P, full


def build_parabolic(g: Geometry, new_id: np.ndarray):
    """Device-order CSR of the stencil (neighbour, 5 coefficients) and the
    full-rank flags."""
    n = g.cell_count
    src, nb, dx, dy = parabolic_stencil(g)
    P, full = fit_coefficients(n, src, dx, dy)
    key = new_id[src]
    order = np.argsort(key, kind="stable")
    ptr = np.zeros(n + 1, dtype=np.int64)
    np.cumsum(np.bincount(key, minlength=n), out=ptr[1:])
    full_dev = np.empty(n, dtype=np.int32)
    full_dev[new_id] = full.astype(np.int32)
    return {
        "ptr": ptr.astype(np.int32),
        "nb": new_id[nb[order]].astype(np.int32),
        "coef": np.ascontiguousarray(P[order]),
        "full": full_dev,
        "rev": reverse_entries(src[order], nb[order], dx[order], dy[order], g.domain_scale),
    }


def reverse_entries(src, nb, dx, dy, scale) -> np.ndarray:
    """For every stencil entry (i -> j at offset d) the entry (j -> i at
    offset -d); the two-ring stencils are symmetric (paths reversed), which
    this checks. Entries are matched on the dedup key of the stencil."""
    kx = np.round(dx / scale, 10)
    ky = np.round(dy / scale, 10)
    keys = np.rec.fromarrays((src, nb, kx + 0.0, ky + 0.0))
    back = np.rec.fromarrays((nb, src, -kx + 0.0, -ky + 0.0))
    order = np.argsort(keys, kind="stable")
    pos = np.searchsorted(keys[order], back)
    pos = np.minimum(pos, len(order) - 1)
    rev = order[pos]
    if len(rev) and not np.array_equal(keys[rev], back):
        raise ValueError("The two-ring stencil is not symmetric; the parabolic adjoint needs it to be.")
    return rev.astype(np.int32)
