"""ORACLE (test infrastructure only): the implicit-Euler chain, restated.

Follows `src/meshcfd/core/solver.py`:

* `lusgs_matrix`      <- `_implicit_preconditioner` assembly (`:1185-1301`):
  scalar first-order Rusanov spectral-radius matrix S (diagonal counts every
  face with neighbour < 0 again, periodic faces included);
* `lusgs_apply`       <- the LU-SGS operator (`:1302-1323`):
  M^-1 v = U^-1 (D (L^-1 v)), L = tril(S), U = triu(S) (both with the
  diagonal), D = max(diag S, 1e-14), natural cell order, per variable;
* `jacobian_operator` <- `_matrix_free_implicit_operator` (`:1532-1567`);
* `solve_gmres`       <- `_solve_gmres` (`:1489-1529`) over SciPy's GMRES;
* `implicit_euler_step` <- `_implicit_euler_step` / `_implicit_euler_solve_once`
  (`:1337-1486`), inviscid 4-variable state.
"""

from __future__ import annotations

import math
from dataclasses import replace

import numpy as np
import scipy.sparse as sp
from scipy.sparse.linalg import LinearOperator, gmres, spsolve_triangular

from .euler import conservative_to_primitive, sound_speed
from .solver import Oracle, OracleError


def lusgs_matrix(primitive, g, gamma, time_step):
    """Scalar matrix S as CSR and its diagonal (`solver.py:1185-1301`)."""
    n = len(primitive)
    c = sound_speed(primitive, gamma)
    h = np.full(n, time_step, dtype=float) if np.ndim(time_step) == 0 else np.asarray(time_step, dtype=float)
    diag = np.ones(n)
    rows, cols, vals = [np.arange(n)], [np.arange(n)], [np.ones(n)]

    def pair(fo, fn, oo, on, length):
        no = g.face_normals[fo]
        nn = g.face_normals[fn]
        ro = np.abs(primitive[oo, 1] * no[:, 0] + primitive[oo, 2] * no[:, 1]) + c[oo]
        rn = np.abs(primitive[on, 1] * nn[:, 0] + primitive[on, 2] * nn[:, 1]) + c[on]
        coef = 0.5 * np.maximum(ro, rn) * length
        co = h[oo] * coef
        cn = h[on] * coef
        np.add.at(diag, oo, co / g.areas[oo])
        np.add.at(diag, on, cn / g.areas[on])
        rows.extend((oo, on))
        cols.extend((on, oo))
        vals.extend((-co / g.areas[oo], -cn / g.areas[on]))

    inner = g.interior_faces
    if len(inner):
        pair(inner, inner, g.owners[inner], g.neighbours[inner], g.face_lengths[inner])
    if len(g.periodic_faces_a):
        fa, fb = g.periodic_faces_a, g.periodic_faces_b
        pair(fa, fb, g.owners[fa], g.owners[fb], g.face_lengths[fa])
    bnd = np.flatnonzero(g.neighbours < 0)
    if len(bnd):
        o = g.owners[bnd]
        n_ = g.face_normals[bnd]
        r = np.abs(primitive[o, 1] * n_[:, 0] + primitive[o, 2] * n_[:, 1]) + c[o]
        np.add.at(diag, o, h[o] * r * g.face_lengths[bnd] / g.areas[o])
    vals[0] = diag.copy()
    S = sp.coo_matrix(
        (np.concatenate(vals), (np.concatenate(rows), np.concatenate(cols))), shape=(n, n)
    ).tocsr()
    return S, diag


def lusgs_apply(S, vector, variables=4):
    """M^-1 v with (S kron I) split into lower/upper triangles (`solver.py:1302-1323`)."""
    block = sp.kron(S, sp.eye(variables, format="csc"), format="csc")
    lower = sp.tril(block, format="csr")
    upper = sp.triu(block, format="csr")
    d = np.maximum(np.repeat(S.diagonal(), variables), 1e-14)
    forward = spsolve_triangular(lower, vector, lower=True)
    return spsolve_triangular(upper, d * forward, lower=False)


def jacobian_operator(oracle1, state, rate, scale_grid, time_step):
    """v -> v - dt*J v / sigma with a one-sided difference of the first-order
    residual (`solver.py:1532-1567`)."""
    shape = state.shape
    base = float(np.linalg.norm(state / scale_grid))
    fd = math.sqrt(np.finfo(float).eps)

    def mv(v):
        d = v.reshape(shape)
        dn = float(np.linalg.norm(d))
        if dn == 0.0:
            return v.copy()
        eps = fd * max(1.0, base) / dn
        pr = oracle1.residual(state + eps * scale_grid * d)
        jv = (pr - rate) / eps
        step = time_step * jv if np.ndim(time_step) == 0 else np.asarray(time_step)[:, None] * jv
        return (d - step / scale_grid).ravel()

    return LinearOperator((state.size, state.size), matvec=mv, dtype=float)


def solve_gmres(op, M, rhs, tol, maxit, m):
    x = np.zeros_like(rhs)
    used = 0
    while used < maxit:
        count = [0]

        def cb(_):
            count[0] += 1

        x, info = gmres(op, rhs, x0=x, M=M, restart=min(m, maxit - used), maxiter=1, rtol=tol, atol=0.0,
                        callback=cb, callback_type="pr_norm")
        used += count[0]
        if info == 0:
            return x
        if info < 0:
            raise OracleError(f"Implicit Euler GMRES broke down (status {info}).")
        if count[0] == 0:
            break
    raise OracleError(f"Implicit Euler GMRES did not converge within {maxit} iterations (relative tolerance {tol:g}).")


def variable_scales(state, primitive, gamma):
    """`solver.py:1393-1413` for the 4-variable state."""
    rho_scale = max(float(np.max(primitive[:, 0])), 1e-12)
    speed = max(float(np.max(np.abs(primitive[:, 1:3]))), float(np.max(sound_speed(primitive, gamma))), 1e-12)
    scales = np.maximum(np.max(np.abs(state), axis=0), 1e-12)
    scales[0] = max(scales[0], rho_scale)
    scales[1:3] = np.maximum(scales[1:3], rho_scale * speed)
    scales[3] = max(scales[3], rho_scale * speed**2, float(np.max(primitive[:, 3])) / (gamma - 1.0))
    return scales


def solve_once(g, settings, initial, time_step):
    """Newton + GMRES + line search (`solver.py:1371-1486`)."""
    gamma = settings.gamma
    full = Oracle(g, settings)
    first = Oracle(g, replace(settings, reconstruction="First order"))
    state = initial.copy()
    primitive = conservative_to_primitive(state, gamma)
    sc = variable_scales(state, primitive, gamma)[None, :]

    def evaluate(candidate):
        rate = full.residual(candidate)
        step = time_step * rate if np.ndim(time_step) == 0 else np.asarray(time_step)[:, None] * rate
        return candidate - initial - step

    for _ in range(12):
        eq = evaluate(state)
        scaled = eq / sc
        norm = float(np.max(np.abs(scaled)))
        upd = float(np.max(np.abs((state - initial) / sc)))
        if norm <= 1e-9 * max(1.0, upd):
            return state
        primitive = conservative_to_primitive(state, gamma)
        S, _ = lusgs_matrix(primitive, g, gamma, time_step)
        M = LinearOperator((state.size, state.size), matvec=lambda v: lusgs_apply(S, v), dtype=float)
        rate1 = first.residual(state)
        op = jacobian_operator(first, state, rate1, sc, time_step)
        corr = solve_gmres(op, M, -scaled.ravel(), settings.gmres_tolerance, settings.gmres_max_iterations,
                           settings.gmres_krylov_dimension)
        damping = 1.0
        accepted = False
        for _ in range(8):
            cand = state + damping * corr.reshape(state.shape) * sc
            try:
                cn = float(np.max(np.abs(evaluate(cand) / sc)))
            except (OracleError, ValueError, FloatingPointError):
                cn = math.inf
            if cn < norm:
                state = cand
                accepted = True
                break
            damping *= 0.5
        if not accepted:
            raise OracleError("Implicit Euler Newton iteration could not reduce its residual.")
    raise OracleError("Implicit Euler Newton iteration did not converge.")


def implicit_euler_step(g, settings, initial, time_step, allow_reduction=True):
    """dt-halving wrapper (`solver.py:1337-1368`); returns (state, dt)."""
    h = float(time_step) if np.ndim(time_step) == 0 else np.asarray(time_step, dtype=float).copy()
    last = None
    for _ in range(12 if allow_reduction else 1):
        try:
            return solve_once(g, settings, initial, h), h
        except OracleError as error:
            last = error
            h = h * 0.5
    if not allow_reduction:
        raise OracleError(f"Implicit Euler could not converge at the configured fixed time step: {last}")
    raise OracleError(f"Implicit Euler could not converge after reducing the time step 12 times; last failure: {last}")


def lusgs_apply_transpose(S, vector, variables=4):
    """M^T v = L^-T (D U^-T v) for the M of `lusgs_apply` (test reference of
    the adjoint preconditioner; no counterpart in the reference)."""
    block = sp.kron(S, sp.eye(variables, format="csc"), format="csc")
    lower = sp.tril(block, format="csr")
    upper = sp.triu(block, format="csr")
    d = np.maximum(np.repeat(S.diagonal(), variables), 1e-14)
    first = spsolve_triangular(upper.T.tocsr(), vector, lower=True)
    return spsolve_triangular(lower.T.tocsr(), d * first, lower=False)
