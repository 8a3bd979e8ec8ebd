"""Seeded mesh generators for the workloads this package is measured on.

Numbering follows the reference's structured-mesh builders
(`benchmarks/common.py:53-113`, `tests/test_solver.py:57-131`): cells are
visited in row-major order, each cell's edges in its CCW vertex order, and the
first cell to meet an edge owns it; the face keeps that cell's CCW node order.
Boundary patches are numbered bottom 0, right 1, top 2, left 3.

All generators are vectorised (NumPy) so that 1M-4M cell meshes build in
seconds; the reference builds its meshes with Python loops.
"""

from __future__ import annotations

import math

import numpy as np

from .model import (
    BoundaryCondition,
    BoundaryCurve,
    Case,
    FiniteVolumeMesh,
    MeshPatch,
    NumericalSettings,
    PeriodicBoundaryPair,
)


def _faces_from_cells(cells: np.ndarray, sizes: np.ndarray):
    """Derive faces, owners and neighbours from CCW cell vertex lists.

    Edges are enumerated cell by cell, edge by edge; the first occurrence of an
    undirected edge creates the face (so face ids grow in first-visit order).
    """
    cell_count, width = cells.shape
    slots = np.arange(width)
    valid = slots[None, :] < sizes[:, None]
    nxt = np.where(slots[None, :] + 1 < sizes[:, None], slots[None, :] + 1, 0)
    start = cells[valid]
    end = np.take_along_axis(cells, nxt, axis=1)[valid]
    edge_cell = np.broadcast_to(np.arange(cell_count)[:, None], cells.shape)[valid]
    lo = np.minimum(start, end)
    hi = np.maximum(start, end)
    node_count = int(cells.max()) + 1
    key = lo.astype(np.int64) * node_count + hi
    order = np.argsort(key, kind="stable")
    sorted_key = key[order]
    first = np.ones(len(order), dtype=bool)
    first[1:] = sorted_key[1:] != sorted_key[:-1]
    group = np.cumsum(first) - 1  # group id per sorted edge
    first_pos = order[first]  # edge index of first occurrence, per group
    counts = np.diff(np.append(np.flatnonzero(first), len(order)))
    if np.any(counts > 2):
        raise ValueError("An edge is shared by more than two cells.")
    # Faces numbered by first visit.
    face_rank = np.empty(len(first_pos), dtype=np.int64)
    face_rank[np.argsort(first_pos, kind="stable")] = np.arange(len(first_pos))
    second = np.full(len(first_pos), -1, dtype=np.int64)
    has_second = counts == 2
    second_edge = order[np.flatnonzero(first)[has_second] + 1]
    second[has_second] = edge_cell[second_edge]
    face_count = len(first_pos)
    faces = np.empty((face_count, 2), dtype=np.int64)
    owners = np.empty(face_count, dtype=np.int64)
    neighbours = np.empty(face_count, dtype=np.int64)
    faces[face_rank, 0] = start[first_pos]
    faces[face_rank, 1] = end[first_pos]
    owners[face_rank] = edge_cell[first_pos]
    neighbours[face_rank] = second
    del group
    return faces, owners, neighbours


def _grid_mesh(
    nx: int,
    ny: int,
    x0: float,
    x1: float,
    y0: float,
    y1: float,
    jitter: float,
    rng: np.random.Generator | None,
    split_fraction: float = 0.0,
):
    hx = (x1 - x0) / nx
    hy = (y1 - y0) / ny
    i = np.arange(nx + 1)
    j = np.arange(ny + 1)
    xs = x0 + hx * i
    ys = y0 + hy * j
    xs[-1] = x1
    ys[-1] = y1
    px = np.broadcast_to(xs[None, :], (ny + 1, nx + 1)).copy()
    py = np.broadcast_to(ys[:, None], (ny + 1, nx + 1)).copy()
    if jitter > 0.0 and nx > 1 and ny > 1:
        offsets = rng.uniform(-jitter, jitter, size=(ny - 1, nx - 1, 2))
        px[1:-1, 1:-1] += offsets[..., 0] * hx
        py[1:-1, 1:-1] += offsets[..., 1] * hy
    points = np.column_stack((px.ravel(), py.ravel()))

    jj, ii = np.meshgrid(np.arange(ny), np.arange(nx), indexing="ij")
    bl = (jj * (nx + 1) + ii).ravel()
    br = bl + 1
    tr = br + nx + 1
    tl = bl + nx + 1
    quad_count = nx * ny
    if split_fraction > 0.0:
        split_count = int(round(split_fraction * quad_count))
        chosen = rng.choice(quad_count, size=split_count, replace=False)
        diagonal = rng.integers(0, 2, size=quad_count)
        split = np.zeros(quad_count, dtype=bool)
        split[chosen] = True
    else:
        split = np.zeros(quad_count, dtype=bool)
        diagonal = np.zeros(quad_count, dtype=np.int64)
    per_quad = np.where(split, 2, 1)
    first_cell = np.concatenate(([0], np.cumsum(per_quad)[:-1]))
    cell_count = int(per_quad.sum())
    cells = np.full((cell_count, 4), -1, dtype=np.int64)
    sizes = np.empty(cell_count, dtype=np.int64)
    q = ~split
    cells[first_cell[q]] = np.column_stack((bl[q], br[q], tr[q], tl[q]))
    sizes[first_cell[q]] = 4
    d0 = split & (diagonal == 0)  # diagonal bl-tr
    c = first_cell[d0]
    cells[c, :3] = np.column_stack((bl[d0], br[d0], tr[d0]))
    cells[c + 1, :3] = np.column_stack((bl[d0], tr[d0], tl[d0]))
    d1 = split & (diagonal == 1)  # diagonal br-tl
    c = first_cell[d1]
    cells[c, :3] = np.column_stack((bl[d1], br[d1], tl[d1]))
    cells[c + 1, :3] = np.column_stack((br[d1], tr[d1], tl[d1]))
    sizes[first_cell[split]] = 3
    sizes[first_cell[split] + 1] = 3

    faces, owners, neighbours = _faces_from_cells(cells, sizes)
    node_i = np.tile(i, ny + 1)
    node_j = np.repeat(j, nx + 1)
    a, b = faces[:, 0], faces[:, 1]
    boundary = neighbours < 0
    patch_ids = np.full(len(faces), -1, dtype=np.int64)
    patch_ids[boundary & (node_j[a] == 0) & (node_j[b] == 0)] = 0
    patch_ids[boundary & (node_i[a] == nx) & (node_i[b] == nx)] = 1
    patch_ids[boundary & (node_j[a] == ny) & (node_j[b] == ny)] = 2
    patch_ids[boundary & (node_i[a] == 0) & (node_i[b] == 0)] = 3
    if np.any(patch_ids[boundary] < 0):
        raise ValueError("Boundary face not on the rectangle's sides.")
    patches = tuple(
        MeshPatch(pid, name, np.flatnonzero(patch_ids == pid))
        for pid, name in enumerate(("bottom", "right", "top", "left"))
    )
    return FiniteVolumeMesh(
        points=points,
        cells=cells,
        cell_sizes=sizes,
        faces=faces,
        face_owners=owners,
        face_neighbours=neighbours,
        face_patch_ids=patch_ids,
        patches=patches,
    )


def rectangle_mesh(
    nx: int,
    ny: int,
    bounds: tuple[float, float, float, float] = (0.0, 1.0, 0.0, 1.0),
    jitter: float = 0.0,
    seed: int | None = None,
    split_fraction: float = 0.0,
) -> FiniteVolumeMesh:
    """Quad grid on a rectangle, interior nodes jittered by U(-jitter*h, jitter*h).

    With ``split_fraction`` > 0 that fraction of quads (sampled without
    replacement) is split into two CCW triangles along a random diagonal; the
    two triangles take consecutive cell ids in the quad's place. The jitter is
    drawn before the split choice from the same generator.
    """
    rng = np.random.default_rng(seed) if (jitter > 0.0 or split_fraction > 0.0) else None
    x0, x1, y0, y1 = bounds
    return _grid_mesh(nx, ny, x0, x1, y0, y1, jitter, rng, split_fraction)


def cell_centroids(mesh: FiniteVolumeMesh) -> np.ndarray:
    """Shoelace centroids in the reference's summation order
    (`initial_conditions.py:56-75`, identical to `solver.py:1655-1672`)."""
    pts = mesh.points
    cells, sizes = mesh.cells, mesh.cell_sizes
    twice = np.zeros(len(cells))
    xs = np.zeros(len(cells))
    ys = np.zeros(len(cells))
    for k in range(cells.shape[1]):
        active = sizes > k
        nk = np.where(k + 1 < sizes, k + 1, 0)
        a = cells[:, k]
        b = cells[np.arange(len(cells)), nk]
        x0 = pts[np.where(active, a, 0), 0]
        y0 = pts[np.where(active, a, 0), 1]
        x1 = pts[np.where(active, b, 0), 0]
        y1 = pts[np.where(active, b, 0), 1]
        cross = x0 * y1 - x1 * y0
        twice = np.where(active, twice + cross, twice)
        xs = np.where(active, xs + (x0 + x1) * cross, xs)
        ys = np.where(active, ys + (y0 + y1) * cross, ys)
    return np.column_stack((xs / (3.0 * twice), ys / (3.0 * twice)))


# ---------------------------------------------------------------------------
# Initial states
# ---------------------------------------------------------------------------


def isentropic_vortex(
    centers: np.ndarray,
    strength: float = 5.0,
    gamma: float = 1.4,
    free_stream=(1.0, 1.0, 1.0, 1.0),
    center=(0.0, 0.0),
) -> np.ndarray:
    """Shu's isentropic vortex sampled at cell centroids -> (C,4) primitive.

    Perturbations: du = -b/(2pi) e^((1-r^2)/2) y, dv = b/(2pi) e^((1-r^2)/2) x,
    dT = -(gamma-1) b^2/(8 gamma pi^2) e^(1-r^2), with T = p/rho and the free
    stream rho = p = 1 so that rho = T^(1/(gamma-1)), p = rho^gamma.
    """
    rho_inf, u_inf, v_inf, p_inf = free_stream
    x = centers[:, 0] - center[0]
    y = centers[:, 1] - center[1]
    r2 = x * x + y * y
    bump = np.exp(0.5 * (1.0 - r2))
    du = -strength / (2.0 * math.pi) * bump * y
    dv = strength / (2.0 * math.pi) * bump * x
    t_inf = p_inf / rho_inf
    dt = -(gamma - 1.0) * strength**2 / (8.0 * gamma * math.pi**2) * bump * bump
    temperature = t_inf + dt
    entropy = p_inf / rho_inf**gamma
    rho = (temperature / (entropy)) ** (1.0 / (gamma - 1.0))
    p = rho * temperature
    return np.column_stack((rho, u_inf + du, v_inf + dv, p))


FOUR_QUADRANT_STATES = (
    (1.5, 0.0, 0.0, 1.5),  # x > xc, y > yc
    (0.5323, 1.206, 0.0, 0.3),  # x < xc, y > yc
    (0.138, 1.206, 1.206, 0.029),  # x < xc, y < yc
    (0.5323, 0.0, 1.206, 0.3),  # x > xc, y < yc
)


def four_quadrant(centers: np.ndarray, split=(0.5, 0.5)) -> np.ndarray:
    x = centers[:, 0]
    y = centers[:, 1]
    right = x > split[0]
    top = y > split[1]
    quadrant = np.where(top, np.where(right, 0, 1), np.where(right, 3, 2))
    return np.asarray(FOUR_QUADRANT_STATES, dtype=float)[quadrant]


# ---------------------------------------------------------------------------
# Workload cases
# ---------------------------------------------------------------------------


def periodic_case(mesh: FiniteVolumeMesh, numerics: NumericalSettings) -> Case:
    """All four sides periodic, paired right<->left (x) and bottom<->top (y),
    as the reference's `make_case` does (`benchmarks/common.py:116-131`)."""
    x = mesh.points[:, 0]
    y = mesh.points[:, 1]
    bcs = [
        BoundaryCondition(p.id, p.name, kind="Periodic") for p in mesh.patches
    ]
    pairs = [
        PeriodicBoundaryPair(1, 3, "x", float(x.max() - x.min())),
        PeriodicBoundaryPair(0, 2, "y", float(y.max() - y.min())),
    ]
    return Case(mesh=mesh, boundary_conditions=bcs, periodic_pairs=pairs, numerics=numerics)


def vortex_workload(n: int = 64, steps: int = 100):
    """Isentropic vortex (strength 5) on the periodic square [-5,5]^2, n x n
    quads, interior nodes jittered by U(-0.2h, 0.2h) from seed 0; Rusanov,
    first order, explicit Euler, CFL 0.5."""
    mesh = rectangle_mesh(n, n, (-5.0, 5.0, -5.0, 5.0), jitter=0.2, seed=0)
    numerics = NumericalSettings(
        reconstruction="First order",
        flux="Rusanov",
        time_integration="Explicit Euler",
        time_step_mode="Global adaptive",
        cfl=0.5,
        max_iterations=steps,
        end_time=1.0e30,
        residual_tolerance=1.0e-300,
        convergence_history_interval=max(steps + 1, 10),
    )
    case = periodic_case(mesh, numerics)
    primitive = isentropic_vortex(cell_centroids(mesh))
    return case, primitive


def riemann_workload(n: int = 1000, steps: int = 200, split_fraction: float = 0.3):
    """Four-quadrant Riemann problem on [0,1]^2: n x n quads, interior jitter
    U(-0.15h, 0.15h) and 30% of quads split into triangles, both from seed 1;
    HLLC + MUSCL linear (Barth-Jespersen), SSP-RK3, CFL 0.8, transmissive
    boundaries; quadrant split point (0.5, 0.5)."""
    mesh = rectangle_mesh(
        n, n, (0.0, 1.0, 0.0, 1.0), jitter=0.15, seed=1, split_fraction=split_fraction
    )
    numerics = NumericalSettings(
        reconstruction="MUSCL linear",
        limiter="Barth-Jespersen",
        flux="HLLC",
        time_integration="Explicit RK3",
        time_step_mode="Global adaptive",
        cfl=0.8,
        max_iterations=steps,
        end_time=1.0e30,
        residual_tolerance=1.0e-300,
        convergence_history_interval=max(steps + 1, 10),
    )
    bcs = [BoundaryCondition(p.id, p.name, kind="Transmissive") for p in mesh.patches]
    case = Case(mesh=mesh, boundary_conditions=bcs, numerics=numerics)
    primitive = four_quadrant(cell_centroids(mesh))
    return case, primitive


def geometric_radii(n_r: int, r_inner: float, r_outer: float, first_spacing: float) -> np.ndarray:
    """n_r + 1 radii from r_inner to r_outer whose spacings grow by a constant
    ratio q from ``first_spacing`` (q found by bisection; uniform if the first
    spacing is already at least the mean spacing)."""
    span = r_outer - r_inner
    if first_spacing * n_r >= span:
        return np.linspace(r_inner, r_outer, n_r + 1)

    def total(q):
        with np.errstate(over="ignore"):
            return first_spacing * (float(np.power(np.float64(q), n_r)) - 1.0) / (q - 1.0)

    lo, hi = 1.0 + 1e-15, 2.0
    while total(hi) < span:
        hi *= 2.0
    for _ in range(200):
        mid = 0.5 * (lo + hi)
        if total(mid) < span:
            lo = mid
        else:
            hi = mid
    q = 0.5 * (lo + hi)
    r = r_inner + first_spacing * (q ** np.arange(n_r + 1) - 1.0) / (q - 1.0)
    r[-1] = r_outer
    return r


def cylinder_ogrid(
    n_theta: int,
    n_r: int,
    r_inner: float = 0.5,
    r_outer: float = 30.0,
    first_spacing: float = 1e-3,
    jitter: float = 0.05,
    seed: int | None = 0,
) -> FiniteVolumeMesh:
    """O-grid of quads around a circle: ``n_theta`` cells around, ``n_r``
    rings outward with geometric radial stretching (first spacing
    ``first_spacing``). The angular wrap shares nodes, so it is an ordinary
    interior seam (no periodic pair). Nodes of the interior rings are
    jittered by U(-jitter, jitter) times the local radial spacing (the smaller
    of the two adjacent ones) and the angular spacing; the wall and
    far-field rings are exact circles. Cells are numbered ring by ring from
    the wall, CCW. Patches: 0 "wall" (inner circle), 1 "farfield"."""
    if n_theta < 3 or n_r < 1:
        raise ValueError("O-grid needs n_theta >= 3 and n_r >= 1")
    r = geometric_radii(n_r, r_inner, r_outer, first_spacing)
    dth = 2.0 * math.pi / n_theta
    theta = dth * np.arange(n_theta)
    rr = np.broadcast_to(r[:, None], (n_r + 1, n_theta)).copy()
    tt = np.broadcast_to(theta[None, :], (n_r + 1, n_theta)).copy()
    if jitter > 0.0 and n_r > 1:
        rng = np.random.default_rng(seed)
        dr = np.diff(r)
        h = np.minimum(dr[:-1], dr[1:])  # interior rings 1..n_r-1
        off = rng.uniform(-jitter, jitter, size=(n_r - 1, n_theta, 2))
        rr[1:-1] += off[..., 0] * h[:, None]
        tt[1:-1] += off[..., 1] * dth
    points = np.column_stack(((rr * np.cos(tt)).ravel(), (rr * np.sin(tt)).ravel()))
    k, j = np.meshgrid(np.arange(n_r), np.arange(n_theta), indexing="ij")
    jn = (j + 1) % n_theta
    a = (k * n_theta + j).ravel()  # ring k, angle j
    b = ((k + 1) * n_theta + j).ravel()  # ring k+1, angle j
    c = ((k + 1) * n_theta + jn).ravel()
    d = (k * n_theta + jn).ravel()
    # (theta, r) -> (x, y) reverses orientation: going outward first is CCW
    cells = np.column_stack((a, b, c, d)).astype(np.int64)
    sizes = np.full(len(cells), 4, dtype=np.int64)
    faces, owners, neighbours = _faces_from_cells(cells, sizes)
    ring = np.repeat(np.arange(n_r + 1), n_theta)
    f0, f1 = faces[:, 0], faces[:, 1]
    boundary = neighbours < 0
    patch_ids = np.full(len(faces), -1, dtype=np.int64)
    patch_ids[boundary & (ring[f0] == 0) & (ring[f1] == 0)] = 0
    patch_ids[boundary & (ring[f0] == n_r) & (ring[f1] == n_r)] = 1
    if np.any(patch_ids[boundary] < 0):
        raise ValueError("Boundary face off the O-grid's circles.")
    patches = tuple(MeshPatch(pid, name, np.flatnonzero(patch_ids == pid)) for pid, name in enumerate(("wall", "farfield")))
    return FiniteVolumeMesh(
        points=points,
        cells=cells,
        cell_sizes=sizes,
        faces=faces,
        face_owners=owners,
        face_neighbours=neighbours,
        face_patch_ids=patch_ids,
        patches=patches,
    )


def free_stream(mach: float, alpha_deg: float, gamma: float = 1.4, density: float = 1.0,
                pressure: float | None = None):
    """(rho, u, v, p) of a free stream at Mach ``mach`` and angle ``alpha_deg``;
    by default p = 1/gamma so that the sound speed is 1."""
    p = 1.0 / gamma if pressure is None else pressure
    c = math.sqrt(gamma * p / density)
    a = math.radians(alpha_deg)
    return (density, mach * c * math.cos(a), mach * c * math.sin(a), p)


def cylinder_case(
    n_theta: int = 2048,
    n_r: int = 2048,
    mach: float = 0.2,
    alpha_deg: float = 0.0,
    reynolds: float | None = 40.0,
    r_outer: float = 30.0,
    first_spacing: float = 1e-3,
    jitter: float = 0.05,
    seed: int = 0,
    numerics: NumericalSettings | None = None,
):
    """Flow past the unit-diameter cylinder on `cylinder_ogrid`.

    Laminar (``reynolds`` given): air with Sutherland viscosity at
    T = 288.15 K, the density chosen so that Re = rho U D / mu, a no-slip wall
    at the free-stream temperature, far field outside; Roe + MUSCL parabolic,
    implicit Euler, local steps, CFL 100, GMRES(30) rtol 1e-3. Inviscid
    (``reynolds=None``): unit free stream (rho = 1, c = 1), slip wall.
    Returns (case, initial primitive state = free stream everywhere)."""
    mesh = cylinder_ogrid(n_theta, n_r, 0.5, r_outer, first_spacing, jitter, seed)
    gamma = 1.4
    if numerics is None:
        numerics = NumericalSettings(
            flux="Roe", reconstruction="MUSCL parabolic", time_integration="Implicit Euler",
            time_step_mode="Local (steady)", cfl=100.0, gmres_krylov_dimension=30, gmres_tolerance=1e-3,
            residual_tolerance=1e-10, end_time=1e30,
        )
    if reynolds is not None:
        numerics.viscosity_model = "Sutherland"
        numerics.prandtl_number = 0.72
        R = numerics.gas_constant
        T = 288.15
        tr = numerics.sutherland_reference_temperature
        mu = numerics.sutherland_reference_viscosity * (T / tr) ** 1.5 * (tr + numerics.sutherland_temperature) / (
            T + numerics.sutherland_temperature)
        c = math.sqrt(gamma * R * T)
        speed = mach * c
        rho = reynolds * mu / (speed * 1.0)
        fs = free_stream(mach, alpha_deg, gamma, density=rho, pressure=rho * R * T)
        wall = BoundaryCondition(0, "wall", "Wall", {"temperature": T}, "Temperature", "No-slip")
    else:
        fs = free_stream(mach, alpha_deg, gamma)
        wall = BoundaryCondition(0, "wall", "Wall", {}, "Temperature", "Slip")
    far = BoundaryCondition(1, "farfield", "Far Field",
                            {"density": fs[0], "velocity_x": fs[1], "velocity_y": fs[2], "pressure": fs[3]})
    case = Case(mesh=mesh, boundary_conditions=[wall, far], numerics=numerics, name="cylinder")
    prim = np.tile(np.asarray(fs, dtype=float), (mesh.cell_count, 1))
    return case, prim


def bump_surface(x, h: float, x0: float = 1.0, chord: float = 1.0, xp=np):
    """Height of a circular-arc bump of height ``h`` over [x0, x0 + chord]
    (zero elsewhere): the arc through (x0, 0), (x0 + chord/2, h),
    (x0 + chord, 0). ``xp`` is numpy or torch (then ``x`` and ``h`` may be
    tensors and the result is differentiable in ``h``)."""
    radius = (0.25 * chord * chord + h * h) / (2.0 * h)
    xc = x0 + 0.5 * chord
    yc = h - radius
    dx = x - xc
    inside = xp.abs(dx) < 0.5 * chord
    arg = radius * radius - xp.where(inside, dx, 0.0 * dx) ** 2
    return xp.where(inside, yc + xp.sqrt(arg), 0.0 * dx)


def bump_nodes(base, h, height: float = 1.0, x0: float = 1.0, chord: float = 1.0, xp=np):
    """Channel nodes over a bump: base nodes (x, y) of the rectangle
    [0, L] x [0, height] moved to y + y_b(x) (1 - y/height), so the bottom
    follows the bump and the top stays straight."""
    x, y = base[:, 0], base[:, 1]
    moved = y + bump_surface(x, h, x0, chord, xp) * (1.0 - y / height)
    return np.column_stack((x, moved)) if xp is np else xp.stack((x, moved), 1)


def bump_channel(nx: int, ny: int, h: float, length: float = 3.0, height: float = 1.0, x0: float = 1.0,
                 chord: float = 1.0) -> FiniteVolumeMesh:
    """Quad grid of the channel [0, length] x [0, height] with a circular-arc
    bump of height ``h`` on the bottom (patches: 0 bottom wall, 1 right,
    2 top, 3 left). The connectivity does not depend on ``h``; the nodes are
    `bump_nodes` of the rectangle's."""
    if not (0.0 < h < 0.5 * height):
        raise ValueError("bump height must be in (0, height/2)")
    base = _grid_mesh(nx, ny, 0.0, length, 0.0, height, 0.0, None)
    base.points = bump_nodes(base.points, h, height, x0, chord)
    base.face_curves = bump_face_curves(base, h, x0, chord)
    return base


def bump_arc(h: float, x0: float = 1.0, chord: float = 1.0) -> dict:
    """The bump's circular arc as a MeshCAD arc descriptor, run
    counter-clockwise from its right end over the top to its left end."""
    radius = (0.25 * chord * chord + h * h) / (2.0 * h)
    xc, yc = x0 + 0.5 * chord, h - radius
    right = math.atan2(-yc, 0.5 * chord)
    return {"type": "arc", "center": [xc, yc], "radius": radius, "start_angle": right,
            "end_angle": math.pi - right}


def bump_face_curves(mesh: FiniteVolumeMesh, h: float, x0: float = 1.0, chord: float = 1.0) -> tuple:
    """Per face: the bump arc with the parameters of its two nodes for bottom
    faces on the bump (so that refinement splits them on the arc), None
    elsewhere."""
    arc = bump_arc(h, x0, chord)
    xc, yc = arc["center"]
    a0, a1 = arc["start_angle"], arc["end_angle"]
    pts = mesh.points
    out = []
    for f, (n0, n1) in enumerate(mesh.faces):
        on_bump = (mesh.face_patch_ids[f] == 0 and x0 - 1e-12 <= pts[n0, 0] <= x0 + chord + 1e-12
                   and x0 - 1e-12 <= pts[n1, 0] <= x0 + chord + 1e-12)
        if not on_bump:
            out.append(None)
            continue
        t = [min(1.0, max(0.0, (math.atan2(pts[n, 1] - yc, pts[n, 0] - xc) - a0) / (a1 - a0))) for n in (n0, n1)]
        out.append(BoundaryCurve(arc, float(t[0]), float(t[1]), 0))
    return tuple(out)


def bump_case(nx: int, ny: int, h: float, mach: float = 0.675, integrator: str = "Implicit Euler",
              cfl: float = 50.0):
    """Inviscid subsonic channel flow over the bump: velocity inlet at Mach
    ``mach`` (rho = 1, T = 1, R = 1, so p = 1), pressure outlet p = 1, slip
    walls; HLLC + MUSCL linear (Barth-Jespersen), local pseudo-time steps.
    Returns (case, uniform initial primitive state)."""
    gamma = 1.4
    speed = mach * math.sqrt(gamma)
    numerics = NumericalSettings(
        flux="HLLC", reconstruction="MUSCL linear", time_integration=integrator,
        time_step_mode="Local (steady)", cfl=cfl, gas_constant=1.0, end_time=1e30,
        gmres_krylov_dimension=30, gmres_tolerance=1e-3, residual_tolerance=1e-9,
    )
    mesh = bump_channel(nx, ny, h)
    bcs = [
        BoundaryCondition(0, "bump", "Wall", {}, "Temperature", "Slip"),
        BoundaryCondition(1, "outlet", "Pressure Outlet", {"pressure": 1.0}),
        BoundaryCondition(2, "top", "Wall", {}, "Temperature", "Slip"),
        BoundaryCondition(3, "inlet", "Velocity Inlet",
                          {"density": 1.0, "temperature": 1.0, "velocity_x": speed, "velocity_y": 0.0}),
    ]
    case = Case(mesh=mesh, boundary_conditions=bcs, numerics=numerics, name=f"bump h={h}")
    prim = np.tile((1.0, speed, 0.0, 1.0), (mesh.cell_count, 1))
    return case, prim
