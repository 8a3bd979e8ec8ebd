"""Adaptive refinement and coarsening of polygon meshes (host side).

Follows the reference's `src/meshcfd/core/adaptation.py`:

* `select_adaptation` (`adaptation.py:40-171`): a distance-weighted
  least-squares gradient of the indicator variable per cell (face neighbours
  and periodic partners), the indicator max(|grad| sqrt(A), largest jump) /
  value scale, refinement marks above the threshold below the maximum level
  (anisotropic "x" / "y" in a quadrilateral's local directions when the
  variation along one is more than twice the other and the split passes the
  quality gate, else "xy"), coarsening of refinement records whose children
  all fall below 0.3 x threshold, then the 2:1 balance and the quality gate
  until both hold;
* `balance_plan` (`adaptation.py:252-451`): periodic partners share marks and
  coarsening, projected levels and per-axis (directional) levels differ by at
  most one across every face, resolving the first violation in cell order
  and scanning again, with the reference's choices at each violation;
* `adapt_mesh` (`adaptation.py:680-795`): leaves in cell order, coarsened
  siblings replaced by their parent at the first child's position, marked
  cells split by `split_polygon` (`adaptation.py:830-938`: quadrilaterals at
  edge midpoints and the vertex mean, curved boundary edges at their
  parametric midpoints; other polygons clipped at their bounding-box
  middles), then `_rebuild_mesh` (`adaptation.py:1032-1190`): every leaf edge
  is cut at the mesh points lying on it, so a coarse cell next to refined
  ones gets hanging-node subfaces, faces numbered at first visit;
* `transfer_cell_values` (`solver.py:1594-1617`): area-weighted averages of
  the source cells, rescaled so each source group keeps its integral.

Geometry is exact in the same floating-point steps as the reference (edge
midpoints 0.5 (a + b), quad centres by summing the corners and dividing by
four, points keyed by rounding to 14 decimals), so adapted meshes carry the
reference's coordinates and numbering. Pure Python over the leaves, like the
reference: meant for meshes of up to a few hundred thousand cells.
"""

from __future__ import annotations

import json
import math
from dataclasses import dataclass

import numpy as np

from .curves import evaluate_boundary_curve
from .model import BoundaryCurve, FiniteVolumeMesh, MeshPatch, RefinementRecord

TOL = 1e-11  # the reference's geometry tolerance (adaptation.py:25)
AXES = ("x", "y")


class AdaptationBalanceError(ValueError):
    """A refinement plan that cannot satisfy the balance constraints
    (reference `AdaptationBalanceError`, adaptation.py:27-28)."""


@dataclass(frozen=True)
class Leaf:
    cell_id: str
    level: int
    vertices: tuple
    sources: tuple
    curves: tuple = ()


# ---------------------------------------------------------------------------
# Mesh access
# ---------------------------------------------------------------------------


def cell_vertices(mesh: FiniteVolumeMesh, cell: int) -> tuple:
    """The cell's vertices as (x, y) float tuples, in its CCW order."""
    ids = mesh.cells[cell, : mesh.cell_sizes[cell]]
    return tuple((float(mesh.points[n, 0]), float(mesh.points[n, 1])) for n in ids)


def all_cell_vertices(mesh: FiniteVolumeMesh) -> list:
    """`cell_vertices` of every cell (the same tuples), in one pass."""
    pts = [tuple(p) for p in mesh.points.tolist()]
    return [tuple(pts[n] for n in row[:size]) for row, size in zip(mesh.cells.tolist(), mesh.cell_sizes.tolist())]


def _cell_ids(mesh):
    return tuple(mesh.cell_ids) if len(mesh.cell_ids) else tuple(f"r{k}" for k in range(mesh.cell_count))


def _cell_levels(mesh):
    return tuple(int(v) for v in mesh.cell_levels) if len(mesh.cell_levels) else (0,) * mesh.cell_count


def cell_areas(mesh: FiniteVolumeMesh) -> np.ndarray:
    """|signed area| of every cell (reference `cell_areas`, adaptation.py:512-513)."""
    return np.array([abs(signed_area(cell_vertices(mesh, c))) for c in range(mesh.cell_count)])


# ---------------------------------------------------------------------------
# Planar helpers (reference adaptation.py:967-1249), same float steps
# ---------------------------------------------------------------------------


def point_key(p):
    return round(p[0], 14), round(p[1], 14)


def edge_key(a, b):
    return tuple(sorted((point_key(a), point_key(b))))


def dist2(a, b):
    return (a[0] - b[0]) ** 2 + (a[1] - b[1]) ** 2


def signed_area(vertices) -> float:
    n = len(vertices)
    total = 0
    for k in range(n):
        a, b = vertices[k], vertices[(k + 1) % n]
        total = total + (a[0] * b[1] - b[0] * a[1])
    return 0.5 * total


def clean_polygon(poly) -> tuple:
    """Drop repeated consecutive points (and a closing repeat)."""
    out = []
    for p in poly:
        if not out or dist2(out[-1], p) > 1e-26:
            out.append(p)
    if len(out) > 1 and dist2(out[0], out[-1]) <= 1e-26:
        out.pop()
    return tuple(out)


def drop_collinear(poly) -> tuple:
    """Remove vertices between two collinear edges, first one found each pass."""
    v = list(clean_polygon(poly))
    while len(v) > 3:
        for k, cur in enumerate(v):
            prev, nxt = v[k - 1], v[(k + 1) % len(v)]
            e0 = (cur[0] - prev[0], cur[1] - prev[1])
            e1 = (nxt[0] - cur[0], nxt[1] - cur[1])
            cross = e0[0] * e1[1] - e0[1] * e1[0]
            if abs(cross) <= TOL * max(math.hypot(*e0) * math.hypot(*e1), 1e-30):
                v.pop(k)
                break
        else:
            break
    return tuple(v)


def clip_polygon(poly, axis: int, value: float, keep_less: bool) -> tuple:
    """Sutherland-Hodgman clip against the line coordinate[axis] = value."""
    if not poly:
        return ()

    def inside(p):
        d = p[axis] - value
        return d <= TOL if keep_less else d >= -TOL

    out = []
    n = len(poly)
    for k in range(n):
        a, b = poly[k], poly[(k + 1) % n]
        ia, ib = inside(a), inside(b)
        if ia:
            out.append(a)
        if ia != ib:
            frac = (value - a[axis]) / (b[axis] - a[axis])
            out.append((a[0] + frac * (b[0] - a[0]), a[1] + frac * (b[1] - a[1])))
    return tuple(out)


def on_segment(p, a, b) -> bool:
    ex, ey = b[0] - a[0], b[1] - a[1]
    ox, oy = p[0] - a[0], p[1] - a[1]
    length = math.hypot(ex, ey)
    if length <= 1e-15:
        return dist2(p, a) <= TOL**2
    cross = abs(ex * oy - ey * ox) / length
    t = (ox * ex + oy * ey) / (length * length)
    return cross <= TOL and -TOL <= t <= 1 + TOL


def segment_parameter(p, a, b) -> float:
    dx, dy = b[0] - a[0], b[1] - a[1]
    return ((p[0] - a[0]) * dx + (p[1] - a[1]) * dy) / (dx * dx + dy * dy)


def aspect_ratio(vertices) -> float:
    """sqrt of the second-moment eigenvalue ratio of the corners
    (adaptation.py:189-197)."""
    pts = np.asarray(drop_collinear(vertices), dtype=float)
    if len(pts) < 3:
        return math.inf
    c = pts - np.mean(pts, axis=0)
    ev = np.linalg.eigvalsh(c.T @ c / len(pts))
    if ev[0] <= TOL**2:
        return math.inf
    return math.sqrt(float(ev[-1] / ev[0]))


def local_directions(corners):
    """Unit mean directions of a quadrilateral's two edge pairs
    (adaptation.py:174-181)."""
    p = np.asarray(corners, dtype=float)
    first = 0.5 * ((p[1] - p[0]) + (p[2] - p[3]))
    second = 0.5 * ((p[3] - p[0]) + (p[2] - p[1]))
    if np.linalg.norm(first) <= TOL or np.linalg.norm(second) <= TOL:
        return np.asarray((1.0, 0.0)), np.asarray((0.0, 1.0))
    return first / np.linalg.norm(first), second / np.linalg.norm(second)


def span_along(vertices, direction) -> float:
    return float(np.ptp(np.asarray(vertices, dtype=float) @ direction))


# ---------------------------------------------------------------------------
# Boundary curves on cell edges
# ---------------------------------------------------------------------------


def curve_midpoint(curve: BoundaryCurve):
    return evaluate_boundary_curve(curve.geometry, 0.5 * (curve.start_parameter + curve.end_parameter))


def _reversed(curve: BoundaryCurve) -> BoundaryCurve:
    return BoundaryCurve(curve.geometry, curve.end_parameter, curve.start_parameter, curve.patch_id)


def cell_edge_curves(mesh: FiniteVolumeMesh, cell: int, by_edge: dict | None) -> tuple:
    """Per edge of the cell (start -> end): its boundary curve oriented along
    the edge, or None; () when the mesh has no curves (adaptation.py:516-545)."""
    if by_edge is None:
        return ()
    ids = [int(n) for n in mesh.cells[cell, : mesh.cell_sizes[cell]]]
    out = []
    for k, a in enumerate(ids):
        b = ids[(k + 1) % len(ids)]
        hit = by_edge.get((min(a, b), max(a, b)))
        if hit is None:
            out.append(None)
            continue
        face, curve = hit
        out.append(curve if (int(mesh.faces[face, 0]), int(mesh.faces[face, 1])) == (a, b) else _reversed(curve))
    return tuple(out)


def _curves_by_edge(mesh: FiniteVolumeMesh):
    if not mesh.face_curves:
        return None
    return {
        (min(int(mesh.faces[f, 0]), int(mesh.faces[f, 1])), max(int(mesh.faces[f, 0]), int(mesh.faces[f, 1]))):
        (f, c)
        for f, c in enumerate(mesh.face_curves)
        if c is not None
    }


def _is_curved(corners, curve_of_edge: dict) -> bool:
    """An edge whose curve midpoint leaves the chord midpoint (adaptation.py:653-671)."""
    n = len(corners)
    for k in range(n):
        a, b = corners[k], corners[(k + 1) % n]
        curve = curve_of_edge.get(edge_key(a, b))
        if curve is None:
            continue
        m = curve_midpoint(curve)
        chord = (0.5 * (a[0] + b[0]), 0.5 * (a[1] + b[1]))
        scale = max(1.0, abs(a[0]), abs(a[1]), abs(b[0]), abs(b[1]))
        if dist2(m, chord) > (scale * 1e-12) ** 2:
            return True
    return False


def child_edge_curves(vertices, curves, children) -> tuple:
    """Curves of the children's edges that lie on a curved parent edge: the
    parent curve's parameters at its end points and midpoint
    (adaptation.py:598-644)."""
    if len(curves) != len(vertices):
        return tuple(() for _ in children)
    segments = []
    n = len(vertices)
    for k in range(n):
        curve = curves[k]
        if curve is None:
            continue
        a, b = vertices[k], vertices[(k + 1) % n]
        params = {
            point_key(a): curve.start_parameter,
            point_key(curve_midpoint(curve)): 0.5 * (curve.start_parameter + curve.end_parameter),
            point_key(b): curve.end_parameter,
        }
        segments.append((curve.geometry, curve.patch_id, params))
    result = []
    for child in children:
        edges = []
        m = len(child)
        for k in range(m):
            ka, kb = point_key(child[k]), point_key(child[(k + 1) % m])
            found = None
            for geometry, patch, params in segments:
                if ka in params and kb in params:
                    found = BoundaryCurve(geometry, params[ka], params[kb], patch)
                    break
            edges.append(found)
        result.append(tuple(edges))
    return tuple(result)


def merged_edge_curves(vertices, leaves) -> tuple:
    """The parent's edge curves from its children's (adaptation.py:548-595)."""
    grouped: dict = {}
    for leaf in leaves:
        if len(leaf.curves) != len(leaf.vertices):
            continue
        m = len(leaf.vertices)
        for k in range(m):
            curve = leaf.curves[k]
            if curve is None:
                continue
            key = (json.dumps(curve.geometry, sort_keys=True, separators=(",", ":")), curve.patch_id)
            geometry, params = grouped.setdefault(key, (curve.geometry, {}))
            params.setdefault(point_key(leaf.vertices[k]), curve.start_parameter)
            params.setdefault(point_key(leaf.vertices[(k + 1) % m]), curve.end_parameter)
    out = []
    n = len(vertices)
    for k in range(n):
        ka, kb = point_key(vertices[k]), point_key(vertices[(k + 1) % n])
        hits = [(g, key[1], p) for key, (g, p) in grouped.items() if ka in p and kb in p]
        if len(hits) != 1:
            out.append(None)
            continue
        g, patch, p = hits[0]
        out.append(BoundaryCurve(g, p[ka], p[kb], patch))
    return tuple(out)


# ---------------------------------------------------------------------------
# Splitting
# ---------------------------------------------------------------------------


def _check_area(parent, pieces, curved: bool = False):
    """Children must tile the parent (adaptation.py:941-964)."""
    parent_area = abs(signed_area(parent))
    child_area = sum(abs(signed_area(p)) for p in pieces)
    if curved:
        orientation = math.copysign(1.0, signed_area(parent))
        if any(signed_area(p) * orientation <= 1e-14 for p in pieces):
            raise ValueError("Curved boundary splitting produced an invalid cell.")
        return pieces
    if not math.isclose(parent_area, child_area, rel_tol=1e-10, abs_tol=max(parent_area, 1.0) * 1e-12):
        raise ValueError("Adaptive splitting requires a convex cell that can be partitioned by axis-aligned cuts.")
    return pieces


def split_polygon(polygon, axis: str, curves=()) -> tuple:
    """Children of a cell split along ``axis`` ("x", "y" or "xy")
    (adaptation.py:830-938). Quadrilaterals split at edge midpoints (a curved
    boundary edge at its parametric midpoint): "x"/"y" across the edge pair
    whose midpoints differ most in y/x, "xy" into four around the corner
    mean; other polygons are clipped at their bounding-box middles."""
    corners = drop_collinear(polygon)
    if len(corners) == 4:
        curve_of_edge = {}
        if len(curves) == len(polygon):
            n = len(polygon)
            for k in range(n):
                if curves[k] is not None:
                    curve_of_edge[edge_key(polygon[k], polygon[(k + 1) % n])] = curves[k]
        mids = []
        for k in range(4):
            a, b = corners[k], corners[(k + 1) % 4]
            c = curve_of_edge.get(edge_key(a, b))
            mids.append(curve_midpoint(c) if c is not None else (0.5 * (a[0] + b[0]), 0.5 * (a[1] + b[1])))
        curved = _is_curved(corners, curve_of_edge)
        if axis in AXES:
            coord = 1 if axis == "x" else 0
            first, opposite = max(((0, 2), (1, 3)), key=lambda pr: abs(mids[pr[0]][coord] - mids[pr[1]][coord]))
            half_a = [mids[first]]
            k = (first + 1) % 4
            while k != (opposite + 1) % 4:
                half_a.append(corners[k])
                k = (k + 1) % 4
            half_a.append(mids[opposite])
            half_b = [mids[opposite]]
            while k != (first + 1) % 4:
                half_b.append(corners[k])
                k = (k + 1) % 4
            half_b.append(mids[first])
            return _check_area(corners, (tuple(half_a), tuple(half_b)), curved)
        if axis == "xy":
            sx = 0
            sy = 0
            for p in corners:
                sx = sx + p[0]
                sy = sy + p[1]
            centre = (sx / 4.0, sy / 4.0)
            pieces = tuple((corners[k], mids[k], centre, mids[(k - 1) % 4]) for k in range(4))
            return _check_area(corners, pieces, curved)
    xs = [p[0] for p in polygon]
    ys = [p[1] for p in polygon]
    mx = 0.5 * (min(xs) + max(xs))
    my = 0.5 * (min(ys) + max(ys))
    if axis == "x":
        pieces = (clip_polygon(polygon, 0, mx, True), clip_polygon(polygon, 0, mx, False))
    elif axis == "y":
        pieces = (clip_polygon(polygon, 1, my, True), clip_polygon(polygon, 1, my, False))
    elif axis == "xy":
        pieces = tuple(clip_polygon(clip_polygon(polygon, 0, mx, left), 1, my, bottom)
                       for left, bottom in ((True, True), (False, True), (True, False), (False, False)))
    else:
        raise ValueError(f"Unsupported refinement split axis '{axis}'.")
    valid = []
    for piece in pieces:
        c = clean_polygon(piece)
        if len(c) >= 3 and abs(signed_area(c)) > 1e-14:
            valid.append(c)
    return _check_area(polygon, tuple(valid))


def split_passes_quality(vertices, axis: str, max_aspect: float = 6.0) -> bool:
    """The quality gate (adaptation.py:200-213): every child keeps >= 10% of
    the parent's area and an aspect ratio <= max_aspect."""
    try:
        children = split_polygon(vertices, axis)
    except ValueError:
        return False
    parent = abs(signed_area(vertices))
    if parent <= TOL:
        return False
    for child in children:
        if abs(signed_area(child)) / parent < 0.10 or aspect_ratio(child) > max_aspect:
            return False
    return True


def gated_axis(vertices, axis: str) -> str:
    return axis if axis == "xy" or split_passes_quality(vertices, axis) else "xy"


# ---------------------------------------------------------------------------
# Directional levels and the balance
# ---------------------------------------------------------------------------


def directional_levels(vertices, base) -> tuple:
    """(x, y) refinement level of a cell against its root cell: the rounded
    log2 of the bounding-box span ratio per axis (adaptation.py:493-509)."""
    out = []
    for axis in (0, 1):
        span = max(p[axis] for p in vertices) - min(p[axis] for p in vertices)
        base_span = max(p[axis] for p in base) - min(p[axis] for p in base)
        out.append(0 if span <= TOL or base_span <= TOL else max(0, round(math.log2(base_span / span))))
    return out[0], out[1]


def _root_vertices(records):
    return {r.parent_id: r.vertices for r in records if "." not in r.parent_id}


def cell_directional_levels(mesh, cell_ids, records) -> tuple:
    """`directional_levels` of every cell against its root: the bounding-box
    spans in NumPy (min and max are exact), the log2 of the ratios in Python
    for the cells whose root box differs from their own."""
    roots = _root_vertices(records)
    n = mesh.cell_count
    if n == 0:
        return ()
    width = mesh.cells.shape[1]
    valid = np.arange(width)[None, :] < mesh.cell_sizes[:, None]
    ids = np.where(valid, mesh.cells, 0)
    xy = mesh.points[ids]  # (n, width, 2)
    big = np.inf
    lo = np.where(valid[:, :, None], xy, big).min(axis=1)
    hi = np.where(valid[:, :, None], xy, -big).max(axis=1)
    span = (hi - lo).tolist()
    base_of = {}
    for rid, verts in roots.items():
        base_of[rid] = tuple(max(p[axis] for p in verts) - min(p[axis] for p in verts) for axis in (0, 1))
    out = []
    for k, cid in enumerate(cell_ids):
        base = base_of.get(cid.split(".", 1)[0])
        if base is None:  # the cell is its own root: ratio 1, level 0
            out.append((0, 0))
            continue
        own = span[k]
        out.append(tuple(0 if own[a] <= TOL or base[a] <= TOL else max(0, round(math.log2(base[a] / own[a])))
                         for a in (0, 1)))
    return tuple(out)


def parent_directional_levels(records) -> dict:
    roots = _root_vertices(records)
    return {r.parent_id: directional_levels(r.vertices, roots.get(r.parent_id.split(".", 1)[0], r.vertices))
            for r in records}


def _least_disruptive_axis(cell, neighbours, projected, levels_dir) -> int:
    """The split axis whose projected directional levels jump least against
    the neighbours (max jump, then total, then the cell's own level;
    adaptation.py:225-249)."""
    own = projected[cell]

    def cost(a):
        jumps = [abs(own[c] + int(c == a) - projected[nb][c]) for nb in neighbours[cell] for c in (0, 1)]
        return max(jumps), sum(jumps), levels_dir[cell][a]

    return min((0, 1), key=cost)


def balance_plan(marks: dict, coarsen: set, neighbours, cell_ids, levels, levels_dir, parent_dir: dict,
                 maximum_level: int, records, periodic_pairs) -> None:
    """Adjust ``marks`` and ``coarsen`` in place until projected levels and
    projected directional levels differ by at most one across every face
    (adaptation.py:252-451); raises `AdaptationBalanceError` when that is
    impossible or the plan cycles.

    Each round fixes the first violation in scan order (cells ascending,
    each cell's larger neighbours in its set's order), as a full rescan
    would; a fix or a periodic adjustment changes the projections of a few
    cells only, so the projections are updated for those cells and the next
    scan starts at the first cell of a pair that holds one of them, which
    finds the same first violation without rescanning the cells before it.
    States are recognised for the cycle test by an order-free sum of item
    hashes kept up to date with every change."""
    by_parent = {r.parent_id: r for r in records}
    parent_of = {child: r.parent_id for r in records for child in r.child_ids}
    n = len(cell_ids)
    index = {cid: k for k, cid in enumerate(cell_ids)}
    kids_of = {p: [index[ch] for ch in r.child_ids if ch in index] for p, r in by_parent.items()}
    nb_lists = [list(neighbours[k]) for k in range(n)]
    mask = (1 << 64) - 1
    state = {"h": 0}
    dirty = set()

    def item_hash(item):
        return hash(item) & mask

    for k, axis in marks.items():
        state["h"] = (state["h"] + item_hash(("m", k, axis))) & mask
    for parent in coarsen:
        state["h"] = (state["h"] + item_hash(("c", parent))) & mask

    def set_mark(k, axis):
        old = marks.get(k)
        if old is not None:
            state["h"] = (state["h"] - item_hash(("m", k, old))) & mask
        marks[k] = axis
        state["h"] = (state["h"] + item_hash(("m", k, axis))) & mask
        dirty.add(k)

    def del_mark(k):
        state["h"] = (state["h"] - item_hash(("m", k, marks.pop(k)))) & mask
        dirty.add(k)

    def drop_coarsen(parent):
        coarsen.remove(parent)
        state["h"] = (state["h"] - item_hash(("c", parent))) & mask
        dirty.update(kids_of.get(parent, ()))

    def project(k):
        parent = parent_of.get(cell_ids[k])
        if parent in coarsen:
            return by_parent[parent].level, parent_dir[parent]
        m = marks.get(k, "")
        return levels[k] + int(k in marks), (levels_dir[k][0] + int("x" in m), levels_dir[k][1] + int("y" in m))

    proj_lev, proj_dir = [0] * n, [None] * n
    for k in range(n):
        proj_lev[k], proj_dir[k] = project(k)

    def fix(lower, higher, axis):
        """Resolve one violation between a coarser ``lower`` and a finer
        ``higher`` cell (axis None: levels; else one direction)."""
        parent = parent_of.get(cell_ids[lower])
        if parent in coarsen:
            drop_coarsen(parent)
            return
        if lower in marks:
            if axis is None or axis in marks[lower]:
                if higher in marks:
                    del_mark(higher)
                    return
                raise AdaptationBalanceError(
                    "The current mesh cannot satisfy 2:1 refinement balance in one adaptation step."
                    if axis is None else
                    "The current mesh cannot satisfy anisotropic transition balance in one adaptation step."
                )
            both = set(marks[lower]) | {axis}
            set_mark(lower, "xy" if len(both) == 2 else next(iter(both)))
            return
        if levels[lower] < maximum_level:
            if axis is None:
                set_mark(lower, AXES[_least_disruptive_axis(lower, neighbours, proj_dir, levels_dir)])
            else:
                set_mark(lower, axis)
            return
        if higher in marks:
            del_mark(higher)
            return
        raise AdaptationBalanceError(
            "The current mesh cannot satisfy 2:1 refinement balance within the configured maximum level."
            if axis is None else
            "The current mesh cannot satisfy anisotropic transition balance within the configured maximum level."
        )

    def first_violation(k0):
        for k in range(k0, n):
            for nb in nb_lists[k]:
                if nb <= k:
                    continue
                d = proj_lev[k] - proj_lev[nb]
                if abs(d) > 1:
                    return k, (k if d < 0 else nb), (nb if d < 0 else k), None
                for ai, axis in enumerate(AXES):
                    d = proj_dir[k][ai] - proj_dir[nb][ai]
                    if abs(d) > 1:
                        return k, (k if d < 0 else nb), (nb if d < 0 else k), axis
        return None

    seen = set()
    changed = True
    restart = 0
    while changed:
        changed = False
        signature = (state["h"], len(marks), len(coarsen))
        if signature in seen:
            raise AdaptationBalanceError("The adaptive refinement plan cycles while balancing mesh transitions.")
        seen.add(signature)
        # periodic partners: coarsen both or neither, refine both the same way
        for a, b, _ in periodic_pairs:
            pa, pb = parent_of.get(cell_ids[a]), parent_of.get(cell_ids[b])
            if pa in coarsen and pb not in coarsen:
                drop_coarsen(pa)
                changed = True
            if pb in coarsen and pa not in coarsen:
                drop_coarsen(pb)
                changed = True
        for a, b, _ in periodic_pairs:
            if a in marks and b not in marks:
                if levels[b] >= maximum_level:
                    del_mark(a)
                else:
                    set_mark(b, marks[a])
                changed = True
            elif b in marks and a not in marks:
                if levels[a] >= maximum_level:
                    del_mark(b)
                else:
                    set_mark(a, marks[b])
                changed = True
            elif a in marks and b in marks and marks[a] != marks[b]:
                if marks[a] != "xy" or marks[b] != "xy":
                    set_mark(a, "xy")
                    set_mark(b, "xy")
                    changed = True
        # projections of the changed cells, and the first pair that may have changed
        for k in dirty:
            proj_lev[k], proj_dir[k] = project(k)
            restart = min(restart, k, min(nb_lists[k], default=k))
        dirty.clear()
        found = first_violation(restart)
        if found is None:
            restart = n
            continue
        restart, lower, higher, axis = found
        fix(lower, higher, axis)
        changed = True
    # final verification (adaptation.py:427-451)
    for k in dirty:
        proj_lev[k], proj_dir[k] = project(k)
    for k in range(n):
        for nb in nb_lists[k]:
            if nb <= k:
                continue
            if abs(proj_lev[k] - proj_lev[nb]) > 1:
                raise AdaptationBalanceError("The projected adaptive mesh violates 2:1 refinement balance.")
            if any(abs(x - y) > 1 for x, y in zip(proj_dir[k], proj_dir[nb])):
                raise AdaptationBalanceError("The projected adaptive mesh violates anisotropic transition balance.")


# ---------------------------------------------------------------------------
# Neighbour structure, indicator and marks
# ---------------------------------------------------------------------------


def periodic_cell_pairs(mesh, pairs) -> list:
    """(owner of face a, owner of face b, shift) per matched periodic face
    pair (adaptation.py:798-815)."""
    from .geometry import match_periodic_faces

    out = []
    for pair in pairs:
        fa, fb, shift = match_periodic_faces(mesh, pair.patch_a_id, pair.patch_b_id)
        s = (float(shift[0]), float(shift[1]))
        out.extend((int(mesh.face_owners[a]), int(mesh.face_owners[b]), s) for a, b in zip(fa, fb))
    return out


def _neighbour_sets(mesh, periodic):
    nbrs = [set() for _ in range(mesh.cell_count)]
    for o, n in zip(mesh.face_owners.tolist(), mesh.face_neighbours.tolist()):
        if n < 0:
            continue
        nbrs[o].add(n)
        nbrs[n].add(o)
    for a, b, _ in periodic:
        nbrs[a].add(b)
        nbrs[b].add(a)
    return nbrs


def active_periodic_pairs(case) -> tuple:
    kinds = {bc.patch_id: bc.kind for bc in case.boundary_conditions}
    return tuple(p for p in case.periodic_pairs
                 if kinds.get(p.patch_a_id) == "Periodic" and kinds.get(p.patch_b_id) == "Periodic")


def indicator(mesh, centers, areas, values, links) -> tuple:
    """Per cell (indicator, gradient): the distance-weighted least-squares
    gradient of ``values`` over the links (lstsq of the sqrt(1/d^2)-weighted
    system, rcond 1e-12; cells with fewer than two samples keep zero), and
    max(|grad| sqrt(A) / scale, largest |jump| / scale) with scale =
    max(range, 1e-8 max|v|, 1e-30) (adaptation.py:91-122)."""
    n = mesh.cell_count
    values = np.asarray(values, dtype=float)
    lk = np.asarray(links, dtype=float).reshape(-1, 4)
    la, lb = lk[:, 0].astype(np.int64), lk[:, 1].astype(np.int64)
    dx = centers[lb, 0] + lk[:, 2] - centers[la, 0]
    dy = centers[lb, 1] + lk[:, 3] - centers[la, 1]
    dv = values[lb] - values[la]
    # each link gives (dx, dy, dv) to a and the negation to b; a cell's
    # samples in link order
    cell = np.column_stack((la, lb)).ravel()
    sx = np.column_stack((dx, -dx)).ravel()
    sy = np.column_stack((dy, -dy)).ravel()
    sv = np.column_stack((dv, -dv)).ravel()
    order = np.argsort(cell, kind="stable")
    cell, sx, sy, sv = cell[order], sx[order], sy[order], sv[order]
    count = np.bincount(cell, minlength=n)
    slot = np.arange(len(cell)) - np.repeat(np.cumsum(count) - count, count)
    grads = np.zeros((n, 2))
    kmax = int(count.max()) if n else 0
    if kmax >= 2:
        A = np.zeros((n, kmax, 2))
        rhs = np.zeros((n, kmax))
        use = count >= 2
        keep = use[cell]
        off = np.column_stack((sx, sy))[keep]
        w = 1.0 / np.maximum(np.sum(off * off, axis=1), TOL**2)
        A[cell[keep], slot[keep]] = off * np.sqrt(w)[:, None]
        rhs[cell[keep], slot[keep]] = sv[keep] * np.sqrt(w)
        # minimum-norm least squares per cell (zero rows change nothing)
        grads[use] = np.einsum("cij,cj->ci", np.linalg.pinv(A[use], rcond=1e-12), rhs[use])
    value_range = float(np.ptp(values))
    scale = max(value_range, float(np.max(np.abs(values))) * 1e-8, 1e-30)
    grad_ind = np.linalg.norm(grads, axis=1) * np.sqrt(areas) / scale
    jump = np.zeros(n)
    np.maximum.at(jump, la, np.abs(dv) / scale)
    return np.maximum(grad_ind, jump), grads


def select_adaptation(mesh, case, centers, areas, values, threshold: float, maximum_level: int):
    """Refinement marks {cell: "x" | "y" | "xy"} and parent ids to coarsen
    (reference `select_adaptation`, adaptation.py:40-171)."""
    cell_ids = _cell_ids(mesh)
    levels = _cell_levels(mesh)
    values = np.asarray(values, dtype=float)
    if not (len(centers) == len(areas) == len(values) == mesh.cell_count):
        raise ValueError("Adaptation indicators must match the mesh cell count.")
    if not np.isfinite(values).all():
        raise ValueError("Adaptation indicator values must be finite.")
    centers = np.asarray(centers, dtype=float)
    links = [(int(o), int(n), 0.0, 0.0) for o, n in zip(mesh.face_owners, mesh.face_neighbours) if n >= 0]
    periodic = periodic_cell_pairs(mesh, active_periodic_pairs(case))
    nbrs = _neighbour_sets(mesh, ())
    for a, b, s in periodic:
        nbrs[a].add(b)
        nbrs[b].add(a)
        links.append((a, b, -s[0], -s[1]))
        links.append((b, a, s[0], s[1]))
    ind, grads = indicator(mesh, centers, np.asarray(areas, dtype=float), values, links)
    marks = {}
    for c in np.flatnonzero(ind > threshold).tolist():
        if levels[c] >= maximum_level:
            continue
        verts = cell_vertices(mesh, c)
        corners = drop_collinear(verts)
        axis = "xy"
        if len(corners) == 4 and aspect_ratio(corners) < 3.0:
            lx, ly = local_directions(corners)
            vx = abs(float(np.dot(grads[c], lx))) * span_along(corners, lx)
            vy = abs(float(np.dot(grads[c], ly))) * span_along(corners, ly)
            if vx > 2.0 * max(vy, 1e-30):
                axis = "x"
            elif vy > 2.0 * max(vx, 1e-30):
                axis = "y"
            if axis != "xy" and not split_passes_quality(verts, axis):
                axis = "xy"
        marks[c] = axis
    index = {cid: k for k, cid in enumerate(cell_ids)}
    coarsen = {r.parent_id for r in mesh.refinement_records
               if all(ch in index and ind[index[ch]] < threshold * 0.3 for ch in r.child_ids)}
    levels_dir = cell_directional_levels(mesh, cell_ids, mesh.refinement_records)
    parent_dir = parent_directional_levels(mesh.refinement_records)
    # balance, then the quality gate on marks the balance added, until stable
    while True:
        balance_plan(marks, coarsen, nbrs, cell_ids, levels, levels_dir, parent_dir, maximum_level,
                     mesh.refinement_records, periodic)
        adjusted = False
        for c, axis in tuple(marks.items()):
            final = gated_axis(cell_vertices(mesh, c), axis)
            if final != axis:
                marks[c] = final
                adjusted = True
        if not adjusted:
            return marks, coarsen


# ---------------------------------------------------------------------------
# Adapting and rebuilding
# ---------------------------------------------------------------------------


def adapt_mesh(mesh: FiniteVolumeMesh, marks: dict, coarsen: set, periodic_pairs=()):
    """The adapted mesh and, per new cell, the old cells it covers
    (reference `adapt_mesh`, adaptation.py:680-795)."""
    if mesh.cell_count == 0:
        raise ValueError("Cannot adapt a mesh without cells.")
    cell_ids = _cell_ids(mesh)
    levels = _cell_levels(mesh)
    if len(cell_ids) != mesh.cell_count or len(levels) != mesh.cell_count:
        raise ValueError("Mesh refinement metadata does not match its cells.")
    marks = dict(marks)
    coarsen = set(coarsen)
    periodic = periodic_cell_pairs(mesh, periodic_pairs)
    nbrs = _neighbour_sets(mesh, periodic)
    maximum_level = max((*levels, *(levels[c] + 1 for c in marks)), default=0)
    balance_plan(marks, coarsen, nbrs, cell_ids, levels,
                 cell_directional_levels(mesh, cell_ids, mesh.refinement_records),
                 parent_directional_levels(mesh.refinement_records), maximum_level, mesh.refinement_records,
                 periodic)
    records = {r.parent_id: r for r in mesh.refinement_records}
    curves = _curves_by_edge(mesh)
    verts = all_cell_vertices(mesh)
    if curves is None:
        old = [Leaf(cell_ids[k], levels[k], verts[k], (k,), ()) for k in range(mesh.cell_count)]
    else:
        # only the owners of curved (boundary) faces have a curved edge
        curved = {int(mesh.face_owners[f]) for f, _ in curves.values()}
        old = [Leaf(cell_ids[k], levels[k], verts[k], (k,),
                    cell_edge_curves(mesh, k, curves) if k in curved else (None,) * len(verts[k]))
               for k in range(mesh.cell_count)]
    position = {leaf.cell_id: k for k, leaf in enumerate(old)}
    merged_children: set = set()
    parent_at_first_child: dict = {}
    for parent in sorted(coarsen):
        rec = records.get(parent)
        if rec is None or any(ch not in position for ch in rec.child_ids):
            continue
        kids = [old[position[ch]] for ch in rec.child_ids]
        if any(kid.level != rec.level + 1 for kid in kids):
            continue
        merged_children.update(rec.child_ids)
        parent_at_first_child[rec.child_ids[0]] = Leaf(
            parent, rec.level, rec.vertices, tuple(s for kid in kids for s in kid.sources),
            merged_edge_curves(rec.vertices, kids))
        for key in tuple(records):
            if key == parent or key.startswith(parent + "."):
                del records[key]
    leaves, new_records = [], []
    for k, leaf in enumerate(old):
        if leaf.cell_id in parent_at_first_child:
            leaves.append(parent_at_first_child[leaf.cell_id])
            continue
        if leaf.cell_id in merged_children:
            continue
        axis = marks.get(k)
        if axis is None:
            leaves.append(leaf)
            continue
        axis = gated_axis(leaf.vertices, axis)
        children = split_polygon(leaf.vertices, axis, leaf.curves)
        if len(children) < 2:
            leaves.append(leaf)
            continue
        kid_curves = child_edge_curves(leaf.vertices, leaf.curves, children)
        kid_ids = tuple(f"{leaf.cell_id}.{j}" for j in range(len(children)))
        new_records.append(RefinementRecord(leaf.cell_id, kid_ids, leaf.vertices, leaf.level))
        leaves.extend(Leaf(cid, leaf.level + 1, poly, leaf.sources, cc)
                      for cid, poly, cc in zip(kid_ids, children, kid_curves))
    if not merged_children and not new_records:
        return mesh, tuple((k,) for k in range(mesh.cell_count))
    rebuilt = rebuild_mesh(mesh, leaves, tuple(records.values()) + tuple(new_records))
    lv = np.asarray(rebuilt.cell_levels)
    inner = rebuilt.face_neighbours >= 0
    bad = np.flatnonzero(inner & (np.abs(lv[rebuilt.face_owners] - lv[np.maximum(rebuilt.face_neighbours, 0)]) > 1))
    if len(bad):
        raise ValueError(f"Adapted face {int(bad[0])} violates the 2:1 refinement balance.")
    return rebuilt, tuple(leaf.sources for leaf in leaves)


class _Grid:
    """Uniform buckets of items with an axis-aligned box each, for the
    candidate searches of `rebuild_mesh`; a query returns the items whose
    box meets the query box, in ascending item order, so a search over the
    candidates visits them in the order a full scan would."""

    def __init__(self, boxes, h: float):
        self.h = h
        self.cells = {}
        for i, (x0, x1, y0, y1) in enumerate(boxes):
            for key in self._keys(x0, x1, y0, y1):
                self.cells.setdefault(key, []).append(i)

    def _keys(self, x0, x1, y0, y1):
        h = self.h
        for ix in range(math.floor(x0 / h), math.floor(x1 / h) + 1):
            for iy in range(math.floor(y0 / h), math.floor(y1 / h) + 1):
                yield ix, iy

    def query(self, x0, x1, y0, y1):
        found = set()
        for key in self._keys(x0, x1, y0, y1):
            found.update(self.cells.get(key, ()))
        return sorted(found)


def _segment_box(a, b):
    """A box holding every point `on_segment` accepts for segment ab (TOL
    across it and up to TOL times its length beyond each end), with a
    factor 2 of margin for rounding."""
    e = 2.0 * TOL * (1.0 + math.hypot(b[0] - a[0], b[1] - a[1]))
    return (min(a[0], b[0]) - e, max(a[0], b[0]) + e, min(a[1], b[1]) - e, max(a[1], b[1]) + e)


def _boundary_patch(a, b, boundary_edges, grid) -> int:
    """The patch of the first old boundary edge under the midpoint of ab."""
    mid = (0.5 * (a[0] + b[0]), 0.5 * (a[1] + b[1]))
    for i in grid.query(mid[0], mid[0], mid[1], mid[1]):
        p, q, patch = boundary_edges[i]
        if on_segment(mid, p, q):
            return patch
    return -1


def _edges_with_other_points(P, V, ia, ib, ea, eb, h):
    """Per leaf edge (endpoints ea, eb; their points ia, ib among the
    distinct points P): whether any other point of P lies in the box that
    `_points_on_edge` searches (the edge's range along its main axis, within
    TOL, and its `_segment_box` across). A superset test in NumPy: edges
    without such a point hold only their endpoints."""
    n_e = len(ea)
    if n_e == 0 or len(P) == 0:
        return np.zeros(n_e, dtype=bool)
    d = eb - ea
    main = np.where(np.abs(d[:, 0]) >= np.abs(d[:, 1]), 0, 1)
    e = 2.0 * TOL * (1.0 + np.hypot(d[:, 0], d[:, 1]))
    lo_box = np.minimum(ea, eb) - e[:, None]
    hi_box = np.maximum(ea, eb) + e[:, None]
    rows = np.arange(n_e)
    lo_box[rows, main] = np.minimum(ea, eb)[rows, main] - TOL
    hi_box[rows, main] = np.maximum(ea, eb)[rows, main] + TOL
    cell = np.floor(P / h).astype(np.int64)
    base = cell.min(axis=0)
    width = int(cell[:, 1].max() - base[1]) + 1
    bid = (cell[:, 0] - base[0]) * width + (cell[:, 1] - base[1])
    order = np.argsort(bid, kind="stable")
    sbid = bid[order]
    c0 = np.floor(lo_box / h).astype(np.int64) - base
    c1 = np.floor(hi_box / h).astype(np.int64) - base
    c0[:, 1] = np.maximum(c0[:, 1], 0)
    c1[:, 1] = np.minimum(c1[:, 1], width - 1)
    nx = np.maximum(c1[:, 0] - c0[:, 0] + 1, 0)
    ny = np.maximum(c1[:, 1] - c0[:, 1] + 1, 0)
    cnt = nx * ny
    edge = np.repeat(rows, cnt)
    off = np.arange(int(cnt.sum())) - np.repeat(np.cumsum(cnt) - cnt, cnt)
    bx = c0[edge, 0] + off // np.maximum(ny[edge], 1)
    by = c0[edge, 1] + off % np.maximum(ny[edge], 1)
    q = bx * width + by
    s0 = np.searchsorted(sbid, q, "left")
    s1 = np.searchsorted(sbid, q, "right")
    m = s1 - s0
    pe = np.repeat(edge, m)
    pp = order[np.repeat(s0, m) + (np.arange(int(m.sum())) - np.repeat(np.cumsum(m) - m, m))]
    inside = np.all((P[pp] >= lo_box[pe]) & (P[pp] <= hi_box[pe]), axis=1) & (pp != ia[pe]) & (pp != ib[pe])
    found = np.zeros(n_e, dtype=bool)
    found[pe[inside]] = True
    return found


def _points_on_edge(a, b, known, rank, point_grid):
    """The distinct points on segment ab in order along it: candidates within
    TOL of its range along its main axis, taken in that axis's sorted order,
    kept by `on_segment`, then sorted by their parameter."""
    axis = 0 if abs(b[0] - a[0]) >= abs(b[1] - a[1]) else 1
    lo = min(a[axis], b[axis]) - TOL
    hi = max(a[axis], b[axis]) + TOL
    x0, x1, y0, y1 = _segment_box(a, b)
    if axis == 0:
        x0, x1 = lo, hi
    else:
        y0, y1 = lo, hi
    cand = [i for i in point_grid.query(x0, x1, y0, y1) if lo <= known[i][axis] <= hi]
    cand.sort(key=lambda i: rank[axis][i])
    on_edge = [known[i] for i in cand if on_segment(known[i], a, b)]
    on_edge.sort(key=lambda p: segment_parameter(p, a, b))
    return on_edge


def rebuild_mesh(old: FiniteVolumeMesh, leaves, records) -> FiniteVolumeMesh:
    """Faces, owners and neighbours of the leaves (reference `_rebuild_mesh`,
    adaptation.py:1032-1190): each leaf edge is cut at every mesh point on
    it, faces are numbered at first visit (owner = first visiting leaf), new
    boundary faces take the patch of the old boundary edge under their
    midpoint or, on a curve, the curve's patch and sub-range.

    The edges whose search box holds no point but their own endpoints (all
    but those next to a hanging node) are cut nowhere and are resolved in
    NumPy; the others go through `_points_on_edge`. Points and faces are
    numbered at first visit over the same sequence of sub-edges."""
    sizes_in = [len(leaf.vertices) for leaf in leaves]
    flat = [p for leaf in leaves for p in leaf.vertices]
    memo = {}  # a vertex shared by several leaves is one tuple, keyed once
    keys = [memo[p] if p in memo else memo.setdefault(p, point_key(p)) for p in flat]
    distinct = {}
    vid = np.empty(len(flat), dtype=np.int64)
    for v, key in enumerate(keys):
        vid[v] = distinct.setdefault(key, len(distinct))
    first_of = {}
    for v, key in enumerate(keys):
        first_of.setdefault(key, v)
    known = tuple(flat[first_of[key]] for key in distinct)
    key_index = distinct
    P = np.asarray(known, dtype=np.float64).reshape(-1, 2)
    V = np.asarray(flat, dtype=np.float64).reshape(-1, 2)
    sizes_arr = np.asarray(sizes_in, dtype=np.int64)
    starts = np.cumsum(sizes_arr) - sizes_arr
    leaf_of = np.repeat(np.arange(len(leaves)), sizes_arr)
    va = np.arange(len(flat))
    vb = va + 1
    last = starts + sizes_arr - 1
    vb[last] = starts
    ea, eb = V[va], V[vb]
    lengths = np.hypot(eb[:, 0] - ea[:, 0], eb[:, 1] - ea[:, 1])
    h = max(float(np.sort(lengths)[len(lengths) // 2]) if len(lengths) else 1.0, 1e-9)
    ia, ib = vid[va], vid[vb]
    # an edge needs the exact search when another point may lie on it, when
    # an endpoint's distinct point is not the endpoint itself, or when it is
    # degenerate
    exact = _edges_with_other_points(P, V, ia, ib, ea, eb, h)
    exact |= np.any(P[ia] != ea, axis=1) | np.any(P[ib] != eb, axis=1)
    exact |= ((ea[:, 0] - eb[:, 0]) ** 2 + (ea[:, 1] - eb[:, 1]) ** 2) <= 1e-26
    # the sub-edges in visit order: (leaf edge, first point, second point)
    sub_edge, sub_i, sub_j = [va], [ia], [ib]
    slow = np.flatnonzero(exact)
    if len(slow):
        order_x = sorted(range(len(known)), key=lambda i: known[i][0])
        order_y = sorted(range(len(known)), key=lambda i: known[i][1])
        rank = ([0] * len(known), [0] * len(known))
        for r, i in enumerate(order_x):
            rank[0][i] = r
        for r, i in enumerate(order_y):
            rank[1][i] = r
        point_grid = _Grid([(p[0], p[0], p[1], p[1]) for p in known], h)
        se, si, sj = [], [], []
        for v in slow.tolist():
            a, b = flat[v], flat[vb[v]]
            on_edge = _points_on_edge(a, b, known, rank, point_grid)
            for p, q in zip(on_edge, on_edge[1:]):
                if dist2(p, q) <= 1e-26:
                    continue
                se.append(v)
                si.append(key_index[point_key(p)])
                sj.append(key_index[point_key(q)])
        keep = ~exact
        sub_edge = np.concatenate((va[keep], np.asarray(se, dtype=np.int64)))
        sub_i = np.concatenate((ia[keep], np.asarray(si, dtype=np.int64)))
        sub_j = np.concatenate((ib[keep], np.asarray(sj, dtype=np.int64)))
        # visit order: by leaf edge, then along it (stable)
        o = np.argsort(sub_edge, kind="stable")
        sub_edge, sub_i, sub_j = sub_edge[o], sub_i[o], sub_j[o]
    else:
        sub_edge, sub_i, sub_j = va, ia, ib
    n_sub = len(sub_edge)
    # points numbered at first visit over (i0, j0, i1, j1, ...)
    seq = np.column_stack((sub_i, sub_j)).ravel()
    uniq, first = np.unique(seq, return_index=True)
    pid_of_known = np.full(len(known), -1, dtype=np.int64)
    by_first = np.argsort(first, kind="stable")
    pid_of_known[uniq[by_first]] = np.arange(len(uniq))
    points = P[uniq[by_first]]
    I, J = pid_of_known[sub_i], pid_of_known[sub_j]
    # faces numbered at first visit; owner the first visitor, neighbour the second
    lo_p, hi_p = np.minimum(I, J), np.maximum(I, J)
    fkey = lo_p * (len(points) + 1) + hi_p
    ukey, ffirst, finv, fcount = np.unique(fkey, return_index=True, return_inverse=True, return_counts=True)
    if np.any(fcount > 2):
        raise ValueError("Adaptation produced a non-manifold face.")
    forder = np.argsort(ffirst, kind="stable")
    face_id = np.empty(len(ukey), dtype=np.int64)
    face_id[forder] = np.arange(len(ukey))
    fid = face_id[finv]
    n_faces = len(ukey)
    first_visit = ffirst[forder]  # sub-edge index of each face's creation
    faces = np.column_stack((I[first_visit], J[first_visit]))
    sub_leaf = leaf_of[sub_edge]
    owners = sub_leaf[first_visit]
    neighbours = np.full(n_faces, -1, dtype=np.int64)
    second = np.ones(n_sub, dtype=bool)
    second[first_visit] = False
    neighbours[fid[second]] = sub_leaf[second]
    signs = np.where(second & ~((I == faces[fid, 0]) & (J == faces[fid, 1])), -1, 1)
    # boundary faces: patch (and curve) from the creating sub-edge
    boundary_edges = []
    for f in np.flatnonzero(old.face_neighbours < 0):
        a, b = old.faces[f]
        boundary_edges.append(((float(old.points[a, 0]), float(old.points[a, 1])),
                               (float(old.points[b, 0]), float(old.points[b, 1])), int(old.face_patch_ids[f])))
    edge_grid = None
    patches_of = np.full(n_faces, -1, dtype=np.int64)
    curves_of = [None] * n_faces
    for f in np.flatnonzero(neighbours < 0).tolist():
        e_sub = int(first_visit[f])
        v = int(sub_edge[e_sub])
        c = int(leaf_of[v])
        leaf = leaves[c]
        k = v - int(starts[c])
        n = len(leaf.vertices)
        curve = leaf.curves[k] if len(leaf.curves) == n else None
        p, q = known[int(sub_i[e_sub])], known[int(sub_j[e_sub])]
        if curve is None:
            if edge_grid is None:
                edge_grid = _Grid([_segment_box(a, b) for a, b, _ in boundary_edges], h)
            patches_of[f] = _boundary_patch(p, q, boundary_edges, edge_grid)
        else:
            a, b = leaf.vertices[k], leaf.vertices[(k + 1) % n]
            t0 = segment_parameter(p, a, b)
            t1 = segment_parameter(q, a, b)
            span = curve.end_parameter - curve.start_parameter
            patches_of[f] = curve.patch_id
            curves_of[f] = BoundaryCurve(curve.geometry, curve.start_parameter + t0 * span,
                                         curve.start_parameter + t1 * span, curve.patch_id)
    # per leaf: its sub-edges' faces, signs and first points, in visit order
    per_leaf = np.bincount(sub_leaf, minlength=len(leaves))
    width = int(per_leaf.max()) if len(per_leaf) else 0
    slot = np.arange(n_sub) - np.repeat(np.cumsum(per_leaf) - per_leaf, per_leaf)
    cells = np.full((len(leaves), width), -1, dtype=np.int64)
    cf = np.full((len(leaves), width), -1, dtype=np.int64)
    cs = np.zeros((len(leaves), width), dtype=np.int64)
    cells[sub_leaf, slot] = I
    cf[sub_leaf, slot] = fid
    cs[sub_leaf, slot] = signs
    patches = tuple(MeshPatch(pt.id, pt.name, np.flatnonzero(patches_of == pt.id)) for pt in old.patches)
    return FiniteVolumeMesh(
        points=np.asarray(points, dtype=np.float64).reshape(-1, 2),
        cells=cells,
        cell_sizes=per_leaf.astype(np.int64),
        faces=faces.astype(np.int64).reshape(-1, 2),
        face_owners=owners.astype(np.int64),
        face_neighbours=neighbours,
        face_patch_ids=patches_of,
        patches=patches,
        cell_faces=cf,
        cell_face_signs=cs,
        cell_ids=tuple(leaf.cell_id for leaf in leaves),
        cell_levels=tuple(leaf.level for leaf in leaves),
        refinement_records=tuple(records),
        face_curves=tuple(curves_of) if any(c is not None for c in curves_of) else (),
    )


# ---------------------------------------------------------------------------
# Solution transfer and the solver's indicator variables
# ---------------------------------------------------------------------------


def transfer_cell_values(values, old_areas, new_areas, source_groups) -> np.ndarray:
    """Area-weighted averages of each new cell's source cells, rescaled so
    that every group of new cells sharing a source group keeps the group's
    integral (reference `_transfer_cell_values`, solver.py:1594-1617)."""
    values = np.asarray(values, dtype=float)
    old_areas = np.asarray(old_areas, dtype=float)
    new_areas = np.asarray(new_areas, dtype=float)
    out = np.asarray([np.average(values[np.asarray(g, dtype=int)], axis=0, weights=old_areas[np.asarray(g, dtype=int)])
                      for g in source_groups], dtype=float)
    outputs: dict = {}
    for k, g in enumerate(source_groups):
        outputs.setdefault(tuple(g), []).append(k)
    for g, ks in outputs.items():
        old_area = float(np.sum(old_areas[np.asarray(g, dtype=int)]))
        new_area = float(np.sum(new_areas[np.asarray(ks, dtype=int)]))
        if not math.isfinite(new_area) or new_area <= 0.0:
            raise ValueError("Adapted cells must have positive finite areas.")
        out[ks] *= old_area / new_area
    return out


INDICATOR_VARIABLES = {"Density": 0, "Velocity X": 1, "Velocity Y": 2, "Pressure": 3}


def indicator_values(primitive, variable: str) -> np.ndarray:
    """The adaptation variable of the primitive state (solver.py:1570-1591)."""
    if variable not in INDICATOR_VARIABLES:
        raise ValueError(f"Unsupported adaptive mesh indicator '{variable}'.")
    return np.asarray(primitive)[:, INDICATOR_VARIABLES[variable]]
