"""MeshCAD boundary-curve descriptors evaluated at a normalised parameter
(reference `src/meshcfd/core/curves.py:9-80`): straight lines, circular arcs
(counter-clockwise from the start angle; a non-positive span wraps by 2 pi),
Catmull-Rom splines through their points with uniform parameter spacing and
clamped end tangents, and spline segments (a sub-range of a spline with
linearly blended end offsets)."""

from __future__ import annotations

import math


def catmull_rom(points, t: float) -> tuple[float, float]:
    """Point at t in [0, 1] of the uniform Catmull-Rom spline through
    ``points`` (a straight segment for two points). Segment k covers
    [k/(n-1), (k+1)/(n-1)]; the controls before the first and after the last
    point repeat the end points."""
    n = len(points)
    if n == 2:
        (x0, y0), (x1, y1) = points
        return x0 + (x1 - x0) * t, y0 + (y1 - y0) * t
    segments = n - 1
    s = t * segments
    k = min(int(s), segments - 1)
    u = s - k
    u2 = u * u
    u3 = u2 * u
    p0, p1, p2, p3 = points[max(0, k - 1)], points[k], points[k + 1], points[min(k + 2, n - 1)]
    out = []
    for a in (0, 1):
        c0, c1, c2, c3 = p0[a], p1[a], p2[a], p3[a]
        # 0.5 * (2 p1 + (p2 - p0) u + (2 p0 - 5 p1 + 4 p2 - p3) u^2 + (-p0 + 3 p1 - 3 p2 + p3) u^3)
        out.append(0.5 * (2.0 * c1 + (-c0 + c2) * u + (2.0 * c0 - 5.0 * c1 + 4.0 * c2 - c3) * u2
                          + (-c0 + 3.0 * c1 - 3.0 * c2 + c3) * u3))
    return out[0], out[1]


def evaluate_boundary_curve(geometry: dict, parameter: float) -> tuple[float, float]:
    """(x, y) of a curve descriptor at ``parameter`` (clamped to [0, 1])."""
    t = min(1.0, max(0.0, parameter))
    kind = geometry.get("type")
    if kind == "line":
        (x0, y0), (x1, y1) = geometry["start"], geometry["end"]
        return x0 + (x1 - x0) * t, y0 + (y1 - y0) * t
    if kind == "arc":
        start = geometry["start_angle"]
        span = geometry["end_angle"] - start
        while span <= 0.0:
            span += 2.0 * math.pi
        angle = start + span * t
        cx, cy = geometry["center"]
        r = geometry["radius"]
        return cx + r * math.cos(angle), cy + r * math.sin(angle)
    if kind == "spline":
        return catmull_rom(geometry["points"], t)
    if kind == "spline_segment":
        a, b = geometry["start_t"], geometry["end_t"]
        x, y = catmull_rom(geometry["points"], a + (b - a) * t)
        so = geometry.get("start_offset", (0.0, 0.0))
        eo = geometry.get("end_offset", (0.0, 0.0))
        return x + so[0] * (1.0 - t) + eo[0] * t, y + so[1] * (1.0 - t) + eo[1] * t
    raise ValueError(f"Unsupported boundary curve type: {kind!r}.")
