"""ORACLE (test infrastructure only): literal restatement of `_build_geometry`.

Follows `src/meshcfd/core/solver.py:1654-2238` statement by statement with
Python loops over cells and faces (so it is only used on small meshes), to pin
the vectorised product builder `meshcfd_amd.geometry.build_geometry` bit for
bit. Periodic matching follows `src/meshcfd/core/periodic.py:70-116`.
"""

from __future__ import annotations

import math

import numpy as np


def _edge_sort_key(edge, axis):
    first, second = edge
    return (
        (first[axis] + second[axis]) / 2.0,
        abs(second[axis] - first[axis]),
        min(first[axis], second[axis]),
        max(first[axis], second[axis]),
    )


def match_periodic_faces(points, faces, patch_faces_a, patch_faces_b):
    """Tangent-sorted pairing of two translated patches (`periodic.py:70-116`)."""
    def edges(ids):
        return [tuple(sorted((tuple(points[faces[f][0]]), tuple(points[faces[f][1]])))) for f in ids]

    ea, eb = edges(patch_faces_a), edges(patch_faces_b)
    pts = [p for e in ea + eb for p in e]

    def bounds(es):
        ps = [p for e in es for p in e]
        return [(min(p[k] for p in ps), max(p[k] for p in ps)) for k in (0, 1)]

    ba, bb = bounds(ea), bounds(eb)
    allb = bounds(ea + eb)
    scale = max(1.0, max(abs(v) for p in pts for v in p), *(hi - lo for lo, hi in allb))
    tol = scale * 1e-9
    for axis in (0, 1):
        ortho = 1 - axis
        if abs(ba[ortho][0] - bb[ortho][0]) > tol or abs(ba[ortho][1] - bb[ortho][1]) > tol:
            continue
        translation = ((bb[axis][0] + bb[axis][1]) - (ba[axis][0] + ba[axis][1])) / 2.0
        if abs(translation) <= tol:
            continue
        shift = (translation, 0.0) if axis == 0 else (0.0, translation)
        moved = [
            tuple(sorted(((a[0] + shift[0], a[1] + shift[1]), (b[0] + shift[0], b[1] + shift[1]))))
            for a, b in ea
        ]
        oa = sorted(zip(patch_faces_a, moved), key=lambda it: _edge_sort_key(it[1], ortho))
        ob = sorted(zip(patch_faces_b, eb), key=lambda it: _edge_sort_key(it[1], ortho))
        return [(fa, fb) for (fa, _), (fb, _) in zip(oa, ob)], shift
    raise ValueError("no periodic translation")


def build_geometry(mesh, periodic_pairs):
    """Per-cell/per-face loops of `solver.py:1655-1700, 2004-2105`.

    ``periodic_pairs`` is a list of (patch_a_id, patch_b_id). Returns a dict
    of the arrays the residual consumes.
    """
    points = [tuple(map(float, p)) for p in mesh.points]
    cells = [tuple(int(v) for v in mesh.cells[i, : mesh.cell_sizes[i]]) for i in range(mesh.cell_count)]
    faces = [tuple(int(v) for v in f) for f in mesh.faces]
    centers, areas = [], []
    for node_ids in cells:
        pts = [points[k] for k in node_ids]
        twice = sum(x0 * y1 - x1 * y0 for (x0, y0), (x1, y1) in zip(pts, pts[1:] + pts[:1]))
        xs = 0.0
        ys = 0.0
        for (x0, y0), (x1, y1) in zip(pts, pts[1:] + pts[:1]):
            cross = x0 * y1 - x1 * y0
            xs += (x0 + x1) * cross
            ys += (y0 + y1) * cross
        centers.append((xs / (3.0 * twice), ys / (3.0 * twice)))
        areas.append(0.5 * abs(twice))
    centers = np.asarray(centers)
    areas = np.asarray(areas)
    face_centers, normals, lengths = [], [], []
    owners = [int(v) for v in mesh.face_owners]
    neighbours = [int(v) for v in mesh.face_neighbours]
    for f, (a, b) in enumerate(faces):
        ax, ay = points[a]
        bx, by = points[b]
        ex, ey = bx - ax, by - ay
        length = math.sqrt(ex * ex + ey * ey)
        n = [ey / length, -ex / length]
        c = (0.5 * (ax + bx), 0.5 * (ay + by))
        o = owners[f]
        if n[0] * (c[0] - centers[o][0]) + n[1] * (c[1] - centers[o][1]) < 0.0:
            n = [-n[0], -n[1]]
        face_centers.append(c)
        normals.append(n)
        lengths.append(length)
    patches = {p.id: [int(f) for f in p.face_ids] for p in mesh.patches}
    mappings = []
    for pa, pb in periodic_pairs:
        matched, shift = match_periodic_faces(points, faces, patches[pa], patches[pb])
        mappings.extend((fa, fb, shift) for fa, fb in matched)
    interior = [f for f in range(len(faces)) if neighbours[f] >= 0]
    src, nbr, gdx, gdy = [], [], [], []
    for f in interior:
        o, nb = owners[f], neighbours[f]
        dx = centers[nb][0] - centers[o][0]
        dy = centers[nb][1] - centers[o][1]
        src += [o, nb]
        nbr += [nb, o]
        gdx += [dx, -dx]
        gdy += [dy, -dy]
    for fa, fb, shift in mappings:
        oa, ob = owners[fa], owners[fb]
        dx = (centers[ob][0] - shift[0]) - centers[oa][0]
        dy = (centers[ob][1] - shift[1]) - centers[oa][1]
        src += [oa, ob]
        nbr += [ob, oa]
        gdx += [dx, -dx]
        gdy += [dy, -dy]
    n = len(cells)
    xx = [0.0] * n
    xy = [0.0] * n
    yy = [0.0] * n
    wx, wy = [], []
    for s, dx, dy in zip(src, gdx, gdy):
        w = 1.0 / max(dx * dx + dy * dy, 1e-300)
        wx.append(w * dx)
        wy.append(w * dy)
        xx[s] += (w * dx) * dx
        xy[s] += (w * dx) * dy
        yy[s] += (w * dy) * dy
    inv = []
    for i in range(n):
        det = xx[i] * yy[i] - xy[i] * xy[i]
        if det > 1e-14 * max(xx[i] * yy[i], 1e-300):
            inv.append((yy[i] / det, -xy[i] / det, xx[i] / det))
        else:
            inv.append((0.0, 0.0, 0.0))
    cx = [inv[s][0] * a + inv[s][1] * b for s, a, b in zip(src, wx, wy)]
    cy = [inv[s][1] * a + inv[s][2] * b for s, a, b in zip(src, wx, wy)]
    lim_cells = list(owners) + [neighbours[f] for f in interior]
    lim_dx = [face_centers[f][0] - centers[owners[f]][0] for f in range(len(faces))] + [
        face_centers[f][0] - centers[neighbours[f]][0] for f in interior
    ]
    lim_dy = [face_centers[f][1] - centers[owners[f]][1] for f in range(len(faces))] + [
        face_centers[f][1] - centers[neighbours[f]][1] for f in interior
    ]
    return {
        "centers": centers,
        "areas": areas,
        "face_centers": np.asarray(face_centers),
        "face_normals": np.asarray(normals),
        "face_lengths": np.asarray(lengths),
        "periodic_faces_a": np.asarray([m[0] for m in mappings], dtype=int),
        "periodic_faces_b": np.asarray([m[1] for m in mappings], dtype=int),
        "gradient_sources": np.asarray(src, dtype=int),
        "gradient_neighbours": np.asarray(nbr, dtype=int),
        "gradient_coefficients_x": np.asarray(cx),
        "gradient_coefficients_y": np.asarray(cy),
        "limiter_cells": np.asarray(lim_cells, dtype=int),
        "limiter_dx": np.asarray(lim_dx),
        "limiter_dy": np.asarray(lim_dy),
    }
