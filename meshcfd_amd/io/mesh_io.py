"""MeshCAD finite-volume mesh JSON, format "meshcad-mesh" version 1: reader
and writer (reference `src/meshcfd/io/mesh_io.py:18-19,133-512`).

The file lists node coordinates, faces as node pairs with owner, neighbour
(-1 on the boundary) and patch (-1 inside), cells as face lists with signs
(+1: the face's node order runs along the cell, -1: against it), named
patches with their faces, and optionally per-face boundary curves and
adaptive-refinement metadata. Reading checks the same conditions as the
reference, with the same messages, and turns each cell's signed faces into
its closed node loop.

Arrays instead of Python tuples: the checks and the node loops are
vectorised (successor edges followed with NumPy), so meshes of millions of
cells load in seconds; only cells whose loop the fast path cannot decide
(a node that starts two of the cell's edges) take the reference's greedy walk.
"""

from __future__ import annotations

import json
import math
from pathlib import Path
from typing import Any

import numpy as np

from ..curves import evaluate_boundary_curve
from ..model import BoundaryCurve, FiniteVolumeMesh, MeshPatch, RefinementRecord

FORMAT_NAME = "meshcad-mesh"
FORMAT_VERSION = 1
NO_NEIGHBOUR = -1


class MeshImportError(ValueError):
    """Raised when a mesh file is unsupported or invalid."""


def _is_int(value) -> bool:
    return isinstance(value, int) and not isinstance(value, bool)


def _int_array(value: Any, description: str) -> np.ndarray:
    if not isinstance(value, list):
        raise MeshImportError(f"{description} must be a list.")
    if not all(_is_int(v) for v in value):
        raise MeshImportError(f"{description} must be an integer.")
    return np.asarray(value, dtype=np.int64)


def _finite(value: Any, description: str) -> float:
    if isinstance(value, bool):
        raise MeshImportError(f"{description} must be a finite number.")
    try:
        number = float(value)
    except (OverflowError, TypeError, ValueError) as error:
        raise MeshImportError(f"{description} must be a finite number.") from error
    if not math.isfinite(number):
        raise MeshImportError(f"{description} must be a finite number.")
    return number


def _curve_geometry(value: Any, description: str) -> dict[str, Any]:
    """A validated copy of a curve descriptor (reference `mesh_io.py:51-130`)."""
    if not isinstance(value, dict):
        raise MeshImportError(f"{description} must be an object.")
    kind = value.get("type")
    if kind == "line":
        out = {"type": kind}
        for key in ("start", "end"):
            point = value.get(key)
            if not isinstance(point, list) or len(point) != 2:
                raise MeshImportError(f"{description} has invalid {key} coordinates.")
            out[key] = [_finite(v, f"{description} {key} coordinate") for v in point]
        return out
    if kind == "arc":
        center = value.get("center")
        if not isinstance(center, list) or len(center) != 2:
            raise MeshImportError(f"{description} has an invalid arc center.")
        radius = _finite(value.get("radius"), f"{description} arc radius")
        if radius <= 0.0:
            raise MeshImportError(f"{description} arc radius must be positive.")
        return {
            "type": kind,
            "center": [_finite(v, f"{description} arc center") for v in center],
            "radius": radius,
            "start_angle": _finite(value.get("start_angle"), f"{description} arc start angle"),
            "end_angle": _finite(value.get("end_angle"), f"{description} arc end angle"),
        }
    if kind in ("spline", "spline_segment"):
        raw = value.get("points")
        if not isinstance(raw, list) or len(raw) < 2 or any(not isinstance(p, list) or len(p) != 2 for p in raw):
            raise MeshImportError(f"{description} has invalid spline points.")
        points = [[_finite(v, f"{description} spline point") for v in p] for p in raw]
        if kind == "spline":
            return {"type": kind, "points": points}
        a = _finite(value.get("start_t"), f"{description} spline start parameter")
        b = _finite(value.get("end_t"), f"{description} spline end parameter")
        if not 0.0 <= a < b <= 1.0:
            raise MeshImportError(f"{description} has invalid spline parameters.")
        out = {"type": kind, "points": points, "start_t": a, "end_t": b}
        for key in ("start_offset", "end_offset"):
            off = value.get(key, [0.0, 0.0])
            if not isinstance(off, list) or len(off) != 2:
                raise MeshImportError(f"{description} has an invalid {key}.")
            out[key] = [_finite(v, f"{description} {key}") for v in off]
        return out
    raise MeshImportError(f"{description} has an unsupported geometry type.")


def _coordinates(raw: Any) -> np.ndarray:
    if not isinstance(raw, list) or not raw:
        raise MeshImportError("Mesh must contain at least one node.")
    try:
        pts = np.asarray(raw, dtype=np.float64)
        ok = pts.ndim == 2 and pts.shape[1] == 2 and all(isinstance(c, list) for c in raw)
    except (TypeError, ValueError, OverflowError):
        ok = False
    if not ok:  # find the first bad node for the message
        for index, c in enumerate(raw):
            if not isinstance(c, list) or len(c) != 2:
                raise MeshImportError(f"Node {index} must have two coordinates.")
            try:
                float(c[0]), float(c[1])
            except (TypeError, ValueError, OverflowError) as error:
                raise MeshImportError(f"Node {index} has invalid coordinates.") from error
        raise MeshImportError("Mesh node coordinates are invalid.")
    bad = ~np.isfinite(pts).all(axis=1)
    if bad.any():
        raise MeshImportError(f"Node {int(np.flatnonzero(bad)[0])} coordinates must be finite.")
    return pts


def _faces(raw: Any, n_points: int) -> np.ndarray:
    if not isinstance(raw, list) or not raw:
        raise MeshImportError("Mesh must contain at least one face.")
    for face_id, nodes in enumerate(raw):
        if not isinstance(nodes, list) or len(nodes) != 2:
            raise MeshImportError(f"Face {face_id} must have exactly two nodes.")
        if not (_is_int(nodes[0]) and _is_int(nodes[1])):
            raise MeshImportError(f"Face {face_id} node must be an integer.")
    faces = np.asarray(raw, dtype=np.int64)
    bad = (faces[:, 0] == faces[:, 1]) | (faces < 0).any(axis=1) | (faces >= n_points).any(axis=1)
    if bad.any():
        raise MeshImportError(f"Face {int(np.flatnonzero(bad)[0])} references invalid nodes.")
    return faces


def _node_loops(faces: np.ndarray, cell_faces: list, cell_signs: list, n_faces: int):
    """Checked cell face lists -> (padded faces, padded signs, padded node
    loops, sizes). A cell's loop starts at its first face's start node and
    follows, from each edge's end node, the edge that starts there."""
    n_cells = len(cell_faces)
    sizes = np.empty(n_cells, dtype=np.int64)
    for cell_id, (ids, signs) in enumerate(zip(cell_faces, cell_signs)):
        _int_array(ids, f"Cell {cell_id} faces")
        _int_array(signs, f"Cell {cell_id} face signs")
        if len(ids) < 3 or len(ids) != len(signs):
            raise MeshImportError(f"Cell {cell_id} must have at least three faces.")
        sizes[cell_id] = len(ids)
    flat_f = np.fromiter((f for ids in cell_faces for f in ids), dtype=np.int64, count=int(sizes.sum()))
    flat_s = np.fromiter((v for sg in cell_signs for v in sg), dtype=np.int64, count=int(sizes.sum()))
    cell_of = np.repeat(np.arange(n_cells), sizes)
    bad = (flat_f < 0) | (flat_f >= n_faces)
    if bad.any():
        raise MeshImportError(f"Cell {int(cell_of[np.flatnonzero(bad)[0]])} references invalid faces.")
    bad = (flat_s != 1) & (flat_s != -1)
    if bad.any():
        raise MeshImportError(f"Cell {int(cell_of[np.flatnonzero(bad)[0]])} face signs must be -1 or 1.")
    start = np.where(flat_s == 1, faces[flat_f, 0], faces[flat_f, 1])
    end = np.where(flat_s == 1, faces[flat_f, 1], faces[flat_f, 0])
    first = np.concatenate(([0], np.cumsum(sizes)[:-1]))
    # successor of each edge: the edge of the same cell starting at its end node
    key_start = cell_of * (faces.max() + 1) + start
    order = np.argsort(key_start, kind="stable")
    ks = key_start[order]
    dup = np.zeros(len(ks), dtype=bool)
    dup[1:] = ks[1:] == ks[:-1]
    complex_cells = set(np.unique(cell_of[order[dup]]).tolist()) if dup.any() else set()
    key_end = cell_of * (faces.max() + 1) + end
    pos = np.searchsorted(ks, key_end)
    found = (pos < len(ks)) & (ks[np.minimum(pos, len(ks) - 1)] == key_end)
    succ = np.where(found, order[np.minimum(pos, len(ks) - 1)], -1)
    width = int(sizes.max())
    chain = np.full((n_cells, width), -1, dtype=np.int64)
    cur = first.copy()
    ok = np.ones(n_cells, dtype=bool)
    chain[:, 0] = cur
    for k in range(1, width):
        active = k < sizes
        nxt = np.where(active, succ[cur], cur)
        ok &= ~(active & (nxt < 0))
        nxt = np.where(nxt < 0, cur, nxt)
        chain[:, k] = np.where(active, nxt, -1)
        cur = nxt
    # every edge used once and the loop closes on the first edge
    closes = succ[cur] == first
    used = np.sort(np.where(chain >= 0, chain, np.iinfo(np.int64).max), axis=1)
    distinct = np.all((used[:, 1:] != used[:, :-1]) | (used[:, 1:] == np.iinfo(np.int64).max), axis=1)
    ok &= closes & distinct
    loops = np.full((n_cells, width), -1, dtype=np.int64)
    valid = chain >= 0
    loops[valid] = start[chain[valid]]
    for c in sorted(complex_cells | set(np.flatnonzero(~ok).tolist())):
        # the reference's greedy walk (`mesh_io.py:322-344`)
        a, n = int(first[c]), int(sizes[c])
        edges = [(int(start[a + i]), int(end[a + i])) for i in range(n)]
        ordered = [edges.pop(0)]
        while edges:
            nxt = next((i for i, e in enumerate(edges) if e[0] == ordered[-1][1]), None)
            if nxt is None:
                break
            ordered.append(edges.pop(nxt))
        if edges or ordered[-1][1] != ordered[0][0]:
            raise MeshImportError(f"Cell {c} faces do not form a closed node loop.")
        loops[c, :n] = [e[0] for e in ordered]
    pad_f = np.full((n_cells, width), -1, dtype=np.int64)
    pad_s = np.zeros((n_cells, width), dtype=np.int64)
    col = np.arange(len(flat_f)) - np.repeat(first, sizes)
    pad_f[cell_of, col] = flat_f
    pad_s[cell_of, col] = flat_s
    return pad_f, pad_s, loops, sizes


def mesh_from_dict(data: dict[str, Any]) -> FiniteVolumeMesh:
    if data.get("format") != FORMAT_NAME or data.get("version") != FORMAT_VERSION:
        raise MeshImportError(
            f"Unsupported mesh format. Expected MeshCAD finite-volume mesh JSON version {FORMAT_VERSION}."
        )
    if data.get("dimension") != 2:
        raise MeshImportError("Only two-dimensional meshes are supported.")
    try:
        coordinates = data["nodes"]["coordinates"]
        face_nodes = data["faces"]["nodes"]
        owners_raw = data["faces"]["owner"]
        neighbours_raw = data["faces"]["neighbour"]
        patch_raw = data["faces"]["patch"]
        cell_faces = data["cells"]["faces"]
        cell_signs = data["cells"]["face_signs"]
        raw_patches = data["patches"]
    except (KeyError, TypeError) as error:
        raise MeshImportError(f"Mesh file is missing required data: {error}.") from error

    points = _coordinates(coordinates)
    faces = _faces(face_nodes, len(points))
    owners = _int_array(owners_raw, "Face owners")
    neighbours = _int_array(neighbours_raw, "Face neighbours")
    patch_ids = _int_array(patch_raw, "Face patch identifiers")
    if not len(faces) == len(owners) == len(neighbours) == len(patch_ids):
        raise MeshImportError("Face connectivity arrays have inconsistent lengths.")

    if not isinstance(raw_patches, list):
        raise MeshImportError("Patches must be a list.")
    patches, seen_ids, seen_names = [], set(), set()
    for item in raw_patches:
        if not isinstance(item, dict):
            raise MeshImportError("Each patch must be a JSON object.")
        pid = item.get("id")
        if not _is_int(pid):
            raise MeshImportError("Patch identifier must be an integer.")
        name = item.get("name")
        if not isinstance(name, str) or not name.strip():
            raise MeshImportError(f"Patch {pid} must have a non-empty name.")
        if pid in seen_ids or name in seen_names:
            raise MeshImportError("Patch identifiers and names must be unique.")
        ids = _int_array(item.get("faces"), f"Patch {name} faces")
        if ((ids < 0) | (ids >= len(faces))).any():
            raise MeshImportError(f"Patch {name} references invalid faces.")
        patches.append(MeshPatch(pid, name, ids))
        seen_ids.add(pid)
        seen_names.add(name)

    if not isinstance(cell_faces, list) or not isinstance(cell_signs, list):
        raise MeshImportError("Cell connectivity must be represented by lists.")
    if not cell_faces or len(cell_faces) != len(cell_signs):
        raise MeshImportError("Mesh must have cells with consistent face connectivity.")
    n_cells = len(cell_faces)
    known = np.asarray(sorted(seen_ids), dtype=np.int64)
    unknown = (patch_ids != NO_NEIGHBOUR) & ~np.isin(patch_ids, known)
    if unknown.any():
        raise MeshImportError("A face references an unknown boundary patch.")
    bad_owner = (owners < 0) | (owners >= n_cells)
    bad_nb = (neighbours != NO_NEIGHBOUR) & ((neighbours < 0) | (neighbours >= n_cells) | (neighbours == owners))
    bad_bd = (neighbours == NO_NEIGHBOUR) != (patch_ids != NO_NEIGHBOUR)
    first_bad = np.flatnonzero(bad_owner | bad_nb | bad_bd)
    if len(first_bad):
        f = int(first_bad[0])
        if bad_owner[f]:
            raise MeshImportError(f"Face {f} references an invalid owner cell.")
        if bad_nb[f]:
            raise MeshImportError(f"Face {f} references an invalid neighbour cell.")
        raise MeshImportError(f"Face {f} has inconsistent boundary data.")

    face_curves = ()
    raw_curves = data.get("boundary_curves")
    if raw_curves is not None:
        if not isinstance(raw_curves, list) or len(raw_curves) != len(faces):
            raise MeshImportError("Boundary curves must contain one entry per face.")
        parsed = []
        for face_id, raw in enumerate(raw_curves):
            if raw is None:
                parsed.append(None)
                continue
            parsed.append(_boundary_curve(raw, face_id, faces, points, neighbours, patch_ids))
        face_curves = tuple(parsed)

    pad_f, pad_s, loops, sizes = _node_loops(faces, cell_faces, cell_signs, len(faces))

    listed = np.concatenate([p.face_ids for p in patches]) if patches else np.empty(0, np.int64)
    for p in patches:
        if (patch_ids[p.face_ids] != p.id).any():
            raise MeshImportError(f"Patch {p.name} contains an unrelated face.")
    if len(np.unique(listed)) != len(listed):
        raise MeshImportError("A boundary face cannot belong to multiple patches.")
    member = np.zeros(len(faces), dtype=bool)
    member[listed] = True
    bad = (patch_ids != NO_NEIGHBOUR) != member
    if bad.any():
        raise MeshImportError(f"Face {int(np.flatnonzero(bad)[0])} has inconsistent patch membership.")

    cell_ids, cell_levels, records = _refinement(data.get("refinement", {}), n_cells)
    return FiniteVolumeMesh(
        points=points,
        cells=loops,
        cell_sizes=sizes,
        faces=faces,
        face_owners=owners,
        face_neighbours=neighbours,
        face_patch_ids=patch_ids,
        patches=tuple(patches),
        cell_faces=pad_f,
        cell_face_signs=pad_s,
        cell_ids=cell_ids,
        cell_levels=cell_levels,
        refinement_records=records,
        face_curves=face_curves,
    )


def _boundary_curve(raw, face_id, faces, points, neighbours, patch_ids) -> BoundaryCurve:
    if not isinstance(raw, dict):
        raise MeshImportError(f"Boundary curve {face_id} must be an object.")
    if neighbours[face_id] != NO_NEIGHBOUR:
        raise MeshImportError(f"Internal face {face_id} cannot have a boundary curve.")
    pid = raw.get("patch_id")
    if not _is_int(pid):
        raise MeshImportError(f"Boundary curve {face_id} patch must be an integer.")
    if pid != patch_ids[face_id]:
        raise MeshImportError(f"Boundary curve {face_id} has an inconsistent patch.")
    params = raw.get("parameters")
    if not isinstance(params, list) or len(params) != 2:
        raise MeshImportError(f"Boundary curve {face_id} must have two parameters.")
    a, b = (_finite(v, f"Boundary curve {face_id} parameter") for v in params)
    if not (0.0 <= a <= 1.0 and 0.0 <= b <= 1.0 and a != b):
        raise MeshImportError(f"Boundary curve {face_id} has invalid parameters.")
    geometry = _curve_geometry(raw.get("geometry"), f"Boundary curve {face_id}")
    nodes = faces[face_id]
    tol = 1.0e-8 * max(1.0, float(np.max(np.abs(points[nodes]))))
    for node, t in zip(nodes, (a, b)):
        try:
            x, y = evaluate_boundary_curve(geometry, t)
        except (ArithmeticError, KeyError, TypeError, ValueError) as error:
            raise MeshImportError(f"Boundary curve {face_id} cannot be evaluated.") from error
        if abs(x - points[node, 0]) > tol or abs(y - points[node, 1]) > tol:
            raise MeshImportError(f"Boundary curve {face_id} does not match its face nodes.")
    return BoundaryCurve(geometry, a, b, int(pid))


def _refinement(raw: Any, n_cells: int):
    if not isinstance(raw, dict):
        raise MeshImportError("Refinement metadata must be an object.")
    ids = raw.get("cell_ids", [])
    levels = raw.get("cell_levels", [])
    if not isinstance(ids, list) or not isinstance(levels, list):
        raise MeshImportError("Cell refinement identifiers and levels must be lists.")
    if not all(isinstance(c, str) and c for c in ids):
        raise MeshImportError("Cell refinement identifiers must be non-empty strings.")
    levels = _int_array(levels, "Cell refinement levels").tolist()
    if (ids or levels) and not len(ids) == len(levels) == n_cells:
        raise MeshImportError("Cell refinement metadata has inconsistent lengths.")
    if len(set(ids)) != len(ids) or any(v < 0 for v in levels):
        raise MeshImportError("Cell refinement metadata contains duplicate or negative values.")
    raw_records = raw.get("records", [])
    if not isinstance(raw_records, list):
        raise MeshImportError("Refinement records must be a list.")
    if raw_records and not ids:
        raise MeshImportError("Refinement records require cell identifiers and levels.")
    records, parents = [], set()
    for index, r in enumerate(raw_records):
        if not isinstance(r, dict):
            raise MeshImportError(f"Refinement record {index} must be an object.")
        parent, children, vertices, level = r.get("parent_id"), r.get("child_ids"), r.get("vertices"), r.get("level")
        if not _is_int(level):
            raise MeshImportError("Refinement parent level must be an integer.")
        if (not isinstance(parent, str) or not parent or parent in parents or not isinstance(children, list)
                or len(children) < 2 or any(not isinstance(c, str) or not c for c in children)
                or len(set(children)) != len(children) or not isinstance(vertices, list) or len(vertices) < 3
                or level < 0):
            raise MeshImportError(f"Refinement record {index} is invalid.")
        parents.add(parent)
        verts = []
        for v in vertices:
            if not isinstance(v, list) or len(v) != 2:
                raise MeshImportError(f"Refinement record {index} has invalid vertices.")
            try:
                p = (float(v[0]), float(v[1]))
            except (OverflowError, TypeError, ValueError) as error:
                raise MeshImportError(f"Refinement record {index} has invalid vertices.") from error
            if not all(math.isfinite(x) for x in p):
                raise MeshImportError(f"Refinement record {index} has non-finite vertices.")
            verts.append(p)
        records.append(RefinementRecord(parent, tuple(children), tuple(verts), level))
    return tuple(ids), tuple(levels), tuple(records)


def load_mesh(path: str | Path) -> FiniteVolumeMesh:
    source = Path(path)
    try:
        with source.open("r", encoding="utf-8") as stream:
            data = json.load(stream)
    except (OSError, UnicodeError, json.JSONDecodeError) as error:
        raise MeshImportError(f"Could not read mesh file '{source}': {error}") from error
    if not isinstance(data, dict):
        raise MeshImportError("Mesh file must contain a JSON object.")
    return mesh_from_dict(data)


def cell_face_lists(mesh: FiniteVolumeMesh) -> tuple[np.ndarray, np.ndarray]:
    """(padded face ids, signs) of each cell in its node-loop order: the face
    along edge (v_i, v_i+1) with sign +1 if the cell owns it (its node order
    is the owner's), -1 if it is the face's neighbour. The lists a file
    carried are returned as they were."""
    if mesh.cell_faces is not None and mesh.cell_face_signs is not None:
        return mesh.cell_faces, mesh.cell_face_signs
    n_cells, width = mesh.cells.shape
    sizes = mesh.cell_sizes
    n_nodes = len(mesh.points)
    col = np.arange(width)
    active = col[None, :] < sizes[:, None]
    a = mesh.cells
    b = np.take_along_axis(a, np.where(col[None, :] + 1 < sizes[:, None], col[None, :] + 1, 0), axis=1)
    cells = np.broadcast_to(np.arange(n_cells)[:, None], a.shape)
    # directed edges owned by a cell: (owner, n0, n1); by a neighbour: (neighbour, n1, n0)
    f = mesh.faces
    nf = len(f)
    fid = np.arange(nf)
    inner = mesh.face_neighbours >= 0
    key_cell = np.concatenate((mesh.face_owners, mesh.face_neighbours[inner]))
    key_a = np.concatenate((f[:, 0], f[inner, 1]))
    key_b = np.concatenate((f[:, 1], f[inner, 0]))
    key_face = np.concatenate((fid, fid[inner]))
    key_sign = np.concatenate((np.ones(nf, np.int64), -np.ones(int(inner.sum()), np.int64)))
    # a directed edge (n_a -> n_b) lies on the left of exactly one cell, so
    # n_a * N + n_b identifies it (a key with the cell in it would overflow
    # 64 bits at a few million cells); the cell is checked after the lookup
    keys = key_a * n_nodes + key_b
    order = np.argsort(keys, kind="stable")
    ks = keys[order]
    if len(ks) > 1 and np.any(ks[1:] == ks[:-1]):
        raise ValueError("Two faces run along the same directed edge.")
    q = a[active] * n_nodes + b[active]
    pos = np.searchsorted(ks, q)
    hit = (pos < len(ks)) & (ks[np.minimum(pos, len(ks) - 1)] == q)
    hit &= key_cell[order[np.minimum(pos, len(ks) - 1)]] == cells[active]
    if not hit.all():
        bad = int(cells[active][np.flatnonzero(~hit)[0]])
        raise ValueError(f"Cell {bad} has an edge that is not one of its faces.")
    faces = np.full(a.shape, -1, dtype=np.int64)
    signs = np.zeros(a.shape, dtype=np.int64)
    faces[active] = key_face[order[pos]]
    signs[active] = key_sign[order[pos]]
    return faces, signs


def mesh_to_dict(mesh: FiniteVolumeMesh) -> dict[str, Any]:
    """The MeshCAD JSON object of a mesh (reference `mesh_io.py:454-512`)."""
    faces, signs = cell_face_lists(mesh)
    sizes = np.asarray(mesh.cell_sizes)
    data = {
        "format": FORMAT_NAME,
        "version": FORMAT_VERSION,
        "dimension": 2,
        "nodes": {"coordinates": np.asarray(mesh.points, dtype=np.float64).tolist()},
        "faces": {
            "nodes": np.asarray(mesh.faces).tolist(),
            "owner": np.asarray(mesh.face_owners).tolist(),
            "patch": np.asarray(mesh.face_patch_ids).tolist(),
            "neighbour": np.asarray(mesh.face_neighbours).tolist(),
        },
        "cells": {
            "faces": [row[:n] for row, n in zip(faces.tolist(), sizes.tolist())],
            "face_signs": [row[:n] for row, n in zip(signs.tolist(), sizes.tolist())],
        },
        "patches": [{"id": p.id, "name": p.name, "faces": np.asarray(p.face_ids).tolist()} for p in mesh.patches],
    }
    if mesh.cell_ids or mesh.cell_levels or mesh.refinement_records:
        data["refinement"] = {
            "cell_ids": list(mesh.cell_ids),
            "cell_levels": list(mesh.cell_levels),
            "records": [
                {"parent_id": r.parent_id, "child_ids": list(r.child_ids),
                 "vertices": [list(v) for v in r.vertices], "level": r.level}
                for r in mesh.refinement_records
            ],
        }
    if mesh.face_curves:
        if len(mesh.face_curves) != len(mesh.faces):
            raise ValueError("Boundary curve metadata does not match mesh faces.")
        data["boundary_curves"] = [
            None if c is None else {"geometry": c.geometry, "parameters": [c.start_parameter, c.end_parameter],
                                    "patch_id": c.patch_id}
            for c in mesh.face_curves
        ]
    return data


def save_mesh(path: str | Path, mesh: FiniteVolumeMesh) -> Path:
    destination = Path(path)
    destination.parent.mkdir(parents=True, exist_ok=True)
    with destination.open("w", encoding="utf-8") as stream:
        json.dump(mesh_to_dict(mesh), stream)
    return destination
