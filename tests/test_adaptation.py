"""Mesh adaptation (host side): the reference's known answers
(tests/test_adaptation.py and tests/test_curved_adaptation.py re-expressed on
`meshcfd_amd.adaptation`), plus invariants every adapted mesh must keep:
areas preserved, every interior face shared by exactly two cells whose face
loops close, 2:1 balance, conservative transfer."""

import math

import numpy as np
import pytest

from meshcfd_amd import adaptation as ad
from meshcfd_amd.curves import evaluate_boundary_curve
from meshcfd_amd.geometry import match_periodic_faces
from meshcfd_amd.io.mesh_io import mesh_from_dict, mesh_to_dict
from meshcfd_amd.model import (
    BoundaryCondition,
    BoundaryCurve,
    Case,
    FiniteVolumeMesh,
    MeshPatch,
    PeriodicBoundaryPair,
    RefinementRecord,
)


def two_cell_mesh():
    """Reference tests/test_adaptation.py:14-55."""
    return mesh_from_dict({
        "format": "meshcad-mesh", "version": 1, "dimension": 2,
        "nodes": {"coordinates": [[0, 0], [1, 0], [2, 0], [0, 1], [1, 1], [2, 1]]},
        "faces": {"nodes": [[0, 1], [1, 4], [4, 3], [3, 0], [1, 2], [2, 5], [5, 4]],
                  "owner": [0, 0, 0, 0, 1, 1, 1], "patch": [0, -1, 2, 3, 0, 1, 2],
                  "neighbour": [-1, 1, -1, -1, -1, -1, -1]},
        "cells": {"faces": [[0, 1, 2, 3], [4, 5, 6, 1]], "face_signs": [[1, 1, 1, 1], [1, 1, 1, -1]]},
        "patches": [{"id": 0, "name": "bottom", "faces": [0, 4]}, {"id": 1, "name": "right", "faces": [5]},
                    {"id": 2, "name": "top", "faces": [2, 6]}, {"id": 3, "name": "left", "faces": [3]}],
    })


def same_mesh(a: FiniteVolumeMesh, b: FiniteVolumeMesh) -> bool:
    keys = ("points", "faces", "face_owners", "face_neighbours", "face_patch_ids", "cell_sizes")
    if any(not np.array_equal(getattr(a, k), getattr(b, k)) for k in keys):
        return False
    cells_a = [tuple(a.cells[c, : a.cell_sizes[c]]) for c in range(a.cell_count)]
    cells_b = [tuple(b.cells[c, : b.cell_sizes[c]]) for c in range(b.cell_count)]
    return (cells_a == cells_b and tuple(a.cell_ids) == tuple(b.cell_ids)
            and tuple(a.cell_levels) == tuple(b.cell_levels)
            and tuple(a.refinement_records) == tuple(b.refinement_records))


def check_invariants(mesh: FiniteVolumeMesh, area=None):
    """Face loops close, interior faces have two different cells, levels are
    2:1 balanced, and the cells tile the original area."""
    areas = ad.cell_areas(mesh)
    assert np.all(areas > 0)
    if area is not None:
        assert areas.sum() == pytest.approx(area, rel=1e-12)
    inner = mesh.face_neighbours >= 0
    assert np.all(mesh.face_owners[inner] != mesh.face_neighbours[inner])
    lv = np.asarray(mesh.cell_levels)
    assert np.all(np.abs(lv[mesh.face_owners[inner]] - lv[mesh.face_neighbours[inner]]) <= 1)
    # every cell's node loop follows its faces: owners traverse faces forward
    for c in range(mesh.cell_count):
        loop = list(mesh.cells[c, : mesh.cell_sizes[c]])
        edges = {(loop[k], loop[(k + 1) % len(loop)]) for k in range(len(loop))}
        fo = np.flatnonzero(mesh.face_owners == c)
        fn = np.flatnonzero(mesh.face_neighbours == c)
        assert len(fo) + len(fn) == len(loop)
        for f in fo:
            assert tuple(mesh.faces[f]) in edges
        for f in fn:
            assert tuple(mesh.faces[f][::-1]) in edges


def test_refinement_creates_hanging_node_subfaces_and_preserves_area():
    original = two_cell_mesh()
    refined, groups = ad.adapt_mesh(original, {0: "xy"}, set())
    assert refined.cell_count == 5
    assert refined.cell_sizes[4] == 5
    assert ad.cell_areas(refined).sum() == pytest.approx(ad.cell_areas(original).sum())
    assert groups == ((0,), (0,), (0,), (0,), (1,))
    assert int(np.sum(refined.face_neighbours >= 0)) == 6
    assert same_mesh(mesh_from_dict(mesh_to_dict(refined)), refined)
    check_invariants(refined, 2.0)


@pytest.mark.parametrize(("axis", "varying"), (("x", 0), ("y", 1)))
def test_directional_refinement_splits_along_marked_axis(axis, varying):
    refined, groups = ad.adapt_mesh(two_cell_mesh(), {0: axis}, set())
    children = [c for c, src in enumerate(groups) if src == (0,)]
    assert len(children) == 2
    centres = [refined.points[refined.cells[c, : refined.cell_sizes[c]]].mean(axis=0) for c in children]
    assert abs(centres[0][varying] - centres[1][varying]) == pytest.approx(0.5)
    assert centres[0][1 - varying] == pytest.approx(centres[1][1 - varying])
    check_invariants(refined, 2.0)


def test_four_way_refinement_uses_bilinear_center_for_skewed_quad():
    polygon = ((0.0, 0.0), (2.0, 0.0), (1.6, 1.0), (0.2, 1.4))
    children = ad.split_polygon(polygon, "xy")
    shared = set(children[0])
    for child in children[1:]:
        shared &= set(child)
    assert shared == {(sum(p[0] for p in polygon) / 4.0, sum(p[1] for p in polygon) / 4.0)}


def test_refined_siblings_can_be_coarsened_back_to_parent():
    original = two_cell_mesh()
    refined, _ = ad.adapt_mesh(original, {0: "x"}, set())
    restored, groups = ad.adapt_mesh(refined, {}, {"r0"})
    assert restored.cell_count == original.cell_count
    assert tuple(restored.cell_levels) == (0, 0)
    assert restored.refinement_records == ()
    assert groups == ((0, 1), (2,))
    assert ad.cell_areas(restored).sum() == pytest.approx(2.0)
    assert same_mesh(mesh_from_dict(mesh_to_dict(restored)), restored)
    # back to the original geometry
    np.testing.assert_array_equal(np.sort(restored.points, axis=0), np.sort(original.points, axis=0))


def test_nested_refinement_can_be_coarsened_to_the_original_cell():
    original = two_cell_mesh()
    refined, _ = ad.adapt_mesh(original, {0: "xy"}, set())
    again, _ = ad.adapt_mesh(refined, {0: "x"}, set())
    partial, _ = ad.adapt_mesh(again, {}, {"r0.0"})
    restored, _ = ad.adapt_mesh(partial, {}, {"r0"})
    assert again.cell_count == 6
    assert partial.cell_count == 5
    assert restored.cell_count == original.cell_count
    assert max(restored.cell_levels) == 0
    assert restored.refinement_records == ()
    assert ad.cell_areas(restored).sum() == pytest.approx(2.0)
    for m in (refined, again, partial, restored):
        check_invariants(m, 2.0)


def test_refinement_propagates_to_restore_two_to_one_balance():
    once, _ = ad.adapt_mesh(two_cell_mesh(), {0: "xy"}, set())
    refined, _ = ad.adapt_mesh(once, {0: "xy"}, set())
    lv = np.asarray(refined.cell_levels)
    assert lv.max() == 2 and np.any(lv == 1)
    check_invariants(refined, 2.0)


def test_refinement_does_not_coarsen_a_cell_into_a_two_to_one_violation():
    refined, _ = ad.adapt_mesh(two_cell_mesh(), {0: "xy", 1: "xy"}, set())
    ids = refined.cell_ids
    target = next(
        first
        for f in range(refined.face_count) if refined.face_neighbours[f] >= 0
        for first, second in ((int(refined.face_owners[f]), int(refined.face_neighbours[f])),
                              (int(refined.face_neighbours[f]), int(refined.face_owners[f])))
        if ids[first].startswith("r1.") and ids[second].startswith("r0.")
    )
    adapted, _ = ad.adapt_mesh(refined, {target: "xy"}, {"r0"})
    assert any(cid.startswith("r0.") for cid in adapted.cell_ids)
    assert max(adapted.cell_levels) == 2
    check_invariants(adapted, 2.0)


def test_anisotropic_refinement_limits_directional_transition_between_neighbours():
    diff, _ = ad.adapt_mesh(two_cell_mesh(), {0: "x", 1: "y"}, set())
    ids = diff.cell_ids
    target, nb = next((int(diff.face_owners[f]), int(diff.face_neighbours[f])) for f in range(diff.face_count)
                      if diff.face_neighbours[f] >= 0 and ids[diff.face_owners[f]].startswith("r0.")
                      and ids[diff.face_neighbours[f]].startswith("r1."))
    nb_id = ids[nb]
    adapted, _ = ad.adapt_mesh(diff, {target: "x"}, set())
    assert any(cid.startswith(nb_id + ".") for cid in adapted.cell_ids)
    dl = ad.cell_directional_levels(adapted, adapted.cell_ids, adapted.refinement_records)
    for f in range(adapted.face_count):
        n = adapted.face_neighbours[f]
        if n >= 0:
            o = adapted.face_owners[f]
            assert all(abs(dl[o][a] - dl[n][a]) <= 1 for a in (0, 1))


def test_anisotropic_transition_uses_each_original_cells_shape_as_its_baseline():
    original = two_cell_mesh()
    x = original.points[:, 0]
    original.points = np.column_stack((np.where(x <= 1.0, 2.0 * x, x + 1.0), original.points[:, 1]))
    refined, _ = ad.adapt_mesh(original, {0: "y"}, set())
    assert refined.cell_count == 3
    assert any(cid.startswith("r1") for cid in refined.cell_ids)


def test_two_to_one_balance_chooses_axis_that_avoids_anisotropic_plan_cycle():
    marks = {1: "x"}
    ad.balance_plan(marks, set(), [{1}, {0}], ("r0", "r1"), (2, 3), ((2, 1), (0, 0)), {}, 4, (), [])
    assert marks == {0: "y", 1: "xy"}


def test_balance_keeps_periodic_coarsening_paired_after_balance_adjustment():
    records = (
        RefinementRecord("r0", ("r0.0", "r0.1"), ((0, 0), (1, 0), (1, 1), (0, 1)), 0),
        RefinementRecord("r1", ("r1.0", "r1.1"), ((2, 0), (3, 0), (3, 1), (2, 1)), 0),
    )
    coarsen = {"r0", "r1"}
    ad.balance_plan({}, coarsen, [{2}, set(), {0}], ("r0.0", "r1.0", "r2.0"), (1, 1, 3),
                    ((1, 0), (1, 0), (3, 0)), {"r0": (0, 0), "r1": (0, 0)}, 3, records, [(0, 1, (0.0, 0.0))])
    assert not coarsen


def test_quality_gate():
    assert not ad.split_passes_quality(((0.0, 0.0), (12.0, 0.0), (12.0, 1.0), (0.0, 1.0)), "y")
    regular = ((0.0, 0.0), (1.0, 0.0), (1.0, 1.0), (0.0, 1.0))
    assert ad.split_passes_quality(regular, "x") and ad.split_passes_quality(regular, "y")


# -- curved boundaries (reference tests/test_curved_adaptation.py) ----------

def curved_periodic_mesh():
    """Reference tests/test_curved_adaptation.py:19-130: 2x2 quads between two
    Catmull-Rom spline sides 2 apart (patches left 0, right 1, bottom 2, top 3)."""
    from meshcfd_amd.io.mesh_io import mesh_from_dict as _unused  # noqa: F401

    nx = ny = 2
    left = {"type": "spline", "points": [[0.0, 0.0], [-0.4, 0.33], [0.4, 0.67], [0.0, 1.0]]}
    right = {"type": "spline", "points": [[x + 2.0, y] for x, y in left["points"]]}
    pts = []
    for j in range(ny + 1):
        a = evaluate_boundary_curve(left, j / ny)
        b = evaluate_boundary_curve(right, j / ny)
        pts.extend((a[0] + (b[0] - a[0]) * i / nx, a[1] + (b[1] - a[1]) * i / nx) for i in range(nx + 1))

    def patch_of(p, q):
        pi, pj = p % (nx + 1), p // (nx + 1)
        qi, qj = q % (nx + 1), q // (nx + 1)
        if pj == qj == 0:
            return 2
        if pi == qi == nx:
            return 1
        if pj == qj == ny:
            return 3
        if pi == qi == 0:
            return 0
        return -1

    faces, owners, nbrs, patch, curves, cells = [], [], [], [], [], []
    face_of = {}
    for j in range(ny):
        for i in range(nx):
            bl = j * (nx + 1) + i
            cell = (bl, bl + 1, bl + nx + 2, bl + nx + 1)
            c = len(cells)
            cells.append(cell)
            for p, q in zip(cell, cell[1:] + cell[:1]):
                key = tuple(sorted((p, q)))
                if key not in face_of:
                    face_of[key] = len(faces)
                    faces.append((p, q))
                    owners.append(c)
                    nbrs.append(-1)
                    pid = patch_of(p, q)
                    patch.append(pid)
                    curve = None
                    if pid in (0, 1):
                        curve = BoundaryCurve(left if pid == 0 else right, (p // (nx + 1)) / ny,
                                              (q // (nx + 1)) / ny, pid)
                    curves.append(curve)
                else:
                    f = face_of[key]
                    nbrs[f] = c
                    patch[f] = -1
                    curves[f] = None
    patch = np.asarray(patch)
    return FiniteVolumeMesh(
        points=np.asarray(pts, dtype=float), cells=np.asarray(cells, dtype=np.int64),
        cell_sizes=np.full(len(cells), 4, dtype=np.int64), faces=np.asarray(faces, dtype=np.int64),
        face_owners=np.asarray(owners), face_neighbours=np.asarray(nbrs), face_patch_ids=patch,
        patches=tuple(MeshPatch(k, n, np.flatnonzero(patch == k)) for k, n in enumerate(("left", "right", "bottom", "top"))),
        face_curves=tuple(curves),
    )


@pytest.mark.parametrize(("axis", "face_count"), (("x", 2), ("y", 4), ("xy", 4)))
def test_refinement_and_coarsening_preserve_curved_periodic_patches(axis, face_count):
    """Reference tests/test_curved_adaptation.py:133-184."""
    mesh = curved_periodic_mesh()
    refined, groups = ad.adapt_mesh(mesh, {c: axis for c in range(4)}, set())
    fa, fb, shift = match_periodic_faces(refined, 0, 1)
    assert tuple(shift) == pytest.approx((2.0, 0.0))
    assert len(fa) == face_count
    assert len(refined.patches[0].face_ids) == face_count and len(refined.patches[1].face_ids) == face_count
    old_areas, new_areas = ad.cell_areas(mesh), ad.cell_areas(refined)
    moved = ad.transfer_cell_values(np.ones((mesh.cell_count, 1)), old_areas, new_areas, groups)
    assert float(np.dot(new_areas, moved[:, 0])) == pytest.approx(float(np.sum(old_areas)))
    assert any(
        evaluate_boundary_curve(c.geometry, 0.5 * (c.start_parameter + c.end_parameter))
        != tuple((refined.points[refined.faces[f][0]][k] + refined.points[refined.faces[f][1]][k]) / 2.0
                 for k in (0, 1))
        for f, c in enumerate(refined.face_curves) if c is not None
    )
    again, _ = ad.adapt_mesh(refined, {c: axis for c in range(refined.cell_count)}, set())
    assert len(match_periodic_faces(again, 0, 1)[0]) == (face_count if axis == "x" else 2 * face_count)
    restored, _ = ad.adapt_mesh(refined, {}, {f"r{c}" for c in range(4)})
    assert restored.cell_count == mesh.cell_count
    assert len(match_periodic_faces(restored, 0, 1)[0]) == 2


def test_arc_boundary_split_uses_its_parametric_midpoint():
    """Reference tests/test_curved_adaptation.py:187-201."""
    geometry = {"type": "arc", "center": [0.5, 0.5], "radius": 2**0.5 / 2,
                "start_angle": -3 * 3.141592653589793 / 4, "end_angle": -3.141592653589793 / 4}
    curve = BoundaryCurve(geometry, 0.0, 1.0, 0)
    children = ad.split_polygon(((0.0, 0.0), (1.0, 0.0), (1.0, 1.0), (0.0, 1.0)), "xy", (curve, None, None, None))
    assert children[0][1] == evaluate_boundary_curve(geometry, 0.5)


def test_periodic_adaptation_marking_keeps_spline_side_cells_paired():
    """Reference tests/test_curved_adaptation.py:204-247."""
    mesh = curved_periodic_mesh()
    bcs = [BoundaryCondition(p.id, p.name, kind="Periodic" if p.id in (0, 1) else "Wall") for p in mesh.patches]
    case = Case(mesh=mesh, boundary_conditions=bcs, periodic_pairs=[PeriodicBoundaryPair(0, 1, "x", 2.0)])
    centres = np.asarray([mesh.points[mesh.cells[c]].mean(axis=0) for c in range(mesh.cell_count)])
    marks, _ = ad.select_adaptation(mesh, case, centres, ad.cell_areas(mesh), np.asarray((1.0, 2.0, 1.5, 4.0)),
                                    1e-8, 1)
    fa, fb, _ = match_periodic_faces(mesh, 0, 1)
    assert marks
    for a, b in zip(fa, fb):
        assert marks.get(int(mesh.face_owners[a])) == marks.get(int(mesh.face_owners[b]))
    assert any(int(mesh.face_owners[a]) in marks for a in fa)
    refined, _ = ad.adapt_mesh(mesh, marks, set())
    assert len(match_periodic_faces(refined, 0, 1)[0])
    assert len(refined.patches[0].face_ids) == len(refined.patches[1].face_ids)


def test_transfer_keeps_integrals_and_constant_fields():
    mesh = two_cell_mesh()
    refined, groups = ad.adapt_mesh(mesh, {0: "xy", 1: "x"}, set())
    vals = np.array([[1.0, 2.0, -3.0, 4.0], [0.5, 0.25, 7.0, 1.0]])
    old_a, new_a = ad.cell_areas(mesh), ad.cell_areas(refined)
    moved = ad.transfer_cell_values(vals, old_a, new_a, groups)
    np.testing.assert_allclose(new_a @ moved, old_a @ vals, rtol=1e-14)
    back, groups2 = ad.adapt_mesh(refined, {}, {"r0", "r1"})
    moved_back = ad.transfer_cell_values(moved, new_a, ad.cell_areas(back), groups2)
    np.testing.assert_allclose(moved_back, vals, rtol=1e-14)


def test_bump_channel_indicator_refines_near_the_bump_and_stays_balanced():
    from meshcfd_amd import meshgen
    from meshcfd_amd.geometry import build_geometry

    case, prim = meshgen.bump_case(24, 8, 0.08)
    g = build_geometry(case)
    # a density field with a sharp front over the bump
    rho = 1.0 + 0.3 * np.tanh((g.centers[:, 0] - 1.5) / 0.05)
    marks, coarsen = ad.select_adaptation(case.mesh, case, g.centers, g.areas, rho, 0.25, 2)
    assert marks and not coarsen
    marked_x = g.centers[list(marks), 0]
    assert np.all(np.abs(marked_x - 1.5) < 0.5)
    refined, groups = ad.adapt_mesh(case.mesh, marks, coarsen)
    check_invariants(refined)
    # refined bump faces move onto the arc, above the original chords
    area = float(ad.cell_areas(refined).sum())
    assert float(g.areas.sum()) - 1e-3 < area < float(g.areas.sum())
    curves = [c for c in refined.face_curves if c is not None]
    assert len(curves) > 4  # the bump's four faces were split


# -- the incremental balance plan against full rescans -----------------------

def _balance_by_full_rescans(marks, coarsen, neighbours, cell_ids, levels, levels_dir, parent_dir,
                             maximum_level, records, periodic_pairs):
    """The balance loop as the reference writes it (adaptation.py:252-451):
    every round recomputes all projections and rescans every face from the
    first cell, fixing the first violation; a plain restatement for
    checking `ad.balance_plan`, which updates projections in place and
    resumes its scan."""
    by_parent = {r.parent_id: r for r in records}
    parent_of = {child: r.parent_id for r in records for child in r.child_ids}
    seen = set()

    def projected():
        lev, dirs = [], []
        for k, cid in enumerate(cell_ids):
            parent = parent_of.get(cid)
            if parent in coarsen:
                lev.append(by_parent[parent].level)
                dirs.append(parent_dir[parent])
            else:
                lev.append(levels[k] + int(k in marks))
                m = marks.get(k, "")
                dirs.append((levels_dir[k][0] + int("x" in m), levels_dir[k][1] + int("y" in m)))
        return lev, dirs

    def fix(lower, higher, axis, proj_dir):
        parent = parent_of.get(cell_ids[lower])
        if parent in coarsen:
            coarsen.remove(parent)
            return
        if lower in marks:
            if axis is None or axis in marks[lower]:
                if higher in marks:
                    del marks[higher]
                    return
                raise ad.AdaptationBalanceError("one step")
            both = set(marks[lower]) | {axis}
            marks[lower] = "xy" if len(both) == 2 else next(iter(both))
            return
        if levels[lower] < maximum_level:
            marks[lower] = (ad.AXES[ad._least_disruptive_axis(lower, neighbours, proj_dir, levels_dir)]
                            if axis is None else axis)
            return
        if higher in marks:
            del marks[higher]
            return
        raise ad.AdaptationBalanceError("maximum level")

    changed = True
    while changed:
        changed = False
        signature = (frozenset(marks.items()), frozenset(coarsen))
        if signature in seen:
            raise ad.AdaptationBalanceError("cycle")
        seen.add(signature)
        for a, b, _ in periodic_pairs:
            pa, pb = parent_of.get(cell_ids[a]), parent_of.get(cell_ids[b])
            if pa in coarsen and pb not in coarsen:
                coarsen.remove(pa)
                changed = True
            if pb in coarsen and pa not in coarsen:
                coarsen.remove(pb)
                changed = True
        for a, b, _ in periodic_pairs:
            if a in marks and b not in marks:
                if levels[b] >= maximum_level:
                    del marks[a]
                else:
                    marks[b] = marks[a]
                changed = True
            elif b in marks and a not in marks:
                if levels[a] >= maximum_level:
                    del marks[b]
                else:
                    marks[a] = marks[b]
                changed = True
            elif a in marks and b in marks and marks[a] != marks[b]:
                if marks[a] != "xy" or marks[b] != "xy":
                    marks[a] = marks[b] = "xy"
                    changed = True
        lev, dirs = projected()
        for k in range(len(cell_ids)):
            for nb in neighbours[k]:
                if nb <= k:
                    continue
                d = lev[k] - lev[nb]
                if abs(d) > 1:
                    fix(k if d < 0 else nb, nb if d < 0 else k, None, dirs)
                    changed = True
                    break
                for ai, axis in enumerate(ad.AXES):
                    d = dirs[k][ai] - dirs[nb][ai]
                    if abs(d) > 1:
                        fix(k if d < 0 else nb, nb if d < 0 else k, axis, dirs)
                        changed = True
                        break
                if changed:
                    break
            if changed:
                break


@pytest.mark.parametrize("seed", range(12))
def test_incremental_balance_plan_equals_full_rescans(seed):
    """On a twice-adapted bump channel (cells at levels 0-2, hanging nodes)
    random refinement marks and coarsening sets are balanced by
    `ad.balance_plan` and by full rescans: the same marks and coarsening
    sets, or the same error type (12 seeds: 1-9 marks changed and 2-15
    coarsening parents dropped per plan by the fixes)."""
    from meshcfd_amd import meshgen
    from meshcfd_amd.geometry import build_geometry

    case, prim = meshgen.bump_case(24, 8, 0.08)
    for _ in range(2):
        g = build_geometry(case)
        x = g.centers[:, 0]
        p = np.tile(prim[:1], (case.mesh.cell_count, 1)).astype(float)
        p[:, 0] *= 1.0 + 0.3 * np.tanh((x - 0.5 * x.max()) / 0.05)
        values = ad.indicator_values(p, case.numerics.adaptation_variable)
        marks, coarsen = ad.select_adaptation(case.mesh, case, g.centers, g.areas, values, 0.02, 3)
        case.mesh, _ = ad.adapt_mesh(case.mesh, marks, coarsen)
    mesh = case.mesh
    cell_ids = ad._cell_ids(mesh)
    levels = ad._cell_levels(mesh)
    nbrs = ad._neighbour_sets(mesh, ())
    levels_dir = ad.cell_directional_levels(mesh, cell_ids, mesh.refinement_records)
    parent_dir = ad.parent_directional_levels(mesh.refinement_records)
    rng = np.random.default_rng(seed)
    n = mesh.cell_count
    marks = {int(k): str(rng.choice(["x", "y", "xy"]))
             for k in rng.choice(n, size=int(rng.integers(1, n // 3)), replace=False)}
    parents = sorted(r.parent_id for r in mesh.refinement_records)
    coarsen = {p for p in parents if rng.random() < 0.3}
    maximum = int(rng.integers(1, 4))
    results = []
    for balance in (ad.balance_plan, _balance_by_full_rescans):
        m, c = dict(marks), set(coarsen)
        try:
            balance(m, c, nbrs, cell_ids, levels, levels_dir, parent_dir, maximum, mesh.refinement_records, [])
            results.append((m, c))
        except ad.AdaptationBalanceError as error:
            results.append(type(error))
    assert results[0] == results[1]


# -- the vectorised rebuild against the per-edge loop ------------------------

def _rebuild_by_full_scans(old, leaves):
    """The reference's `_rebuild_mesh` loop (adaptation.py:1032-1190) as a
    plain restatement: every leaf edge takes the known points within TOL of
    its range along its main axis (in that axis's sorted order) that lie on
    it, sorted along it; points and faces numbered at first visit; a new
    boundary face takes the first old boundary edge under its midpoint, or
    its curve. For checking `ad.rebuild_mesh`."""
    distinct = {}
    for leaf in leaves:
        for p in leaf.vertices:
            distinct.setdefault(ad.point_key(p), p)
    known = tuple(distinct.values())
    edges = []
    for f in np.flatnonzero(old.face_neighbours < 0):
        a, b = old.faces[f]
        edges.append(((float(old.points[a, 0]), float(old.points[a, 1])),
                      (float(old.points[b, 0]), float(old.points[b, 1])), int(old.face_patch_ids[f])))
    points, pid, faces, owners, neighbours, patches, curves = [], {}, [], [], [], [], []
    cell_faces, cell_signs, cell_nodes, face_of_edge = [], [], [], {}

    def point_id(p):
        key = ad.point_key(p)
        if key not in pid:
            pid[key] = len(points)
            points.append(p)
        return pid[key]

    for c, leaf in enumerate(leaves):
        fids, signs, loop = [], [], []
        n = len(leaf.vertices)
        lc = leaf.curves if len(leaf.curves) == n else (None,) * n
        for k in range(n):
            a, b = leaf.vertices[k], leaf.vertices[(k + 1) % n]
            axis = 0 if abs(b[0] - a[0]) >= abs(b[1] - a[1]) else 1
            lo, hi = min(a[axis], b[axis]) - ad.TOL, max(a[axis], b[axis]) + ad.TOL
            on_edge = [p for p in sorted(known, key=lambda p: p[axis])
                       if lo <= p[axis] <= hi and ad.on_segment(p, a, b)]
            on_edge.sort(key=lambda p: ad.segment_parameter(p, a, b))
            for p, q in zip(on_edge, on_edge[1:]):
                if ad.dist2(p, q) <= 1e-26:
                    continue
                i, j = point_id(p), point_id(q)
                loop.append(i)
                key = (min(i, j), max(i, j))
                f = face_of_edge.get(key)
                if f is None:
                    f = face_of_edge[key] = len(faces)
                    faces.append((i, j))
                    owners.append(c)
                    neighbours.append(-1)
                    if lc[k] is None:
                        mid = (0.5 * (p[0] + q[0]), 0.5 * (p[1] + q[1]))
                        patches.append(next((e[2] for e in edges if ad.on_segment(mid, e[0], e[1])), -1))
                        curves.append(None)
                    else:
                        cv = lc[k]
                        t0, t1 = ad.segment_parameter(p, a, b), ad.segment_parameter(q, a, b)
                        span = cv.end_parameter - cv.start_parameter
                        patches.append(cv.patch_id)
                        curves.append(BoundaryCurve(cv.geometry, cv.start_parameter + t0 * span,
                                                    cv.start_parameter + t1 * span, cv.patch_id))
                    sign = 1
                else:
                    neighbours[f] = c
                    patches[f] = -1
                    curves[f] = None
                    sign = 1 if faces[f] == (i, j) else -1
                fids.append(f)
                signs.append(sign)
        cell_faces.append(fids)
        cell_signs.append(signs)
        cell_nodes.append(loop)
    width = max(len(x) for x in cell_nodes)
    pad = lambda rows, fill: np.array([r + [fill] * (width - len(r)) for r in rows], dtype=np.int64)
    return {"points": np.asarray(points, dtype=float), "cells": pad(cell_nodes, -1),
            "cell_sizes": np.array([len(x) for x in cell_nodes]), "faces": np.asarray(faces, dtype=np.int64),
            "face_owners": np.asarray(owners), "face_neighbours": np.asarray(neighbours),
            "face_patch_ids": np.asarray(patches), "cell_faces": pad(cell_faces, -1),
            "cell_face_signs": pad(cell_signs, 0),
            "face_curves": repr(tuple(curves) if any(c is not None for c in curves) else ())}


@pytest.mark.parametrize("which", ["bump", "curved periodic"])
def test_vectorised_rebuild_equals_the_per_edge_loop(which, monkeypatch):
    """Every mesh `ad.adapt_mesh` rebuilds equals the per-edge loop's, array
    for array: a bump channel refined twice (768 -> 848 -> 960 cells, curved
    faces on the bump) then coarsened and refined at once (128 parents
    merged, 80 cells split), and the reference's curved periodic mesh split
    in x, y and both."""
    from meshcfd_amd import meshgen
    from meshcfd_amd.geometry import build_geometry

    seen = []
    real = ad.rebuild_mesh

    def spy(old, leaves, records):
        out = real(old, leaves, records)
        seen.append((out, _rebuild_by_full_scans(old, leaves)))
        return out

    monkeypatch.setattr(ad, "rebuild_mesh", spy)
    if which == "bump":
        case, prim = meshgen.bump_case(48, 16, 0.08)
        for rnd in range(3):
            g = build_geometry(case)
            x = g.centers[:, 0]
            p = np.tile(prim[:1], (case.mesh.cell_count, 1)).astype(float)
            front = 0.5 * x.max() if rnd < 2 else 0.8 * x.max()  # the third round coarsens
            p[:, 0] *= 1.0 + 0.3 * np.tanh((x - front) / 0.05)
            values = ad.indicator_values(p, case.numerics.adaptation_variable)
            marks, coarsen = ad.select_adaptation(case.mesh, case, g.centers, g.areas, values, 0.02, 3)
            case.mesh, _ = ad.adapt_mesh(case.mesh, marks, coarsen)
    else:
        mesh = curved_periodic_mesh()
        for axis in ("x", "y", "xy"):
            ad.adapt_mesh(mesh, {c: axis for c in range(4)}, set())
    assert seen
    for out, want in seen:
        for name, value in want.items():
            got = repr(out.face_curves) if name == "face_curves" else getattr(out, name)
            assert (got == value) if name == "face_curves" else np.array_equal(np.asarray(got), value), name
