"""Adapted meshes against the reference's own output on the same inputs
(tests/golden/adaptation_cases.json, written by
tests/golden/make_adaptation_golden.py): node coordinates bit for bit, the
same faces, owners, neighbours and patches in the same numbering, the same
cell face lists, refinement ids, levels and records, the same boundary-curve
parameters, the same source groups, and select_adaptation's marks."""

import json
from pathlib import Path

import numpy as np
import pytest

from meshcfd_amd import adaptation as ad
from meshcfd_amd.io.mesh_io import mesh_from_dict, mesh_to_dict
from meshcfd_amd.model import BoundaryCondition, Case, PeriodicBoundaryPair

CASES = json.loads((Path(__file__).resolve().parent / "golden" / "adaptation_cases.json").read_text())


def assert_same(got: dict, want: dict):
    np.testing.assert_array_equal(np.asarray(got["nodes"]["coordinates"], dtype=float),
                                  np.asarray(want["nodes"]["coordinates"], dtype=float))
    for key in ("nodes", "owner", "neighbour", "patch"):
        assert [list(x) if isinstance(x, (list, tuple)) else x for x in got["faces"][key]] == want["faces"][key], key
    assert got["cells"]["faces"] == want["cells"]["faces"]
    assert got["cells"]["face_signs"] == want["cells"]["face_signs"]
    assert [(p["id"], p["name"], list(p["faces"])) for p in got["patches"]] == \
        [(p["id"], p["name"], list(p["faces"])) for p in want["patches"]]
    gr, wr = got.get("refinement", {}), want.get("refinement", {})
    assert list(gr.get("cell_ids", [])) == wr.get("cell_ids", [])
    assert list(gr.get("cell_levels", [])) == wr.get("cell_levels", [])
    norm = lambda rs: [(r["parent_id"], list(r["child_ids"]), [list(map(float, v)) for v in r["vertices"]], r["level"])
                       for r in rs]
    assert norm(gr.get("records", [])) == norm(wr.get("records", []))
    gc, wc = got.get("boundary_curves"), want.get("boundary_curves")
    assert (gc is None) == (wc is None)
    if wc is not None:
        assert len(gc) == len(wc)
        for a, b in zip(gc, wc):
            assert (a is None) == (b is None)
            if b is not None:
                assert a["patch_id"] == b["patch_id"] and a["geometry"] == b["geometry"]
                assert [float(x) for x in a["parameters"]] == [float(x) for x in b["parameters"]]


def _as_json(d):
    return json.loads(json.dumps(d))


@pytest.mark.parametrize("case", [c for c in CASES if "steps" in c], ids=lambda c: c["name"])
def test_adapt_mesh_reproduces_the_reference(case):
    mesh = mesh_from_dict(case["input"])
    pairs = [PeriodicBoundaryPair(*p) for p in case["periodic"]]
    for step in case["steps"]:
        marks = {int(k): v for k, v in step["marks"].items()}
        mesh, groups = ad.adapt_mesh(mesh, marks, set(step["coarsen"]), periodic_pairs=pairs)
        assert [list(g) for g in groups] == step["groups"]
        assert_same(_as_json(mesh_to_dict(mesh)), step["mesh"])


def test_select_adaptation_reproduces_the_reference_on_the_bump():
    case = next(c for c in CASES if c["name"] == "bump select")
    for step in case["select"]:
        mesh = mesh_from_dict(step["input"])
        bcs = [BoundaryCondition(p.id, p.name, "Wall") for p in mesh.patches]
        marks, coarsen = ad.select_adaptation(mesh, Case(mesh=mesh, boundary_conditions=bcs),
                                              np.asarray(step["centers"]), np.asarray(step["areas"]),
                                              np.asarray(step["values"]), step["threshold"], step["maximum_level"])
        assert {str(k): v for k, v in sorted(marks.items())} == step["marks"]
        assert sorted(coarsen) == step["coarsen"]
        adapted, groups = ad.adapt_mesh(mesh, marks, coarsen)
        assert [list(g) for g in groups] == step["groups"]
        assert_same(_as_json(mesh_to_dict(adapted)), step["mesh"])
