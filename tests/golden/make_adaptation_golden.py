"""Writes adaptation_cases.json: meshes adapted by the reference's own
`adapt_mesh` / `select_adaptation` (src/meshcfd/core/adaptation.py), the one
part of adaptation whose exact output (coordinates, numbering, refinement
records, curve parameters) the reference's tests do not pin.

Run in the development container from the repository root:
    python tests/golden/make_adaptation_golden.py
It imports the reference from /root/reference/src with a stand-in `numba`
module whose `njit` returns the function unchanged (adaptation never calls a
compiled kernel). Each case stores the input mesh (MeshCAD mesh JSON), the
marks or indicator, and the reference's output mesh and source groups.
"""

import json
import sys
import types
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parents[2]
sys.path.insert(0, str(ROOT))
numba = types.ModuleType("numba")
numba.njit = lambda *a, **k: (a[0] if a and callable(a[0]) else (lambda f: f))
sys.modules["numba"] = numba
sys.path.insert(0, "/root/reference/src")

from meshcfd.core import adaptation as ref  # noqa: E402
from meshcfd.core.model import BoundaryCondition as RBC  # noqa: E402
from meshcfd.core.model import Case as RCase  # noqa: E402
from meshcfd.core.model import PeriodicBoundaryPair as RPair  # noqa: E402
from meshcfd.io import mesh_io as rio  # noqa: E402

from meshcfd_amd import meshgen  # noqa: E402
from meshcfd_amd.io.mesh_io import mesh_to_dict  # noqa: E402

TWO_CELL = {
    "format": "meshcad-mesh", "version": 1, "dimension": 2,
    "nodes": {"coordinates": [[0, 0], [1, 0], [2, 0], [0, 1], [1, 1], [2, 1]]},
    "faces": {"nodes": [[0, 1], [1, 4], [4, 3], [3, 0], [1, 2], [2, 5], [5, 4]],
              "owner": [0, 0, 0, 0, 1, 1, 1], "patch": [0, -1, 2, 3, 0, 1, 2],
              "neighbour": [-1, 1, -1, -1, -1, -1, -1]},
    "cells": {"faces": [[0, 1, 2, 3], [4, 5, 6, 1]], "face_signs": [[1, 1, 1, 1], [1, 1, 1, -1]]},
    "patches": [{"id": 0, "name": "bottom", "faces": [0, 4]}, {"id": 1, "name": "right", "faces": [5]},
                {"id": 2, "name": "top", "faces": [2, 6]}, {"id": 3, "name": "left", "faces": [3]}],
}


def adapt_chain(mesh_dict, steps, periodic=()):
    """Apply the reference's adapt_mesh for each (marks, coarsen) in turn."""
    out = []
    mesh = rio.mesh_from_dict(mesh_dict)
    pairs = [RPair(*p) for p in periodic]
    for marks, coarsen in steps:
        mesh, groups = ref.adapt_mesh(mesh, {int(k): v for k, v in marks.items()}, set(coarsen), periodic_pairs=pairs)
        out.append({"marks": marks, "coarsen": sorted(coarsen), "mesh": rio.mesh_to_dict(mesh),
                    "groups": [list(g) for g in groups]})
    return out


def main():
    cases = []
    cases.append({"name": "two-cell chain", "input": TWO_CELL, "periodic": [], "steps": adapt_chain(TWO_CELL, [
        ({"0": "xy"}, []), ({"0": "xy"}, []), ({"2": "x", "5": "y"}, []), ({}, ["r0.0"]), ({}, ["r0"])])})
    cases.append({"name": "two-cell anisotropic", "input": TWO_CELL, "periodic": [], "steps": adapt_chain(TWO_CELL, [
        ({"0": "x", "1": "y"}, []), ({"0": "x"}, [])])})
    # jittered periodic rectangle: periodic partners and non-axis-aligned quads
    rect = meshgen.rectangle_mesh(6, 4, (0.0, 1.5, 0.0, 1.0), jitter=0.15, seed=7)
    rect_dict = mesh_to_dict(rect)
    periodic = [(1, 3, "x", 1.5), (0, 2, "y", 1.0)]
    cases.append({"name": "periodic jittered", "input": rect_dict, "periodic": periodic,
                  "steps": adapt_chain(rect_dict, [({"0": "xy", "7": "x", "11": "y"}, []),
                                                   ({"0": "xy", "3": "xy"}, [])], periodic)})
    # the bump channel with its arc: select_adaptation on a density front, two levels
    bump = meshgen.bump_channel(12, 4, 0.08)
    bump_dict = mesh_to_dict(bump)
    rmesh = rio.mesh_from_dict(bump_dict)
    rcase = RCase(mesh=rmesh)
    rcase.boundary_conditions = [RBC(p.id, p.name, "Wall") for p in rmesh.patches]
    steps = []
    mesh = rmesh
    for level in (1, 2):
        pts = np.asarray(mesh.points)
        centers = np.asarray([pts[list(c)].mean(axis=0) for c in mesh.cells])
        areas = np.asarray(ref.cell_areas(mesh))
        values = 1.0 + 0.3 * np.tanh((centers[:, 0] - 1.45) / 0.08) + 0.1 * centers[:, 1]
        marks, coarsen = ref.select_adaptation(mesh, rcase, centers, areas, values, 0.2, 2)
        mesh_in = rio.mesh_to_dict(mesh)
        mesh, groups = ref.adapt_mesh(mesh, marks, coarsen)
        steps.append({"input": mesh_in, "centers": centers.tolist(), "areas": areas.tolist(),
                      "values": values.tolist(), "threshold": 0.2, "maximum_level": 2,
                      "marks": {str(k): v for k, v in sorted(marks.items())}, "coarsen": sorted(coarsen),
                      "mesh": rio.mesh_to_dict(mesh), "groups": [list(g) for g in groups]})
    cases.append({"name": "bump select", "select": steps})
    path = Path(__file__).resolve().parent / "adaptation_cases.json"
    path.write_text(json.dumps(cases, separators=(",", ":")))
    print(f"wrote {path}: {len(cases)} cases")


if __name__ == "__main__":
    main()
