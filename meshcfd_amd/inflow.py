"""Pressure-inlet, mass-flow-inlet and mass-flow-outlet records on the device.

The device computes these boundaries' exterior states in two kernels (one
pass over the records, one workgroup per mass-flow-inlet patch) before the
flux and time-step kernels read them; `mcfd_mesh_set_inflow` uploads what
they need:

* per record (the boundary records of kinds 7-9, contiguous after the far
  field): its flux face, its parameters and n.d as the reference evaluates
  it (`normals @ direction` per patch, `solver.py:4423, 4507`);
* per mass-flow-inlet patch: a summation program that adds the patch's
  per-face flows in NumPy's pairwise order, so the bisection of
  `_mass_flow_inlet_boundary_states` (`solver.py:4521-4545`) compares the
  same sums as the reference's `np.sum`.
"""

from __future__ import annotations

import ctypes as C

import numpy as np

from . import native
from .geometry import Geometry, SolverError
from .layout import FACE_MASS_FLOW_INLET, FACE_MASS_FLOW_OUTLET, FACE_PRESSURE_INLET, INFLOW_KINDS, DeviceLayout

BLOCK = 8192  # NumPy's reduction buffer: np.sum adds blocks of this size in turn
PAIRWISE_LEAF = 128  # pairwise summation recurses down to blocks of at most this size
MAX_PATCH_FACES = 4096


class McfdInflowDesc(C.Structure):
    _fields_ = [
        ("first_record", C.c_int32),
        ("n_records", C.c_int32),
        ("n_patches", C.c_int32),
        ("n_leaves", C.c_int32),
        ("n_program", C.c_int32),
        ("reserved", C.c_int32),
        ("face", C.c_void_p),
        ("params", C.c_void_p),
        ("projection", C.c_void_p),
        ("patch_rec", C.c_void_p),
        ("leaf_ptr", C.c_void_p),
        ("leaf", C.c_void_p),
        ("prog_ptr", C.c_void_p),
        ("program", C.c_void_p),
    ]


native.EXPORTS["mcfd_mesh_set_inflow"] = (C.c_int, [C.c_void_p, C.c_void_p])
native.ERROR_MESSAGES.update({
    12: "Pressure inlet '{name}' cannot prescribe a supersonic outflow.",
    13: "Pressure inlet '{name}' needs Mach number to define a supersonic inflow.",
    14: "Pressure inlet '{name}' total conditions are incompatible with the interior characteristic state.",
    15: "Pressure inlet '{name}' produced an invalid state.",
    16: "Mass-flow inlet '{name}' has invalid interior pressure.",
    17: "Mass-flow inlet '{name}' cannot satisfy the configured flow rate.",
    18: "Mass-flow outlet '{name}' has invalid interior density.",
})


def _pairwise(start: int, n: int, leaves: list, program: list) -> None:
    """NumPy's pairwise_sum over values start .. start+n-1 (n >= 1): blocks
    of at most 128 are leaves; larger ones split at n/2 rounded down to a
    multiple of 8, left half first."""
    if n <= PAIRWISE_LEAF:
        program.append(len(leaves))
        leaves.append((start, n))
        return
    half = n // 2
    half -= half % 8
    _pairwise(start, half, leaves, program)
    _pairwise(start + half, n - half, leaves, program)
    program.append(-1)


def summation_program(n: int) -> tuple[list, list]:
    """Leaves (start, size) and token program (t >= 0 push leaf t, -1 add
    the top two) of np.sum over n contiguous float64 values: the blocks of
    8192 are summed pairwise and added to the running total in turn (the
    reduction's initial 0.0 is dropped: 0.0 + x == x for the positive flows
    summed here)."""
    if n <= 0:
        raise ValueError("empty patch")
    leaves: list = []
    program: list = []
    for b, start in enumerate(range(0, n, BLOCK)):
        _pairwise(start, min(BLOCK, n - start), leaves, program)
        if b:
            program.append(-1)
    return leaves, program


def evaluate_program(values: np.ndarray, leaves, program) -> float:
    """Host evaluation of a summation program (mirrors the device kernel)."""
    sums = []
    for start, m in leaves:
        a = values[start:start + m]
        if m < 8:
            res = 0.0
            for x in a:
                res += float(x)
        else:
            r = [float(x) for x in a[:8]]
            end = m - m % 8
            for i in range(8, end, 8):
                for j in range(8):
                    r[j] += float(a[i + j])
            res = ((r[0] + r[1]) + (r[2] + r[3])) + ((r[4] + r[5]) + (r[6] + r[7]))
            for i in range(end, m):
                res += float(a[i])
        sums.append(res)
    stack = []
    for t in program:
        if t >= 0:
            stack.append(sums[t])
        else:
            b = stack.pop()
            a = stack.pop()
            stack.append(a + b)
    return stack[0]


def has_inflow(layout: DeviceLayout) -> bool:
    return bool(np.isin(layout.bc_kind, INFLOW_KINDS).any())


def inflow_desc(g: Geometry, layout: DeviceLayout):
    """(McfdInflowDesc, arrays to keep alive) for the layout's inflow records."""
    kinds = layout.bc_kind
    rec = np.flatnonzero(np.isin(kinds, INFLOW_KINDS))
    n = len(rec)
    first = int(rec[0]) if n else 0
    if n and not np.array_equal(rec, first + np.arange(n)):
        raise SolverError("Inflow boundary records are not contiguous.")
    n_pi, n_mi, n_mo = (int(np.sum(kinds == k)) for k in INFLOW_KINDS)
    if (n_pi, n_mi, n_mo) != (len(g.pressure_inlet_faces), len(g.mass_flow_inlet_faces),
                              len(g.mass_flow_outlet_faces)):
        raise SolverError("Inflow records do not match the geometry.")
    # flux face of each record: face_bc maps flux face -> record
    face = np.full(n, -1, dtype=np.int32)
    fb = layout.face_bc
    sel = (fb >= first) & (fb < first + n)
    face[fb[sel] - first] = np.flatnonzero(sel).astype(np.int32)
    params = np.zeros((n, 4))
    proj = np.zeros(n)
    params[:n_pi] = g.pressure_inlet_params
    proj[:n_pi] = g.pressure_inlet_projection
    params[n_pi:n_pi + n_mi] = g.mass_flow_inlet_params
    proj[n_pi:n_pi + n_mi] = g.mass_flow_inlet_projection
    params[n_pi + n_mi:, :2] = g.mass_flow_outlet_params
    patch_rec, leaf_ptr, leaf, prog_ptr, program = [], [0], [], [0], []
    patches = g.mass_flow_inlet_patch
    for k in (np.unique(patches) if len(patches) else ()):
        members = np.flatnonzero(patches == k)
        if not np.array_equal(members, members[0] + np.arange(len(members))):
            raise SolverError("Mass-flow inlet patch faces are not contiguous.")
        if len(members) > MAX_PATCH_FACES:
            raise SolverError(f"Mass-flow inlet patches of more than {MAX_PATCH_FACES} faces are not supported "
                              "on the GPU.")
        patch_rec.append(n_pi + int(members[0]))
        leaves, prog = summation_program(len(members))
        leaf.extend(leaves)
        program.extend(prog)
        leaf_ptr.append(len(leaf))
        prog_ptr.append(len(program))
    n_patches = len(patch_rec)
    if n_patches:
        patch_rec.append(n_pi + n_mi)
    arrays = dict(
        face=face,
        params=np.ascontiguousarray(params.reshape(-1)),
        projection=np.ascontiguousarray(proj),
        patch_rec=np.asarray(patch_rec or [0], dtype=np.int32),
        leaf_ptr=np.asarray(leaf_ptr, dtype=np.int32),
        leaf=np.asarray(leaf or [(0, 0)], dtype=np.int32).reshape(-1),
        prog_ptr=np.asarray(prog_ptr, dtype=np.int32),
        program=np.asarray(program or [0], dtype=np.int32),
    )
    desc = McfdInflowDesc(first, n, n_patches, len(leaf), len(program), 0,
                          *(C.c_void_p(a.ctypes.data) for a in arrays.values()))
    return desc, arrays


def upload(lib, handle, g: Geometry, layout: DeviceLayout) -> None:
    desc, keep = inflow_desc(g, layout)
    native.check(lib.mcfd_mesh_set_inflow(handle, C.byref(desc)))
    del keep


def record_name(g: Geometry, layout: DeviceLayout, record: int) -> str:
    """Patch name of inflow record ``record`` (0-based within the inflow
    records) for error messages."""
    n_pi, n_mi = len(g.pressure_inlet_faces), len(g.mass_flow_inlet_faces)
    names = g.inflow_patch_names or ((), (), ())
    try:
        if record < n_pi:
            return names[0][int(g.pressure_inlet_patch[record])]
        if record < n_pi + n_mi:
            return names[1][int(g.mass_flow_inlet_patch[record - n_pi])]
        return names[2][int(g.mass_flow_outlet_patch[record - n_pi - n_mi])]
    except (IndexError, ValueError):
        return "?"


__all__ = ["McfdInflowDesc", "inflow_desc", "upload", "summation_program", "evaluate_program", "has_inflow",
           "record_name", "FACE_PRESSURE_INLET", "FACE_MASS_FLOW_INLET", "FACE_MASS_FLOW_OUTLET"]
