"""MeshCFD case files, format "meshcfd-case" versions 1-13: reader and writer
(reference `src/meshcfd/io/project_io.py:31-1018`).

A case file holds the case name, the mesh (embedded as a MeshCAD mesh object,
or, from version 2, a path relative to the case file), the boundary
conditions, periodic pairs, numerical settings, initial-condition blocks,
and the last solution with its time, step count, CPU time and convergence
history. Reading and writing check the same conditions as the reference,
with the same messages (every failure is a ValueError), so a case saved by
either program loads in the other.
"""

from __future__ import annotations

import json
import math
import os
from pathlib import Path
from typing import Any

import numpy as np

from ..expressions import ExpressionError, compile_expression, evaluate_initial_conditions
from ..geometry import match_periodic_faces, shoelace
from ..model import (
    BOUNDARY_TYPES,
    FLUX_SCHEMES,
    LIMITER_SCHEMES,
    RECONSTRUCTION_SCHEMES,
    TIME_INTEGRATION_SCHEMES,
    BoundaryCondition,
    Case,
    InitialConditionBlock,
    InitialConditionSettings,
    InitialState,
    NumericalSettings,
    PeriodicBoundaryPair,
)
from .mesh_io import load_mesh, mesh_from_dict, mesh_to_dict

PROJECT_FORMAT = "meshcfd-case"
PROJECT_VERSION = 13
WALL_VELOCITY_EXPRESSION_KEYS = ("wall_velocity_x_expression", "wall_velocity_y_expression")
ADAPTATION_VARIABLES = ("Density", "Velocity X", "Velocity Y", "Pressure", "Turbulent k",
                        "Specific dissipation rate")
CHOICES = (
    ("reconstruction", RECONSTRUCTION_SCHEMES),
    ("limiter", LIMITER_SCHEMES),
    ("flux", FLUX_SCHEMES),
    ("time_integration", TIME_INTEGRATION_SCHEMES),
    ("time_step_mode", ("Global adaptive", "Global fixed", "Local (steady)")),
    ("implicit_preconditioner", ("ILU", "LU-SGS")),
    ("viscosity_model", ("Inviscid", "Constant", "Sutherland")),
    ("turbulence_model", ("None", "Menter k-omega SST")),
    ("turbulence_wall_treatment", ("Wall-resolved", "Wall functions")),
    ("turbulence_omega_formulation", ("Standard SST", "Regularized k-log-omega")),
)


# -- scalar checks (reference `project_io.py:924-994`) ------------------------
def _finite(value: Any, label: str) -> float:
    if isinstance(value, bool):
        raise ValueError(f"{label} must be numeric.")
    try:
        number = float(value)
    except (OverflowError, TypeError, ValueError) as error:
        raise ValueError(f"{label} must be numeric.") from error
    if not math.isfinite(number):
        raise ValueError(f"{label} must be finite.")
    return number


def _non_negative(value: Any, label: str) -> float:
    number = _finite(value, label)
    if number < 0.0:
        raise ValueError(f"{label} must be non-negative.")
    return number


def _positive(value: Any, label: str) -> float:
    if isinstance(value, bool):
        raise ValueError(f"{label} must be a positive number.")
    try:
        number = float(value)
    except (OverflowError, TypeError, ValueError) as error:
        raise ValueError(f"{label} must be a positive number.") from error
    if not math.isfinite(number) or number <= 0.0:
        raise ValueError(f"{label} must be a positive finite number.")
    return number


def _count(value: Any, label: str) -> int:
    if isinstance(value, bool) or not isinstance(value, int) or value < 0:
        raise ValueError(f"{label} must be a non-negative integer.")
    return value


def _positive_int(value: Any, label: str) -> int:
    if isinstance(value, bool) or not isinstance(value, int) or value <= 0:
        raise ValueError(f"{label} must be a positive integer.")
    return value


def _thermal(value: Any) -> str:
    if value not in ("Temperature", "Heat Flux"):
        raise ValueError("Wall thermal condition must be Temperature or Heat Flux.")
    return value


def _mechanical(value: Any) -> str:
    if value not in ("Slip", "No-slip", "Prescribed velocity"):
        raise ValueError("Wall mechanical condition must be Slip, No-slip, or Prescribed velocity.")
    return value


def _wall_expressions(values: dict) -> None:
    for key in WALL_VELOCITY_EXPRESSION_KEYS:
        if key not in values:
            continue
        if not isinstance(values[key], str):
            raise ValueError(f"Boundary expression '{key}' must be a string.")
        try:
            compile_expression(values[key])
        except ExpressionError as error:
            raise ValueError(f"Invalid boundary expression '{key}': {error}") from error


# numerical settings: (field, reader); readers take (value, label)
NUMERIC_FIELDS = (
    ("reconstruction", None, None),
    ("limiter", None, None),
    ("flux", None, None),
    ("time_integration", None, None),
    ("time_step_mode", None, None),
    ("cfl", _positive, "CFL"),
    ("fixed_time_step", _positive, "Fixed time step"),
    ("gmres_tolerance", _positive, "GMRES tolerance"),
    ("gmres_max_iterations", _positive_int, "Maximum GMRES iterations"),
    ("gmres_krylov_dimension", _positive_int, "GMRES Krylov dimension"),
    ("implicit_preconditioner", None, None),
    ("max_iterations", _positive_int, "Maximum iterations"),
    ("end_time", _positive, "End time"),
    ("residual_tolerance", _positive, "Residual tolerance"),
    ("convergence_history_interval", _positive_int, "Convergence history interval"),
    ("gamma", _positive, "Gamma"),
    ("gas_constant", _positive, "Gas constant"),
    ("viscosity_model", None, None),
    ("dynamic_viscosity", _positive, "Dynamic viscosity"),
    ("sutherland_reference_viscosity", _positive, "Sutherland reference viscosity"),
    ("sutherland_reference_temperature", _positive, "Sutherland reference temperature"),
    ("sutherland_temperature", _positive, "Sutherland temperature"),
    ("prandtl_number", _positive, "Prandtl number"),
    ("acceleration_x", _finite, "Acceleration X"),
    ("acceleration_y", _finite, "Acceleration Y"),
    ("adaptive_mesh", None, None),
    ("adaptation_interval", _positive_int, "Adaptation interval"),
    ("adaptation_max_level", _positive_int, "Maximum adaptation level"),
    ("adaptation_refine_threshold", _positive, "Adaptation refinement threshold"),
    ("adaptation_variable", None, None),
    ("turbulence_model", None, None),
    ("turbulence_wall_treatment", None, None),
    ("turbulence_omega_formulation", None, None),
)


def _check_sst(s: NumericalSettings) -> None:
    if (s.turbulence_model == "Menter k-omega SST" and s.turbulence_omega_formulation == "Regularized k-log-omega"
            and s.turbulence_wall_treatment != "Wall-resolved"):
        raise ValueError("The regularized k-log-omega formulation requires wall-resolved treatment.")


def _numerics_from_dict(raw: Any) -> NumericalSettings:
    if not isinstance(raw, dict):
        raise ValueError("Numerical settings must be an object.")
    defaults = NumericalSettings()
    values = {}
    for name, reader, label in NUMERIC_FIELDS:
        value = raw.get(name, getattr(defaults, name))
        values[name] = reader(value, label) if reader else value
    s = NumericalSettings(**values)
    if not isinstance(s.adaptive_mesh, bool):
        raise ValueError("Adaptive mesh setting must be boolean.")
    if s.adaptation_max_level > 12:
        raise ValueError("Maximum adaptation level must not exceed 12.")
    for name, choices in CHOICES:
        if getattr(s, name) not in choices:
            raise ValueError(f"Unsupported numerical setting for {name}.")
    if s.gamma <= 1.0:
        raise ValueError("Gamma must be greater than 1.")
    if s.adaptation_variable not in ADAPTATION_VARIABLES:
        raise ValueError("Unsupported adaptive mesh indicator variable.")
    _check_sst(s)
    return s


def _check_numerics_for_saving(s: NumericalSettings) -> None:
    """The writer's checks (reference `project_io.py:38-113`)."""
    _positive(s.end_time, "End time")
    if not isinstance(s.adaptive_mesh, bool):
        raise ValueError("Adaptive mesh setting must be boolean.")
    _positive_int(s.adaptation_interval, "Adaptation interval")
    _positive_int(s.convergence_history_interval, "Convergence history interval")
    if s.time_step_mode not in ("Global adaptive", "Global fixed", "Local (steady)"):
        raise ValueError("Unsupported time-stepping mode.")
    _positive(s.fixed_time_step, "Fixed time step")
    if _positive_int(s.adaptation_max_level, "Maximum adaptation level") > 12:
        raise ValueError("Maximum adaptation level must not exceed 12.")
    _positive(s.adaptation_refine_threshold, "Adaptation refinement threshold")
    if s.adaptation_variable not in ADAPTATION_VARIABLES:
        raise ValueError("Unsupported adaptive mesh indicator variable.")
    if s.viscosity_model not in ("Inviscid", "Constant", "Sutherland"):
        raise ValueError("Unsupported viscosity model.")
    if s.turbulence_model not in ("None", "Menter k-omega SST"):
        raise ValueError("Unsupported turbulence model.")
    if s.turbulence_wall_treatment not in ("Wall-resolved", "Wall functions"):
        raise ValueError("Unsupported turbulence wall treatment.")
    if s.turbulence_omega_formulation not in ("Standard SST", "Regularized k-log-omega"):
        raise ValueError("Unsupported SST omega formulation.")
    _check_sst(s)
    for name, label in (("prandtl_number", "Prandtl number"), ("dynamic_viscosity", "Dynamic viscosity"),
                        ("sutherland_reference_viscosity", "Sutherland reference viscosity"),
                        ("sutherland_reference_temperature", "Sutherland reference temperature"),
                        ("sutherland_temperature", "Sutherland temperature")):
        _positive(getattr(s, name), label)
    _finite(s.acceleration_x, "Acceleration X")
    _finite(s.acceleration_y, "Acceleration Y")


# -- states, history, periodic pairs, initial conditions ----------------------
STATE_LABELS = (("density", _positive, "density"), ("velocity_x", _finite, "velocity X"),
                ("velocity_y", _finite, "velocity Y"), ("pressure", _positive, "pressure"))


def _solution_to_list(states, mesh) -> list | None:
    if states is None:
        return None
    if mesh is None:
        raise ValueError("A final solution requires a mesh.")
    if len(states) != mesh.cell_count:
        raise ValueError("Final solution state count must match the mesh cell count.")
    out = []
    for st in states:
        item = {key: reader(getattr(st, key), f"Final solution {label}") for key, reader, label in STATE_LABELS}
        item["turbulence_k"] = _non_negative(st.turbulence_k, "Final solution turbulence k")
        item["turbulence_omega"] = _positive(st.turbulence_omega, "Final solution turbulence omega")
        out.append(item)
    return out


def _solution_from_list(raw: Any, mesh) -> tuple | None:
    if raw is None:
        return None
    if mesh is None:
        raise ValueError("A final solution requires a mesh.")
    if not isinstance(raw, list):
        raise ValueError("Final solution must be a list of cell states or null.")
    if len(raw) != mesh.cell_count:
        raise ValueError("Final solution state count must match the mesh cell count.")
    states = []
    for i, item in enumerate(raw):
        if not isinstance(item, dict):
            raise ValueError(f"Final solution state {i} must be an object.")
        values = {key: reader(item.get(key), f"Final solution {label} at cell {i}")
                  for key, reader, label in STATE_LABELS}
        values["turbulence_k"] = _non_negative(item.get("turbulence_k", 1e-6),
                                               f"Final solution turbulence k at cell {i}")
        values["turbulence_omega"] = _positive(item.get("turbulence_omega", 1.0),
                                               f"Final solution turbulence omega at cell {i}")
        states.append(InitialState(**values))
    return tuple(states)


def _history_checked(history: list, steps: int) -> list:
    prev = (0, 0.0, 0.0)
    for step, time, cpu, _ in history:
        if step > steps:
            raise ValueError("Convergence history step exceeds solution steps.")
        if step <= prev[0] or time < prev[1] or cpu < prev[2]:
            raise ValueError("Convergence history values must be nondecreasing.")
        prev = (step, time, cpu)
    return history


def _history_entry(step, time, cpu, residuals) -> tuple:
    if not isinstance(residuals, (list, tuple)) or len(residuals) not in (4, 6):
        raise ValueError("Convergence history must contain four or six residuals.")
    return (_positive_int(step, "Convergence history step"), _non_negative(time, "Convergence history time"),
            _non_negative(cpu, "Convergence history CPU time"),
            tuple(_non_negative(r, "Convergence residual") for r in residuals))


def _history_to_list(history: Any, steps: int) -> list:
    if not isinstance(history, list):
        raise ValueError("Convergence history must be a list.")
    checked = []
    for entry in history:
        if not isinstance(entry, tuple) or len(entry) != 4:
            raise ValueError("Each convergence history entry must have four values.")
        if not isinstance(entry[3], tuple):
            raise ValueError("Convergence history must contain four or six residuals.")
        checked.append(_history_entry(*entry))
    return [{"step": s, "time": t, "cpu_time": c, "residuals": list(r)} for s, t, c, r in
            _history_checked(checked, steps)]


def _history_from_list(raw: Any, steps: int) -> list:
    if not isinstance(raw, list):
        raise ValueError("Convergence history must be a list.")
    out = []
    for entry in raw:
        if not isinstance(entry, dict):
            raise ValueError("Each convergence history entry must be an object.")
        step = _positive_int(entry.get("step"), "Convergence history step")
        time = _non_negative(entry.get("time"), "Convergence history time")
        cpu = _non_negative(entry.get("cpu_time"), "Convergence history CPU time")
        res = entry.get("residuals")
        if not isinstance(res, list) or len(res) not in (4, 6):
            raise ValueError("Convergence history must contain four or six residuals.")
        out.append((step, time, cpu, tuple(_non_negative(r, "Convergence residual") for r in res)))
    return _history_checked(out, steps)


def _pairs_from_list(raw: Any) -> list:
    if not isinstance(raw, list):
        raise ValueError("Periodic boundary pairs must be a list.")
    pairs = []
    for item in raw:
        if not isinstance(item, dict):
            raise ValueError("Each periodic boundary pair must be an object.")
        a, b = item.get("patch_a_id"), item.get("patch_b_id")
        if any(isinstance(p, bool) or not isinstance(p, int) for p in (a, b)):
            raise ValueError("Periodic patch identifiers must be integers.")
        axis = item.get("axis")
        if axis not in ("x", "y"):
            raise ValueError("Periodic axis must be 'x' or 'y'.")
        pairs.append(PeriodicBoundaryPair(a, b, axis, _positive(item.get("pitch"), "Periodic pitch")))
    return pairs


def _check_pairs(mesh, conditions, pairs) -> None:
    if not pairs and not any(c.kind == "Periodic" for c in conditions):
        return
    if mesh is None:
        raise ValueError("Periodic boundary conditions require a mesh.")
    kinds = {c.patch_id: c.kind for c in conditions}
    paired = set()
    for pair in pairs:
        if pair.patch_a_id == pair.patch_b_id:
            raise ValueError("A periodic pair must contain two different patches.")
        if pair.patch_a_id in paired or pair.patch_b_id in paired:
            raise ValueError("A patch can belong to only one periodic pair.")
        paired.update((pair.patch_a_id, pair.patch_b_id))
        if any(kinds.get(p) != "Periodic" for p in (pair.patch_a_id, pair.patch_b_id)):
            raise ValueError("Both patches in a periodic pair must use the Periodic type.")
        try:
            _, _, shift = match_periodic_faces(mesh, pair.patch_a_id, pair.patch_b_id)
        except (ValueError, KeyError) as error:
            raise ValueError(f"Invalid periodic patch pair: {error}") from error
        axis = "x" if shift[0] != 0.0 else "y"
        pitch = abs(float(shift[0] if axis == "x" else shift[1]))
        if pair.axis != axis or not math.isclose(pair.pitch, pitch, rel_tol=1e-9, abs_tol=max(pitch, 1.0) * 1e-9):
            raise ValueError("Stored periodic axis or pitch does not match the mesh.")
    if paired != {c.patch_id for c in conditions if c.kind == "Periodic"}:
        raise ValueError("Each Periodic boundary must belong to exactly one pair.")


def _compile_block(block: InitialConditionBlock) -> None:
    try:
        compile_expression(block.domain, boolean=True)
        for key in ("density", "velocity_x", "velocity_y", "pressure"):
            compile_expression(getattr(block, key))
    except ExpressionError as error:
        raise ValueError(f"Invalid expression in initial-condition block '{block.name}': {error}") from error


BASE_LABELS = (("density", _positive, "Density"), ("velocity_x", _finite, "Velocity X"),
               ("velocity_y", _finite, "Velocity Y"), ("pressure", _positive, "Pressure"),
               ("turbulence_k", _non_negative, "Initial turbulence k"),
               ("turbulence_omega", _positive, "Initial turbulence omega"))


def _initial_from_dict(raw: Any) -> InitialConditionSettings:
    if not isinstance(raw, dict):
        raise ValueError("Initial conditions must be an object.")
    base = raw.get("base_state", {})
    if not isinstance(base, dict):
        raise ValueError("Initial-condition base state must be an object.")
    d = InitialState()
    state = InitialState(**{k: reader(base.get(k, getattr(d, k)), label) for k, reader, label in BASE_LABELS})
    blocks_raw = raw.get("blocks", [])
    if not isinstance(blocks_raw, list):
        raise ValueError("Initial-condition blocks must be a list.")
    blocks = []
    for index, item in enumerate(blocks_raw):
        if not isinstance(item, dict):
            raise ValueError("Each initial-condition block must be an object.")
        name = item.get("name")
        if not isinstance(name, str) or not name.strip():
            raise ValueError(f"Initial-condition block {index + 1} needs a name.")
        exprs = {k: item.get(k) for k in ("domain", "density", "velocity_x", "velocity_y", "pressure")}
        if not all(isinstance(v, str) for v in exprs.values()):
            raise ValueError(f"Initial-condition block '{name}' has an invalid expression.")
        block = InitialConditionBlock(name=name, **exprs)
        _compile_block(block)
        blocks.append(block)
    return InitialConditionSettings(base_state=state, blocks=blocks)


def _initial_to_dict(settings: InitialConditionSettings) -> dict:
    st = settings.base_state
    for k, reader, label in BASE_LABELS:
        reader(getattr(st, k), label)
    blocks = []
    for block in settings.blocks:
        if not isinstance(block.name, str) or not block.name.strip():
            raise ValueError("Initial-condition block name must not be empty.")
        _compile_block(block)
        blocks.append({k: getattr(block, k) for k in ("name", "domain", "density", "velocity_x", "velocity_y",
                                                      "pressure")})
    return {"base_state": {k: getattr(st, k) for k, _, _ in BASE_LABELS}, "blocks": blocks}


# -- the case ------------------------------------------------------------------
def case_to_dict(case: Case, case_path: str | Path | None = None) -> dict[str, Any]:
    s = case.numerics
    _check_numerics_for_saving(s)
    for c in case.boundary_conditions:
        _mechanical(c.mechanical_condition)
        _thermal(c.thermal_condition)
        _wall_expressions(c.values)
    _check_pairs(case.mesh, case.boundary_conditions, case.periodic_pairs)
    mesh = case.mesh
    if mesh is None:
        mesh_data = None
    elif case.mesh_path is None or mesh.refinement_records or any(level > 0 for level in mesh.cell_levels):
        mesh_data = mesh_to_dict(mesh)
    else:
        if case_path is None:
            raise ValueError("A case file path is required to save the mesh reference.")
        try:
            mesh_data = Path(os.path.relpath(Path(case.mesh_path).resolve(),
                                             Path(case_path).resolve().parent)).as_posix()
        except ValueError as error:
            raise ValueError("The mesh file and case file must be on the same drive to save a relative mesh "
                             "reference.") from error
    initial = _initial_to_dict(case.initial_conditions)
    solution = _solution_to_list(case.final_solution, mesh)
    history = _history_to_list(case.convergence_history, case.solution_steps)
    cpu = _non_negative(case.solution_cpu_time, "Solution CPU time")
    if history and cpu < history[-1]["cpu_time"]:
        raise ValueError("Solution CPU time must include the convergence history.")
    if case.final_solution is None:
        if case.solution_time != 0.0 or case.solution_steps != 0 or cpu != 0.0 or history:
            raise ValueError("Solution time and steps require a final solution.")
        time, steps = 0.0, 0
    else:
        time = _non_negative(case.solution_time, "Solution time")
        steps = _count(case.solution_steps, "Solution steps")
    return {
        "format": PROJECT_FORMAT,
        "version": PROJECT_VERSION,
        "name": case.name,
        "mesh": mesh_data,
        "boundary_conditions": [
            {"patch_id": c.patch_id, "patch_name": c.patch_name, "kind": c.kind, "values": dict(c.values),
             "mechanical_condition": c.mechanical_condition, "thermal_condition": c.thermal_condition}
            for c in case.boundary_conditions
        ],
        "periodic_pairs": [{"patch_a_id": p.patch_a_id, "patch_b_id": p.patch_b_id, "axis": p.axis,
                            "pitch": p.pitch} for p in case.periodic_pairs],
        "numerics": {name: getattr(s, name) for name, _, _ in NUMERIC_FIELDS},
        "initial_conditions": initial,
        "final_solution": solution,
        "solution_time": time,
        "solution_steps": steps,
        "solution_cpu_time": cpu,
        "convergence_history": history,
    }


def _conditions_from_list(raw: Any) -> list[BoundaryCondition]:
    if not isinstance(raw, list):
        raise ValueError("Boundary conditions must be a list.")
    out, ids, names = [], set(), set()
    for item in raw:
        if not isinstance(item, dict):
            raise ValueError("Each boundary condition must be an object.")
        pid, pname, kind = item.get("patch_id"), item.get("patch_name"), item.get("kind")
        values = item.get("values", {})
        if isinstance(pid, bool) or not isinstance(pid, int):
            raise ValueError("Boundary patch identifiers must be integers.")
        if not isinstance(pname, str) or not pname.strip():
            raise ValueError("Boundary patch names must not be empty.")
        if not isinstance(kind, str) or not kind:
            raise ValueError(f"Boundary condition for {pname} needs a type.")
        if kind not in BOUNDARY_TYPES:
            raise ValueError(f"Unsupported boundary type '{kind}'.")
        if pid in ids or pname in names:
            raise ValueError("Boundary patch identifiers and names must be unique.")
        if not isinstance(values, dict):
            raise ValueError(f"Boundary values for {pname} must be an object.")
        checked = {}
        for key, value in values.items():
            if not isinstance(key, str):
                raise ValueError("Boundary value names must be strings.")
            if key in WALL_VELOCITY_EXPRESSION_KEYS:
                _wall_expressions({key: value})
                checked[key] = value
                continue
            if isinstance(value, bool):
                raise ValueError(f"Boundary value '{key}' must be numeric.")
            try:
                number = float(value)
            except (OverflowError, TypeError, ValueError) as error:
                raise ValueError(f"Boundary value '{key}' must be numeric.") from error
            if not math.isfinite(number):
                raise ValueError(f"Boundary value '{key}' must be finite.")
            checked[key] = number
        out.append(BoundaryCondition(pid, pname, kind, checked,
                                     thermal_condition=_thermal(item.get("thermal_condition", "Temperature")),
                                     mechanical_condition=_mechanical(item.get("mechanical_condition", "No-slip"))))
        ids.add(pid)
        names.add(pname)
    return out


def case_from_dict(data: dict[str, Any], case_path: str | Path | None = None) -> Case:
    version = data.get("version")
    if data.get("format") != PROJECT_FORMAT or version not in range(1, PROJECT_VERSION + 1) or isinstance(
            version, bool):
        raise ValueError(f"Unsupported case format. Expected MeshCFD case version 1 through {PROJECT_VERSION}.")
    name = data.get("name")
    if not isinstance(name, str) or not name.strip():
        raise ValueError("Case name must not be empty.")
    raw_mesh = data.get("mesh")
    mesh, mesh_path = None, None
    if raw_mesh is not None:
        if isinstance(raw_mesh, dict):
            mesh = mesh_from_dict(raw_mesh)
        elif version >= 2:
            if not isinstance(raw_mesh, str) or not raw_mesh.strip():
                raise ValueError("Case mesh reference must be a non-empty path or null.")
            ref = Path(raw_mesh)
            if ref.is_absolute():
                raise ValueError("Case mesh reference must be relative.")
            if case_path is None:
                raise ValueError("A case file path is required to resolve the mesh reference.")
            mesh_path = (Path(case_path).resolve().parent / ref).resolve()
            mesh = load_mesh(mesh_path)
        else:
            raise ValueError("Case mesh data must be an object or null.")
    conditions = _conditions_from_list(data.get("boundary_conditions", []))
    numerics = _numerics_from_dict(data.get("numerics", {}))
    initial = _initial_from_dict(data.get("initial_conditions", {}))
    solution = _solution_from_list(data.get("final_solution"), mesh)
    time = _non_negative(data.get("solution_time", 0.0), "Solution time")
    steps = _count(data.get("solution_steps", 0), "Solution steps")
    cpu = _non_negative(data.get("solution_cpu_time", 0.0), "Solution CPU time")
    history = _history_from_list(data.get("convergence_history", []), steps)
    if solution is None and (time != 0.0 or steps != 0 or cpu != 0.0 or history):
        raise ValueError("Solution time and steps require a final solution.")
    if history and cpu < history[-1][2]:
        raise ValueError("Solution CPU time must include the convergence history.")
    if mesh is not None:
        valid = {p.id: p.name for p in mesh.patches}
        if any(c.patch_id not in valid or valid[c.patch_id] != c.patch_name for c in conditions):
            raise ValueError("A boundary condition refers to a patch not in the mesh.")
        configured = {c.patch_id for c in conditions}
        conditions.extend(BoundaryCondition(p.id, p.name) for p in mesh.patches if p.id not in configured)
    pairs = _pairs_from_list(data.get("periodic_pairs", []))
    _check_pairs(mesh, conditions, pairs)
    return Case(mesh=mesh, boundary_conditions=conditions, periodic_pairs=pairs, numerics=numerics, name=name,
                mesh_path=mesh_path, initial_conditions=initial, final_solution=solution, solution_time=time,
                solution_steps=steps, solution_cpu_time=cpu, convergence_history=history)


def load_case(path: str | Path) -> Case:
    source = Path(path)
    try:
        with source.open("r", encoding="utf-8") as stream:
            data = json.load(stream)
    except (OSError, UnicodeError, json.JSONDecodeError) as error:
        raise ValueError(f"Could not read case file '{source}': {error}") from error
    if not isinstance(data, dict):
        raise ValueError("Case file must contain a JSON object.")
    return case_from_dict(data, source)


def save_case(path: str | Path, case: Case) -> None:
    out = Path(path)
    with out.open("w", encoding="utf-8", newline="\n") as stream:
        json.dump(case_to_dict(case, out), stream, ensure_ascii=False, indent=2, allow_nan=False)
        stream.write("\n")


def initial_primitive(case: Case) -> np.ndarray:
    """(C, 4) primitive initial state of a case: its final solution when it
    has one (a resumed run), else its initial-condition blocks evaluated at
    the cell centroids."""
    if case.final_solution is not None:
        return np.array([(s.density, s.velocity_x, s.velocity_y, s.pressure) for s in case.final_solution])
    centers, _ = shoelace(case.mesh)
    return evaluate_initial_conditions(centers, case.initial_conditions)[:, :4]
