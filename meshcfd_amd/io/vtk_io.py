"""ParaView output of cell-centred solutions (reference
`src/meshcfd/io/vtk_io.py:30-244`): one ASCII VTU file per solution (polygon
cells, point coordinates with z = 0, cell fields), a PVD collection that
indexes snapshots by time and can be reopened and extended, and the
convergence history as a CSV table.

Cell fields: density, velocity (3 components), pressure, the turbulence
fields the states carry, and the derived Mach number, total pressure, total
temperature, total density and y+ (`derive_solution_fields`). Numbers are
written with 17 significant digits ("%.17g"), so the files read back bit for
bit.
"""

from __future__ import annotations

import csv
import math
import xml.etree.ElementTree as ET
from collections.abc import Iterable
from pathlib import Path

import numpy as np

from ..geometry import SolverError, build_geometry
from ..model import Case, InitialState

VTK_POLYGON = 7

CONVERGENCE_COLUMNS = (
    "Iteration",
    "PhysicalTime",
    "CPUTime",
    "MassResidual",
    "MomentumXResidual",
    "MomentumYResidual",
    "EnergyResidual",
    "TurbulentKResidual",
    "TurbulentOmegaResidual",
)


def _state_arrays(states) -> tuple[np.ndarray, np.ndarray]:
    """(C, 4) primitive (rho, u, v, p) and (C, 2) turbulence (k, omega) of a
    sequence of InitialState or of a (C, 4) / (C, 6) array."""
    if isinstance(states, np.ndarray):
        arr = np.asarray(states, dtype=np.float64)
        prim = arr[:, :4]
        turb = arr[:, 4:6] if arr.shape[1] >= 6 else np.tile((1e-6, 1.0), (len(arr), 1))
        return prim, turb
    prim = np.array([(s.density, s.velocity_x, s.velocity_y, s.pressure) for s in states], dtype=np.float64)
    turb = np.array([(getattr(s, "turbulence_k", 1e-6), getattr(s, "turbulence_omega", 1.0)) for s in states],
                    dtype=np.float64)
    return prim.reshape(-1, 4), turb.reshape(-1, 2)


def derive_solution_fields(case: Case, states) -> dict[str, np.ndarray]:
    """Mach number, total pressure / temperature / density and y+ per cell
    (reference `solver.py:2814-2950`, laminar part). y+ is non-zero in cells
    that own a no-slip wall face: with the closest such face at distance d
    from the centroid (along its normal), the tangential speed |u_t|
    relative to the wall, tau = mu |u_t| / d, u_tau = sqrt(tau / rho) and
    y+ = u_tau d rho / mu."""
    s = case.numerics
    prim, _ = _state_arrays(states)
    if len(prim) != case.mesh.cell_count:
        raise SolverError("Solution state count must match the mesh cell count.")
    if s.gamma <= 1.0 or s.gas_constant <= 0.0:
        raise SolverError("Derived fields require gamma > 1 and a positive gas constant.")
    if not np.isfinite(prim).all() or np.any(prim[:, (0, 3)] <= 0.0):
        raise SolverError("Derived fields require finite states with positive density and pressure.")
    rho, u, v, p = prim.T
    speed2 = u**2 + v**2
    mach2 = speed2 * rho / (s.gamma * p)
    factor = 1.0 + 0.5 * (s.gamma - 1.0) * mach2
    temperature = p / (rho * s.gas_constant)
    t0 = temperature * factor
    p0 = p * factor ** (s.gamma / (s.gamma - 1.0))
    fields = {
        "mach_number": np.sqrt(mach2),
        "total_pressure": p0,
        "total_temperature": t0,
        "total_density": p0 / (s.gas_constant * t0),
        "y_plus": np.zeros(len(prim)),
    }
    if s.viscosity_model == "Inviscid":
        return fields
    g = build_geometry(case)
    t = np.maximum(temperature, 1e-12)
    if s.viscosity_model == "Sutherland":
        tr = s.sutherland_reference_temperature
        mu = s.sutherland_reference_viscosity * (t / tr) ** 1.5 * (tr + s.sutherland_temperature) / (
            t + s.sutherland_temperature)
    else:
        mu = np.full_like(t, s.dynamic_viscosity)
    if not np.isfinite(mu).all() or np.any(mu <= 0.0):
        raise SolverError("y+ requires positive finite molecular viscosity.")
    walls = np.asarray(g.wall_faces, dtype=np.int64)
    walls = walls[(g.face_kind_code[walls] == 1) & g.wall_no_slip[walls]]
    if len(walls) == 0:
        return fields
    mesh = case.mesh
    a = mesh.points[mesh.faces[walls, 0]]
    b = mesh.points[mesh.faces[walls, 1]]
    edge = b - a
    length = np.linalg.norm(edge, axis=1)
    if np.any(length <= 1e-14):
        raise SolverError("Wall face has invalid length.")
    normal = np.column_stack((edge[:, 1], -edge[:, 0])) / length[:, None]
    owners = g.owners[walls]
    dist = np.abs(np.sum((0.5 * (a + b) - g.centers[owners]) * normal, axis=1))
    if np.any(dist <= 1e-14):
        raise SolverError("Wall face has invalid cell-center distance.")
    # the closest wall face of each owner (first one on ties, in face order)
    order = np.lexsort((np.arange(len(walls)), dist, owners))
    first = np.ones(len(order), dtype=bool)
    first[1:] = owners[order][1:] != owners[order][:-1]
    pick = order[first]
    o = owners[pick]
    rel = prim[o, 1:3] - g.wall_velocities[walls[pick]]
    n = normal[pick]
    tang = rel - np.sum(rel * n, axis=1)[:, None] * n
    tau = mu[o] * np.linalg.norm(tang, axis=1) / dist[pick]
    u_tau = np.sqrt(tau / rho[o])
    fields["y_plus"][o] = u_tau * dist[pick] * rho[o] / mu[o]
    return fields


def _array(parent: ET.Element, vtk_type: str, name: str, values, components: int = 1) -> None:
    attrs = {"type": vtk_type, "Name": name, "format": "ascii"}
    if components != 1:
        attrs["NumberOfComponents"] = str(components)
    node = ET.SubElement(parent, "DataArray", attrs)
    arr = np.asarray(values)
    if np.issubdtype(arr.dtype, np.integer):
        node.text = " ".join(str(int(x)) for x in arr.ravel().tolist())
    else:
        node.text = " ".join(f"{x:.17g}" for x in arr.astype(np.float64).ravel().tolist())


def export_vtu(path: str | Path, case: Case, states, time: float, step: int) -> Path:
    """One solution as an ASCII VTU file (reference `vtk_io.py:73-145`)."""
    destination = Path(path)
    if destination.suffix.lower() != ".vtu":
        raise ValueError("A VTU export filename must end in '.vtu'.")
    if case.mesh is None:
        raise ValueError("Import a mesh before exporting a solution.")
    prim, turb = _state_arrays(states)
    mesh = case.mesh
    if len(prim) != mesh.cell_count:
        raise ValueError("Solution state count must match the mesh cell count.")
    if not math.isfinite(time) or time < 0:
        raise ValueError("Solution time must be finite and non-negative.")
    if isinstance(step, bool) or not isinstance(step, (int, np.integer)) or step < 0:
        raise ValueError("Solution step must be a non-negative integer.")
    fields = derive_solution_fields(case, prim)
    root = ET.Element("VTKFile", {"type": "UnstructuredGrid", "version": "0.1", "byte_order": "LittleEndian"})
    piece = ET.SubElement(ET.SubElement(root, "UnstructuredGrid"), "Piece",
                          {"NumberOfPoints": str(len(mesh.points)), "NumberOfCells": str(mesh.cell_count)})
    fd = ET.SubElement(piece, "FieldData")
    _array(fd, "Float64", "TimeValue", [float(time)])
    _array(fd, "Int64", "Iteration", np.asarray([int(step)], dtype=np.int64))
    pts = np.column_stack((np.asarray(mesh.points, dtype=np.float64), np.zeros(len(mesh.points))))
    _array(ET.SubElement(piece, "Points"), "Float64", "Points", pts, 3)
    cells = ET.SubElement(piece, "Cells")
    sizes = np.asarray(mesh.cell_sizes, dtype=np.int64)
    mask = np.arange(mesh.cells.shape[1])[None, :] < sizes[:, None]
    _array(cells, "Int32", "connectivity", mesh.cells[mask].astype(np.int64))
    _array(cells, "Int32", "offsets", np.cumsum(sizes))
    _array(cells, "UInt8", "types", np.full(mesh.cell_count, VTK_POLYGON, dtype=np.int64))
    cd = ET.SubElement(piece, "CellData", {"Scalars": "Pressure"})
    _array(cd, "Float64", "Density", prim[:, 0])
    _array(cd, "Float64", "Velocity", np.column_stack((prim[:, 1], prim[:, 2], np.zeros(len(prim)))), 3)
    _array(cd, "Float64", "Pressure", prim[:, 3])
    _array(cd, "Float64", "TurbulentK", turb[:, 0])
    _array(cd, "Float64", "TurbulenceOmega", turb[:, 1])
    for name, key in (("MachNumber", "mach_number"), ("TotalPressure", "total_pressure"),
                      ("TotalTemperature", "total_temperature"), ("TotalDensity", "total_density"),
                      ("YPlus", "y_plus")):
        _array(cd, "Float64", name, fields[key])
    destination.parent.mkdir(parents=True, exist_ok=True)
    ET.ElementTree(root).write(destination, encoding="utf-8", xml_declaration=True)
    return destination


class VTKSeriesWriter:
    """VTU snapshots named <stem>_<step:08d>.vtu beside a PVD collection
    that lists them by time (reference `vtk_io.py:148-220`). Reopening an
    existing collection appends to it unless ``append`` is False; a snapshot
    of a step already listed replaces it. The collection is rewritten
    atomically (temporary file, then rename) after each snapshot."""

    def __init__(self, path: str | Path, append: bool = True) -> None:
        self.path = Path(path)
        if self.path.suffix.lower() != ".pvd":
            raise ValueError("A VTK time-series filename must end in '.pvd'.")
        self.entries: list[tuple[float, str]] = []
        if append and self.path.exists():
            self._read()

    def add_snapshot(self, case: Case, states, time: float, step: int) -> Path:
        name = f"{self.path.stem}_{step:08d}.vtu"
        out = export_vtu(self.path.with_name(name), case, states, time, step)
        self.entries = [e for e in self.entries if e[1] != name]
        self.entries.append((float(time), name))
        self.entries.sort()
        self._write()
        return out

    def _read(self) -> None:
        try:
            root = ET.parse(self.path).getroot()
        except (ET.ParseError, OSError) as error:
            raise ValueError(f"Could not read VTK collection '{self.path}': {error}") from error
        if root.attrib.get("type") != "Collection":
            raise ValueError(f"'{self.path}' is not a VTK collection file.")
        coll = root.find("Collection")
        if coll is None:
            raise ValueError(f"'{self.path}' does not contain a VTK collection.")
        for ds in coll.findall("DataSet"):
            name, ts = ds.attrib.get("file"), ds.attrib.get("timestep")
            if name is None or ts is None:
                raise ValueError(f"'{self.path}' contains an invalid dataset entry.")
            try:
                t = float(ts)
            except ValueError as error:
                raise ValueError(f"'{self.path}' contains an invalid dataset time.") from error
            if not math.isfinite(t):
                raise ValueError(f"'{self.path}' contains a non-finite dataset time.")
            self.entries.append((t, name))

    def _write(self) -> None:
        root = ET.Element("VTKFile", {"type": "Collection", "version": "0.1", "byte_order": "LittleEndian"})
        coll = ET.SubElement(root, "Collection")
        for t, name in self.entries:
            ET.SubElement(coll, "DataSet", {"timestep": f"{t:.17g}", "group": "", "part": "0", "file": name})
        self.path.parent.mkdir(parents=True, exist_ok=True)
        tmp = self.path.with_name(self.path.name + ".tmp")
        ET.ElementTree(root).write(tmp, encoding="utf-8", xml_declaration=True)
        tmp.replace(self.path)


def export_convergence_csv(path: str | Path, history: Iterable[tuple[int, float, float, tuple[float, ...]]]) -> Path:
    """Convergence history as a CSV table (reference `vtk_io.py:30-70`):
    iteration, physical time, CPU time and four (or six) residuals; missing
    turbulence residuals are empty fields."""
    destination = Path(path)
    if destination.suffix.lower() != ".csv":
        raise ValueError("A convergence-history filename must end in '.csv'.")
    rows = list(history)
    if not rows:
        raise ValueError("There is no convergence history to export.")

    def cell(value):
        if isinstance(value, str):
            return value
        if isinstance(value, bool) or not math.isfinite(float(value)):
            raise ValueError("Convergence history contains a non-finite numeric value.")
        if isinstance(value, (int, np.integer)):
            return int(value)
        return format(float(value), ".17g")

    destination.parent.mkdir(parents=True, exist_ok=True)
    with destination.open("w", encoding="utf-8", newline="") as stream:
        writer = csv.writer(stream, lineterminator="\n")
        writer.writerow(CONVERGENCE_COLUMNS)
        for iteration, physical_time, cpu_time, residuals in rows:
            if len(residuals) not in (4, 6):
                raise ValueError("Each convergence sample must contain four or six residuals.")
            values = [iteration, physical_time, cpu_time, *residuals]
            values += [""] * (len(CONVERGENCE_COLUMNS) - len(values))
            writer.writerow(cell(v) for v in values)
    return destination


def states_from_primitive(primitive: np.ndarray) -> tuple[InitialState, ...]:
    """InitialState tuple of a (C, 4) primitive array."""
    return tuple(InitialState(*map(float, row)) for row in np.asarray(primitive, dtype=np.float64))
