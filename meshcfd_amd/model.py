"""Case, mesh and numerical-settings types of the residual hot path.

The field names and string values follow the reference data model
(`src/meshcfd/core/model.py:9-34` enums, `:60-74` FiniteVolumeMesh,
`:77-93` BoundaryCondition/PeriodicBoundaryPair, `:95-129` NumericalSettings),
so a caller of the reference solver passes the same settings here.

Meshes are held as NumPy arrays instead of nested tuples: cells are a padded
``(C, max_vertices)`` int array plus a per-cell vertex count, which lets the
geometry builder run vectorised at millions of cells.
"""

from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np

BOUNDARY_TYPES = (
    "Wall",
    "Velocity Inlet",
    "Pressure Inlet",
    "Pressure Outlet",
    "Mass Flow Inlet",
    "Mass Flow Outlet",
    "Symmetry",
    "Far Field",
    "Periodic",
    # No counterpart in the reference: exterior state = interior reconstructed
    # state, so the face flux is the numerical flux F(W_L, W_L).
    "Transmissive",
)
RECONSTRUCTION_SCHEMES = ("First order", "MUSCL linear", "MUSCL parabolic")
LIMITER_SCHEMES = ("Barth-Jespersen", "Venkatakrishnan", "Minmod", "Van Leer")
FLUX_SCHEMES = ("HLLC", "Roe", "Rusanov")
TIME_INTEGRATION_SCHEMES = (
    "Explicit Euler",
    "Explicit RK2",
    "Explicit RK3",
    "Explicit RK4",
    "Implicit Euler",
)
TIME_STEP_MODES = ("Global adaptive", "Global fixed", "Local (steady)")

# Integer codes shared with the C-ABI (include/mcfd.h).
LIMITER_CODES = {"Barth-Jespersen": 0, "Minmod": 1, "Van Leer": 2, "Venkatakrishnan": 3}
FLUX_CODES = {"Rusanov": 0, "HLLC": 1, "Roe": 2}
RECONSTRUCTION_CODES = {"First order": 0, "MUSCL linear": 1, "MUSCL parabolic": 2}


@dataclass
class MeshPatch:
    id: int
    name: str
    face_ids: np.ndarray


@dataclass(frozen=True)
class BoundaryCurve:
    """A boundary face's piece of a MeshCAD curve (reference `model.py`
    BoundaryCurve): the curve descriptor (line / arc / spline /
    spline_segment, evaluated by `curves.evaluate_boundary_curve`) and the
    parameters of the face's two nodes."""

    geometry: dict
    start_parameter: float
    end_parameter: float
    patch_id: int


@dataclass(frozen=True)
class RefinementRecord:
    """One refined parent cell (reference `model.py` RefinementRecord)."""

    parent_id: str
    child_ids: tuple
    vertices: tuple
    level: int


@dataclass
class FiniteVolumeMesh:
    """Polygon mesh (reference `model.py:60-74`) with array storage.

    ``cells[i, :cell_sizes[i]]`` lists the CCW vertex ids of cell ``i``.
    ``faces[f] = (n0, n1)`` in the owner's CCW order; ``face_neighbours[f]``
    is ``-1`` on a boundary face.
    """

    points: np.ndarray  # (N, 2) float64
    cells: np.ndarray  # (C, V) int64, padded with -1
    cell_sizes: np.ndarray  # (C,) int64
    faces: np.ndarray  # (F, 2) int64
    face_owners: np.ndarray  # (F,) int64
    face_neighbours: np.ndarray  # (F,) int64
    face_patch_ids: np.ndarray  # (F,) int64, -1 on interior faces
    patches: tuple[MeshPatch, ...]
    # the file's cell -> face lists (padded with -1) and signs (+1: the face's
    # node order runs along the cell, -1: against it), kept so that a mesh
    # read from a MeshCAD file is written back as it was; None for meshes
    # built in memory (the writer derives them)
    cell_faces: np.ndarray | None = None
    cell_face_signs: np.ndarray | None = None
    cell_ids: tuple = ()
    cell_levels: tuple = ()
    refinement_records: tuple = ()
    face_curves: tuple = ()  # per face: BoundaryCurve or None; () for none

    @property
    def cell_count(self) -> int:
        return int(self.cells.shape[0])

    @property
    def face_count(self) -> int:
        return int(self.faces.shape[0])


@dataclass
class BoundaryCondition:
    patch_id: int
    patch_name: str
    kind: str = "Wall"
    values: dict = field(default_factory=dict)
    thermal_condition: str = "Temperature"
    mechanical_condition: str = "No-slip"


@dataclass(frozen=True)
class PeriodicBoundaryPair:
    patch_a_id: int
    patch_b_id: int
    axis: str
    pitch: float


@dataclass
class NumericalSettings:
    """Reference `model.py:95-129`, same defaults."""

    reconstruction: str = "MUSCL linear"
    limiter: str = "Barth-Jespersen"
    flux: str = "HLLC"
    time_integration: str = "Explicit RK2"
    time_step_mode: str = "Global adaptive"
    cfl: float = 0.5
    fixed_time_step: float = 0.001
    gmres_tolerance: float = 1e-6
    gmres_max_iterations: int = 300
    gmres_krylov_dimension: int = 30
    implicit_preconditioner: str = "LU-SGS"
    max_iterations: int = 1000
    end_time: float = 1.0
    residual_tolerance: float = 1e-6
    convergence_history_interval: int = 10
    gamma: float = 1.4
    gas_constant: float = 287.05
    viscosity_model: str = "Inviscid"
    dynamic_viscosity: float = 1.8e-5
    sutherland_reference_viscosity: float = 1.716e-5
    sutherland_reference_temperature: float = 273.15
    sutherland_temperature: float = 110.4
    prandtl_number: float = 0.72
    acceleration_x: float = 0.0
    acceleration_y: float = 0.0
    # carried by case files with the reference's defaults; solve_case adapts
    # the mesh with these settings (adaptation.py) and rejects a turbulence
    # model other than "None" (the SST model is outside this package)
    adaptive_mesh: bool = False
    adaptation_interval: int = 20
    adaptation_max_level: int = 3
    adaptation_refine_threshold: float = 0.25
    adaptation_variable: str = "Density"
    turbulence_model: str = "None"
    turbulence_wall_treatment: str = "Wall-resolved"
    turbulence_omega_formulation: str = "Standard SST"


@dataclass
class InitialConditionBlock:
    """A region (boolean expression of the centroid x, y) and the field
    expressions set there (reference `model.py` InitialConditionBlock)."""

    name: str = "Block 1"
    domain: str = "True"
    density: str = "1.225"
    velocity_x: str = "0"
    velocity_y: str = "0"
    pressure: str = "101325"


@dataclass
class InitialConditionSettings:
    base_state: "InitialState" = None
    blocks: list = field(default_factory=list)

    def __post_init__(self):
        if self.base_state is None:
            self.base_state = InitialState()


@dataclass
class Case:
    """A mesh with its boundary conditions, periodic pairs and numerics; the
    remaining fields are what a case file carries besides (the reference
    `model.py` Case)."""

    mesh: FiniteVolumeMesh
    boundary_conditions: list[BoundaryCondition]
    periodic_pairs: list[PeriodicBoundaryPair] = field(default_factory=list)
    numerics: NumericalSettings = field(default_factory=NumericalSettings)
    name: str = "case"
    mesh_path: object = None
    initial_conditions: InitialConditionSettings = field(default_factory=InitialConditionSettings)
    final_solution: object = None  # tuple of InitialState, or None
    solution_time: float = 0.0
    solution_steps: int = 0
    solution_cpu_time: float = 0.0
    convergence_history: list = field(default_factory=list)


@dataclass(frozen=True)
class InitialState:
    density: float = 1.225
    velocity_x: float = 0.0
    velocity_y: float = 0.0
    pressure: float = 101325.0
    # carried (and written to VTU files) with the reference's defaults; the
    # turbulence model itself is outside this package
    turbulence_k: float = 1e-6
    turbulence_omega: float = 1.0
