"""Batched cases: many meshes of uneven size advanced as one device mesh.

Cases are concatenated cell-major with per-case offsets. Device cells are
ordered case by case (a Hilbert order inside each case), so every halo block
of the fused stage kernel lies in one case or straddles two consecutive ones,
and each case's cells form one contiguous range. Per-case CFL steps are
segmented minima over those ranges (`mcfd_batch_time_steps`), broadcast to the
cells, so the same stage kernels run with per-cell steps.

No counterpart in the reference, which runs one case per process.
"""

from __future__ import annotations

import math
from dataclasses import dataclass, replace

import numpy as np

from . import meshgen
from .geometry import Geometry, build_geometry
from .layout import build_layout, hilbert_order
from .model import BoundaryCondition, Case, NumericalSettings


@dataclass
class BatchSpec:
    kind: str  # "vortex" or "shock tube"
    n: int  # quads along x
    m: int  # quads along y
    mach: float
    strength: float  # vortex strength (vortex cases)


def batch_specs(count: int = 256, seed: int = 2, cells=(2e3, 2e5)) -> list[BatchSpec]:
    """Cases alternate vortex / shock tube; cell counts log-uniform in
    [cells[0], cells[1]] and inflow Mach uniform in [0.1, 2.0] (seed 2).
    A mesh of q quads has 1.3 q cells after its 30% triangle split, so the
    quad counts aim at target / 1.3."""
    rng = np.random.default_rng(seed)
    target = np.exp(rng.uniform(np.log(cells[0]), np.log(cells[1]), count))
    mach = rng.uniform(0.1, 2.0, count)
    quads = target / 1.3
    specs = []
    for k in range(count):
        if k % 2 == 0:
            n = max(4, int(round(math.sqrt(quads[k]))))
            specs.append(BatchSpec("vortex", n, n, float(mach[k]), 5.0))
        else:
            m = max(2, int(round(math.sqrt(quads[k] / 10.0))))
            specs.append(BatchSpec("shock tube", 10 * m, m, float(mach[k]), 0.0))
    return specs


def _numerics():
    return NumericalSettings(
        reconstruction="MUSCL linear",
        limiter="Barth-Jespersen",
        flux="HLLC",
        time_integration="Explicit RK3",
        time_step_mode="Global adaptive",
        cfl=0.5,
        end_time=1.0e30,
    )


def build_case(spec: BatchSpec, seed: int):
    """One instance: a jittered quad mesh with 30% triangles (as the
    four-quadrant workload) and its initial primitive state.

    vortex: periodic [-5,5]^2, strength `strength`, free stream rho = p = 1,
    velocity (M*sqrt(gamma), 0). shock tube: [0,10]x[0,1], Sod states
    (1, 0, 0, 1) | (0.125, 0, 0, 0.1) at x = 5 with both sides moving at
    M*sqrt(gamma) in x, transmissive boundaries.
    """
    gamma = 1.4
    numerics = _numerics()
    c_inf = math.sqrt(gamma)
    if spec.kind == "vortex":
        mesh = meshgen.rectangle_mesh(spec.n, spec.m, (-5.0, 5.0, -5.0, 5.0), jitter=0.15, seed=seed,
                                      split_fraction=0.3)
        case = meshgen.periodic_case(mesh, numerics)
        prim = meshgen.isentropic_vortex(
            meshgen.cell_centroids(mesh), strength=spec.strength, free_stream=(1.0, spec.mach * c_inf, 0.0, 1.0)
        )
    else:
        mesh = meshgen.rectangle_mesh(spec.n, spec.m, (0.0, 10.0, 0.0, 1.0), jitter=0.15, seed=seed,
                                      split_fraction=0.3)
        bcs = [BoundaryCondition(p.id, p.name, kind="Transmissive") for p in mesh.patches]
        case = Case(mesh=mesh, boundary_conditions=bcs, numerics=numerics)
        x = meshgen.cell_centroids(mesh)[:, 0]
        left = x < 5.0
        u = spec.mach * c_inf
        prim = np.column_stack((np.where(left, 1.0, 0.125), np.full_like(x, u), np.zeros_like(x),
                                np.where(left, 1.0, 0.1)))
    return case, prim


def concat_geometries(geoms: list[Geometry]) -> tuple[Geometry, np.ndarray]:
    """One Geometry of disjoint components; returns it and the cell offsets."""
    cell_off = np.concatenate(([0], np.cumsum([g.cell_count for g in geoms])))
    face_off = np.concatenate(([0], np.cumsum([len(g.owners) for g in geoms])))

    def cat(name, shift=None):
        parts = []
        for k, g in enumerate(geoms):
            a = getattr(g, name)
            if shift == "cell":
                a = np.where(a >= 0, a + cell_off[k], a)
            elif shift == "face":
                a = a + face_off[k]
            parts.append(a)
        return np.concatenate(parts)

    cell_fields = ("owners", "neighbours", "interior_owners", "interior_neighbours", "gradient_sources",
                   "gradient_neighbours", "limiter_cells")
    face_fields = ("interior_faces", "wall_faces", "symmetry_faces", "periodic_faces_a", "periodic_faces_b",
                   "velocity_inlet_faces", "pressure_outlet_faces", "far_field_faces", "transmissive_faces",
                   "diffusive_boundary_faces", "pressure_inlet_faces", "mass_flow_inlet_faces",
                   "mass_flow_outlet_faces")
    kw = {}
    for f in Geometry.__dataclass_fields__:
        if f == "domain_scale":
            kw[f] = max(g.domain_scale for g in geoms)
        elif f == "inflow_patch_names":
            kw[f] = tuple(sum((tuple(g.inflow_patch_names[k]) if g.inflow_patch_names else () for g in geoms), ())
                          for k in range(3))
        elif f in ("pressure_inlet_patch", "mass_flow_inlet_patch", "mass_flow_outlet_patch"):
            # patch ordinals continue across instances
            shift = np.concatenate(([0], np.cumsum([int(getattr(g, f).max()) + 1 if len(getattr(g, f)) else 0
                                                    for g in geoms])))
            kw[f] = np.concatenate([getattr(g, f) + shift[k] for k, g in enumerate(geoms)]).astype(np.int64)
        elif f in cell_fields:
            kw[f] = cat(f, "cell")
        elif f in face_fields:
            kw[f] = cat(f, "face")
        else:
            kw[f] = cat(f)
    return Geometry(**kw), cell_off


def case_major_order(geoms: list[Geometry], cell_off: np.ndarray) -> np.ndarray:
    return np.concatenate([cell_off[k] + hilbert_order(g.centers) for k, g in enumerate(geoms)])


@dataclass
class Batch:
    cases: list
    primitives: list
    geometry: Geometry
    cell_offsets: np.ndarray  # mesh-order offsets
    layout: object
    case_ptr: np.ndarray  # device-order cell ranges per case (= cell_offsets, case-major order)
    numerics: NumericalSettings

    @property
    def n_cases(self) -> int:
        return len(self.cases)


def build_batch(specs: list[BatchSpec], base_seed: int = 100) -> Batch:
    cases, prims, geoms = [], [], []
    for k, spec in enumerate(specs):
        case, prim = build_case(spec, base_seed + k)
        cases.append(case)
        prims.append(prim)
        geoms.append(build_geometry(case))
    geometry, cell_off = concat_geometries(geoms)
    order = case_major_order(geoms, cell_off)
    layout = build_layout(geometry, order=order)
    return Batch(cases, prims, geometry, cell_off, layout, cell_off.copy(), replace(cases[0].numerics))


def batch_case(batch: Batch) -> Case:
    """A Case object standing for the whole batch (numerics shared by all
    instances); the geometry is supplied separately."""
    c = batch.cases[0]
    return Case(mesh=c.mesh, boundary_conditions=c.boundary_conditions, numerics=batch.numerics, name="batch")
