"""The GPU path on adapted meshes (hanging-node subfaces, cells of five and
six vertices, arc-snapped bump faces): residuals and steps equal the oracle
bit for bit; solve_case with adaptive_mesh equals an oracle loop that adapts
at the same steps (same host adaptation, same transfer); the reference's
curved periodic vortex run (tests/test_curved_adaptation.py:250-287)."""

import math

import numpy as np
import pytest

from meshcfd_amd import adaptation as ad
from meshcfd_amd import meshgen
from meshcfd_amd.geometry import build_geometry, match_periodic_faces
from oracle.euler import conservative_to_primitive, primitive_to_conservative
from oracle.solver import Oracle


def adapted_bump(levels=2, nx=30, ny=10):
    """The Mach 0.675 bump channel adapted ``levels`` times around a circular
    density front over the bump (HLLC + MUSCL linear BJ, explicit SSP-RK3, CFL 0.8)."""
    case, prim = meshgen.bump_case(nx, ny, 0.06, integrator="Explicit RK3", cfl=0.8)
    case.numerics.time_step_mode = "Global adaptive"
    for _ in range(levels):
        g = build_geometry(case)
        r = np.hypot(g.centers[:, 0] - 1.5, g.centers[:, 1] - 0.3)
        rho = 1.0 + 0.3 * np.tanh((0.3 - r) / 0.05)
        marks, coarsen = ad.select_adaptation(case.mesh, case, g.centers, g.areas, rho, 0.2, levels)
        case.mesh, _ = ad.adapt_mesh(case.mesh, marks, coarsen)
    g = build_geometry(case)
    x, y = g.centers[:, 0], g.centers[:, 1]
    speed = 0.675 * math.sqrt(1.4)
    prim = np.column_stack((1.0 + 0.05 * np.sin(3 * x) * np.cos(2 * y), speed * (1.0 + 0.02 * np.cos(4 * x)),
                            0.03 * np.sin(5 * x + y), 1.0 + 0.04 * np.cos(3 * y + x)))
    return case, g, prim


def test_adapted_bump_has_hanging_faces():
    case, g, _ = adapted_bump()
    sizes = case.mesh.cell_sizes
    assert sizes.max() >= 5 and max(case.mesh.cell_levels) == 2
    assert np.all(g.areas > 0)


@pytest.mark.gpu
@pytest.mark.parametrize("fused", [True, False])
def test_gpu_residual_and_steps_on_an_adapted_bump_equal_the_oracle(cuda_device, fused):
    from meshcfd_amd.device import DeviceMesh

    case, g, prim = adapted_bump()
    print(f"case: adapted bump {case.mesh.cell_count} cells fused={fused}", flush=True)
    dm = DeviceMesh(case, cuda_device, geometry=g, fused=fused, block_cells=64)
    o = Oracle(g, case.numerics)
    u = primitive_to_conservative(prim, 1.4)
    np.testing.assert_array_equal(dm.to_mesh_order(dm.residual(dm.to_device_order(u))), o.residual(u))
    dm.set_state(u)
    t = dm.run(10)
    want, t_ref = o.run(prim, 10)
    np.testing.assert_array_equal(dm.state(), want)
    assert t == t_ref
    dm.close()


@pytest.mark.gpu
def test_gpu_solve_case_adapts_like_an_oracle_loop(cuda_device):
    """adaptive_mesh every 3 steps, up to 2 levels, density indicator:
    solve_case's meshes and final state equal an oracle loop that steps on
    the CPU and adapts at the same steps with the same host code."""
    from meshcfd_amd.solver import solve_case

    case, g, prim = adapted_bump(levels=0)
    s = case.numerics
    s.adaptive_mesh = True
    s.adaptation_interval = 3
    s.adaptation_max_level = 2
    s.adaptation_refine_threshold = 0.02
    s.adaptation_variable = "Density"
    s.max_iterations = 9
    s.end_time = 1e30
    seen = []
    ocase, _, _ = adapted_bump(levels=0)
    ocase.numerics = s
    print("case: solve_case with adaptation", flush=True)
    res = solve_case(case, prim, progress_callback=seen.append, device=cuda_device)
    # oracle loop
    u = primitive_to_conservative(prim, 1.4)
    t = 0.0
    adapted_at = []
    for step in range(1, 10):
        og = build_geometry(ocase)
        u, dt, _, _ = Oracle(og, s).step(u, t)
        t += dt
        if step % 3 == 0:
            values = conservative_to_primitive(u, 1.4)[:, 0]
            marks, coarsen = ad.select_adaptation(ocase.mesh, ocase, og.centers, og.areas, values, 0.02, 2)
            if marks or coarsen:
                new, groups = ad.adapt_mesh(ocase.mesh, marks, coarsen, periodic_pairs=ocase.periodic_pairs)
                u = ad.transfer_cell_values(u, ad.cell_areas(ocase.mesh), ad.cell_areas(new), groups)
                ocase.mesh = new
                adapted_at.append(step)
    print("adapted at", adapted_at, "cells", case.mesh.cell_count, ocase.mesh.cell_count)
    assert adapted_at and [p.step for p in seen if p.mesh is not None] == adapted_at
    assert case.mesh.cell_count == ocase.mesh.cell_count > 300
    np.testing.assert_array_equal(case.mesh.points, ocase.mesh.points)
    np.testing.assert_array_equal(res.primitive, conservative_to_primitive(u, 1.4))
    assert res.time == t and res.steps == 9


@pytest.mark.gpu
def test_gpu_vortex_runs_after_adapting_a_curved_periodic_mesh(cuda_device):
    """Reference tests/test_curved_adaptation.py:250-287."""
    from meshcfd_amd.model import BoundaryCondition, Case, InitialState, NumericalSettings, PeriodicBoundaryPair
    from meshcfd_amd.solver import solve_case
    from tests.test_adaptation import curved_periodic_mesh

    mesh = curved_periodic_mesh()
    numerics = NumericalSettings(reconstruction="First order", time_integration="Explicit Euler", cfl=0.1,
                                 end_time=0.01, max_iterations=1, adaptive_mesh=True, adaptation_interval=1,
                                 adaptation_max_level=1, adaptation_refine_threshold=1e-12,
                                 adaptation_variable="Velocity X")
    case = Case(mesh=mesh, boundary_conditions=[BoundaryCondition(p.id, p.name, kind="Periodic") for p in mesh.patches],
                periodic_pairs=[PeriodicBoundaryPair(0, 1, "x", 2.0), PeriodicBoundaryPair(2, 3, "y", 1.0)],
                numerics=numerics)
    states = []
    for c in range(mesh.cell_count):
        x, y = mesh.points[mesh.cells[c]].mean(axis=0)
        r2 = (x - 1.0) ** 2 + (y - 0.5) ** 2
        e = math.exp(0.5 * (1.0 - r2))
        temperature = 1.0 - (0.4 / (8.0 * 1.4 * math.pi**2)) * math.exp(1.0 - r2)
        rho = temperature ** (1.0 / 0.4)
        states.append(InitialState(rho, 1.0 - e * (y - 0.5) / (2.0 * math.pi), 1.0 + e * (x - 1.0) / (2.0 * math.pi),
                                   rho**1.4))
    print("case: curved periodic vortex with adaptation", flush=True)
    result = solve_case(case, tuple(states), device=cuda_device)
    assert result.steps == 1
    assert case.mesh.cell_count > mesh.cell_count
    assert len(match_periodic_faces(case.mesh, 0, 1)[0]) > 0
    assert len(match_periodic_faces(case.mesh, 2, 3)[0]) > 0
