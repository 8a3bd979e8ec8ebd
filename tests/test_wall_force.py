"""Wall force objective with viscous traction (`steady_adjoint.WallForce`,
`ViscousWallTraction`) on the CPU, pinned without retyping its formulas:

* against the oracle's viscous residual: in a channel periodic in x with
  walls below and above, the interior and periodic faces cancel in
  sum_i A_i R_i, so sum_i A_i R_visc(i) is the walls' viscous flux, minus the
  viscous force on the walls (both components, any state);
* against Couette shear: u = U y, v = 0, uniform T gives the wall shear
  mu U exactly (the weighted least squares and the wall correction are exact
  for linear fields), so the viscous force is mu U times the wall length.
"""

from types import SimpleNamespace

import numpy as np
import pytest
import torch

from meshcfd_amd import meshgen
from meshcfd_amd.geometry import build_geometry
from meshcfd_amd.model import BoundaryCondition, Case, NumericalSettings, PeriodicBoundaryPair
from meshcfd_amd.steady_adjoint import WallForce
from oracle.euler import primitive_to_conservative
from oracle.solver import divergence_operator
from oracle.viscous import viscous_residual


def _top(kind, mechanical, thermal):
    if kind == "Symmetry":
        return BoundaryCondition(2, "top", "Symmetry")
    other_mech = "Slip" if mechanical == "No-slip" else "No-slip"
    if thermal == "Temperature":
        return BoundaryCondition(2, "top", "Wall", {"heat_flux": -0.2}, "Heat Flux", other_mech)
    return BoundaryCondition(2, "top", "Wall", {"temperature": 0.9}, "Temperature", other_mech)


def _channel(mechanical="No-slip", thermal="Temperature", viscosity="Constant", flux_value=0.3, top="Symmetry"):
    numerics = NumericalSettings(viscosity_model=viscosity, dynamic_viscosity=0.02, gas_constant=1.0,
                                 sutherland_reference_temperature=1.0, sutherland_reference_viscosity=0.02,
                                 sutherland_temperature=0.4, end_time=1e30)
    mesh = meshgen.rectangle_mesh(12, 8, (0.0, 1.0, 0.0, 1.0), jitter=0.2, seed=3, split_fraction=0.3)
    values = {"temperature": 1.2} if thermal == "Temperature" else {"heat_flux": flux_value}
    bcs = [
        BoundaryCondition(0, "bottom", "Wall", values, thermal, mechanical),
        BoundaryCondition(1, "right", "Periodic"),
        _top(top, mechanical, thermal),
        BoundaryCondition(3, "left", "Periodic"),
    ]
    case = Case(mesh=mesh, boundary_conditions=bcs, numerics=numerics,
                periodic_pairs=[PeriodicBoundaryPair(1, 3, "x", 1.0)], name="channel")
    g = build_geometry(case)
    dm = SimpleNamespace(geometry=g, n=g.cell_count, layout=SimpleNamespace(cell_order=np.arange(g.cell_count)),
                         device="cpu", settings=numerics)
    return case, g, dm


@pytest.mark.parametrize("mechanical", ["No-slip", "Slip"])
@pytest.mark.parametrize("thermal", ["Temperature", "Heat Flux"])
@pytest.mark.parametrize("viscosity", ["Constant", "Sutherland"])
def test_viscous_wall_force_equals_the_oracle_wall_flux(mechanical, thermal, viscosity):
    """Bottom wall (mechanical, thermal) and top wall with the other
    mechanical and thermal condition."""
    case, g, dm = _channel(mechanical, thermal, viscosity, top="Wall")
    x, y = g.centers[:, 0], g.centers[:, 1]
    rng = np.random.default_rng(2)
    n = g.cell_count
    prim = np.column_stack((1.0 + 0.2 * np.sin(2 * np.pi * x) * y + 0.05 * rng.random(n),
                            0.4 * y + 0.1 * np.cos(2 * np.pi * x) + 0.05 * rng.random(n),
                            0.1 * np.sin(2 * np.pi * x) * y * (1 - y) + 0.02 * rng.random(n),
                            1.0 + 0.1 * y * y + 0.05 * rng.random(n)))
    u = primitive_to_conservative(prim, case.numerics.gamma)
    r = viscous_residual(prim, g, case.numerics, divergence_operator(g))
    wall_flux = np.sum(g.areas[:, None] * r[:, 1:3], axis=0)
    force = WallForce(dm, 0.0)
    _, fv = force.components(torch.as_tensor(u))
    print(mechanical, thermal, viscosity, "force", fv.numpy(), "oracle", -wall_flux)
    np.testing.assert_allclose(fv.numpy(), -wall_flux, rtol=1e-11, atol=1e-13 * np.max(np.abs(wall_flux)))
    assert np.min(np.abs(wall_flux)) > 1e-3


def test_couette_shear():
    case, g, dm = _channel()
    n = g.cell_count
    speed = 0.7
    prim = np.column_stack((np.ones(n), speed * g.centers[:, 1], np.zeros(n), np.ones(n)))
    u = torch.as_tensor(primitive_to_conservative(prim, case.numerics.gamma))
    force = WallForce(dm, 0.0)
    fp, fv = force.components(u)
    assert float(fv[0]) == pytest.approx(0.02 * speed * 1.0, rel=1e-12)
    assert abs(float(fv[1])) <= 1e-12
    # uniform pressure 1 on the bottom wall (outward normal -y, length 1)
    assert float(fp[1]) == pytest.approx(-1.0, rel=1e-13)
    # the objective is the force on (cos a, sin a), and d/da is its rotation
    drag = WallForce(dm, 30.0)
    a = np.radians(30.0)
    total = fp + fv
    assert float(drag(u)) == pytest.approx(float(total[0]) * np.cos(a) + float(total[1]) * np.sin(a), rel=1e-13)
    assert drag.d_alpha(u) == pytest.approx(-float(total[0]) * np.sin(a) + float(total[1]) * np.cos(a), rel=1e-13)
