"""Pressure-inlet and mass-flow boundaries: the oracle against the
reference's known answers (`tests/test_solver.py:953-1116`) and the GPU
residual, time steps and steps against the oracle.

GPU tolerance: per component max |gpu - ref| / max |ref| <= 1e-12 for
residuals and time steps (the mass-flow bisection compares a patch sum whose
rounding differs between NumPy's pairwise sum and the device's tree sum; the
pressure-inlet state goes through pow), 1e-10 after 10 steps."""


import numpy as np
import pytest

from meshcfd_amd import meshgen
from meshcfd_amd.geometry import SolverError, build_geometry
from meshcfd_amd.model import BoundaryCondition, Case, NumericalSettings
from oracle import euler
from oracle.inflow import InflowError, mass_flow_inlet_states, mass_flow_outlet_states, pressure_inlet_states
from oracle.solver import Oracle

T0 = 1.0 + 0.2**2 / 7.0  # total temperature of (rho, u, p, R) = (1, 0.2, 1, 1)


def channel(inlet_kind, outlet_kind="Mass Flow Outlet", n=8, flux="HLLC", reconstruction="MUSCL linear",
            integrator="Explicit RK3", inlet_values=None, outlet_values=None, viscous=False):
    extra = {"viscosity_model": "Constant", "dynamic_viscosity": 0.02} if viscous else {}
    numerics = NumericalSettings(flux=flux, reconstruction=reconstruction, time_integration=integrator,
                                 gas_constant=1.0, cfl=0.5, end_time=1e30, **extra)
    mesh = meshgen.rectangle_mesh(2 * n, n, (0.0, 2.0, 0.0, 1.0), jitter=0.2, seed=4, split_fraction=0.3)
    if inlet_values is None:
        inlet_values = ({"pressure": T0**3.5, "temperature": T0, "velocity_x": 1.0, "velocity_y": 0.1}
                        if inlet_kind == "Pressure Inlet" else
                        {"mass_flow": 0.2, "temperature": T0, "velocity_x": 1.0, "velocity_y": 0.1})
    if outlet_values is None:
        outlet_values = {"mass_flow": 0.2} if outlet_kind == "Mass Flow Outlet" else {"pressure": 0.98}
    bcs = [
        BoundaryCondition(0, "bottom", "Wall"),
        BoundaryCondition(1, "right", outlet_kind, dict(outlet_values)),
        BoundaryCondition(2, "top", "Symmetry"),
        BoundaryCondition(3, "left", inlet_kind, dict(inlet_values)),
    ]
    return Case(mesh=mesh, boundary_conditions=bcs, numerics=numerics, name=f"{inlet_kind}/{outlet_kind}")


def flow_state(case, seed=2):
    g = build_geometry(case)
    x, y = g.centers[:, 0], g.centers[:, 1]
    rng = np.random.default_rng(seed)
    n = len(x)
    rho = 1.0 + 0.05 * np.sin(3 * x) * np.cos(2 * y) + 0.02 * rng.random(n)
    u = 0.2 + 0.05 * np.cos(2 * y) + 0.02 * rng.random(n)
    v = 0.02 * np.sin(3 * x)
    p = 1.0 + 0.03 * np.cos(2 * (x - y)) + 0.02 * rng.random(n)
    return np.column_stack((rho, u, v, p))


def test_pressure_inlet_reproduces_the_matching_subsonic_state():
    """Reference tests/test_solver.py:953-992."""
    case = channel("Pressure Inlet", inlet_values={"pressure": T0**3.5, "temperature": T0,
                                                   "velocity_x": 1.0, "velocity_y": 0.0})
    g = build_geometry(case)
    faces = g.pressure_inlet_faces
    interior = np.tile((1.0, 0.2, 0.0, 1.0), (len(faces), 1))
    args = (g.face_normals[faces], g.pressure_inlet_params, g.pressure_inlet_patch, 1.4, 1.0, ("left",))
    states = pressure_inlet_states(interior, *args)
    np.testing.assert_allclose(states, interior, rtol=1e-12, atol=1e-12)
    # the inlet is the left side (outward normal -x): u = 2 enters supersonically
    with pytest.raises(InflowError, match="Pressure inlet 'left' needs Mach number to define a supersonic inflow"):
        pressure_inlet_states(np.tile((1.0, 2.0, 0.0, 1.0), (len(faces), 1)), *args)
    with pytest.raises(InflowError, match="Pressure inlet 'left' cannot prescribe a supersonic outflow"):
        pressure_inlet_states(np.tile((1.0, -2.0, 0.0, 1.0), (len(faces), 1)), *args)


def test_pressure_inlet_state_satisfies_invariant_and_isentropic_relations():
    case = channel("Pressure Inlet")
    g = build_geometry(case)
    faces = g.pressure_inlet_faces
    rng = np.random.default_rng(1)
    interior = np.column_stack((1.0 + 0.1 * rng.random(len(faces)), 0.1 + 0.1 * rng.random(len(faces)),
                                0.05 * rng.random(len(faces)), 1.0 + 0.1 * rng.random(len(faces))))
    n = g.face_normals[faces]
    st = pressure_inlet_states(interior, n, g.pressure_inlet_params, g.pressure_inlet_patch, 1.4, 1.0)
    p0, t0, d = g.pressure_inlet_params[:, 0], g.pressure_inlet_params[:, 1], g.pressure_inlet_params[:, 2:4]
    temp = st[:, 3] / st[:, 0]
    np.testing.assert_allclose(st[:, 3] / p0, (temp / t0) ** 3.5, rtol=1e-13)
    speed = np.hypot(st[:, 1], st[:, 2])
    mach = speed / np.sqrt(1.4 * temp)
    np.testing.assert_allclose(temp, t0 / (1.0 + 0.2 * mach**2), rtol=1e-13)
    np.testing.assert_allclose(st[:, 1:3], speed[:, None] * d, rtol=1e-13, atol=1e-15)
    # incoming invariant of the state equals the interior's outgoing one
    c_in = np.sqrt(1.4 * interior[:, 3] / interior[:, 0])
    out_inv = interior[:, 1] * n[:, 0] + interior[:, 2] * n[:, 1] + 2.0 * c_in / 0.4
    in_inv = st[:, 1] * n[:, 0] + st[:, 2] * n[:, 1] + 2.0 * np.sqrt(1.4 * temp) / 0.4
    np.testing.assert_allclose(in_inv, out_inv, rtol=1e-12)


def test_mass_flow_inlet_satisfies_the_integrated_flow_rate():
    """Reference tests/test_solver.py:995-1029."""
    case = channel("Mass Flow Inlet", inlet_values={"mass_flow": 0.25, "temperature": 2.0,
                                                    "velocity_x": 1.0, "velocity_y": 0.0})
    g = build_geometry(case)
    faces = g.mass_flow_inlet_faces
    n, w = g.face_normals[faces], g.face_lengths[faces]
    st = mass_flow_inlet_states(np.tile((1.0, 0.0, 0.0, 1.0), (len(faces), 1)), n, w, g.mass_flow_inlet_params,
                                g.mass_flow_inlet_patch, 1.4, 1.0)
    flow = np.sum(st[:, 0] * -(st[:, 1:3] * n).sum(axis=1) * w)
    assert flow == pytest.approx(0.25, rel=1e-12)
    assert g.mass_flow_inlet_params[0, 0] == 0.25
    np.testing.assert_allclose(st[:, 3], 1.0)
    cp = 1.4 / 0.4
    np.testing.assert_allclose(st[:, 3] / st[:, 0], 2.0 - (st[:, 1] ** 2 + st[:, 2] ** 2) / (2 * cp), rtol=1e-13)


def test_mass_flow_outlet_keeps_density_pressure_and_tangential_velocity():
    """Reference tests/test_solver.py:1032-1052."""
    case = channel("Pressure Inlet", outlet_values={"mass_flow": 0.6})
    g = build_geometry(case)
    faces = g.mass_flow_outlet_faces
    n, w = g.face_normals[faces], g.face_lengths[faces]
    interior = np.tile((2.0, 0.4, 0.3, 1.2), (len(faces), 1))
    st = mass_flow_outlet_states(interior, n, g.mass_flow_outlet_params)
    assert np.sum(st[:, 0] * (st[:, 1:3] * n).sum(axis=1) * w) == pytest.approx(0.6)
    np.testing.assert_allclose(st[:, (0, 3)], interior[:, (0, 3)])
    t = np.column_stack((-n[:, 1], n[:, 0]))
    np.testing.assert_allclose((st[:, 1:3] * t).sum(1), (interior[:, 1:3] * t).sum(1), rtol=1e-13, atol=1e-15)


@pytest.mark.parametrize("kind", ["Pressure Inlet", "Mass Flow Inlet"])
def test_inlets_reject_directions_out_of_the_domain(kind):
    """Reference tests/test_solver.py:1070-1090."""
    values = ({"pressure": 101325.0, "temperature": 288.15} if kind == "Pressure Inlet"
              else {"mass_flow": 0.2, "temperature": 288.15})
    values.update({"velocity_x": -1.0, "velocity_y": 0.0})
    with pytest.raises(SolverError, match="direction must point into the domain"):
        build_geometry(channel(kind, inlet_values=values))


@pytest.mark.parametrize("inlet", ["Pressure Inlet", "Mass Flow Inlet"])
def test_inlets_with_mass_flow_outlet_conserve_and_step(inlet):
    """One oracle RK3 step runs (reference :1093-1116) and the residual is
    conservative up to the boundary fluxes."""
    case = channel(inlet)
    prim = flow_state(case)
    o = Oracle(build_geometry(case), case.numerics)
    u1, _ = o.run(prim, 1)
    assert np.all(np.isfinite(u1))


# --- GPU parity ---------------------------------------------------------------


def _rel(gpu, ref):
    scale = np.max(np.abs(ref), axis=0)
    scale = np.where(scale > 0, scale, 1.0)
    return np.max(np.abs(gpu - ref), axis=0) / scale


CASES = [("Pressure Inlet", "Mass Flow Outlet"), ("Mass Flow Inlet", "Mass Flow Outlet"),
         ("Mass Flow Inlet", "Pressure Outlet"), ("Pressure Inlet", "Pressure Outlet")]


@pytest.mark.gpu
@pytest.mark.parametrize("flux", ["HLLC", "Roe", "Rusanov"])
@pytest.mark.parametrize("kinds", CASES)
def test_gpu_inflow_residual_and_dt_match_oracle(cuda_device, kinds, flux):
    from meshcfd_amd.device import DeviceMesh

    case = channel(kinds[0], kinds[1], flux=flux)
    print(f"case: {case.name} {flux}", flush=True)
    dm = DeviceMesh(case, cuda_device, fused=True, block_cells=64)
    prim = flow_state(case)
    u = euler.primitive_to_conservative(prim, 1.4)
    o = Oracle(dm.geometry, case.numerics)
    gpu = dm.to_mesh_order(dm.residual(dm.to_device_order(u)))
    dm.check_status()
    err = _rel(gpu, o.residual(u))
    print("residual rel err", err)
    assert np.all(err <= 1e-12), err
    dm.time_steps(dm.to_device_order(u))
    dm.check_status()
    dt = dm.to_mesh_order(dm.dt_cell)
    ref = o.stable_time_steps(prim)
    assert np.max(np.abs(dt - ref) / ref) <= 1e-12
    dm.close()


@pytest.mark.gpu
@pytest.mark.parametrize("integrator", ["Explicit Euler", "Explicit RK3"])
@pytest.mark.parametrize("kinds", CASES[:2])
def test_gpu_inflow_steps_match_oracle(cuda_device, kinds, integrator):
    from meshcfd_amd.device import DeviceMesh

    case = channel(kinds[0], kinds[1], integrator=integrator)
    print(f"case: {case.name} {integrator} steps", flush=True)
    dm = DeviceMesh(case, cuda_device, fused=True, block_cells=64)
    prim = flow_state(case)
    dm.set_state(euler.primitive_to_conservative(prim, 1.4))
    dm.run(10)
    gpu = dm.state()
    ref, _ = Oracle(dm.geometry, case.numerics).run(prim, 10)
    err = _rel(gpu, ref)
    print("rel err", err)
    assert np.all(err <= 1e-10), err
    dm.close()


# --- host-side device layout of the inflow records ------------------------------


@pytest.mark.parametrize("n", [1, 5, 7, 8, 9, 63, 64, 127, 128, 129, 130, 255, 256, 300, 1000, 4096, 8192, 8193,
                               20000])
def test_summation_program_matches_numpy_sum(n):
    """The device's summation program reproduces np.sum bit for bit
    (NumPy's pairwise summation in blocks of 8192)."""
    from meshcfd_amd.inflow import evaluate_program, summation_program

    rng = np.random.default_rng(n)
    values = rng.random(n) * 10.0 ** rng.uniform(-4, 4, n)
    values = values * 1.0000001  # a fresh contiguous array, like the reference's product
    leaves, program = summation_program(n)
    assert evaluate_program(values, leaves, program) == float(np.sum(values))
    if n <= 4096:  # the device's LDS bounds
        assert len(leaves) <= 128


def test_summation_program_catches_a_reordered_sum():
    """A plain left-to-right sum differs from np.sum on such data, so the
    check above can fail."""
    rng = np.random.default_rng(3)
    values = rng.random(1000) * 10.0 ** rng.uniform(-4, 4, 1000)
    total = 0.0
    for x in values:
        total += float(x)
    assert total != float(np.sum(values))


@pytest.mark.parametrize("kinds", [("Pressure Inlet", "Mass Flow Outlet"), ("Mass Flow Inlet", "Pressure Outlet")])
def test_inflow_records_follow_the_far_field_in_the_layout(kinds):
    from meshcfd_amd.inflow import inflow_desc
    from meshcfd_amd.layout import INFLOW_KINDS, build_layout

    case = channel(*kinds)
    g = build_geometry(case)
    lay = build_layout(g)
    desc, arrays = inflow_desc(g, lay)
    rec = np.flatnonzero(np.isin(lay.bc_kind, INFLOW_KINDS))
    assert desc.first_record == rec[0] and desc.n_records == len(rec)
    # each record's flux face carries that record, and its normal is the mesh face's
    faces = arrays["face"]
    assert np.array_equal(lay.face_bc[faces], desc.first_record + np.arange(desc.n_records))
    mesh_faces = np.concatenate((g.pressure_inlet_faces, g.mass_flow_inlet_faces, g.mass_flow_outlet_faces))
    assert np.array_equal(lay.face_nx[faces], g.face_normals[mesh_faces, 0])
    # the time-step items of these faces point at their records
    assert np.isin(rec, lay.dts_bc).all()


@pytest.mark.gpu
@pytest.mark.parametrize("kinds", CASES[:2])
def test_gpu_inflow_implicit_step_matches_oracle(cuda_device, kinds):
    """Implicit Euler with local steps (CFL 20): the finite-difference
    Jacobian re-evaluates the inflow states at every perturbed state, as the
    reference's residual does."""
    from meshcfd_amd.device import DeviceMesh
    from oracle import implicit as oimp

    case = channel(kinds[0], kinds[1], integrator="Implicit Euler")
    case.numerics.cfl = 20.0
    case.numerics.time_step_mode = "Local (steady)"
    case.numerics.gmres_tolerance = 1e-6
    print(f"case: {case.name} implicit", flush=True)
    dm = DeviceMesh(case, cuda_device, fused=True, block_cells=64)
    prim = flow_state(case)
    u0 = euler.primitive_to_conservative(prim, 1.4)
    dm.set_state(u0)
    dm.step_implicit(0.0)
    dm.check_status()
    gpu = dm.state()
    o = Oracle(dm.geometry, case.numerics)
    ref, _ = oimp.implicit_euler_step(dm.geometry, case.numerics, u0, o.stable_time_steps(prim))
    err = _rel(gpu, ref)
    print("implicit rel err", err)
    assert np.all(err <= 1e-8), err
    dm.close()


@pytest.mark.parametrize("kinds", CASES)
def test_viscous_oracle_takes_inflow_states_as_boundary_values(kinds):
    """Viscous channels with pressure / mass-flow inlets: the inlets' exterior
    states from the owner cells are the viscous boundary values
    (`solver.py:2749-2772`), so the viscous residual changes with the inlet's
    total conditions; mass-flow outlets are not in that map."""
    from oracle.viscous import inlet_states

    case = channel(kinds[0], kinds[1], viscous=True)
    o = Oracle(build_geometry(case), case.numerics)
    prim = flow_state(case)
    r = o.residual(euler.primitive_to_conservative(prim, 1.4))
    assert np.all(np.isfinite(r))
    g = o.g
    states = inlet_states(prim, g, case.numerics)
    inlet_faces = g.pressure_inlet_faces if kinds[0] == "Pressure Inlet" else g.mass_flow_inlet_faces
    assert set(int(f) for f in inlet_faces) <= set(states)
    assert not set(int(f) for f in g.mass_flow_outlet_faces) & set(states)


@pytest.mark.gpu
@pytest.mark.parametrize("kinds", CASES)
def test_gpu_viscous_inflow_residual_and_dt_match_oracle(cuda_device, kinds):
    from meshcfd_amd.device import DeviceMesh

    case = channel(kinds[0], kinds[1], viscous=True)
    print(f"case: viscous {case.name}", flush=True)
    dm = DeviceMesh(case, cuda_device, block_cells=64)
    prim = flow_state(case)
    u = euler.primitive_to_conservative(prim, 1.4)
    o = Oracle(dm.geometry, case.numerics)
    gpu = dm.to_mesh_order(dm.residual(dm.to_device_order(u)))
    dm.check_status()
    err = _rel(gpu, o.residual(u))
    print("residual rel err", err)
    assert np.all(err <= 1e-12), err
    dm.time_steps(dm.to_device_order(u))
    dm.check_status()
    dt = dm.to_mesh_order(dm.dt_cell)
    ref = o.stable_time_steps(prim)
    assert np.max(np.abs(dt - ref) / ref) <= 1e-12
    dm.close()


@pytest.mark.gpu
def test_gpu_inflow_errors_name_the_patch(cuda_device):
    """A supersonic inflow state at the pressure inlet is reported with the
    reference's message (the kernels finish; the host raises)."""
    from meshcfd_amd.device import DeviceMesh

    case = channel("Pressure Inlet", "Pressure Outlet", reconstruction="First order")
    dm = DeviceMesh(case, cuda_device)
    prim = flow_state(case)
    prim[:, 1] = 3.0  # Mach > 2 everywhere, entering through the left side
    u = dm.to_device_order(euler.primitive_to_conservative(prim, 1.4))
    dm.residual(u)
    with pytest.raises(SolverError, match="Pressure inlet 'left' needs Mach number to define a supersonic inflow"):
        dm.check_status()
    dm.close()


# --- adjoint of the pressure-inlet and mass-flow boundaries -------------------


def _vjp_setup(cuda_device, case, seed=5):
    from meshcfd_amd.adjoint import residual_vjp
    from meshcfd_amd.device import DeviceMesh

    dm = DeviceMesh(case, cuda_device, block_cells=64)
    # flow_state's v depends on x alone, so cells with equal centroid x (the
    # straight outlet column) tie in v, where the limiter's extrema are not
    # differentiable; a 1% random factor removes the ties
    prim = flow_state(case, seed=seed)
    prim = prim * (1.0 + 0.01 * np.random.default_rng(seed + 100).random(prim.shape))
    u = euler.primitive_to_conservative(prim, 1.4)
    lam = np.random.default_rng(11).normal(size=u.shape)
    ubar = dm.to_mesh_order(residual_vjp(dm, dm.to_device_order(u), dm.to_device_order(lam)))
    dm.check_status()
    return dm, u, lam, ubar


def _inflow_owners(g):
    faces = np.concatenate((g.pressure_inlet_faces, g.mass_flow_inlet_faces, g.mass_flow_outlet_faces))
    return np.unique(g.owners[faces])


@pytest.mark.gpu
@pytest.mark.parametrize("flux,viscous", [("HLLC", False), ("Rusanov", False), ("Roe", False), ("HLLC", True)])
@pytest.mark.parametrize("reconstruction", ["First order", "MUSCL linear"])
@pytest.mark.parametrize("kinds", CASES)
def test_gpu_inflow_vjp_matches_central_difference_of_oracle(cuda_device, kinds, flux, reconstruction, viscous):
    """lam^T (dR/dU) v of the device VJP against central differences of the
    oracle residual (step 1e-7), for v over all cells and for v on the owner
    cells of the inlet and outlet faces alone (where the inflow states'
    derivatives are most of the product): 1e-6 of |fd|. The exterior states
    are bisection roots (64 / 80 halvings, resolved to rounding), so the
    differences see their implicit-function derivatives."""
    case = channel(kinds[0], kinds[1], flux=flux, reconstruction=reconstruction, viscous=viscous)
    print(f"case: inflow VJP {case.name} {flux} {reconstruction} viscous={viscous}", flush=True)
    dm, u, lam, ubar = _vjp_setup(cuda_device, case)
    o = Oracle(dm.geometry, case.numerics)
    rng = np.random.default_rng(13)
    scale = np.max(np.abs(u), axis=0)
    owners = _inflow_owners(dm.geometry)
    for where in ("all", "inflow owners"):
        v = rng.normal(size=u.shape) * scale
        if where != "all":
            mask = np.zeros(len(u), dtype=bool)
            mask[owners] = True
            v[~mask] = 0.0
        eps = 1e-7
        fd = float(np.sum(lam * (o.residual(u + eps * v) - o.residual(u - eps * v)) / (2 * eps)))
        ad = float(np.sum(ubar * v))
        print(where, "fd", fd, "adjoint", ad)
        assert abs(ad - fd) <= 1e-6 * max(abs(fd), 1.0), (where, fd, ad)
    dm.close()


@pytest.mark.gpu
@pytest.mark.parametrize("viscous", [False, True])
@pytest.mark.parametrize("kinds", CASES)
def test_gpu_inflow_vjp_entries_of_the_boundary_cells(cuda_device, kinds, viscous):
    """Every VJP entry of the inlet and outlet faces' owner cells equals a
    central difference of the oracle residual (first order, HLLC; step 1e-6
    of the entry: the mass-flow bisection's rounding makes differences at
    1e-7 scatter by a few 1e-6), to 1e-6: a wrong term of one boundary
    kind's chain fails here even where a random product would average it
    out."""
    case = channel(kinds[0], kinds[1], n=4, reconstruction="First order", viscous=viscous)
    print(f"case: inflow VJP entries {case.name} viscous={viscous}", flush=True)
    dm, u, lam, ubar = _vjp_setup(cuda_device, case, seed=7)
    o = Oracle(dm.geometry, case.numerics)
    for i in _inflow_owners(dm.geometry):
        for k in range(4):
            e = np.zeros_like(u)
            e[i, k] = 1e-6 * max(abs(u[i, k]), 1.0)
            fd = np.sum(lam * (o.residual(u + e) - o.residual(u - e))) / (2 * e[i, k])
            assert abs(ubar[i, k] - fd) <= 1e-6 * max(abs(fd), 1.0), (i, k, ubar[i, k], fd)
    dm.close()


@pytest.mark.gpu
@pytest.mark.parametrize("inlet", ["Pressure Inlet", "Mass Flow Inlet"])
def test_gpu_inflow_steady_adjoint_matches_reconverged_differences(cuda_device, inlet):
    """A channel with a pressure or mass-flow inlet and a pressure outlet
    (HLLC, first order) under a body force a_x: d(outlet total-pressure
    loss)/d(a_x) from one adjoint solve through the inflow VJP,
    lambda . dR/d(a_x) with dR/d(a_x) = (0, rho, 0, rho u) per cell, equals
    central differences of re-converged steady states (a_x = 0.05 +- 1e-3)
    to 1e-5 of the derivative."""
    import torch

    from meshcfd_amd.device import DeviceMesh
    from meshcfd_amd.steady_adjoint import OutletTotalPressureLoss, SteadyAdjoint, objective_gradient, solve_steady

    def steady(ax, warm=None):
        case = channel(inlet, "Pressure Outlet", reconstruction="First order", integrator="Implicit Euler")
        case.numerics.time_step_mode = "Local (steady)"
        case.numerics.cfl = 50.0
        case.numerics.gmres_tolerance = 1e-3
        case.numerics.acceleration_x = ax
        dm = DeviceMesh(case, cuda_device, block_cells=64)
        dm.set_state(euler.primitive_to_conservative(flow_state(case), 1.4) if warm is None else warm)
        hist = solve_steady(dm, tol=1e-11, max_steps=400, raise_cfl=2.0)
        print(f"  steady {inlet} a_x={ax}: {len(hist) - 1} steps, rms {hist[-1]:.3e}", flush=True)
        assert hist[-1] <= 1e-11, hist[-5:]
        loss = OutletTotalPressureLoss(dm, T0**3.5)
        return dm, loss, float(loss(dm.u).item())

    print(f"case: inflow steady adjoint {inlet}", flush=True)
    ax, h = 0.05, 1e-3
    dm, loss, j0 = steady(ax)
    u_star = dm.state()
    sa = SteadyAdjoint(dm, cfl=1e3, rtol=1e-4)
    _, dj_du = objective_gradient(loss, dm.u)
    lam, ahist = sa.solve(dj_du, tol=1e-11, max_iterations=80)
    assert ahist[-1] <= 1e-11, ahist[-5:]
    grad = float(torch.sum(lam[:, 1] * dm.u[:, 0] + lam[:, 3] * dm.u[:, 1]).item())
    dm.close()
    jp = steady(ax + h, u_star)
    jm = steady(ax - h, u_star)
    for d in (jp[0], jm[0]):
        d.close()
    fd = (jp[2] - jm[2]) / (2 * h)
    print("adjoint", grad, "fd", fd, "rel", abs(grad - fd) / abs(fd))
    assert abs(grad - fd) <= 1e-5 * abs(fd)
